/*
 * sparse.c — sparse LU for the CPU oracle (test infrastructure and CPU
 * baseline only; see oracle.h).
 *
 * Left-looking Gilbert-Peierls LU (the algorithm of KLU without block
 * triangular form): column ordering by minimum degree on the pattern of
 * A + A^T, threshold partial pivoting (diagonal preferred when within 0.1 of
 * the column maximum) on the first factorization, then refactorization with
 * the same pivots and patterns; a refactorization whose pivot falls below
 * 1e-8 of its column is redone with pivoting.  Written from the published
 * algorithm (Gilbert & Peierls 1988; Davis & Palamadai Natarajan 2010), not
 * from the HIP library, which uses static pivots and an entry-level schedule.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sparse.h"

/* ---- minimum degree on A + A^T (explicit elimination graph) -------------- */
typedef struct {
    int* v;
    int len, cap;
} ivec;

static void iv_push(ivec* a, int x) {
    if (a->len == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->v = (int*)realloc(a->v, sizeof(int) * a->cap);
    }
    a->v[a->len++] = x;
}

static int iv_has(const ivec* a, int x) {
    for (int i = 0; i < a->len; i++)
        if (a->v[i] == x) return 1;
    return 0;
}

static void iv_del(ivec* a, int x) {
    for (int i = 0; i < a->len; i++)
        if (a->v[i] == x) { a->v[i] = a->v[--a->len]; return; }
}

static void min_degree(int n, const int* colptr, const int* rowind, int* order) {
    ivec* adj = (ivec*)calloc(n, sizeof(ivec));
    for (int j = 0; j < n; j++)
        for (int p = colptr[j]; p < colptr[j + 1]; p++) {
            int i = rowind[p];
            if (i == j) continue;
            if (!iv_has(&adj[i], j)) iv_push(&adj[i], j);
            if (!iv_has(&adj[j], i)) iv_push(&adj[j], i);
        }
    char* done = (char*)calloc(n, 1);
    for (int k = 0; k < n; k++) {
        int best = -1, bd = 0x7fffffff;
        for (int v = 0; v < n; v++)
            if (!done[v] && adj[v].len < bd) { bd = adj[v].len; best = v; }
        order[k] = best;
        done[best] = 1;
        ivec* nb = &adj[best];
        for (int a = 0; a < nb->len; a++) iv_del(&adj[nb->v[a]], best);
        for (int a = 0; a < nb->len; a++)
            for (int b = a + 1; b < nb->len; b++) {
                int x = nb->v[a], y = nb->v[b];
                if (!iv_has(&adj[x], y)) { iv_push(&adj[x], y); iv_push(&adj[y], x); }
            }
        free(nb->v);
        nb->v = NULL;
        nb->len = nb->cap = 0;
    }
    for (int v = 0; v < n; v++) free(adj[v].v);
    free(adj);
    free(done);
}

/* ---- factor storage ----------------------------------------------------- */
static void grow(int** ri, double** vx, int* cap, int need) {
    if (need <= *cap) return;
    int c = *cap ? *cap : 64;
    while (c < need) c *= 2;
    *ri = (int*)realloc(*ri, sizeof(int) * c);
    *vx = (double*)realloc(*vx, sizeof(double) * c);
    *cap = c;
}

void sp_free(sp_lu* f) {
    free(f->Q); free(f->pinv); free(f->Lp); free(f->Li); free(f->Lx);
    free(f->Up); free(f->Ui); free(f->Ux); free(f->x); free(f->stack); free(f->mark);
    free(f->pos);
    memset(f, 0, sizeof(*f));
}

int sp_analyze(sp_lu* f, int n, const int* colptr, const int* rowind) {
    memset(f, 0, sizeof(*f));
    f->n = n;
    f->Q = (int*)malloc(sizeof(int) * n);
    min_degree(n, colptr, rowind, f->Q);
    f->pinv = (int*)malloc(sizeof(int) * n);
    f->Lp = (int*)calloc(n + 1, sizeof(int));
    f->Up = (int*)calloc(n + 1, sizeof(int));
    f->x = (double*)calloc(n, sizeof(double));
    f->stack = (int*)malloc(sizeof(int) * 2 * n);
    f->mark = (int*)malloc(sizeof(int) * n);
    f->pos = (int*)malloc(sizeof(int) * n);
    f->factored = 0;
    return 0;
}

/* reach of column b's pattern in the graph of L (pivot order), topological
 * order written to top..n-1 of xi; returns top */
static int reach(sp_lu* f, const int* colptr, const int* rowind, int col, int* xi) {
    int n = f->n, top = n;
    int* stk = f->stack;       /* node stack in [0,n), position stack in pos[] */
    for (int p = colptr[col]; p < colptr[col + 1]; p++) {
        int r0 = rowind[p];
        if (f->mark[r0] == f->stamp) continue;
        int head = 0;
        stk[0] = r0;
        while (head >= 0) {
            int r = stk[head];
            int j = f->pinv[r];            /* column of L where r is pivot, or -1 */
            if (f->mark[r] != f->stamp) {
                f->mark[r] = f->stamp;
                f->pos[r] = (j < 0) ? 0 : f->Lp[j];
            }
            int done = 1;
            if (j >= 0) {
                int end = f->Lp[j + 1];
                while (f->pos[r] < end) {
                    int i = f->Li[f->pos[r]++];
                    if (f->mark[i] == f->stamp) continue;
                    stk[++head] = i;
                    done = 0;
                    break;
                }
            }
            if (done) {
                head--;
                xi[--top] = r;
            }
        }
    }
    return top;
}

int sp_factor(sp_lu* f, const int* colptr, const int* rowind, const double* vals) {
    int n = f->n;
    int* xi = (int*)malloc(sizeof(int) * n);
    int lcap = 0, ucap = 0, lnz = 0, unz = 0;
    for (int i = 0; i < n; i++) { f->pinv[i] = -1; f->mark[i] = -1; }
    f->stamp = 0;
    for (int k = 0; k < n; k++) {
        int col = f->Q[k];
        f->stamp = k;
        f->Lp[k] = lnz;          /* end of column k-1, read by the reach */
        f->Up[k] = unz;
        int top = reach(f, colptr, rowind, col, xi);
        for (int p = top; p < n; p++) f->x[xi[p]] = 0.0;
        for (int p = colptr[col]; p < colptr[col + 1]; p++) f->x[rowind[p]] = vals[p];
        /* triangular solve in topological order */
        for (int p = top; p < n; p++) {
            int r = xi[p], j = f->pinv[r];
            if (j < 0) continue;
            double xj = f->x[r];
            for (int q = f->Lp[j] + 1; q < f->Lp[j + 1]; q++) f->x[f->Li[q]] -= f->Lx[q] * xj;
        }
        /* pivot among non-pivotal rows */
        int ipiv = -1;
        double amax = -1.0;
        for (int p = top; p < n; p++) {
            int r = xi[p];
            if (f->pinv[r] >= 0) continue;
            double a = fabs(f->x[r]);
            if (a > amax) { amax = a; ipiv = r; }
        }
        if (ipiv < 0 || amax == 0.0 || !isfinite(amax)) { free(xi); return -1; }
        if (f->pinv[col] < 0 && fabs(f->x[col]) >= 0.1 * amax) ipiv = col;
        double piv = f->x[ipiv];
        /* U(:,k): pivotal rows (in pivot order) + diagonal last */
        grow(&f->Ui, &f->Ux, &ucap, unz + (n - top) + 1);
        f->Up[k] = unz;
        for (int p = top; p < n; p++) {
            int r = xi[p];
            if (f->pinv[r] >= 0) { f->Ui[unz] = f->pinv[r]; f->Ux[unz++] = f->x[r]; }
        }
        f->Ui[unz] = k;
        f->Ux[unz++] = piv;
        /* L(:,k): pivot row first (unit), then the rest scaled */
        grow(&f->Li, &f->Lx, &lcap, lnz + (n - top) + 1);
        f->Lp[k] = lnz;
        f->Li[lnz] = ipiv;
        f->Lx[lnz++] = 1.0;
        for (int p = top; p < n; p++) {
            int r = xi[p];
            if (f->pinv[r] < 0 && r != ipiv) { f->Li[lnz] = r; f->Lx[lnz++] = f->x[r] / piv; }
        }
        f->pinv[ipiv] = k;
        for (int p = top; p < n; p++) f->x[xi[p]] = 0.0;
    }
    f->Lp[n] = lnz;
    f->Up[n] = unz;
    free(xi);
    f->factored = 1;
    return 0;
}

/* numeric refactorization on the stored patterns and pivots */
int sp_refactor(sp_lu* f, const int* colptr, const int* rowind, const double* vals) {
    if (!f->factored) return sp_factor(f, colptr, rowind, vals);
    int n = f->n;
    for (int k = 0; k < n; k++) {
        int col = f->Q[k];
        for (int p = colptr[col]; p < colptr[col + 1]; p++) f->x[rowind[p]] = vals[p];
        /* U entries are stored in the topological order of the first
         * factorization's reach */
        for (int q = f->Up[k]; q < f->Up[k + 1] - 1; q++) {
            int j = f->Ui[q];
            int r = f->Li[f->Lp[j]];                /* pivot row of column j */
            double xj = f->x[r];
            f->Ux[q] = xj;
            for (int t = f->Lp[j] + 1; t < f->Lp[j + 1]; t++) f->x[f->Li[t]] -= f->Lx[t] * xj;
        }
        int rp = f->Li[f->Lp[k]];
        double piv = f->x[rp];
        double cmax = fabs(piv);
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; t++) {
            double a = fabs(f->x[f->Li[t]]);
            if (a > cmax) cmax = a;
        }
        if (piv == 0.0 || !isfinite(piv) || fabs(piv) < 1e-8 * cmax) {
            for (int i = 0; i < n; i++) f->x[i] = 0.0;
            f->factored = 0;
            return sp_factor(f, colptr, rowind, vals);
        }
        f->Ux[f->Up[k + 1] - 1] = piv;
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; t++) f->Lx[t] = f->x[f->Li[t]] / piv;
        f->x[rp] = 0.0;
        for (int q = f->Up[k]; q < f->Up[k + 1] - 1; q++) f->x[f->Li[f->Lp[f->Ui[q]]]] = 0.0;
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; t++) f->x[f->Li[t]] = 0.0;
    }
    return 0;
}

/* solve A x = b in place (b indexed by rows, x by columns) */
void sp_solve(const sp_lu* f, double* b) {
    int n = f->n;
    double* y = f->x;      /* scratch, zero on entry and exit */
    /* L y = P b : column k of L has its pivot row first */
    for (int k = 0; k < n; k++) {
        double v = b[f->Li[f->Lp[k]]];
        y[k] = v;
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; t++) b[f->Li[t]] -= f->Lx[t] * v;
    }
    /* U z = y (columns: entries above the diagonal then the diagonal last) */
    for (int k = n - 1; k >= 0; k--) {
        double v = y[k] / f->Ux[f->Up[k + 1] - 1];
        y[k] = v;
        for (int q = f->Up[k]; q < f->Up[k + 1] - 1; q++) y[f->Ui[q]] -= f->Ux[q] * v;
    }
    for (int k = 0; k < n; k++) { b[f->Q[k]] = y[k]; }
    for (int k = 0; k < n; k++) y[k] = 0.0;
}
