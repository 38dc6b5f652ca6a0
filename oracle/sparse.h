/* sparse.h — sparse LU of the CPU oracle (test infrastructure only). */
#ifndef CEDAR_ORACLE_SPARSE_H
#define CEDAR_ORACLE_SPARSE_H

typedef struct {
    int n;
    int* Q;              /* column order: factor column k is A(:, Q[k])   */
    int* pinv;           /* row r is the pivot of factor column pinv[r]   */
    int *Lp, *Li;        /* L by columns, pivot row first (unit diagonal) */
    double* Lx;
    int *Up, *Ui;        /* U by columns, diagonal last; Ui = pivot index */
    double* Ux;
    double* x;           /* dense work vector, all zero between calls     */
    int *stack, *mark, *pos;
    int stamp, factored;
} sp_lu;

int  sp_analyze(sp_lu* f, int n, const int* colptr, const int* rowind);
int  sp_factor(sp_lu* f, const int* colptr, const int* rowind, const double* vals);
int  sp_refactor(sp_lu* f, const int* colptr, const int* rowind, const double* vals);
void sp_solve(const sp_lu* f, double* b);
void sp_free(sp_lu* f);

#endif
