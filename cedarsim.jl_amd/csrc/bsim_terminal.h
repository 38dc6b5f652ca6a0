// bsim_terminal.h — terminal evaluation of one BSIM-CMG instance (host and
// device): internal-node solve and Schur condensation onto d, g, s, e.
#pragma once
#include "va_runtime.h"
#include "../_gen/bsimcmg.h"

namespace cedar {

using va_bsimcmg::Card;
using va_bsimcmg::Inst;
using D2 = VaDual<2>;
constexpr int kBsimNodes = va_bsimcmg::n_nodes;
static_assert(kBsimNodes == 6, "BSIM-CMG with internal di, si");

// One model evaluation at node voltages V with the partials taken along
// nodes a and b (a dual number with two partials: the generated code's cost
// and live state grow with the number of partials, and each partial is
// computed independently, so three two-partial passes give exactly the
// partials of one six-partial pass at far lower register pressure).
// The bias-independent part of the model (geometry, binning, temperature)
// ran once per instance (va_bsimcmg::setup) and left its values in S[k *
// stride]; each evaluation runs only the bias-dependent rest.
CEDAR_VA_HD inline __attribute__((noinline)) void eval_pair(const Card& P, const Inst& Ip, double temp,
                                                            const double* S, long long stride,
                                                            const double* V, int a, int b, D2* I, D2* Q) {
    D2 Vd[kBsimNodes];
    for (int k = 0; k < kBsimNodes; k++) Vd[k] = D2(V[k]);
    Vd[a].d[0] = 1.0;
    Vd[b].d[1] = 1.0;
    va_bsimcmg::eval_bias<D2>(P, Ip, temp, S, stride, Vd, I, Q);
}

// The once-per-instance part; a zero-partial dual keeps the arithmetic of
// dual-typed temporaries identical to a full evaluation.
CEDAR_VA_HD inline __attribute__((noinline)) void setup_instance(const Card& P, const Inst& Ip, double temp,
                                                                 double* S, long long stride) {
    va_bsimcmg::setup<VaDual<0>>(P, Ip, temp, S, stride);
}

// One instance: returns the Newton iterations used on the internal nodes (or
// -1 when they did not converge).
CEDAR_VA_HD inline int bsimcmg_terminal(const Card& P, const Inst& Ip, double temp,
                                        const double* S, long long stride, const double* vt, double* out) {
    double V[kBsimNodes] = {vt[0], vt[1], vt[2], vt[3], vt[0], vt[2]};
    D2 I[kBsimNodes], Q[kBsimNodes];
    int it = 0, conv = 0;
    for (; it < 20; it++) {
        eval_pair(P, Ip, temp, S, stride, V, 4, 5, I, Q);
        if (conv) break;
        const double a = I[4].d[0], b = I[4].d[1], c = I[5].d[0], d = I[5].d[1];
        const double f0 = I[4].v, f1 = I[5].v;
        const double det = a * d - b * c;
        const double dx0 = (d * f0 - b * f1) / det, dx1 = (a * f1 - c * f0) / det;
        V[4] -= dx0;
        V[5] -= dx1;
        if (!(dx0 == dx0) || !(dx1 == dx1)) return -1;
        const double tol = 1e-15 * (1.0 + (V[4] < 0 ? -V[4] : V[4]) + (V[5] < 0 ? -V[5] : V[5]));
        if ((dx0 < 0 ? -dx0 : dx0) <= tol && (dx1 < 0 ? -dx1 : dx1) <= tol) conv = 1;
    }
    if (!conv) return -1;
    // the (di, si) pass at the solution: currents, charges and the internal
    // columns of G and C
    const double gii[2][2] = {{I[4].d[0], I[4].d[1]}, {I[5].d[0], I[5].d[1]}};
    const double det = gii[0][0] * gii[1][1] - gii[0][1] * gii[1][0];
    const double inv[2][2] = {{gii[1][1] / det, -gii[0][1] / det}, {-gii[1][0] / det, gii[0][0] / det}};
    double gti[4][2], cti[4][2], M[4][2];
    const double cii[2][2] = {{Q[4].d[0], Q[4].d[1]}, {Q[5].d[0], Q[5].d[1]}};
    for (int t = 0; t < 4; t++) {
        gti[t][0] = I[t].d[0];
        gti[t][1] = I[t].d[1];
        cti[t][0] = Q[t].d[0];
        cti[t][1] = Q[t].d[1];
        for (int r = 0; r < 2; r++) M[t][r] = -(gti[t][0] * inv[0][r] + gti[t][1] * inv[1][r]);
    }
    double* oI = out;
    double* oQ = out + 4;
    double* oG = out + 8;
    double* oC = out + 24;
    for (int t = 0; t < 4; t++) {
        oI[t] = I[t].v;
        oQ[t] = Q[t].v + M[t][0] * Q[4].v + M[t][1] * Q[5].v;
    }
    // terminal columns, two per pass
    for (int pass = 0; pass < 2; pass++) {
        const int j0 = 2 * pass;
        eval_pair(P, Ip, temp, S, stride, V, j0, j0 + 1, I, Q);
        for (int c = 0; c < 2; c++) {
            const int j = j0 + c;
            const double s0 = -(inv[0][0] * I[4].d[c] + inv[0][1] * I[5].d[c]);
            const double s1 = -(inv[1][0] * I[4].d[c] + inv[1][1] * I[5].d[c]);
            const double c4 = Q[4].d[c] + cii[0][0] * s0 + cii[0][1] * s1;
            const double c5 = Q[5].d[c] + cii[1][0] * s0 + cii[1][1] * s1;
            for (int t = 0; t < 4; t++) {
                oG[4 * t + j] = I[t].d[c] + gti[t][0] * s0 + gti[t][1] * s1;
                const double cij = Q[t].d[c] + cti[t][0] * s0 + cti[t][1] * s1;
                oC[4 * t + j] = cij + M[t][0] * c4 + M[t][1] * c5;
            }
        }
    }
    return it;
}

}  // namespace cedar
