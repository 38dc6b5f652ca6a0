// bsimcmg.hip — batched BSIM-CMG 107 device evaluation on gfx950.
//
// The model code is generated at build time by vagen from the model's
// Verilog-A source (bsimcmg.va, its body at VerilogAParser.jl/cmc_models/
// bsimcmg107/bsimcmg_body.include:988-3990 in the reference tree, which the
// reference compiles through make_spice_device, src/vasim.jl:663-875).  One
// lane evaluates one instance in FP64 with a forward-mode dual number that
// carries the partial derivatives with respect to all six node voltages
// (d, g, s, e and the internal di, si).
//
// Per instance:
//   1. the internal nodes di, si are solved so that their resistive KCL
//      holds (Newton on the 2x2 block of dI/dV);
//   2. the six-node conductance G = dI/dV and capacitance C = dQ/dV are
//      condensed onto the four terminals (Schur complement over di, si):
//         S = -G_ii^-1 G_it   (dV_i/dV_t),   M = -G_ti G_ii^-1
//         I_t' = I_t,  G' = G_tt + G_ti S
//         Q_t' = Q_t + M Q_i,  C' = C_tt + C_ti S + M (C_it + C_ii S)
//      (charge on an internal node reaches the terminals through the
//      resistive network, hence the M terms).
// Output per instance: I[4], Q[4], G[4x4], C[4x4] (row = terminal equation,
// column = terminal voltage), terminals in the order d, g, s, e.
//
// The model card (~900 parameters) is shared: each workgroup stages it in
// LDS once; per-instance parameters (L, NFIN, ...) live in the lane.
#include <cstring>

#include "kernels.h"
#include "va_runtime.h"
#include "../_gen/bsimcmg.h"

namespace cedar {

namespace {

using namespace va_bsimcmg;
constexpr int NN = n_nodes;
static_assert(NN == 6, "BSIM-CMG with internal di, si");
using D6 = VaDual<NN>;
constexpr int kBsimThreads = 256;
constexpr int kOutPer = 40;

}  // namespace

// One instance: returns the Newton iterations used on the internal nodes (or
// a negative value when they did not converge).
CEDAR_VA_HD inline int bsimcmg_terminal(const Card& P, const Inst& Ip, double temp,
                                        const double* vt, double* out) {
    double V[NN] = {vt[0], vt[1], vt[2], vt[3], vt[0], vt[2]};
    D6 Vd[NN], I[NN], Q[NN];
    int it = 0, conv = 0;
    for (; it < 20; it++) {
        for (int k = 0; k < NN; k++) Vd[k] = D6::var(V[k], k);
        eval<D6>(P, Ip, temp, Vd, I, Q);
        if (conv) break;
        const double a = I[4].d[4], b = I[4].d[5], c = I[5].d[4], d = I[5].d[5];
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
    // condensation
    const double gii[2][2] = {{I[4].d[4], I[4].d[5]}, {I[5].d[4], I[5].d[5]}};
    const double det = gii[0][0] * gii[1][1] - gii[0][1] * gii[1][0];
    const double inv[2][2] = {{gii[1][1] / det, -gii[0][1] / det}, {-gii[1][0] / det, gii[0][0] / det}};
    double S[2][4], M[4][2];
    for (int j = 0; j < 4; j++)
        for (int r = 0; r < 2; r++)
            S[r][j] = -(inv[r][0] * I[4].d[j] + inv[r][1] * I[5].d[j]);
    for (int t = 0; t < 4; t++)
        for (int r = 0; r < 2; r++)
            M[t][r] = -(I[t].d[4] * inv[0][r] + I[t].d[5] * inv[1][r]);
    double* oI = out;
    double* oQ = out + 4;
    double* oG = out + 8;
    double* oC = out + 24;
    for (int t = 0; t < 4; t++) {
        oI[t] = I[t].v;
        oQ[t] = Q[t].v + M[t][0] * Q[4].v + M[t][1] * Q[5].v;
        for (int j = 0; j < 4; j++) {
            oG[4 * t + j] = I[t].d[j] + I[t].d[4] * S[0][j] + I[t].d[5] * S[1][j];
            const double cij = Q[t].d[j] + Q[t].d[4] * S[0][j] + Q[t].d[5] * S[1][j];
            const double c4 = Q[4].d[j] + Q[4].d[4] * S[0][j] + Q[4].d[5] * S[1][j];
            const double c5 = Q[5].d[j] + Q[5].d[4] * S[0][j] + Q[5].d[5] * S[1][j];
            oC[4 * t + j] = cij + M[t][0] * c4 + M[t][1] * c5;
        }
    }
    return it;
}

namespace {

__global__ __launch_bounds__(kBsimThreads) void k_bsimcmg(BsimArgs A) {
    __shared__ Card card;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(A.card);   // a Card
        uint32_t* dst = reinterpret_cast<uint32_t*>(&card);
        for (int w = threadIdx.x; w < (int)(sizeof(Card) / 4); w += kBsimThreads) dst[w] = src[w];
    }
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * kBsimThreads;
    for (int64_t i = (int64_t)blockIdx.x * kBsimThreads + threadIdx.x; i < A.n; i += stride) {
        Inst ip = *reinterpret_cast<const Inst*>(A.inst);
        ip.L = A.L[i];
        ip.given_L = 1;
        ip.NFIN = A.nfin[i];
        ip.given_NFIN = 1;
        inst_defaults(card, ip);
        const double vt[4] = {A.v[4 * i], A.v[4 * i + 1], A.v[4 * i + 2], A.v[4 * i + 3]};
        const int it = bsimcmg_terminal(card, ip, A.temp, vt, A.out + (size_t)kOutPer * i);
        if (A.iters) A.iters[i] = it;
    }
}

}  // namespace

int bsim_card_size() { return (int)sizeof(Card); }
int bsim_inst_size() { return (int)sizeof(Inst); }
int bsim_out_per() { return kOutPer; }

// Host: defaults for a card whose given parameters are set, and the
// instance template (given flags only for parameters the caller fixed).
void bsim_card_defaults(void* card) { card_defaults(*reinterpret_cast<Card*>(card)); }

// Set a card or instance parameter by name; returns 0, or -1 when the name
// is unknown.  is_inst selects the table.
int bsim_set_param(void* card, void* inst, const char* name, double value) {
    for (int k = 0; k < n_card; k++)
        if (!std::strcmp(card_table[k].name, name)) {
            char* b = reinterpret_cast<char*>(card);
            if (card_table[k].is_int) *reinterpret_cast<int*>(b + card_table[k].off) = (int)value;
            else *reinterpret_cast<double*>(b + card_table[k].off) = value;
            b[card_table[k].given_off] = 1;
            return 0;
        }
    for (int k = 0; k < n_inst; k++)
        if (!std::strcmp(inst_table[k].name, name)) {
            char* b = reinterpret_cast<char*>(inst);
            if (inst_table[k].is_int) *reinterpret_cast<int*>(b + inst_table[k].off) = (int)value;
            else *reinterpret_cast<double*>(b + inst_table[k].off) = value;
            b[inst_table[k].given_off] = 1;
            return 0;
        }
    return -1;
}

int bsim_max_blocks() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bsimcmg, kBsimThreads, 0) != hipSuccess) return 0;
    return nb;
}

hipError_t launch_bsimcmg(const BsimArgs& a, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_bsimcmg, dim3(nblocks), dim3(kBsimThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace cedar
