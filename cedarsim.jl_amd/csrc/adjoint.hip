// adjoint.hip — batched discrete adjoint of the BDF transient on gfx950.
//
// Objective per instance: energy delivered by one voltage source,
//     E = sum_m w_m P(u_m),  P(u) = -(V(a) - V(b)) * I_src,
// trapezoid weights on the accepted time points; parameters: every MOSFET
// width.  Two persistent kernels, one wavefront per instance:
//
//   k_adj_forward  the adaptive transient of tran_device.h with a recording
//                  policy: every accepted time and segment count goes to a
//                  step log, E is integrated, and every K accepted steps the
//                  three states the predictor needs are checkpointed.
//   k_adj_reverse  walks the segments [sK, (s+1)K) from the last one back:
//                  restores checkpoint s, recomputes the segment's states by
//                  replaying the log (same predictor and Newton), then for
//                  m descending solves
//                     J_m^T l_m = -( w_m dP/du(u_m) + a_{m+1,1} C^T l_{m+1}
//                                    + a_{m+2,2} C^T l_{m+2} )
//                  with transposed level-scheduled solves on J_m's LU, and
//                  accumulates dE/dW_k += (l_m[d] - l_m[s]) * I_ch,k / W_k.
// Step 0 is the consistent-initialisation system (charge rows pinned), so
// the adjoint needs CS_INIT_UIC.  C^T l is carried between steps in the
// registers of the lanes that own the unknowns; each device's gradient is
// accumulated by one lane in a fixed order, so results are bit-reproducible.
#include "tran_device.h"

namespace cedar {

namespace {

struct AdjRec {
    double* t_log;
    int32_t* npts_log;
    long long cap;
    double* ckpt;
    long long ckcap;
    int K, n, va, vb, vj;
    double E, Pprev;
    long long M;

    template <int NR>
    __device__ __forceinline__ bool on_accept(const DPlan*, const double* u, double tn,
                                              double tprev, int npts, long long m,
                                              const double* Uh0, const double* Uh1,
                                              const double* Uh2) {
        if (m >= cap) return false;
        const int tid = threadIdx.x;
        if (tid == 0) {
            t_log[m] = tn;
            npts_log[m] = npts;
        }
        const double Pn = -((va >= 0 ? u[va] : 0.0) - (vb >= 0 ? u[vb] : 0.0)) * u[vj];
        if (m > 0) E += 0.5 * (tn - tprev) * (Pn + Pprev);
        Pprev = Pn;
        M = m;
        if (m % K == 0) {
            const long long c = m / K;
            if (c >= ckcap) return false;
            double* cp = ckpt + (size_t)c * 3 * n;
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) {
                    cp[i] = Uh0[r];
                    cp[n + i] = Uh1[r];
                    cp[2 * n + i] = Uh2[r];
                }
            }
        }
        return true;
    }
    __device__ __forceinline__ void on_finish(int, long long) {}
};

struct LdsR {
    double *F, *inv, *u, *x, *qx, *st, *lam, *ctl;
};

__device__ __forceinline__ void step_coeffs(const double* tl, const int32_t* npl, long long q,
                                            int max_order, double& a0, double& a1, double& a2) {
    const int k = bdf_order(npl[q], max_order);
    bdf_coeffs(k, tl[q], tl[q - 1], q >= 2 ? tl[q - 2] : 0.0, a0, a1, a2);
}

template <int NR>
__device__ void run_reverse(const DPlan* __restrict__ P, const double* __restrict__ p,
                            const double* __restrict__ u0, const AdjArgs& A, int inst,
                            const LdsR& L, double* __restrict__ Glin, double* __restrict__ Clin,
                            double* __restrict__ linbuf, double* __restrict__ seg) {
    const int tid = threadIdx.x;
    const int n = P->n;
    const cs_tran_opts& o = A.b.o;
    const long long M = A.n_steps[inst];
    const int K = A.K;
    const double* tl = A.t_log + (size_t)inst * A.cap;
    const int32_t* npl = A.npts_log + (size_t)inst * A.cap;
    const double* ck = A.ckpt + (size_t)inst * A.ckcap * 3 * A.b.n_max;
    double* grad = A.grad + A.g_off[inst];
    const int cid = A.b.inst_cid[inst];
    const int va = A.obj[3 * cid], vb = A.obj[3 * cid + 1], vj = A.obj[3 * cid + 2];
    const double gmin = p[P->p_gmin];
    dev_lin_setup<NT>(P, p, linbuf, Glin, Clin);
    for (int d = tid; d < P->n_mos; d += NT) grad[d] = 0.0;

    double ctl1[NR], ctl2[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) ctl1[r] = ctl2[r] = 0.0;

    for (long long s = M / K; s >= 0; s--) {
        const long long m0 = s * K;
        const long long m1 = (m0 + K < M + 1) ? m0 + K : M + 1;
        double U0[NR], U1[NR], U2[NR], Q0[NR], Q1[NR], up[NR];
        const double* cp = ck + (size_t)s * 3 * n;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            U0[r] = U1[r] = U2[r] = Q0[r] = Q1[r] = up[r] = 0.0;
            if (i < n) {
                U0[r] = cp[i];
                U1[r] = cp[n + i];
                U2[r] = cp[2 * n + i];
                seg[i] = U0[r];
                L.u[i] = U1[r];
            }
        }
        __syncthreads();
        dev_charges<NT>(P, Clin, L.u, L.qx);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < n) { Q1[r] = L.qx[i]; L.u[i] = U0[r]; }
        }
        __syncthreads();
        dev_charges<NT>(P, Clin, L.u, L.qx);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < n) Q0[r] = L.qx[i];
        }
        // ---- recompute the segment's states by replaying the log ----------
        for (long long m = m0 + 1; m < m1; m++) {
            const int npts = npl[m];
            const int k = bdf_order(npts, o.max_order);
            const double tn = tl[m], T0 = tl[m - 1];
            const double T1 = m >= 2 ? tl[m - 2] : 0.0, T2 = m >= 3 ? tl[m - 3] : 0.0;
            double a0, a1, a2, w0, w1, w2;
            bdf_coeffs(k, tn, T0, T1, a0, a1, a2);
            const int np = npts < k + 1 ? npts : k + 1;
            pred_weights(np, tn, T0, T1, T2, w0, w1, w2);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) {
                    double v = U0[r];
                    if (np == 2) v = w0 * U0[r] + w1 * U1[r];
                    else if (np == 3) v = w0 * U0[r] + w1 * U1[r] + w2 * U2[r];
                    up[r] = v;
                    L.u[i] = v;
                    L.qx[i] = a1 * Q0[r] + (k == 2 ? a2 * Q1[r] : 0.0);
                }
            }
            dev_sources<NT>(P, p, tn, CS_MODE_TRAN, L.st);
            __syncthreads();
            for (int it = 0; it < 30; it++) {
                dev_mos<NT>(P, p, L.u, L.st);
                __syncthreads();
                dev_assemble<NT>(P, Glin, Clin, a0, L.u, L.qx, L.st, L.F, L.x, nullptr);
                __syncthreads();
                dev_lu<NT>(P, L.F, L.inv);
                dev_fwd<NT>(P, L.F, L.x);
                dev_bwd<NT>(P, L.F, L.inv, L.x);
                double ss = 0.0;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const int i = tid + NT * r;
                    if (i < n) {
                        const double d = L.x[P->pcol[i]];
                        const double ui = L.u[i] + d;
                        L.u[i] = ui;
                        const double q = d / (o.reltol * fabs(ui) + o.abstol);
                        ss += q * q;
                    }
                }
                const double nr = sqrt(block_reduce<NT>(ss, nullptr, OpSum()) / n);
                __syncthreads();
                if (!(nr > o.newton_tol)) break;
            }
            dev_charges<NT>(P, Clin, L.u, L.qx);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) {
                    U2[r] = U1[r]; U1[r] = U0[r]; U0[r] = L.u[i];
                    Q1[r] = Q0[r]; Q0[r] = L.qx[i];
                    seg[(size_t)(m - m0) * n + i] = U0[r];
                }
            }
        }
        // ---- backward over the segment ------------------------------------
        for (long long m = m1 - 1; m >= m0; m--) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) L.u[i] = seg[(size_t)(m - m0) * n + i];
            }
            __syncthreads();
            const double tm = tl[m];
            double w = 0.0;
            if (m > 0) w += 0.5 * (tm - tl[m - 1]);
            if (m < M) w += 0.5 * (tl[m + 1] - tm);
            double c1 = 0.0, c2 = 0.0, b0, b1, b2;
            if (m + 1 <= M) { step_coeffs(tl, npl, m + 1, o.max_order, b0, b1, b2); c1 = b1; }
            if (m + 2 <= M) { step_coeffs(tl, npl, m + 2, o.max_order, b0, b1, b2); c2 = b2; }
            double am0 = 0.0;
            if (m > 0) { step_coeffs(tl, npl, m, o.max_order, am0, b1, b2); }
            dev_sources<NT>(P, p, tm, CS_MODE_TRAN, L.st);
            dev_mos<NT>(P, p, L.u, L.st);
            __syncthreads();
            dev_assemble<NT>(P, Glin, Clin, am0, L.u, nullptr, L.st, L.F, L.x,
                             m == 0 ? (u0 ? u0 : L.u) : nullptr);
            __syncthreads();
            dev_lu<NT>(P, L.F, L.inv);
            const double ua = va >= 0 ? L.u[va] : 0.0, ub = vb >= 0 ? L.u[vb] : 0.0;
            const double uj = L.u[vj];
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) {
                    double dP = 0.0;
                    if (i == va) dP -= uj;
                    if (i == vb) dP += uj;
                    if (i == vj) dP -= (ua - ub);
                    L.x[P->pcol[i]] = -w * dP - c1 * ctl1[r] - c2 * ctl2[r];
                }
            }
            __syncthreads();
            dev_tfwd<NT>(P, L.F, L.inv, L.x);
            dev_tbwd<NT>(P, L.F, L.x);
            for (int k = tid; k < n; k += NT) L.lam[P->rownat[k]] = L.x[k];
            __syncthreads();
            // C^T lambda by permuted column (column access to F)
            for (int j = tid; j < n; j += NT) {
                double sacc = 0.0;
                for (int q = P->tcol_ptr[j]; q < P->tcol_ptr[j + 1]; q++)
                    sacc += Clin[P->tcol_pos[q]] * L.lam[P->rownat[P->tcol_row[q]]];
                L.ctl[P->colnat[P->diag[j]]] = sacc;
            }
            // dE/dW
            for (int d = tid; d < P->n_mos; d += NT) {
                const int4 dv = reinterpret_cast<const int4*>(P->mos)[d];
                const bool pin0 = (m == 0);
                double ld = 0.0, ls = 0.0;
                if (dv.x >= 0 && !(pin0 && P->is_diff[dv.x])) ld = L.lam[dv.x];
                if (dv.z >= 0 && !(pin0 && P->is_diff[dv.z])) ls = L.lam[dv.z];
                const double vd = dv.x >= 0 ? L.u[dv.x] : 0.0, vs = dv.z >= 0 ? L.u[dv.z] : 0.0;
                const double ich = L.st[4 * d] - gmin * (vd - vs);
                grad[d] += (ld - ls) * ich / p[P->p_mos_w + d];
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) { ctl2[r] = ctl1[r]; ctl1[r] = L.ctl[i]; }
            }
        }
    }
    __syncthreads();
}

template <int NR>
__global__ __launch_bounds__(NT, 4) void k_adj_forward(AdjArgs A) {
    extern __shared__ double lds[];
    __shared__ int s_idx;
    const BatchArgs& B = A.b;
    Lds L;
    L.F = lds;
    L.inv = L.F + B.nnz_max;
    L.u = L.inv + B.n_max;
    L.x = L.u + B.n_max;
    L.qx = L.x + B.n_max;
    L.st = L.qx + B.n_max;
    double* Glin = B.glin + (size_t)blockIdx.x * B.nnz_max;
    double* Clin = B.clin + (size_t)blockIdx.x * B.nnz_max;
    double* linbuf = B.linbuf + (size_t)blockIdx.x * B.lin_max;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(B.queue, 1);
        __syncthreads();
        const int qi = __builtin_amdgcn_readfirstlane(s_idx);
        __syncthreads();
        if (qi >= B.n_inst) break;
        const int inst = __builtin_amdgcn_readfirstlane(B.order[qi]);
        const int cid = __builtin_amdgcn_readfirstlane(B.inst_cid[inst]);
        const DPlan* P = B.plans + cid;
        AdjRec rec;
        rec.t_log = A.t_log + (size_t)inst * A.cap;
        rec.npts_log = A.npts_log + (size_t)inst * A.cap;
        rec.cap = A.cap;
        rec.ckpt = A.ckpt + (size_t)inst * A.ckcap * 3 * B.n_max;
        rec.ckcap = A.ckcap;
        rec.K = A.K;
        rec.n = P->n;
        rec.va = A.obj[3 * cid];
        rec.vb = A.obj[3 * cid + 1];
        rec.vj = A.obj[3 * cid + 2];
        rec.E = 0.0;
        rec.Pprev = 0.0;
        rec.M = 0;
        run_instance<NR>(rec, P, B.params + B.p_off[inst], B.u0 + B.u_off[inst], nullptr,
                         B.u_final + B.u_off[inst], B.stats + inst, B.o, L, Glin, Clin, linbuf);
        if (threadIdx.x == 0) {
            A.energy[inst] = rec.E;
            A.n_steps[inst] = rec.M;
        }
    }
}

template <int NR>
__global__ __launch_bounds__(NT, 4) void k_adj_reverse(AdjArgs A) {
    extern __shared__ double lds[];
    __shared__ int s_idx;
    const BatchArgs& B = A.b;
    LdsR L;
    L.F = lds;
    L.inv = L.F + B.nnz_max;
    L.u = L.inv + B.n_max;
    L.x = L.u + B.n_max;
    L.qx = L.x + B.n_max;
    L.lam = L.qx + B.n_max;
    L.ctl = L.lam + B.n_max;
    L.st = L.ctl + B.n_max;
    double* Glin = B.glin + (size_t)blockIdx.x * B.nnz_max;
    double* Clin = B.clin + (size_t)blockIdx.x * B.nnz_max;
    double* linbuf = B.linbuf + (size_t)blockIdx.x * B.lin_max;
    double* seg = A.seg + (size_t)blockIdx.x * A.K * B.n_max;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(A.queue2, 1);
        __syncthreads();
        const int qi = __builtin_amdgcn_readfirstlane(s_idx);
        __syncthreads();
        if (qi >= B.n_inst) break;
        const int inst = __builtin_amdgcn_readfirstlane(B.order[qi]);
        const int cid = __builtin_amdgcn_readfirstlane(B.inst_cid[inst]);
        if (B.stats[inst].retcode != CS_OK) continue;
        const DPlan* P = B.plans + cid;
        run_reverse<NR>(P, B.params + B.p_off[inst], B.u0 + B.u_off[inst], A, inst, L, Glin,
                        Clin, linbuf, seg);
    }
}

}  // namespace

size_t adj_lds_bytes(const BatchArgs& b) {
    return sizeof(double) * ((size_t)b.nnz_max + 6 * (size_t)b.n_max + b.stamp_max);
}

int adj_max_blocks(int nr, size_t lds) {
    int a = 0, b = 0;
    switch (nr) {
    case 1:
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_adj_forward<1>, NT, lds);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_adj_reverse<1>, NT, lds);
        break;
    case 2:
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_adj_forward<2>, NT, lds);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_adj_reverse<2>, NT, lds);
        break;
    case 4:
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_adj_forward<4>, NT, lds);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_adj_reverse<4>, NT, lds);
        break;
    default:
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_adj_forward<8>, NT, lds);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_adj_reverse<8>, NT, lds);
        break;
    }
    return a < b ? a : b;
}

hipError_t launch_adj(int nr, bool reverse, const AdjArgs& a, int nblocks, size_t lds,
                      hipStream_t s) {
#define CEDAR_ADJ_LAUNCH(R)                                                               \
    if (reverse) hipLaunchKernelGGL(k_adj_reverse<R>, dim3(nblocks), dim3(NT), lds, s, a); \
    else hipLaunchKernelGGL(k_adj_forward<R>, dim3(nblocks), dim3(NT), lds, s, a);
    switch (nr) {
    case 1: CEDAR_ADJ_LAUNCH(1) break;
    case 2: CEDAR_ADJ_LAUNCH(2) break;
    case 4: CEDAR_ADJ_LAUNCH(4) break;
    default: CEDAR_ADJ_LAUNCH(8) break;
    }
#undef CEDAR_ADJ_LAUNCH
    return hipGetLastError();
}

}  // namespace cedar
