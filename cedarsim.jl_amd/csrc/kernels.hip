// kernels.hip — gfx950 kernels of the Newton-transient engine.
//
// Batched path (k_tran_batch): a persistent grid of one-wavefront workgroups
// walks a work queue of circuit instances (largest first).  A workgroup runs
// its instance's whole transient on the device: consistent initialisation
// (DC op or UIC), then per step predictor -> Newton (eval -> assemble ->
// LU refactor -> solves -> update/WRMS) -> local-error test -> accept/reject
// -> output interpolation, exactly the rules of rules.h.  The per-unknown
// history lives in registers of the lane that owns the unknown; F, the
// vectors and the stamp buffer live in LDS; the per-instance linear stamp
// values (Glin, Clin) in a per-workgroup global workspace.
//
// Single-instance kernels (k_*): one workgroup each, the pieces the plain
// per-instance path captures into a hipGraph per Newton iteration and the
// DAE interface (cs_residual / cs_jacobian / cs_lu_*) calls directly.
#include "device.h"
#include "kernels.h"

namespace cedar {

// ---------------------------------------------------------------------------
// single-instance kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSingleThreads) void k_lin_setup(const DPlan* dp, InstDev w) {
    const DPlan* P = dp;
    dev_lin_setup<kSingleThreads>(P, w.p, w.linbuf, w.Glin, w.Clin);
}

__global__ __launch_bounds__(kSingleThreads) void k_eval_assemble(const DPlan* dp, InstDev w) {
    __shared__ double red[kSingleThreads];
    const DPlan* P = dp;
    const StepDev sp = *w.sp;
    dev_sources<kSingleThreads>(P, w.p, sp.t, sp.mode, w.st);
    dev_mos<kSingleThreads>(P, w.p, w.u, w.st);
    __syncthreads();
    const double a0 = sp.kind == STEP_TRAN ? sp.a0 : 0.0;
    const double m = dev_assemble<kSingleThreads>(P, w.Glin, w.Clin, a0, w.u,
                                                  sp.kind == STEP_TRAN ? w.qx : nullptr, w.st,
                                                  w.F, w.x, sp.kind == STEP_UIC ? w.u0 : nullptr);
    const double r = block_reduce<kSingleThreads>(m, red, OpMax());
    if (threadIdx.x == 0) w.norm[NRM_RES] = r;
}

__global__ __launch_bounds__(kSingleThreads) void k_lu(const DPlan* dp, InstDev w) {
    const DPlan* P = dp;
    dev_lu<kSingleThreads>(P, w.F, w.inv);
}

__global__ __launch_bounds__(kSingleThreads) void k_solve(const DPlan* dp, InstDev w) {
    const DPlan* P = dp;
    dev_fwd<kSingleThreads>(P, w.F, w.x);
    dev_bwd<kSingleThreads>(P, w.F, w.inv, w.x);
}

__global__ __launch_bounds__(kSingleThreads) void k_solve_t(const DPlan* dp, InstDev w) {
    const DPlan* P = dp;
    dev_tfwd<kSingleThreads>(P, w.F, w.inv, w.x);
    dev_tbwd<kSingleThreads>(P, w.F, w.x);
}

// u += d (DC/UIC: skipped once the residual met the tolerance).  norm: WRMS(d), max|d|, done flag.
__global__ __launch_bounds__(kSingleThreads) void k_update(const DPlan* dp, InstDev w) {
    __shared__ double red[kSingleThreads];
    const DPlan* P = dp;
    const StepDev sp = *w.sp;
    const int tid = threadIdx.x;
    if (sp.kind != STEP_TRAN) {
        const double res = w.norm[NRM_RES];
        if (res <= sp.tol) {
            if (tid == 0) w.norm[NRM_DONE] = 1.0;
            return;
        }
    }
    double ss = 0.0, mx = 0.0;
    for (int i = tid; i < P->n; i += kSingleThreads) {
        const double d = w.x[P->pcol[i]];
        const double ui = w.u[i] + d;
        w.u[i] = ui;
        const double r = d / (sp.rtol * fabs(ui) + sp.atol);
        ss += r * r;
        const double a = fabs(d);
        if (a != a || a > mx) mx = a;
    }
    const double s = block_reduce<kSingleThreads>(ss, red, OpSum());
    const double m = block_reduce<kSingleThreads>(mx, red, OpMax());
    if (tid == 0) {
        w.norm[NRM_WRMS] = sqrt(s / P->n);
        w.norm[NRM_MAXD] = m;
        w.norm[NRM_DONE] = 0.0;
    }
}

__global__ __launch_bounds__(kSingleThreads) void k_charges(const DPlan* dp, InstDev w,
                                                            const double* v, double* q) {
    const DPlan* P = dp;
    dev_charges<kSingleThreads>(P, w.Clin, v, q);
}

void launch_lin_setup(const DPlan* dp, const InstDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_lin_setup, dim3(1), dim3(kSingleThreads), 0, s, dp, w);
}
void launch_eval_assemble(const DPlan* dp, const InstDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_eval_assemble, dim3(1), dim3(kSingleThreads), 0, s, dp, w);
}
void launch_lu(const DPlan* dp, const InstDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_lu, dim3(1), dim3(kSingleThreads), 0, s, dp, w);
}
void launch_solve(const DPlan* dp, const InstDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_solve, dim3(1), dim3(kSingleThreads), 0, s, dp, w);
}
void launch_solve_t(const DPlan* dp, const InstDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_solve_t, dim3(1), dim3(kSingleThreads), 0, s, dp, w);
}
void launch_update(const DPlan* dp, const InstDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_update, dim3(1), dim3(kSingleThreads), 0, s, dp, w);
}
void launch_charges(const DPlan* dp, const InstDev& w, const double* v, double* q,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_charges, dim3(1), dim3(kSingleThreads), 0, s, dp, w, v, q);
}

// ---------------------------------------------------------------------------
// batched persistent transient kernel
// ---------------------------------------------------------------------------
namespace {

constexpr int NT = kBatchThreads;

struct Lds {
    double *F, *inv, *u, *x, *qx, *st;
};

template <int NR>
__device__ void run_instance(const DPlan* __restrict__ P, const double* __restrict__ p,
                             const double* __restrict__ u0, double* __restrict__ wave,
                             double* __restrict__ ufin, cs_tran_stats* __restrict__ stp,
                             const cs_tran_opts& o, const Lds& L, double* __restrict__ Glin,
                             double* __restrict__ Clin, double* __restrict__ linbuf) {
    const int tid = threadIdx.x;
    const int n = P->n;
    const StepDefaults sd = step_defaults(o);
    dev_lin_setup<NT>(P, p, linbuf, Glin, Clin);

    double Uh0[NR], Uh1[NR], Uh2[NR], Qh0[NR], Qh1[NR], up[NR];
    long long steps = 0, rej = 0, nit = 0;
    int rc = CS_OK;
    const bool uic = (o.init == CS_INIT_UIC) && u0 != nullptr;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int i = tid + NT * r;
        Uh0[r] = Uh1[r] = Uh2[r] = Qh0[r] = Qh1[r] = up[r] = 0.0;
        if (i < n) L.u[i] = uic ? u0[i] : 0.0;
    }
    __syncthreads();

    // ---- consistent initial point -------------------------------------
    {
        const int mode = uic ? CS_MODE_TRAN : CS_MODE_DCOP;
        const int maxit = uic ? 50 : 100;
        dev_sources<NT>(P, p, o.t_start, mode, L.st);
        bool ok = false;
        for (int it = 0; it < maxit; it++) {
            dev_mos<NT>(P, p, L.u, L.st);
            __syncthreads();
            const double lm = dev_assemble<NT>(P, Glin, Clin, 0.0, L.u, nullptr, L.st, L.F, L.x,
                                               uic ? u0 : nullptr);
            const double fn = block_reduce<NT>(lm, nullptr, OpMax());
            if (fn <= o.dc_abstol) { ok = true; break; }
            if (!(fn == fn) || isinf(fn)) break;
            __syncthreads();
            dev_lu<NT>(P, L.F, L.inv);
            dev_fwd<NT>(P, L.F, L.x);
            dev_bwd<NT>(P, L.F, L.inv, L.x);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) L.u[i] += L.x[P->pcol[i]];
            }
            nit++;
            __syncthreads();
        }
        if (!ok) rc = CS_E_NOCONV;
    }

    double t = o.t_start;
    if (rc == CS_OK) {
        __syncthreads();
        dev_charges<NT>(P, Clin, L.u, L.qx);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int i = tid + NT * r;
            if (i < n) { Uh0[r] = L.u[i]; Qh0[r] = L.qx[i]; }
        }
        double T0 = o.t_start, T1 = 0.0, T2 = 0.0;
        int npts = 1;
        double h = sd.dt_init;
        int kout = 0;
        const int n_probe = P->n_probe;
        if (wave) {
            while (kout < o.n_out && o.t_start + kout * o.dt_out <= o.t_start + 1e-9 * o.dt_out) {
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const int i = tid + NT * r;
                    if (i < n) {
                        const int sl = P->probe_slot[i];
                        if (sl >= 0) wave[(size_t)kout * n_probe + sl] = Uh0[r];
                    }
                }
                kout++;
            }
        }
        __syncthreads();
        while (t < o.t_stop) {
            if (steps + rej >= o.max_steps) { rc = CS_E_MAXSTEPS; break; }
            double lb = INFINITY;
            for (int s = tid; s < P->n_src; s += NT) {
                const int4 d = P->src[s];
                const double b = wave_next_bp(d.x, p + d.y, d.z, t);
                if (b < lb) lb = b;
            }
            double tb = block_reduce<NT>(lb, nullptr, OpMin());
            const bool srcbp = tb < o.t_stop;
            if (!srcbp) tb = o.t_stop;
            double hh = h < sd.dt_max ? h : sd.dt_max;
            int hit;
            const double tn = clip_step(t, hh, tb, hit);
            const int k = bdf_order(npts, o.max_order);
            double a0, a1, a2, w0, w1, w2;
            bdf_coeffs(k, tn, T0, T1, a0, a1, a2);
            const int np = npts < k + 1 ? npts : k + 1;
            pred_weights(np, tn, T0, T1, T2, w0, w1, w2);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) {
                    double v = Uh0[r];
                    if (np == 2) v = w0 * Uh0[r] + w1 * Uh1[r];
                    else if (np == 3) v = w0 * Uh0[r] + w1 * Uh1[r] + w2 * Uh2[r];
                    up[r] = v;
                    L.u[i] = v;
                    L.qx[i] = a1 * Qh0[r] + (k == 2 ? a2 * Qh1[r] : 0.0);
                }
            }
            dev_sources<NT>(P, p, tn, CS_MODE_TRAN, L.st);
            __syncthreads();
            bool conv = false;
            for (int it = 0; it < o.newton_maxit; it++) {
                dev_mos<NT>(P, p, L.u, L.st);
                __syncthreads();
                dev_assemble<NT>(P, Glin, Clin, a0, L.u, L.qx, L.st, L.F, L.x, nullptr);
                __syncthreads();
                dev_lu<NT>(P, L.F, L.inv);
                dev_fwd<NT>(P, L.F, L.x);
                dev_bwd<NT>(P, L.F, L.inv, L.x);
                nit++;
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
                if (!(nr == nr)) break;
                if (nr <= o.newton_tol) { conv = true; break; }
            }
            if (!conv) {
                rej++;
                h = 0.25 * hh;
                if (h < sd.dt_min) { rc = CS_E_DTMIN; break; }
                continue;
            }
            double fac = 2.0;
            if (npts >= 2) {
                const double c = lte_coeff(k, tn, T0, T1, T2, a0, a1, a2);
                double ss = 0.0;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const int i = tid + NT * r;
                    if (i < n) {
                        const double ui = L.u[i];
                        const double est = c * (ui - up[r]);
                        const double m = fabs(ui) > fabs(Uh0[r]) ? fabs(ui) : fabs(Uh0[r]);
                        const double q = est / (o.reltol * m + o.abstol);
                        ss += q * q;
                    }
                }
                const double err = sqrt(block_reduce<NT>(ss, nullptr, OpSum()) / n);
                if (!(err <= 1.0)) {
                    rej++;
                    h = hh * reject_factor(err, k);
                    if (h < sd.dt_min) { rc = CS_E_DTMIN; break; }
                    continue;
                }
                fac = accept_factor(err, k);
            }
            // accept
            dev_charges<NT>(P, Clin, L.u, L.qx);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int i = tid + NT * r;
                if (i < n) {
                    Uh2[r] = Uh1[r]; Uh1[r] = Uh0[r]; Uh0[r] = L.u[i];
                    Qh1[r] = Qh0[r]; Qh0[r] = L.qx[i];
                }
            }
            T2 = T1; T1 = T0; T0 = tn;
            if (npts < 3) npts++;
            steps++;
            if (wave) {
                while (kout < o.n_out) {
                    const double tq = o.t_start + kout * o.dt_out;
                    if (tq > tn + 1e-9 * o.dt_out) break;
                    double l0, l1, l2 = 0.0;
                    if (npts >= 3) {
                        l0 = (tq - T1) * (tq - T2) / ((T0 - T1) * (T0 - T2));
                        l1 = (tq - T0) * (tq - T2) / ((T1 - T0) * (T1 - T2));
                        l2 = (tq - T0) * (tq - T1) / ((T2 - T0) * (T2 - T1));
                    } else {
                        const double s = (tq - T1) / (T0 - T1);
                        l0 = s;
                        l1 = 1.0 - s;
                    }
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const int i = tid + NT * r;
                        if (i < n) {
                            const int sl = P->probe_slot[i];
                            if (sl >= 0)
                                wave[(size_t)kout * n_probe + sl] =
                                    l0 * Uh0[r] + l1 * Uh1[r] + l2 * Uh2[r];
                        }
                    }
                    kout++;
                }
            }
            t = tn;
            h = hh * fac;
            if (hit && srcbp) {
                npts = 1;
                h = 0.1 * h;
                if (h < sd.dt_min) h = sd.dt_min;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int i = tid + NT * r;
        if (i < n) ufin[i] = (rc == CS_OK || steps > 0) ? Uh0[r] : L.u[i];
    }
    if (tid == 0) {
        stp->steps = steps;
        stp->rejected = rej;
        stp->newton_iters = nit;
        stp->t_end = t;
        stp->retcode = rc;
        stp->pad = 0;
    }
    __syncthreads();
}

template <int NR>
__global__ __launch_bounds__(NT, 4) void k_tran_batch(BatchArgs A) {
    extern __shared__ double lds[];
    __shared__ int s_idx;
    Lds L;
    L.F = lds;
    L.inv = L.F + A.nnz_max;
    L.u = L.inv + A.n_max;
    L.x = L.u + A.n_max;
    L.qx = L.x + A.n_max;
    L.st = L.qx + A.n_max;
    double* Glin = A.glin + (size_t)blockIdx.x * A.nnz_max;
    double* Clin = A.clin + (size_t)blockIdx.x * A.nnz_max;
    double* linbuf = A.linbuf + (size_t)blockIdx.x * A.lin_max;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(A.queue, 1);
        __syncthreads();
        // wave-uniform queue index: the plan and instance pointers then live
        // in scalar registers
        const int qi = __builtin_amdgcn_readfirstlane(s_idx);
        __syncthreads();
        if (qi >= A.n_inst) break;
        const int inst = __builtin_amdgcn_readfirstlane(A.order[qi]);
        const int cid = __builtin_amdgcn_readfirstlane(A.inst_cid[inst]);
        const DPlan* P = A.plans + cid;
        run_instance<NR>(P, A.params + A.p_off[inst], A.u0 ? A.u0 + A.u_off[inst] : nullptr,
                         A.wave ? A.wave + A.w_off[inst] : nullptr, A.u_final + A.u_off[inst],
                         A.stats + inst, A.o, L, Glin, Clin, linbuf);
    }
}

}  // namespace

size_t batch_lds_bytes(const BatchArgs& a) {
    return sizeof(double) * ((size_t)a.nnz_max + 4 * (size_t)a.n_max + a.stamp_max);
}

template <int NR>
static int max_blocks_nr(size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_batch<NR>, NT, lds) != hipSuccess)
        return 0;
    return nb;
}

int batch_max_blocks(int nr, size_t lds) {
    switch (nr) {
    case 1: return max_blocks_nr<1>(lds);
    case 2: return max_blocks_nr<2>(lds);
    case 4: return max_blocks_nr<4>(lds);
    default: return max_blocks_nr<8>(lds);
    }
}

hipError_t launch_tran_batch(int nr, const BatchArgs& a, int nblocks, size_t lds, hipStream_t s) {
    switch (nr) {
    case 1: hipLaunchKernelGGL(k_tran_batch<1>, dim3(nblocks), dim3(NT), lds, s, a); break;
    case 2: hipLaunchKernelGGL(k_tran_batch<2>, dim3(nblocks), dim3(NT), lds, s, a); break;
    case 4: hipLaunchKernelGGL(k_tran_batch<4>, dim3(nblocks), dim3(NT), lds, s, a); break;
    default: hipLaunchKernelGGL(k_tran_batch<8>, dim3(nblocks), dim3(NT), lds, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace cedar
