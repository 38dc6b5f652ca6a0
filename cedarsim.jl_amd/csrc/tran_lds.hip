// tran_lds.hip — the LDS-resident batched transient kernel (k_tran_lds).
//
// One 64-lane wavefront (one workgroup) owns one circuit instance at a time
// and walks a persistent work queue.  When it starts an instance it copies
// the circuit's 16-bit plan image (compact.h) into LDS — skipped when the
// previous instance used the same circuit — and computes the instance's
// linear stamps G, C and source data into LDS.  From then on a Newton
// iteration touches only LDS and registers:
//
//   eval      lane m evaluates MOSFETs m, m+64, ...: terminals (pivot
//             columns), beta = kp*W/L and type*Vth stay in registers;
//             (id, dI/dVd, dI/dVg) go to the stamp buffer;
//   assemble  lane k owns pivot rows k, k+64, ...: F entries gathered from
//             fixed contributor lists (no atomics, fixed summation order),
//             residual -> x;
//   A levels  LU tasks and forward-substitution tasks, one barrier per level;
//   B levels  backward substitution, one barrier per level;
//   update    u[k] += x[k] (pivot order: no permutation), WRMS norm by a
//             cross-lane reduction.
//
// Step control, predictor, local-error test and output interpolation follow
// rules.h exactly as the global-plan kernel (tran_device.h) does; history
// (u and charges of the last three accepted points) lives in the registers
// of the lane that owns the row/column.
#include "compact.h"
#include "device.h"
#include "kernels.h"

namespace cedar {

namespace {

constexpr int W = 64;

// Cold, register-hungry math (waveforms with fmod/sin/exp, pow in the step
// factors) stays out of line so it does not raise the register count of the
// Newton loop; it runs once per step.
__device__ __noinline__ double cold_source(int kind, const double* d, int len, double dc, double t,
                                           int mode) {
    return source_value(kind, d, len, dc, t, mode);
}
__device__ __noinline__ double cold_next_bp(int kind, const double* d, int len, double t) {
    return wave_next_bp(kind, d, len, t);
}
__device__ __noinline__ double cold_reject(double err, int k) { return reject_factor(err, k); }
__device__ __noinline__ double cold_accept(double err, int k) { return accept_factor(err, k); }

__device__ __forceinline__ int lane_id() { return (int)threadIdx.x; }

__device__ __forceinline__ double sv(const double* st, uint32_t r) {
    const double v = st[r >> 1];
    return (r & 1) ? -v : v;
}

// Per-lane MOSFET constants of NM device slots.
template <int NM>
struct DevRegs {
    double beta[NM], tpvth[NM];
    uint32_t dg[NM], sx[NM];   // d | g << 16;  s | (type > 0) << 16
};

struct View {
    double *F, *G, *C, *st, *u, *x, *inv, *wd;
    const uint16_t* B;
};

template <int NM>
__device__ __forceinline__ void eval_mos(const CPlan& c, const View& v, const DevRegs<NM>& dr,
                                         double gmin) {
    const int nm = c.n_mos;
#pragma unroll
    for (int j = 0; j < NM; j++) {
        const int m = lane_id() + W * j;
        if (m < nm) {
            const uint32_t d = dr.dg[j] & 0xFFFF, g = dr.dg[j] >> 16, s = dr.sx[j] & 0xFFFF;
            const double vd = d != CP_GROUND ? v.u[d] : 0.0;
            const double vg = g != CP_GROUND ? v.u[g] : 0.0;
            const double vs = s != CP_GROUND ? v.u[s] : 0.0;
            const double tp = (dr.sx[j] >> 16) ? 1.0 : -1.0;
            double id, dd, dgv, ds;
            sqlaw_core(tp, dr.beta[j], dr.tpvth[j], gmin, vd, vg, vs, id, dd, dgv, ds);
            v.st[m] = id;
            v.st[nm + m] = dd;
            v.st[2 * nm + m] = dgv;
        }
    }
}

__device__ __forceinline__ void eval_sources(const CPlan& c, const View& v, double t, int mode) {
    for (int s = lane_id(); s < c.n_src; s += W) {
        const uint16_t* d = v.B + c.o_src + 4 * s;
        const int kind = d[0], off = d[1], len = d[2];
        const double* wd = v.wd + off;
        v.st[3 * c.n_mos + s] = cold_source(kind, wd, len, wd[len], t, mode);
    }
}

// F = G + a0*C + nonlinear stamps; x = -residual (qx added per row).  With
// pin, rows that carry charge become u_k = pin_k.  Returns max |residual|.
template <int NR>
__device__ __forceinline__ double assemble(const CPlan& c, const View& v, double a0,
                                           const double (&qx)[NR], bool pin,
                                           const double (&pinv)[NR], const uint32_t (&flag)[NR]) {
    const uint16_t* B = v.B;
    double lmax = 0.0;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int k = lane_id() + W * r;
        if (k < c.n) {
            const uint4 rd = reinterpret_cast<const uint4*>(B + c.o_rdesc)[k];
            const int e0 = rd.x & 0xFFFF, e1 = rd.x >> 16, dg = rd.y & 0xFFFF;
            double res;
            if (pin && (flag[r] & 1)) {
                for (int e = e0; e < e1; e++) v.F[e] = (e == dg) ? 1.0 : 0.0;
                res = v.u[k] - pinv[r];
            } else {
                int q = rd.y >> 16;
                for (int e = e0; e < e1; e++) {
                    const int qe = B[c.o_jptr + e + 1];
                    double s = 0.0;
                    for (; q < qe; q++) s += sv(v.st, B[c.o_jref + q]);
                    v.F[e] = s;
                }
                double acc = 0.0;
                const int l0 = rd.z & 0xFFFF, l1 = rd.z >> 16;
                for (int l = l0; l < l1; l++) {
                    const double jl = v.G[l] + a0 * v.C[l];
                    v.F[B[c.o_lpos + l]] += jl;
                    acc += jl * v.u[B[c.o_lcol + l]];
                }
                acc += qx[r];
                const int r0 = rd.w & 0xFFFF, r1 = rd.w >> 16;
                for (int q2 = r0; q2 < r1; q2++) acc += sv(v.st, B[c.o_rref + q2]);
                res = acc;
            }
            v.x[k] = -res;
            const double a = fabs(res);
            if (a != a || a > lmax) lmax = a;
        }
    }
    return lmax;
}

// LU with the forward substitution folded in (A levels), then the backward
// substitution (B levels).  On return x holds the Newton step.
__device__ __forceinline__ void factor_solve(const CPlan& c, const View& v) {
    const uint16_t* B = v.B;
    const uint2* ft = reinterpret_cast<const uint2*>(B + c.o_ftask);
    const uint2* xt = reinterpret_cast<const uint2*>(B + c.o_xtask);
    const uint32_t* up = reinterpret_cast<const uint32_t*>(B + c.o_upd);
    double* F = v.F;
    double* x = v.x;
    double* inv = v.inv;
    for (int l = 0; l < c.nA; l++) {
        const uint2 ad = reinterpret_cast<const uint2*>(B + c.o_adesc)[l];
        const int fb = ad.x & 0xFFFF, nf = (ad.x >> 16) - fb;
        const int xb = ad.y & 0xFFFF, nx = (ad.y >> 16) - xb;
        for (int q = lane_id(); q < nf + nx; q += W) {
            if (q < nf) {
                const uint2 t = ft[fb + q];
                const int e = t.x & 0xFFFF, ub = t.x >> 16, ue = t.y & 0xFFFF;
                const uint32_t ctl = t.y >> 16;
                double acc = F[e];
                for (int w = ub; w < ue; w++) {
                    const uint32_t ab = up[w];
                    acc -= F[ab & 0xFFFF] * F[ab >> 16];
                }
                if (ctl < CP_PLAIN_U) {
                    F[e] = acc * inv[ctl];
                } else {
                    F[e] = acc;
                    if (ctl & CP_DIAG) inv[ctl & 0x7FFF] = 1.0 / acc;
                }
            } else {
                const uint2 t = xt[xb + q - nf];
                const int k = t.x & 0xFFFF, ub = t.x >> 16, ue = t.y & 0xFFFF;
                double acc = x[k];
                for (int w = ub; w < ue; w++) {
                    const uint32_t ab = up[w];
                    acc -= F[ab & 0xFFFF] * x[ab >> 16];
                }
                x[k] = acc;
            }
        }
        __syncthreads();
    }
    for (int l = 0; l < c.nB; l++) {
        const uint32_t bd = reinterpret_cast<const uint32_t*>(B + c.o_bdesc)[l];
        const int xb = bd & 0xFFFF, nx = (bd >> 16) - xb;
        for (int q = lane_id(); q < nx; q += W) {
            const uint2 t = xt[xb + q];
            const int k = t.x & 0xFFFF, ub = t.x >> 16, ue = t.y & 0xFFFF;
            double acc = x[k];
            for (int w = ub; w < ue; w++) {
                const uint32_t ab = up[w];
                acc -= F[ab & 0xFFFF] * x[ab >> 16];
            }
            x[k] = acc * inv[k];
        }
        __syncthreads();
    }
}

// charge of pivot row k: sum over its linear entries C * u
__device__ __forceinline__ double row_charge(const CPlan& c, const View& v, int k) {
    const uint16_t* B = v.B;
    const uint32_t lz = reinterpret_cast<const uint32_t*>(B + c.o_rdesc)[4 * k + 2];
    double acc = 0.0;
    for (int l = lz & 0xFFFF; l < (int)(lz >> 16); l++) acc += v.C[l] * v.u[B[c.o_lcol + l]];
    return acc;
}

__device__ __forceinline__ double lin_value(const DPlan* __restrict__ P, const double* __restrict__ p,
                                            int32_t ref) {
    const int s = ref >> 1;
    const int kd = P->lin_kind[s];
    double val = 1.0;
    if (kd != 0) {
        const double pv = p[P->lin_idx[s]];
        val = kd == 1 ? 1.0 / pv : kd == 2 ? pv : -pv;
    }
    return (ref & 1) ? -val : val;
}

template <int NR>
__device__ void run_lds(const CPlan& c, const DPlan* __restrict__ P, const double* __restrict__ p,
                        const double* __restrict__ u0, double* __restrict__ wave,
                        double* __restrict__ ufin, cs_tran_stats* __restrict__ stp,
                        const cs_tran_opts& o, const View& v) {
    constexpr int NM = 2 * NR;
    const int tid = lane_id();
    const int n = c.n;
    const uint16_t* B = v.B;
    const StepDefaults sd = step_defaults(o);
    const bool uic = (o.init == CS_INIT_UIC) && u0 != nullptr;

    // ---- instance set-up: linear stamps, source data, device constants --
    for (int l = tid; l < c.nlin; l += W) {
        const int e = B[c.o_lpos + l];
        double g = 0.0, cc = 0.0;
        for (int q = P->g_ptr[e]; q < P->g_ptr[e + 1]; q++) g += lin_value(P, p, P->g_ref[q]);
        for (int q = P->c_ptr[e]; q < P->c_ptr[e + 1]; q++) cc += lin_value(P, p, P->c_ref[q]);
        v.G[l] = g;
        v.C[l] = cc;
    }
    for (int s = tid; s < c.n_src; s += W) {
        const int4 d = P->src[s];
        const int off = B[c.o_src + 4 * s + 1];
        for (int j = 0; j < d.z; j++) v.wd[off + j] = p[d.y + j];
        v.wd[off + d.z] = p[d.w];
    }
    DevRegs<NM> dr;
    const double gmin = p[P->p_gmin];
#pragma unroll
    for (int j = 0; j < NM; j++) {
        const int m = tid + W * j;
        dr.beta[j] = 0.0;
        dr.tpvth[j] = 0.0;
        dr.dg[j] = 0;
        dr.sx[j] = 0;
        if (m < c.n_mos) {
            const int4 dv = reinterpret_cast<const int4*>(c.dev)[m];
            const double tp = (double)dv.w;
            dr.beta[j] = p[P->p_mos_kp + m] * p[P->p_mos_w + m] / p[P->p_mos_l + m];
            dr.tpvth[j] = tp * p[P->p_mos_vth + m];
            dr.dg[j] = (uint32_t)dv.x | ((uint32_t)dv.y << 16);
            dr.sx[j] = (uint32_t)dv.z | ((dv.w > 0 ? 1u : 0u) << 16);
        }
    }
    double Uh0[NR], Uh1[NR], Uh2[NR], Qh0[NR], Qh1[NR], up[NR], qx[NR], pinv[NR];
    uint32_t flag[NR];
    int32_t cn[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int k = tid + W * r;
        Uh0[r] = Uh1[r] = Uh2[r] = Qh0[r] = Qh1[r] = up[r] = qx[r] = pinv[r] = 0.0;
        flag[r] = 0;
        cn[r] = 0;
        if (k < n) {
            cn[r] = c.colnat[k];
            flag[r] = c.diffrow[k] ? 1u : 0u;
            pinv[r] = uic ? u0[cn[r]] : 0.0;
            v.u[k] = pinv[r];
        }
    }
    __syncthreads();

    long long steps = 0, rej = 0, nit = 0;
    int rc = CS_OK;
    // ---- consistent initial point ---------------------------------------
    {
        const int mode = uic ? CS_MODE_TRAN : CS_MODE_DCOP;
        const int maxit = uic ? 50 : 100;
        eval_sources(c, v, o.t_start, mode);
        bool ok = false;
        for (int it = 0; it < maxit; it++) {
            eval_mos<NM>(c, v, dr, gmin);
            __syncthreads();
            const double lm = assemble<NR>(c, v, 0.0, qx, uic, pinv, flag);
            const double fn = block_reduce<W>(lm, nullptr, OpMax());
            if (fn <= o.dc_abstol) { ok = true; break; }
            if (!(fn == fn) || __builtin_isinf(fn)) break;
            __syncthreads();
            factor_solve(c, v);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int k = tid + W * r;
                if (k < n) v.u[k] += v.x[k];
            }
            nit++;
            __syncthreads();
        }
        if (!ok) rc = CS_E_NOCONV;
    }

    double t = o.t_start;
    if (rc == CS_OK) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int k = tid + W * r;
            if (k < n) {
                Uh0[r] = v.u[k];
                Qh0[r] = row_charge(c, v, k);
            }
        }
        double T0 = o.t_start, T1 = 0.0, T2 = 0.0;
        int npts = 1;
        double h = sd.dt_init;
        int kout = 0;
        const int n_probe = P->n_probe;
        int32_t ps[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int k = tid + W * r;
            ps[r] = (wave && k < n) ? c.probe[k] : -1;
        }
        if (wave) {
            while (kout < o.n_out && o.t_start + kout * o.dt_out <= o.t_start + 1e-9 * o.dt_out) {
#pragma unroll
                for (int r = 0; r < NR; r++)
                    if (ps[r] >= 0) wave[(size_t)kout * n_probe + ps[r]] = Uh0[r];
                kout++;
            }
        }
        __syncthreads();
        while (t < o.t_stop) {
            if (steps + rej >= o.max_steps) { rc = CS_E_MAXSTEPS; break; }
            double lb = INFINITY;
            for (int s = tid; s < c.n_src; s += W) {
                const uint16_t* d = B + c.o_src + 4 * s;
                const double b = cold_next_bp(d[0], v.wd + d[1], d[2], t);
                if (b < lb) lb = b;
            }
            double tb = block_reduce<W>(lb, nullptr, OpMin());
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
                const int kk = tid + W * r;
                if (kk < n) {
                    double val = Uh0[r];
                    if (np == 2) val = w0 * Uh0[r] + w1 * Uh1[r];
                    else if (np == 3) val = w0 * Uh0[r] + w1 * Uh1[r] + w2 * Uh2[r];
                    up[r] = val;
                    v.u[kk] = val;
                    qx[r] = a1 * Qh0[r] + (k == 2 ? a2 * Qh1[r] : 0.0);
                }
            }
            eval_sources(c, v, tn, CS_MODE_TRAN);
            __syncthreads();
            bool conv = false;
            for (int it = 0; it < o.newton_maxit; it++) {
                eval_mos<NM>(c, v, dr, gmin);
                __syncthreads();
                assemble<NR>(c, v, a0, qx, false, pinv, flag);
                __syncthreads();
                factor_solve(c, v);
                nit++;
                double ss = 0.0;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const int kk = tid + W * r;
                    if (kk < n) {
                        const double d = v.x[kk];
                        const double ui = v.u[kk] + d;
                        v.u[kk] = ui;
                        const double q = d / (o.reltol * fabs(ui) + o.abstol);
                        ss += q * q;
                    }
                }
                const double nr = sqrt(block_reduce<W>(ss, nullptr, OpSum()) / n);
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
                const double cl = lte_coeff(k, tn, T0, T1, T2, a0, a1, a2);
                double ss = 0.0;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const int kk = tid + W * r;
                    if (kk < n) {
                        const double ui = v.u[kk];
                        const double est = cl * (ui - up[r]);
                        const double m = fabs(ui) > fabs(Uh0[r]) ? fabs(ui) : fabs(Uh0[r]);
                        const double q = est / (o.reltol * m + o.abstol);
                        ss += q * q;
                    }
                }
                const double err = sqrt(block_reduce<W>(ss, nullptr, OpSum()) / n);
                if (!(err <= 1.0)) {
                    rej++;
                    h = hh * cold_reject(err, k);
                    if (h < sd.dt_min) { rc = CS_E_DTMIN; break; }
                    continue;
                }
                fac = cold_accept(err, k);
            }
            // accept
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int kk = tid + W * r;
                if (kk < n) {
                    Uh2[r] = Uh1[r]; Uh1[r] = Uh0[r]; Uh0[r] = v.u[kk];
                    Qh1[r] = Qh0[r]; Qh0[r] = row_charge(c, v, kk);
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
                    for (int r = 0; r < NR; r++)
                        if (ps[r] >= 0)
                            wave[(size_t)kout * n_probe + ps[r]] = l0 * Uh0[r] + l1 * Uh1[r] + l2 * Uh2[r];
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
        const int k = tid + W * r;
        if (k < n) ufin[cn[r]] = (rc == CS_OK || steps > 0) ? Uh0[r] : v.u[k];
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
__global__ __launch_bounds__(W) void k_tran_lds(LdsArgs A) {
    extern __shared__ double lds[];
    uint16_t* blob = reinterpret_cast<uint16_t*>(lds + A.dbl_max);
    int cur = -1;
    for (;;) {
        int qi = 0;
        if (threadIdx.x == 0) qi = atomicAdd(A.queue, 1);
        qi = __builtin_amdgcn_readfirstlane(__shfl(qi, 0, W));
        if (qi >= A.n_inst) break;
        const int inst = __builtin_amdgcn_readfirstlane(A.order[qi]);
        const int cid = __builtin_amdgcn_readfirstlane(A.inst_cid[inst]);
        // the plan header is uniform: keep it in scalar registers (a vector
        // load followed by readfirstlane, since the compiler cannot prove the
        // header is never written and so will not use scalar loads itself)
        CPlan c;
        {
            const CPlan cv = A.cplans[cid];
            const uint32_t* src = reinterpret_cast<const uint32_t*>(&cv);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&c);
#pragma unroll
            for (int w = 0; w < (int)(sizeof(CPlan) / 4); w++)
                dst[w] = __builtin_amdgcn_readfirstlane(src[w]);
        }
        if (cid != cur) {
            // the plan image: words is a multiple of 8, the image 16-byte aligned
            const uint4* src = reinterpret_cast<const uint4*>(c.blob);
            uint4* dst = reinterpret_cast<uint4*>(blob);
            for (int q = threadIdx.x; q < c.words / 8; q += W) dst[q] = src[q];
            cur = cid;
            __syncthreads();
        }
        View v;
        v.F = lds;
        v.G = v.F + c.nnz;
        v.C = v.G + c.nlin;
        v.st = v.C + c.nlin;
        v.u = v.st + c.stamp_n;
        v.x = v.u + c.n;
        v.inv = v.x + c.n;
        v.wd = v.inv + c.n;
        v.B = blob;
        run_lds<NR>(c, A.plans + cid, A.params + A.p_off[inst],
                    A.u0 ? A.u0 + A.u_off[inst] : nullptr,
                    A.wave ? A.wave + A.w_off[inst] : nullptr, A.u_final + A.u_off[inst],
                    A.stats + inst, A.o, v);
    }
}

}  // namespace

size_t lds_kernel_bytes(int32_t dbl_max, int32_t words_max) {
    return 8 * (size_t)dbl_max + 2 * (size_t)words_max;
}

int lds_max_blocks(int nr, size_t lds) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (nr) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_lds<1>, W, lds); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_lds<2>, W, lds); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tran_lds<4>, W, lds); break;
        default: break;
    }
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_tran_lds(int nr, const LdsArgs& a, int nblocks, size_t lds, hipStream_t s) {
    switch (nr) {
        case 1: hipLaunchKernelGGL(k_tran_lds<1>, dim3(nblocks), dim3(W), lds, s, a); break;
        case 2: hipLaunchKernelGGL(k_tran_lds<2>, dim3(nblocks), dim3(W), lds, s, a); break;
        case 4: hipLaunchKernelGGL(k_tran_lds<4>, dim3(nblocks), dim3(W), lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cedar
