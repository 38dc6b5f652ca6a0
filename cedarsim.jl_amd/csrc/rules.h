// rules.h — numeric rules shared by the plain (host-driven) path and the
// batched device kernel, so both take exactly the same steps.
//
//   waveforms   src/spectre_env.jl:15-69 (pwl_at_time / find_t_in_ts: a
//               breakpoint belongs to the segment that starts there),
//               :153-166 (pulse = pwl over t mod period), :169-176 (spsin),
//               :71-77 (breakpoints at every corner, per period)
//   MOSFET      square-law Verilog-A module models/sqlaw.va
//   integrator  variable-step BDF1/BDF2 with a polynomial predictor and a
//               local-error estimate c*(u - u_pred) (the role of IDA,
//               src/sweeps.jl:456), rules stated in the transient loop.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/cedarhip.h"

namespace cedar {

#define CEDAR_HD __host__ __device__ __forceinline__

// 1-based i = searchsortedfirst(ts, t), +1 when ts[i] == t
CEDAR_HD double pwl_at(const double* d, int npts, double t) {
    int i = 1;
    while (i <= npts && d[2 * (i - 1)] < t) i++;
    if (i <= npts && d[2 * (i - 1)] == t) i++;
    if (i <= 1) return d[1];
    if (i > npts) return d[2 * (npts - 1) + 1];
    double t0 = d[2 * (i - 2)], y0 = d[2 * (i - 2) + 1];
    double t1 = d[2 * (i - 1)], y1 = d[2 * (i - 1) + 1];
    if (y0 == y1) return y1;
    if (t1 == t0) return (y0 + y1) / 2;
    double slope = (y1 - y0) / (t1 - t0);
    return y0 + (t - t0) * slope;
}

// sine of an angle given in degrees, reduced in degrees first (Julia sind)
CEDAR_HD double sin_deg(double x) {
    double r = fmod(x, 360.0);
    if (r == 0.0 || r == 180.0 || r == -180.0) return 0.0;
    return sin(r * (3.14159265358979323846 / 180.0));
}

// pwl_at on the four corners of one pulse period, with scalars only
CEDAR_HD double pulse_at(double v1, double v2, double c0, double c1, double c2, double c3,
                         double t) {
    // i = 1 + #(corners < t), +1 when the next corner equals t
    int i = 1 + (c0 < t) + (c1 < t) + (c2 < t) + (c3 < t);
    const double ci = i == 1 ? c0 : i == 2 ? c1 : i == 3 ? c2 : c3;
    if (i <= 4 && ci == t) i++;
    if (i <= 1) return v1;
    if (i > 4) return v1;
    const double ta = i == 2 ? c0 : i == 3 ? c1 : c2;
    const double tb = i == 2 ? c1 : i == 3 ? c2 : c3;
    const double ya = i == 2 ? v1 : v2;
    const double yb = i == 4 ? v1 : v2;
    if (ya == yb) return yb;
    if (tb == ta) return (ya + yb) / 2;
    const double slope = (yb - ya) / (tb - ta);
    return ya + (t - ta) * slope;
}

CEDAR_HD double wave_value(int kind, const double* d, int len, double t) {
    if (kind == CS_WAVE_DC) return d[0];
    if (kind == CS_WAVE_PULSE) {
        const double c0 = d[2], c1 = d[2] + d[3], c2 = d[2] + d[3] + d[5];
        const double c3 = d[2] + d[3] + d[5] + d[4];
        return pulse_at(d[0], d[1], c0, c1, c2, c3, fmod(t, d[6]));
    }
    if (kind == CS_WAVE_PWL) return pwl_at(d, len / 2, t);
    // CS_WAVE_SIN: vo va freq td theta phase ncycles
    if (d[3] < t && t < d[6] / d[2])
        return d[0] + d[1] * exp(-(t - d[3]) * d[4]) * sin_deg(360.0 * d[2] * (t - d[3]) + d[5]);
    return d[0] + d[1] * sin_deg(d[5]);
}

// smallest breakpoint strictly after t (a few ulps of slack), +Inf if none
CEDAR_HD double wave_next_bp(int kind, const double* d, int len, double t) {
    double thr = t + 8.0 * 2.220446049250313e-16 * fabs(t) + 1e-300;
    double best = INFINITY;
    if (kind == CS_WAVE_PULSE) {
        const double c0 = d[2], c1 = d[2] + d[3], c2 = d[2] + d[3] + d[5];
        const double c3 = d[2] + d[3] + d[5] + d[4];
        const double per = d[6];
        auto cand = [&](double b) {
            if (b > thr && b < best) best = b;
        };
        if (__builtin_isinf(per)) {
            cand(c0); cand(c1); cand(c2); cand(c3);
        } else {
            const double m = floor(t / per);
#pragma unroll
            for (int r = -1; r <= 1; r++) {
                const double base = (m + r) * per;
                if (c0 < per) cand(base + c0);
                if (c1 < per) cand(base + c1);
                if (c2 < per) cand(base + c2);
                if (c3 < per) cand(base + c3);
            }
        }
    } else if (kind == CS_WAVE_PWL) {
        for (int j = 0; j < len / 2; j++) {
            double b = d[2 * j];
            if (b > thr && b < best) best = b;
        }
    } else if (kind == CS_WAVE_SIN) {
        if (d[3] > thr) best = d[3];
    }
    return best;
}

// value of a source in a mode: DC mode uses the explicit dc value when it is
// not NaN, else the waveform at t = 0 ($time is 0 in :dcop)
CEDAR_HD double source_value(int kind, const double* d, int len, double dc, double t, int mode) {
    if (mode == CS_MODE_DCOP) return (dc == dc) ? dc : wave_value(kind, d, len, 0.0);
    return wave_value(kind, d, len, t);
}

// Square-law MOSFET (models/sqlaw.va): current d->s and dI/dVd, dI/dVg,
// dI/dVs.  Source and drain swap when type*Vds < 0.  gmin is in parallel
// with d-s.  dI/dVs = -(dI/dVd + dI/dVg) holds exactly.
// sqlaw_core takes beta = kp*w/l and tp*vth (tp = +-1) precomputed, the
// form the batched kernel keeps in registers; the arithmetic is the same.
CEDAR_HD void sqlaw_core(double tp, double beta, double tpvth, double gmin, double vd, double vg,
                         double vs, double& id, double& dd, double& dg, double& ds) {
    const double vds = tp * (vd - vs);
    const bool fwd = vds >= 0.0;
    const double vov = fwd ? tp * (vg - vs) - tpvth : tp * (vg - vd) - tpvth;
    const double vx = fwd ? vds : -vds;
    double I = 0.0, gm = 0.0, gds = 0.0;
    if (vov > 0.0) {
        if (vx < vov) {
            I = beta * (vov * vx - 0.5 * vx * vx);
            gm = beta * vx;
            gds = beta * (vov - vx);
        } else {
            I = 0.5 * beta * vov * vov;
            gm = beta * vov;
        }
    }
    if (fwd) {
        id = tp * I;
        dg = gm;
        dd = gds;
        ds = -(gm + gds);
    } else {
        id = -tp * I;
        dg = -gm;
        dd = gm + gds;
        ds = -gds;
    }
    id += gmin * (vd - vs);
    dd += gmin;
    ds -= gmin;
}

CEDAR_HD void sqlaw_eval(int type, double vth, double kp, double w, double l, double gmin,
                         double vd, double vg, double vs, double& id, double& dd,
                         double& dg, double& ds) {
    const double tp = (double)type;
    sqlaw_core(tp, kp * w / l, tp * vth, gmin, vd, vg, vs, id, dd, dg, ds);
}

// ---- BDF step rules --------------------------------------------------------
// Order: BDF2 when the segment holds >= 3 points and max_order >= 2.
CEDAR_HD int bdf_order(int npts, int max_order) { return (npts >= 3 && max_order >= 2) ? 2 : 1; }

// d/dt at tn of the interpolant through (tn, T0[, T1]): a0, a1, a2
CEDAR_HD void bdf_coeffs(int k, double tn, double T0, double T1, double& a0, double& a1,
                         double& a2) {
    if (k == 1) {
        a0 = 1.0 / (tn - T0);
        a1 = -a0;
        a2 = 0.0;
    } else {
        a0 = 1.0 / (tn - T0) + 1.0 / (tn - T1);
        a1 = (tn - T1) / ((T0 - tn) * (T0 - T1));
        a2 = (tn - T0) / ((T1 - tn) * (T1 - T0));
    }
}

// Lagrange weights of the predictor through np points (T0, T1, T2) at tn.
CEDAR_HD void pred_weights(int np, double tn, double T0, double T1, double T2, double& l0,
                           double& l1, double& l2) {
    if (np == 1) {
        l0 = 1.0; l1 = 0.0; l2 = 0.0;
    } else if (np == 2) {
        const double s = (tn - T0) / (T0 - T1);
        l0 = 1.0 + s; l1 = -s; l2 = 0.0;
    } else {
        l0 = (tn - T1) * (tn - T2) / ((T0 - T1) * (T0 - T2));
        l1 = (tn - T0) * (tn - T2) / ((T1 - T0) * (T1 - T2));
        l2 = (tn - T0) * (tn - T1) / ((T2 - T0) * (T2 - T1));
    }
}

// Local-error constant c: est = c * (u - u_pred) for BDF-k with a predictor
// through T0..Tk:  c = sum_j a_j (t_j - tn)^(k+1) / (a0 * prod_j (tn - T_j)).
CEDAR_HD double lte_coeff(int k, double tn, double T0, double T1, double T2, double a0,
                          double a1, double a2) {
    if (k == 1) {
        const double d1 = T0 - tn;
        return (a1 * d1 * d1) / (a0 * (tn - T0) * (tn - T1));
    }
    const double d1 = T0 - tn, d2 = T1 - tn;
    const double num = a1 * d1 * d1 * d1 + a2 * d2 * d2 * d2;
    return num / (a0 * (tn - T0) * (tn - T1) * (tn - T2));
}

struct StepDefaults {
    double dt_init, dt_min, dt_max;
};
CEDAR_HD StepDefaults step_defaults(const cs_tran_opts& o) {
    const double span = o.t_stop - o.t_start;
    StepDefaults s;
    s.dt_init = o.dt_init > 0 ? o.dt_init : span * 1e-9;
    s.dt_min = o.dt_min > 0 ? o.dt_min : span * 1e-15;
    s.dt_max = o.dt_max > 0 ? o.dt_max : span / 10.0;
    return s;
}

// Cut the step against the next breakpoint tb: land on it when it is within
// one step, halve when it would leave less than a step.  Returns tn.
CEDAR_HD double clip_step(double t, double& hh, double tb, int& hit) {
    hit = 0;
    if (tb - t <= hh * (1.0 + 1e-9)) {
        hh = tb - t;
        hit = 1;
    } else if (tb - t < 2.0 * hh) {
        hh = 0.5 * (tb - t);
    }
    double tn = hit ? tb : t + hh;
    hh = tn - t;
    return tn;
}

// step factor after an LTE test with error err (in tolerance units)
CEDAR_HD double reject_factor(double err, int k) {
    double f = (err == err) ? 0.9 * pow(err, -1.0 / (double)(k + 1)) : 0.1;
    return f < 0.1 ? 0.1 : f;
}
CEDAR_HD double accept_factor(double err, int k) {
    double f = (err > 0.0) ? 0.9 * pow(err, -1.0 / (double)(k + 1)) : 2.0;
    return f > 2.0 ? 2.0 : f;
}

}  // namespace cedar
