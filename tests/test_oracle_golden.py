"""Pins the CPU oracle against the reference's own known answers and closed forms.

Each test names the reference test it restates.  The reference cannot run here
(Julia is absent), so these are the recorded answers of test/basic.jl,
test/transients.jl, test/sweep.jl and test/sensitivity.jl, plus closed forms
and brute-force checks for what those leave open (square-law regions, the
BDF integrator's order, the dense LU).
"""
import json
import math
import os

import numpy as np
import pytest

import cedarhip as ch
from cedarhip import (C, CurrentSource, Gnd, I, L, Named, R, SquareLawMOSFET, V, VoltageSource,
                      net, pulse, pwl, spsin, vccs, vcvs)
from cedarhip.workloads import inverter
from oracle import oracle as O


GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                   "reference_answers.json")))


def dcop(fn):
    flat = ch.lower(fn)
    u, st, rc = O.dcop(flat, flat.params())
    assert rc == 0
    return flat, u


def tran(fn, tspan, n_out=101, u0=None, **kw):
    flat = ch.lower(fn)
    o = ch.grid_options(tspan, n_out=n_out, init=1 if u0 is not None else 0, **kw)
    w, uf, st, rc = O.tran(flat, flat.params(), u0, o)
    assert rc == 0, rc
    return flat, o, w, st


# ---- test/basic.jl -------------------------------------------------------------
def test_simple_vr():
    """test/basic.jl:21-43: I = V/R = 2.5 A."""
    def VR():
        vcc = Named(net, "vcc")(); gnd = Named(net, "gnd")()
        Named(V(5.0), "V")(vcc, gnd); Named(R(2), "R")(vcc, gnd); Gnd()(gnd)
    flat, u = dcop(VR)
    g = GOLD["simple_vr"]
    assert u[flat.index("vcc")] == pytest.approx(g["v_vcc"], abs=1e-7)
    # current through R = V/R; the source supplies it (branch current -2.5)
    assert u[flat.index("V.I")] == pytest.approx(g["i_source"], abs=1e-7)


def test_simple_ir():
    """test/basic.jl:81-106: I(-5) into R(2) gives V = 10."""
    def IR():
        icc = Named(net, "icc")(); gnd = Named(net, "gnd")()
        Named(I(-5.0), "I")(icc, gnd); Named(R(2), "R")(icc, gnd); Gnd()(gnd)
    flat, u = dcop(IR)
    assert u[flat.index("icc")] == pytest.approx(GOLD["simple_ir"]["v_icc"], abs=1e-7)


def test_vrc_step():
    """test/basic.jl:111-141: capacitor starts uncharged; C.I[1] = V/R,
    C.V[end] = V over (0, 1) s with tau = 2 ms."""
    def VRC():
        vcc = Named(net, "vcc")(); vrc = Named(net, "vrc")(); gnd = Named(net, "gnd")()
        Named(V(5.0), "V")(vcc, gnd); Named(R(2000.0), "R")(vcc, vrc)
        Named(C(1e-6), "C")(vrc, gnd); Gnd()(gnd)
    flat, o, w, st = tran(VRC, (0.0, 1.0), n_out=1001, u0=np.array([5.0, 0.0, 0.0]),
                          reltol=1e-8, abstol=1e-8)
    vc = w[:, flat.index("vrc")]
    g = GOLD["vrc"]
    assert vc[0] == pytest.approx(g["vc_start"], abs=1e-7)
    assert vc[-1] == pytest.approx(g["vc_end"], abs=1e-7)
    # C.I[1] = (V - Vc)/R at t = 0: the source current at t=0 is -V/R
    assert w[0, flat.index("V.I")] == pytest.approx(g["i_source_start"], rel=1e-7)
    # whole waveform against the closed form; a per-step local-error controller
    # at reltol 1e-8 accumulates a global error of a few ppm here
    t = o.t_start + o.dt_out * np.arange(o.n_out)
    assert np.max(np.abs(vc - 5.0 * (1 - np.exp(-t / 2e-3)))) < 5e-5


def test_spectre_sources():
    """test/basic.jl:168-199 (the isource / vsource / pwl lines)."""
    def src():
        n = {k: Named(net, f"node_{k}")() for k in (1, 2, 3, 4)}
        gnd = Named(net, "0")()
        Named(I(2.2e-6), "I1")(gnd, n[1]); Named(R(1000.0), "R1")(n[1], gnd)
        Named(CurrentSource(tran=pwl([0, 1e-3, 0.5, 2e-3, 1, 1.75e-3])), "I2")(gnd, n[2])
        Named(R(2e3), "R2")(n[2], gnd)
        Named(V(1.5), "V3")(gnd, n[3]); Named(R(1e3), "R3")(n[3], gnd)
        Named(VoltageSource(tran=pwl([0, 1, 0.5, 2, 1, 5])), "V4")(gnd, n[4])
        Named(R(4e3), "R4")(n[4], gnd)
        Gnd()(gnd)
    flat, o, w, st = tran(src, (0.0, 1.0), n_out=11, reltol=1e-8, abstol=1e-8)
    g = GOLD["spectre_sources"]
    assert np.allclose(w[:, flat.index("node_1")], g["node_1"])
    assert np.allclose(w[:, flat.index("node_3")], g["node_3"])
    assert w[-1, flat.index("node_2")] == pytest.approx(g["node_2_end"])
    assert w[-1, flat.index("node_4")] == pytest.approx(g["node_4_end"])
    # R4.I[end] = V/R = -1.25 mA; R3.I = -1.5 mA
    assert w[-1, flat.index("node_4")] / 4e3 == pytest.approx(g["i_r4_end"])


def test_spice_controlled_sources():
    """test/basic.jl:201-226: B5 (as a VCVS), E6, E8, G7, G9."""
    def src():
        gnd = Named(net, "0")()
        n = {k: Named(net, f"node_{k}")() for k in (1, 5, 6, 7, 8, 9)}
        Named(V(1.0), "V1")(gnd, n[1]); Named(R(1e3), "R1")(n[1], gnd)
        Named(vcvs(2.0), "B5")(gnd, n[5], n[1], gnd); Named(R(1e3), "R5")(n[5], gnd)
        Named(vcvs(2.0), "E6")(gnd, n[6], gnd, n[5]); Named(R(1e3), "R6")(n[6], gnd)
        Named(vcvs(2.0), "E8")(gnd, n[8], gnd, n[5]); Named(R(1e3), "R8")(n[8], gnd)
        Named(vccs(2.0), "G7")(gnd, n[7], gnd, n[5]); Named(R(1e3), "R7")(n[7], gnd)
        Named(vccs(2.0), "G9")(gnd, n[9], gnd, n[5]); Named(R(1e3), "R9")(n[9], gnd)
        Gnd()(gnd)
    flat, u = dcop(src)
    g = GOLD["spice_controlled"]
    for k in ("node_5", "node_6", "node_8", "node_7", "node_9"):
        assert u[flat.index(k)] == pytest.approx(g[k])


# ---- test/transients.jl ----------------------------------------------------------
def test_pwl_current_through_resistor():
    """test/transients.jl:17-63: vout = pwl_val(t) * i_max * r at 10*deftol."""
    def ckt():
        vout = Named(net, "vout")(); gnd = Named(net, "gnd")()
        Named(CurrentSource(tran=pwl([1e-3, 0.0, 9e-3, -2.0])), "I")(vout, gnd)
        Named(R(2.0), "R")(vout, gnd); Gnd()(gnd)
    flat, o, w, st = tran(ckt, (0.0, 10e-3), n_out=201, reltol=1e-8, abstol=1e-8)
    t = o.t_start + o.dt_out * np.arange(o.n_out)
    ref = np.clip((t - 1e-3) / 8e-3, 0, 1) * 2 * 2
    assert np.allclose(w[:, flat.index("vout")], ref, atol=1e-7, rtol=1e-7)


def test_pwl_breakpoint_semantics():
    """test/transients.jl:66-96 and src/spectre_env.jl:15-21: a breakpoint belongs
    to the segment that starts there; slopes 0 before, 5e8 at 100 ns."""
    g = GOLD["pwl_breakpoint_slopes"]
    ts = np.array(g["ts"])
    ys = np.array(g["ys"])
    f = lambda t: O.pwl_at_time(ts, ys, t)
    h = 1e-12
    assert (f(50e-9 + h) - f(50e-9)) / h == pytest.approx(0.0, abs=1e-6)
    assert (f(100e-9 + h) - f(100e-9)) / h == pytest.approx(g["slope_at_100ns"], rel=1e-6)
    assert (f(110e-9 + h) - f(110e-9)) / h == pytest.approx(0.0, abs=1e-6)
    assert (f(200e-9 + h) - f(200e-9)) / h == pytest.approx(g["slope_at_200ns"], rel=1e-6)
    # infinitely steep step: the value at the step is the mean of both sides
    assert O.pwl_at_time(np.array([0.0, 1.0, 1.0, 2.0]), np.array([0.0, 0.0, 5.0, 5.0]), 1.0) == 2.5
    assert O.pwl_at_time(ts, ys, -1.0) == 0.0 and O.pwl_at_time(ts, ys, 1.0) == 0.0


def test_pulse_waveform_and_breakpoints():
    """pulse(v1,v2,td,tr,tf,pw,period) = pwl over t mod period
    (src/spectre_env.jl:153-166); corners repeat every period."""
    d = [0.0, 1.8, 0.0, 100e-12, 100e-12, 0.9e-9, 2e-9]
    w = lambda t: O.wave_value(1, d, t)
    assert w(0.0) == 0.0
    assert w(50e-12) == pytest.approx(0.9)
    assert w(0.5e-9) == 1.8
    assert w(1.05e-9) == pytest.approx(0.9)
    assert w(1.5e-9) == 0.0
    assert w(2e-9 + 50e-12) == pytest.approx(0.9)
    bps = []
    t = 0.0
    for _ in range(8):
        t = O.wave_next_bp(1, d, t)
        bps.append(t)
    assert np.allclose(bps, [0.1e-9, 1.0e-9, 1.1e-9, 2e-9, 2.1e-9, 3.0e-9, 3.1e-9, 4e-9])


def test_butterworth_sine():
    """test/transients.jl:129-179: third-order Butterworth driven by sin(t),
    against the analytic inverse Laplace transform."""
    def ckt():
        vin = Named(net, "vin")(); n1 = Named(net, "n1")(); vout = Named(net, "vout")()
        gnd = Named(net, "gnd")()
        Named(VoltageSource(tran=spsin(0.0, 1.0, 1 / (2 * math.pi))), "V")(vin, gnd)
        Named(L(1.5), "L1")(vin, n1); Named(C(4 / 3), "C2")(n1, gnd)
        Named(L(0.5), "L3")(n1, vout); Named(R(1.0), "R4")(vout, gnd)
        Gnd()(gnd)
    flat = ch.lower(ckt)
    o = ch.grid_options((0.0, 100.0), n_out=1001, reltol=1e-10, abstol=1e-10, init=1)
    w, uf, st, rc = O.tran(flat, flat.params(), np.zeros(flat.n), o)
    assert rc == 0
    t = o.t_start + o.dt_out * np.arange(o.n_out)
    ref = (np.exp(-t) - np.sin(t) - np.cos(t)) / 2 + 2 * np.sin(np.sqrt(3) * t / 2) / (np.sqrt(3) * np.sqrt(np.exp(t)))
    g = GOLD["butterworth"]
    assert np.max(np.abs(w[:, flat.index("vout")] - ref)) < g["tol"]
    half = w[len(t) // 2:, flat.index("vout")]
    assert np.sqrt(np.mean(half ** 2)) == pytest.approx(g["rms_second_half"], abs=0.1)


# ---- test/sweep.jl, test/sensitivity.jl -------------------------------------------
@pytest.mark.parametrize("r1,r2", [(100.0, 100.0), (300.0, 1700.0), (2000.0, 100.0)])
def test_dc_sweep_current(r1, r2):
    """test/sweep.jl:326-340: I(V) = -1/(R1+R2)."""
    def ckt():
        vcc = Named(net, "vcc")(); out = Named(net, "out")(); gnd = Named(net, "gnd")()
        Named(V(1.0), "V")(vcc, gnd); Named(R(r1), "R1")(vcc, out); Named(R(r2), "R2")(out, gnd)
        Gnd()(gnd)
    flat, u = dcop(ckt)
    assert u[flat.index("V.I")] == pytest.approx(-1 / (r1 + r2), rel=1e-7, abs=1e-7)


def test_divider_sensitivity():
    """test/sensitivity.jl:40-41: dR1 = -dR2 = -out/2 for R1 = R2 = 1 (by central FD
    on the oracle's DC solution)."""
    def ckt():
        vcc = Named(net, "vcc")(); out = Named(net, "out")(); gnd = Named(net, "gnd")()
        Named(V(1.0), "V")(vcc, gnd); Named(R(1.0), "R1")(vcc, out); Named(R(1.0), "R2")(out, gnd)
        Gnd()(gnd)
    flat = ch.lower(ckt)
    p = flat.params()
    lay = flat.layout()
    io = flat.index("out")
    out0 = O.dcop(flat, p)[0][io]
    grads = []
    for k in range(2):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[lay["res_r"] + k] += h
        pm[lay["res_r"] + k] -= h
        grads.append((O.dcop(flat, pp)[0][io] - O.dcop(flat, pm)[0][io]) / (2 * h))
    assert grads[0] == pytest.approx(-grads[1], rel=1e-8)
    assert grads[0] == pytest.approx(-out0 / 2, rel=1e-8)


# ---- square-law MOSFET (models/sqlaw.va): closed form + brute force ------------------
BETA_N = 200e-6 * 1e-6 / 0.18e-6


def test_sqlaw_regions_closed_form():
    i = lambda vd, vg, vs: O.sqlaw(1, 0.4, 200e-6, 1e-6, 0.18e-6, 0.0, vd, vg, vs)[0]
    assert i(1.0, 0.3, 0.0) == 0.0                                   # cutoff
    assert i(0.1, 1.0, 0.0) == pytest.approx(BETA_N * (0.6 * 0.1 - 0.005), rel=1e-15)
    assert i(1.5, 1.0, 0.0) == pytest.approx(BETA_N / 2 * 0.36, rel=1e-15)
    # source/drain symmetry: swapping d and s flips the current
    for vd, vg, vs in ((0.3, 1.2, 0.1), (1.7, 0.9, 0.2)):
        assert i(vs, vg, vd) == pytest.approx(-i(vd, vg, vs), rel=1e-15)
    # PMOS is the NMOS with all voltages and the current negated (|Vth| equal)
    ip = lambda vd, vg, vs: O.sqlaw(-1, -0.4, 200e-6, 1e-6, 0.18e-6, 0.0, vd, vg, vs)[0]
    for vd, vg, vs in ((0.3, 1.2, 0.1), (1.7, 0.9, 0.2), (-0.2, 0.5, 0.4)):
        assert ip(-vd, -vg, -vs) == pytest.approx(-i(vd, vg, vs), rel=1e-15)


def test_sqlaw_derivatives_fd():
    rng = np.random.default_rng(0)
    for _ in range(200):
        typ = rng.choice([1, -1])
        v = rng.uniform(-0.3, 2.1, 3)
        args = (typ, 0.4 * typ, 200e-6, 1e-6, 0.18e-6, 1e-12)
        i0, dd, dg, ds = O.sqlaw(*args, *v)
        assert dd + dg + ds == pytest.approx(0.0, abs=1e-18)
        h = 1e-7
        for k, d in enumerate((dd, dg, ds)):
            vp, vm = v.copy(), v.copy()
            vp[k] += h
            vm[k] -= h
            fd = (O.sqlaw(*args, *vp)[0] - O.sqlaw(*args, *vm)[0]) / (2 * h)
            assert fd == pytest.approx(d, rel=1e-5, abs=1e-9)


def test_inverter_dc_closed_form():
    """Inverter at Vin = 0.6 V: NMOS saturated, PMOS triode; Vout solves
    beta_n/2 (Vin-0.4)^2 = beta_p ((1.8-Vin-0.4) Vsd - Vsd^2/2)."""
    bn, bp = 200e-6 / 0.18, 80e-6 * 2 / 0.18
    vin = 0.6
    idn = bn / 2 * (vin - 0.4) ** 2
    vov = 1.8 - vin - 0.4
    vsd = vov - math.sqrt(vov ** 2 - 2 * idn / bp)
    def ckt():
        vdd = Named(net, "vdd")(); a = Named(net, "in")(); y = Named(net, "out")()
        gnd = Named(net, "gnd")()
        Named(V(1.8), "VDD")(vdd, gnd); Named(V(vin), "VIN")(a, gnd)
        Named(SquareLawMOSFET(1, 0.4, 200e-6, 1e-6, 0.18e-6), "MN")(y, a, gnd, gnd)
        Named(SquareLawMOSFET(-1, -0.4, 80e-6, 2e-6, 0.18e-6), "MP")(y, a, vdd, vdd)
        Gnd()(gnd)
    # default gmin (1e-12 S across d-s) keeps the all-zero start non-singular;
    # its current is ~1e-12 A against ~1e-5 A of channel current
    flat = ch.lower(ckt)
    u, st, rc = O.dcop(flat, flat.params())
    assert rc == 0
    assert u[flat.index("out")] == pytest.approx(1.8 - vsd, rel=1e-6)


# ---- the integrator: BDF order and the dense LU ---------------------------------
def test_bdf2_converges_on_rc():
    """Order test on a smooth problem.  With per-step local-error control the
    global error of BDF-k scales like tol^(k/(k+1)): a 100x tighter tolerance
    must cut the error ~21.5x for BDF2 (backward Euler would give 10x)."""
    def ckt():
        a = Named(net, "a")(); b = Named(net, "b")(); gnd = Named(net, "gnd")()
        Named(VoltageSource(tran=spsin(0.0, 1.0, 1e3)), "V")(a, gnd)
        Named(R(1e3), "R")(a, b); Named(C(1e-7), "C")(b, gnd); Gnd()(gnd)
    flat = ch.lower(ckt)
    wc, tau = 2 * math.pi * 1e3, 1e-4
    errs = []
    for rt in (1e-6, 1e-8, 1e-10):
        o = ch.grid_options((0.0, 2e-3), n_out=201, reltol=rt, abstol=rt, init=1)
        w, uf, st, rc = O.tran(flat, flat.params(), np.zeros(flat.n), o)
        t = o.t_start + o.dt_out * np.arange(o.n_out)
        # exact: vb = (sin(wt) - wt cos(wt) + wt e^{-t/tau}) / (1 + (w tau)^2)
        x = wc * tau
        ref = (np.sin(wc * t) - x * np.cos(wc * t) + x * np.exp(-t / tau)) / (1 + x * x)
        errs.append(np.max(np.abs(w[:, flat.index("b")] - ref)))
    assert errs[0] > errs[1] > errs[2]
    assert 15.0 < errs[1] / errs[2] < 35.0
    assert errs[2] < 1e-6


def test_dense_lu_against_numpy():
    rng = np.random.default_rng(1)
    for n in (1, 2, 5, 17, 60):
        A = rng.standard_normal((n, n)) + n * np.eye(n) * rng.choice([0, 1])
        b = rng.standard_normal(n)
        assert np.allclose(O.lu_solve(A, b), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_inverter_transient_runs():
    """Config 0 on the oracle: completes, output swings rail to rail."""
    flat = ch.lower(inverter())
    o = ch.grid_options((0.0, 20e-9), dt_out=10e-12)
    w, uf, st, rc = O.tran(flat, flat.params(), None, o)
    assert rc == 0
    out = w[:, flat.index("out")]
    assert out.max() > 1.79 and out.min() < 0.01
