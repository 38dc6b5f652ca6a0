"""Supernodal multifrontal path (large circuits) on the GPU.

Factor and solve: normwise backward error below 1e-14 and agreement with
scipy's SuperLU on the same Jacobian to 1e-9 (the DC Jacobian's
conditioning);
Jacobian / residual against the level-scheduled path of the same circuit
(exact: the same contributor lists in the same order); DC operating point
and transient against the oracle (waveforms within reltol 1e-6 / abstol
1e-9 V, identical step counts); reruns bit-identical; the full 450 x 450
power grid through DC op, a short transient and size-independent checks.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import cedarhip as ch
from cedarhip import lib as L
from cedarhip.workloads import rc_mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-mf] {msg}", flush=True)


@pytest.mark.parametrize("nx,ninv", [(12, 10), (60, 200), (110, 400)])
def test_factor_solve_matches_sparse_lu(engine, nx, ninv):
    f = rc_mesh(nx, nx, ninv, seed=3)
    s = engine.load(f, solver=L.SOLVER_SUPERNODAL)
    bi = s.big_info
    _say(f"mesh {nx}x{nx} + {ninv} inverters: n {s.n}, fronts {bi.n_super}, levels {bi.n_levels}, "
         f"max front {bi.max_front}, max pivots {bi.max_pivots}")
    rng = np.random.default_rng(nx)
    p = f.params()
    u = rng.uniform(0, 1, s.n)
    for gamma in (0.0, 1e13):
        A = s.jac_matrix(None, u, p, gamma, 0.3e-9)
        s.lu_refactor(A.data)
        b = rng.standard_normal(s.n)
        x = s.lu_solve(b)
        xr = spl.spsolve(A.tocsc(), b)
        err = np.max(np.abs(x - xr)) / np.max(np.abs(xr))
        # normwise backward error (solver independent) and agreement with
        # SuperLU up to the conditioning of the DC (gamma = 0) Jacobian
        anorm = abs(A).sum(axis=1).max()
        berr = np.max(np.abs(A @ x - b)) / (anorm * np.max(np.abs(x)) + np.max(np.abs(b)))
        _say(f"  gamma {gamma:g}: backward error {berr:.2e}, vs SuperLU {err:.2e}")
        assert berr < 1e-14
        assert err < 1e-9


@pytest.mark.parametrize("gemm", ["mfma", "valu"])
def test_panel_step_forms_agree_bitwise(engine, monkeypatch, gemm):
    """A panel step of the large fronts as one launch (each workgroup factors
    the diagonal block and solves its own panel rows and columns) and as
    three launches (diagonal block, solves, trailing update) perform the same
    operations in the same order: the factors and solutions are the same
    bits.  The panel form is forced on every level."""
    f = rc_mesh(110, 110, 400, seed=5)
    monkeypatch.setenv("CEDAR_MF_FRONT", "panel")
    monkeypatch.setenv("CEDAR_MF_GEMM", gemm)
    rng = np.random.default_rng(11)
    p = f.params()
    xs = {}
    for form in ("one", "split"):
        if form == "split":
            monkeypatch.setenv("CEDAR_MF_PANEL", "split")
        else:
            monkeypatch.delenv("CEDAR_MF_PANEL", raising=False)
        s = engine.load(f, solver=L.SOLVER_SUPERNODAL)
        if form == "one":
            u = rng.uniform(0, 1, s.n)
            b = rng.standard_normal(s.n)
        A = s.jac_matrix(None, u, p, 1e13, 0.3e-9)
        s.lu_refactor(A.data)
        xs[form] = s.lu_solve(b)
        _say(f"panel form {form} ({gemm}): max front {s.big_info.max_front}, "
             f"max |x| {np.max(np.abs(xs[form])):.3e}")
    assert np.array_equal(xs["one"], xs["split"])
    berr = np.max(np.abs(A @ xs["one"] - b)) / (abs(A).sum(axis=1).max() * np.max(np.abs(xs["one"])) +
                                                np.max(np.abs(b)))
    assert berr < 1e-14


def test_jacobian_residual_match_level_path(engine):
    f = rc_mesh(20, 20, 40, seed=5)
    s = engine.load(f, solver=L.SOLVER_SUPERNODAL)
    s2 = engine.load(f, solver=L.SOLVER_LEVEL)
    assert s.supernodal and not s2.supernodal
    rng = np.random.default_rng(1)
    p = f.params()
    u = rng.uniform(0, 1, s.n)
    du = rng.standard_normal(s.n) * 1e8
    A = s.jac_matrix(None, u, p, 3e11, 0.2e-9)
    A2 = s2.jac_matrix(None, u, p, 3e11, 0.2e-9)
    assert abs(A - A2).max() == 0.0
    r = s.f(None, du, u, p, 0.2e-9)
    r2 = s2.f(None, du, u, p, 0.2e-9)
    assert np.max(np.abs(r - r2)) <= 1e-12 * np.max(np.abs(r2))


def test_dcop_and_transient_match_oracle(engine):
    f = rc_mesh(16, 16, 20, seed=7)
    s = engine.load(f, solver=L.SOLVER_SUPERNODAL)
    p = f.params()
    dc = s.dcop(p)
    ud, _, rc = O.dcop(f, p)
    assert rc == 0
    assert np.max(np.abs(np.asarray(dc.u) - ud)) <= 1e-9
    o = ch.TranOptions(t_start=0.0, t_stop=0.6e-9, reltol=1e-6, abstol=1e-9, n_out=61, dt_out=1e-11)
    sol = s.tran(o, p)
    w, uf, st, rc = O.tran(f, p, None, o)
    _say(f"16x16 mesh transient: {sol.stats.steps} steps (oracle {st.steps}), "
         f"{sol.stats.newton_iters} Newton iterations")
    assert sol.retcode == "Success" and rc == 0
    assert sol.stats.steps == st.steps and sol.stats.rejected == st.rejected
    wb = np.asarray(sol.wave)
    wr = np.asarray(w).reshape(wb.shape)
    assert np.all(np.abs(wb - wr) <= 1e-6 * np.abs(wr) + 1e-9)
    # bit-identical rerun
    sol2 = s.tran(o, p)
    assert np.array_equal(np.asarray(sol2.wave), wb)


def test_power_grid_full_size(engine):
    """450 x 450 RC mesh with 10^4 inverters (222,502 unknowns): DC op (its
    KCL residual), a factor / solve at the operating point (backward error),
    a 20 ps transient; factor and solve times are reported."""
    f = rc_mesh(450, 450, 10_000, seed=0)
    s = engine.load(f)
    assert s.supernodal
    bi = s.big_info
    _say(f"power grid: n {s.n}, fronts {bi.n_super}, levels {bi.n_levels}, max front {bi.max_front}, "
         f"factor {bi.flops:.3g} flop")
    p = f.params()
    dc = s.dcop(p)
    assert dc.retcode == "Success"
    u = np.asarray(dc.u)
    r = s.f(None, np.zeros(s.n), u, p, 0.0, mode=L.MODE_DCOP)
    _say(f"  DC residual max {np.max(np.abs(r)):.2e}")
    assert np.max(np.abs(r)) <= 1e-9
    # solve accuracy at the operating point
    A = s.jac_matrix(None, u, p, 1e12, 0.0)
    s.lu_refactor(A.data)
    b = np.random.default_rng(0).standard_normal(s.n)
    x = s.lu_solve(b)
    fm, sm = s.lu_times()
    anorm = abs(A).sum(axis=1).max()
    berr = np.max(np.abs(A @ x - b)) / (anorm * np.max(np.abs(x)) + np.max(np.abs(b)))
    _say(f"  factor {fm:.2f} ms, solve {sm:.2f} ms, backward error {berr:.2e}")
    assert berr < 1e-14
    o = ch.TranOptions(t_start=0.0, t_stop=20e-12, reltol=1e-6, abstol=1e-9)
    sol = s.tran(o, p)
    _say(f"  20 ps transient: {sol.retcode}, {sol.stats.steps} steps, {sol.stats.newton_iters} iterations")
    assert sol.retcode == "Success"


def test_full_span_matches_oracle(engine):
    """The power-grid configuration's whole 0-5 ns span (five pulse periods,
    breakpoints at every corner) on the supernodal path against the oracle:
    identical step counts, waveforms within reltol 1e-6 / abstol 1e-9 V."""
    f = rc_mesh(12, 12, 12, seed=2)
    s = engine.load(f, solver=L.SOLVER_SUPERNODAL)
    p = f.params()
    o = ch.TranOptions(t_start=0.0, t_stop=5e-9, reltol=1e-6, abstol=1e-9, n_out=51, dt_out=1e-10)
    sol = s.tran(o, p)
    w, uf, st, rc = O.tran(f, p, None, o)
    _say(f"12x12 mesh 0-5 ns: {sol.stats.steps} steps (oracle {st.steps}), {sol.stats.newton_iters} iterations")
    assert sol.retcode == "Success" and rc == 0
    assert sol.stats.steps == st.steps and sol.stats.rejected == st.rejected
    wb = np.asarray(sol.wave)
    wr = np.asarray(w).reshape(wb.shape)
    assert np.all(np.abs(wb - wr) <= 1e-6 * np.abs(wr) + 1e-9)
