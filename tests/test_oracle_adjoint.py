"""Pins the oracle's discrete adjoint and its sparse LU.

The reference's own sensitivity check is forward sensitivity on a divider
(test/sensitivity.jl:40-41, pinned in test_oracle_golden.py); for the
adjoint there is no reference vector, so it is pinned here against central
finite differences (h = 1e-6 relative) of the same discrete scheme: the
frozen accepted-step sequence replayed with Newton converged tightly.
"""
import numpy as np
import pytest

import cedarhip as ch
from cedarhip import C, Gnd, Named, SquareLawMOSFET, V, VoltageSource, net, pulse
from cedarhip.workloads import NMOS, PMOS, ring_initial_state, ring_oscillator
from oracle import oracle as O

OPTS = dict(t_start=0.0, reltol=1e-6, abstol=1e-9, init=1, newton_tol=1e-9, newton_maxit=30)


def chain():
    vdd = Named(net, "vdd")(); gnd = Named(net, "gnd")(); a = Named(net, "in")()
    s = [Named(net, f"s{k}")() for k in range(4)]
    Named(V(1.8), "VDD")(vdd, gnd)
    Named(VoltageSource(tran=pulse(0, 1.8, 20e-12, 50e-12, 50e-12, 300e-12, 1e-9)), "VIN")(a, gnd)
    prev = a
    for k in range(4):
        Named(SquareLawMOSFET(**NMOS), f"MN{k}")(s[k], prev, gnd, gnd)
        Named(SquareLawMOSFET(**PMOS), f"MP{k}")(s[k], prev, vdd, vdd)
        Named(C(5e-15), f"C{k}")(s[k], gnd)
        prev = s[k]
    Gnd()(gnd)


def _widths(flat, seed):
    rng = np.random.default_rng(seed)
    p = flat.params()
    lay = flat.layout()
    nm = len(flat.arrays["mos_w"])
    p[lay["mos_w"]:lay["mos_w"] + nm] = rng.uniform(0.5e-6, 4e-6, nm)
    return p, lay, nm


def _fd(flat, p, lay, k, u0, tl, npl, t_stop, rel=1e-6):
    hW = rel * p[lay["mos_w"] + k]
    pp, pm = p.copy(), p.copy()
    pp[lay["mos_w"] + k] += hW
    pm[lay["mos_w"] + k] -= hW
    Ep, r1 = O.replay_energy(flat, pp, u0, tl, npl, 0, t_stop=t_stop, **OPTS)
    Em, r2 = O.replay_energy(flat, pm, u0, tl, npl, 0, t_stop=t_stop, **OPTS)
    assert r1 == 0 and r2 == 0
    return (Ep - Em) / (2 * hW)


def test_adjoint_matches_fd_inverter_chain():
    flat = ch.lower(chain)
    p, lay, nm = _widths(flat, 0)
    u0 = np.zeros(flat.n)
    u0[flat.index("vdd")] = 1.8
    for k in range(4):
        u0[flat.index(f"s{k}")] = 1.8 if k % 2 == 0 else 0.0
    E, g, st, rc, tl, npl = O.adjoint(flat, p, u0, 0, t_stop=1e-9, **OPTS)
    assert rc == 0 and E > 0
    for k in range(nm):
        fd = _fd(flat, p, lay, k, u0, tl, npl, 1e-9)
        assert g[k] == pytest.approx(fd, rel=1e-5), k


def test_adjoint_matches_fd_ring():
    flat = ch.lower(ring_oscillator(5))
    p, lay, nm = _widths(flat, 3)
    u0 = ring_initial_state(flat, np.random.default_rng(3))
    E, g, st, rc, tl, npl = O.adjoint(flat, p, u0, 0, t_stop=0.3e-9, **OPTS)
    assert rc == 0
    Er, rr = O.replay_energy(flat, p, u0, tl, npl, 0, t_stop=0.3e-9, **OPTS)
    assert rr == 0 and Er == E            # the replay is the same computation
    for k in (0, 3, 6, 9):
        fd = _fd(flat, p, lay, k, u0, tl, npl, 0.3e-9)
        assert g[k] == pytest.approx(fd, rel=1e-5), k


def test_sparse_matches_dense_on_short_horizon():
    """The sparse LU follows the same pivots as the dense one on a ring, so
    short runs agree to roundoff (long runs may take different step sequences
    once a pivot choice differs)."""
    flat = ch.lower(ring_oscillator(21))
    u0 = ring_initial_state(flat, np.random.default_rng(9))
    o = ch.grid_options((0.0, 0.3e-9), dt_out=5e-12, init=1)
    try:
        O.set_sparse(False)
        wd, _, sd, rcd = O.tran(flat, flat.params(), u0, o)
        O.set_sparse(True)
        ws, _, ss, rcs = O.tran(flat, flat.params(), u0, o)
    finally:
        O.set_sparse(False)
    assert rcd == 0 and rcs == 0 and sd.steps == ss.steps
    assert np.max(np.abs(wd - ws)) < 1e-12


def test_sparse_lu_random_systems():
    """Sparse LU with threshold pivoting against numpy on random sparse
    matrices, including ones that need off-diagonal pivots."""
    rng = np.random.default_rng(2)
    for n in (3, 10, 40):
        def net_fn(n=n):
            nodes = [Named(net, f"n{k}")() for k in range(n)]
            gnd = Named(net, "gnd")()
            Named(V(1.0), "V")(nodes[0], gnd)
            for k in range(1, n):
                Named(ch.R(float(rng.uniform(1, 10))), f"R{k}")(nodes[k - 1], nodes[k])
                Named(ch.R(float(rng.uniform(10, 100))), f"Rg{k}")(nodes[k], gnd)
            Gnd()(gnd)
        flat = ch.lower(net_fn)
        p = flat.params()
        O.set_sparse(True)
        try:
            us, _, rs = O.dcop(flat, p)
        finally:
            O.set_sparse(False)
        ud, _, rd = O.dcop(flat, p)
        assert rs == 0 and rd == 0
        assert np.allclose(us, ud, rtol=1e-12, atol=1e-15)
