"""BSIM-CMG device evaluation on the GPU (cs_bsimcmg_eval) against the oracle.

Tolerance (device stamps within 1e-12 relative): for each instance and each
output block (I, Q, G, C) the error is measured against the block's largest
entry.  The condensed currents and conductances also carry the rounding of
the internal-node solve through the 1e-3 ohm series resistors (1e3 S): a
one-ulp difference in V_di moves I_d by ulp(V)*1e3 and the Schur complement
G_tt + G_ti S cancels 1e3-sized terms, so those two blocks get the
additional floor 16*eps*g_int (times 1 + |V| for currents), where g_int is
the largest internal conductance.  Measured on the C1 distribution the
blocks agree far inside that bound.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _say(msg):
    print(f"[gpu-bsim] {msg}", flush=True)


def c1_sample(n, seed):
    rng = np.random.default_rng(seed)
    L = rng.uniform(20e-9, 100e-9, n)
    nf = rng.integers(1, 9, n).astype(np.float64)
    v = np.zeros((n, 4))
    v[:, 0] = rng.uniform(0.0, 1.0, n)
    v[:, 1] = rng.uniform(-0.2, 1.0, n)
    return L, nf, v


def _need(engine):
    import cedarhip as ch
    if not ch.bsimcmg_available(engine):
        pytest.skip("library built without the generated BSIM-CMG model")
    if O.va_lib() is None:
        pytest.skip("oracle BSIM-CMG library not built")
    return ch


def _block_err(a, b, floor=0.0):
    """max over entries of |a-b| / (max|b| of the block + floor), per instance."""
    ax = tuple(range(1, a.ndim))
    den = np.abs(b).max(axis=ax) + np.asarray(floor)
    den = np.where(den > 0, den, 1.0)
    return np.abs(a - b).max(axis=ax) / den


def _compare(r, L, nf, v, tag):
    I, Q, G, C, it, (v6, I6, Q6, G6, C6) = O.bsim_terminal(L, nf, v, full=True)
    assert np.all(r.iters > 0), f"{tag}: device internal-node Newton failed"
    assert np.all(it > 0)
    assert np.all(np.abs(r.iters - it) <= 1)
    gint = np.abs(G6[:, 4:6, 4:6]).max(axis=(1, 2))
    vmax = np.abs(v).max(axis=1)
    fI = 16 * EPS * gint * (1 + vmax)
    fG = 16 * EPS * gint
    # |err| <= 1e-12 * max|block| + floor, written as a ratio against 1e-12
    eI = _block_err(r.I, I, fI / 1e-12)
    eG = _block_err(r.G, G, fG / 1e-12)
    eQ = _block_err(r.Q, Q)
    eC = _block_err(r.C, C)
    same = np.mean(np.all(r.I == I, axis=1) & np.all(r.G == G, axis=(1, 2)))
    _say(f"{tag}: max rel err I {eI.max():.2e} Q {eQ.max():.2e} G {eG.max():.2e} "
         f"C {eC.max():.2e}; bitwise-equal I,G rows {same:.3f}")
    assert eI.max() <= 1e-12
    assert eG.max() <= 1e-12
    assert eQ.max() <= 1e-12
    assert eC.max() <= 1e-12


def test_bsimcmg_matches_oracle_c1_distribution(engine):
    ch = _need(engine)
    L, nf, v = c1_sample(2048, 11)
    _say("eval 2048 instances, seed 11")
    r = ch.bsimcmg_eval(L, nf, v, engine=engine)
    _compare(r, L, nf, v, "c1-2048")


def test_bsimcmg_edges_and_reverse_bias(engine):
    ch = _need(engine)
    # Vds = 0, reverse Vds, strong inversion, accumulation, extreme L and NFIN
    L = np.array([20e-9, 100e-9, 20e-9, 100e-9, 50e-9, 50e-9, 20e-9, 100e-9])
    nf = np.array([1, 8, 8, 1, 3, 3, 1, 8], np.float64)
    v = np.array([[0.0, 1.0, 0.0, 0.0], [0.0, -0.2, 0.0, 0.0], [1.0, 1.0, 0.0, 0.0],
                  [1.0, -0.2, 0.0, 0.0], [0.0, 0.8, 0.7, 0.0], [0.2, 0.6, 0.9, 0.0],
                  [0.5, 0.5, 0.5, 0.0], [1.0, 0.3, 0.0, -0.2]])
    _say("eval 8 edge instances")
    r = ch.bsimcmg_eval(L, nf, v, engine=engine)
    _compare(r, L, nf, v, "edges")


def test_bsimcmg_full_size_invariants(engine):
    """1,048,576 instances: every internal solve converges and the condensed
    stamps conserve current and charge; a seeded subset matches the oracle."""
    ch = _need(engine)
    n = 1 << 20
    L, nf, v = c1_sample(n, 2024)
    _say(f"eval {n} instances")
    r = ch.bsimcmg_eval(L, nf, v, engine=engine)
    _say(f"kernel {r.kernel_ms:.2f} ms")
    assert np.all(r.iters > 0)
    for a in (r.I, r.Q, r.G, r.C):
        assert np.all(np.isfinite(a))
    gs = np.abs(r.G).max(axis=(1, 2))
    cs = np.abs(r.C).max(axis=(1, 2))
    qs = np.abs(r.Q).max(axis=1)
    floor = 16 * EPS * 1e3
    assert np.all(np.abs(r.I.sum(axis=1)) <= 1e-12 * np.abs(r.I).max(axis=1) + 2 * floor)
    assert np.all(np.abs(r.Q.sum(axis=1)) <= 1e-12 * qs + 1e-35)
    assert np.all(np.abs(r.G.sum(axis=1)).max(axis=1) <= 1e-12 * gs + floor)
    assert np.all(np.abs(r.C.sum(axis=1)).max(axis=1) <= 1e-9 * cs + 1e-35)
    idx = np.random.default_rng(5).choice(n, 512, replace=False)
    idx.sort()

    class Sub:
        pass
    s = Sub()
    s.I, s.Q, s.G, s.C, s.iters = r.I[idx], r.Q[idx], r.G[idx], r.C[idx], r.iters[idx]
    _compare(s, L[idx], nf[idx], v[idx], "c1-full-subset")
