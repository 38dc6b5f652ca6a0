"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

  inverter()          plain CMOS inverter: square-law NMOS (Vth 0.4 V, KP 200 uA/V^2,
                      W/L 1u/0.18u) + PMOS (Vth -0.4 V, KP 80 uA/V^2, W/L 2u/0.18u),
                      VDD 1.8 V, PULSE(0 1.8 0 100p 100p 0.9n 2n) input, 10 fF load.
  ring_oscillator(N)  N-stage ring of the same inverters, 5 fF per stage, VDD 1.8 V.
  ring_mc(...)        Monte Carlo batch: N uniform odd in [11, 101], per-transistor
                      Vth offset ~ N(0, 30 mV), start from a perturbed alternating
                      state (stage k at VDD*(k odd) + 1 mV * N(0,1)).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .netlist import (C, FlatCircuit, Gnd, Named, SquareLawMOSFET, V, VoltageSource, lower, net,
                      pulse)

VDD = 1.8
NMOS = dict(type=1, vth=0.4, kp=200e-6, w=1e-6, l=0.18e-6)
PMOS = dict(type=-1, vth=-0.4, kp=80e-6, w=2e-6, l=0.18e-6)


def inverter(cload=10e-15, period=2e-9, pw=0.9e-9, tr=100e-12, tf=100e-12):
    def circuit():
        vdd = Named(net, "vdd")()
        vin = Named(net, "in")()
        out = Named(net, "out")()
        gnd = Named(net, "gnd")()
        Named(V(VDD), "VDD")(vdd, gnd)
        Named(VoltageSource(tran=pulse(0.0, VDD, 0.0, tr, tf, pw, period)), "VIN")(vin, gnd)
        Named(SquareLawMOSFET(**NMOS), "MN")(out, vin, gnd, gnd)
        Named(SquareLawMOSFET(**PMOS), "MP")(out, vin, vdd, vdd)
        Named(C(cload), "CL")(out, gnd)
        Gnd()(gnd)
    return circuit


def ring_oscillator(nstage: int, cstage=5e-15, wn=None, wp=None):
    """Stage k drives node s{k} from node s{k-1}; devices are ordered
    NMOS_0, PMOS_0, NMOS_1, PMOS_1, ... so mos index 2k is stage k's NMOS."""
    if nstage < 3 or nstage % 2 == 0:
        raise ValueError("a ring oscillator needs an odd stage count >= 3")

    def circuit():
        vdd = Named(net, "vdd")()
        gnd = Named(net, "gnd")()
        s = [Named(net, f"s{k}")() for k in range(nstage)]
        Named(V(VDD), "VDD")(vdd, gnd)
        for k in range(nstage):
            a, y = s[k - 1], s[k]
            n_ = dict(NMOS)
            p_ = dict(PMOS)
            if wn is not None:
                n_["w"] = wn
            if wp is not None:
                p_["w"] = wp
            Named(SquareLawMOSFET(**n_), f"MN{k}")(y, a, gnd, gnd)
            Named(SquareLawMOSFET(**p_), f"MP{k}")(y, a, vdd, vdd)
            Named(C(cstage), f"C{k}")(y, gnd)
        Gnd()(gnd)
    return circuit


def ring_initial_state(flat: FlatCircuit, rng: np.random.Generator, noise=1e-3) -> np.ndarray:
    u0 = np.zeros(flat.n)
    u0[flat.index("vdd")] = VDD
    k = 0
    while f"s{k}" in flat.names:
        u0[flat.index(f"s{k}")] = VDD * (k % 2) + noise * rng.standard_normal()
        k += 1
    return u0


def ring_widths(n_inst=1024, nstage=51, seed=0, wmin=0.5e-6, wmax=4e-6):
    """Config: adjoint sensitivity batch.  n_inst copies of an nstage ring with
    every transistor width W ~ U[wmin, wmax]; seeded perturbed initial state.
    Returns (flat per instance, params per instance, u0 per instance)."""
    rng = np.random.default_rng(seed)
    f = lower(ring_oscillator(nstage))
    lay = f.layout()
    nm = len(f.arrays["mos_w"])
    flats, params, u0s = [], [], []
    for _ in range(n_inst):
        p = f.params().copy()
        p[lay["mos_w"]:lay["mos_w"] + nm] = rng.uniform(wmin, wmax, nm)
        flats.append(f)
        params.append(p)
        u0s.append(ring_initial_state(f, rng))
    return flats, params, u0s


def ring_mc(n_inst=4096, seed=0, nmin=11, nmax=101, sigma_vth=0.030
            ) -> Tuple[List[FlatCircuit], List[np.ndarray], List[np.ndarray], np.ndarray]:
    """Config: batched Monte Carlo of uneven ring oscillators.
    Returns (flat per instance, params per instance, u0 per instance, stages)."""
    rng = np.random.default_rng(seed)
    odd = np.arange(nmin, nmax + 1, 2)
    stages = rng.choice(odd, size=n_inst)
    cache = {}
    flats, params, u0s = [], [], []
    for i, N in enumerate(stages):
        N = int(N)
        if N not in cache:
            cache[N] = lower(ring_oscillator(N))
        f = cache[N]
        p = f.params().copy()
        lay = f.layout()
        nm = len(f.arrays["mos_vth"])
        p[lay["mos_vth"]:lay["mos_vth"] + nm] += sigma_vth * rng.standard_normal(nm)
        flats.append(f)
        params.append(p)
        u0s.append(ring_initial_state(f, rng))
    return flats, params, u0s, stages
