"""Network descriptions of the benchmark workloads (generator form: populations and
projections that the library expands on the device).

Populations are laid out Poisson -> E -> I in id order, as the reference's
ModelMaker::make_brunel does (maker.h:14-16).  Weights are absolute potentials
in volts (J = 0.1 mV), delays in seconds.
"""
from __future__ import annotations

import math

import numpy as np

from mgbrain_amd import native as N

DT = 1e-4
f32 = np.float32


def lifb_consts(v_thresh=0.020):
    """gen_life_const(v_init 0, v_rest 0, v_reset 10 mV, cm 0, tau_m 0, t_ref 2 ms, 1, 1,
    V_th, 0) at dt = 0.1 ms (constgen.h:9-107); LIFB reads DT, V_INIT, V_RESET,
    V_THRESH and TAU_REFRAC only."""
    c = np.zeros(30, dtype=f32)
    c[0], c[3], c[4], c[8] = DT, 0.010, v_thresh, 0.002
    c[10], c[11] = 1.0, 1.0
    e = np.exp(f32(-DT), dtype=f32)
    c[12], c[13], c[15], c[16] = e, e, e, e
    c[21], c[22], c[23], c[24], c[25] = 0.951229453, 0.967216074, 0.0799999982, 0.0599999987, 0.90483743
    return c


def brunel_fixed_k(scale=1.0, K=10000, rate=0.0015, seed=0, plastic_ee=False):
    """80k E + 20k I LIFB + 10k Poisson (x scale), fixed in-degree K from all
    sources (split by source size), J = 0.1 mV, g = 5, delays on the 1.0-2.0 ms grid.
    plastic_ee marks the E->E projection plastic (pair-based STDP)."""
    nP, nE, nI = int(10000 * scale), int(80000 * scale), int(20000 * scale)
    ntot = nP + nE + nI
    Ks = [int(round(K * n / ntot)) for n in (nP, nE, nI)]
    J, wi = float(f32(1e-4)), float(f32(-5e-4))
    dmin, dmax = float(f32(0.001)), float(f32(0.002))
    pops = [dict(num=nP, type=N.POISSON, rate=rate), dict(num=nE, type=N.LIFB), dict(num=nI, type=N.LIFB)]
    projs = []
    for tgt in (1, 2):
        for src, k, w in ((0, Ks[0], J), (1, Ks[1], J), (2, Ks[2], wi)):
            projs.append(dict(src=src, tar=tgt, kind=N.CONN_FIXED_INDEGREE, k=k, wmin=w, wmax=w, dmin=dmin,
                              dmax=dmax, index=len(projs), seed=seed,
                              plastic=1 if (plastic_ee and src == 1 and tgt == 1) else 0))
    desc = (f"Brunel fixed in-degree: {nE} E + {nI} I LIFB + {nP} Poisson @ {rate / DT:.0f} Hz, "
            f"K={K}, J=0.1 mV, g=5, delay U{{1.0..2.0}} ms, dt=0.1 ms")
    return pops, projs, desc, ntot


def brunel_bernoulli(scale=1.0, p=0.1, rate=0.002, seed=0):
    """8000 E + 2000 I LIFB + 1000 Poisson (x scale), Bernoulli p, J = 0.1 mV, g = 5,
    delay 1.4 ms."""
    nP, nE, nI = int(1000 * scale), int(8000 * scale), int(2000 * scale)
    J, wi, d = float(f32(1e-4)), float(f32(-5e-4)), float(f32(0.0014))
    pops = [dict(num=nP, type=N.POISSON, rate=rate), dict(num=nE, type=N.LIFB), dict(num=nI, type=N.LIFB)]
    projs = []
    for s, t, w in ((0, 1, J), (0, 2, J), (1, 1, J), (1, 2, J), (2, 1, wi), (2, 2, wi)):
        projs.append(dict(src=s, tar=t, kind=N.CONN_BERNOULLI, p=p, w=w, d=d, index=len(projs), seed=seed))
    desc = (f"Brunel Bernoulli p={p}: {nE} E + {nI} I LIFB + {nP} Poisson @ {rate / DT:.0f} Hz, "
            f"J=0.1 mV, g=5, delay 1.4 ms, dt=0.1 ms")
    return pops, projs, desc, nP + nE + nI


def stdp_consts(J=1e-4):
    """Pair-based STDP of the E->E synapses: A+ = 0.01, A- = 0.0105 in units of
    W_max = 2J (the reference's A are absolute weight increments, mgcore.cu:351-361),
    tau+ = tau- = 20 ms = 200 steps, W in [0, 2J].  Order: {A_LTP, A_LTD, TAU_LTP,
    TAU_LTD, W_MAX, W_MIN} (constgen.h:140-159)."""
    wmax = 2 * J
    return np.array([0.01 * wmax, -0.0105 * wmax, 200.0, 200.0, wmax, 0.0], dtype=f32)


# ---- batched parameter / seed sweep ---------------------------------------------

SWEEP_G = (3.0, 4.0, 5.0, 6.0)
SWEEP_NU_HZ = (10.0, 15.0, 20.0, 25.0)


def sweep_params(n_nets=64, n_lo=2000, n_hi=50000, size_seed=2024):
    """Network b of the sweep: N LIF neurons drawn log-uniform on [n_lo, n_hi]
    (80 % E / 20 % I) plus 0.1 N Poisson inputs, g = SWEEP_G[b % 4], Poisson rate
    SWEEP_NU_HZ[(b // 4) % 4], seed b."""
    rng = np.random.default_rng(size_seed)
    sizes = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), n_nets))
    out = []
    for b in range(n_nets):
        n = int(round(sizes[b]))
        out.append(dict(n=n, g=SWEEP_G[b % 4], nu=SWEEP_NU_HZ[(b // 4) % 4], seed=b))
    return out


def brunel_sweep_net(n, g, nu, seed, p=0.1, k_cap=5000, v_thresh=0.020):
    """One network of the sweep in generator form (the keyword arguments of
    native.Simulator.generated).  Connectivity: Bernoulli with p_eff = min(p,
    k_cap / sources), so the expected in-degree is at most k_cap."""
    nE = int(round(0.8 * n))
    nI = n - nE
    nP = int(round(0.1 * n))
    p_eff = min(p, k_cap / float(nP + n))
    J = float(f32(1e-4))
    wi = float(f32(-g * 1e-4))
    dmin, dmax = float(f32(0.001)), float(f32(0.002))
    pops = [dict(num=nP, type=N.POISSON, rate=nu * DT), dict(num=nE, type=N.LIFB), dict(num=nI, type=N.LIFB)]
    projs = []
    for s, t, w in ((0, 1, J), (0, 2, J), (1, 1, J), (1, 2, J), (2, 1, wi), (2, 2, wi)):
        projs.append(dict(src=s, tar=t, kind=N.CONN_BERNOULLI, p=p_eff, wmin=w, wmax=w, dmin=dmin, dmax=dmax,
                          index=len(projs), seed=seed))
    return dict(pops=pops, projs=projs, lifconst=lifb_consts(v_thresh), step=DT, dt=DT, seed=seed,
                net_seed=seed)


def sweep_nets(n_nets=64, n_lo=2000, n_hi=50000, size_seed=2024):
    params = sweep_params(n_nets, n_lo, n_hi, size_seed)
    return params, [brunel_sweep_net(q["n"], q["g"], q["nu"], q["seed"]) for q in params]


def net_size(net):
    return sum(p["num"] for p in net["pops"])
