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


def gen_life_const(v_init, v_rest, v_reset, cm, tau_m, tau_refrac, tau_syn_E, tau_syn_I, v_thresh, i_offset,
                   dt=1e-4):
    """ConstGen::gen_life_const (constgen.h:9-107) in float32 (expf via numpy float32)."""
    dt = f32(dt)
    v_init, v_rest, v_reset, cm, tau_m, tau_refrac = map(f32, (v_init, v_rest, v_reset, cm, tau_m, tau_refrac))
    tau_syn_E, tau_syn_I, v_thresh, i_offset = map(f32, (tau_syn_E, tau_syn_I, v_thresh, i_offset))
    rm = f32(1.0)
    if abs(float(cm)) > 0.00001:
        rm = f32(tau_m / cm)
    Cm = np.exp(f32(-dt / tau_m), dtype=f32) if tau_m > 0 else f32(0.0)
    CE = np.exp(f32(-dt / tau_syn_E), dtype=f32) if tau_syn_E > 0 else f32(0.0)
    CI = np.exp(f32(-dt / tau_syn_I), dtype=f32) if tau_syn_I > 0 else f32(0.0)
    io = f32(0.0)  # the reference reads an uninitialised _i_offset; all callers pass 0
    v_tmp = f32(io * rm + v_rest)
    v_tmp = f32(v_tmp * f32(f32(1) - Cm))
    C_E = f32(f32(rm * tau_syn_E) / f32(tau_syn_E - tau_m))
    C_I = f32(f32(rm * tau_syn_I) / f32(tau_syn_I - tau_m))
    C_E = f32(C_E * f32(CE - Cm))
    C_I = f32(C_I * f32(CI - Cm))
    c = np.zeros(30, dtype=f32)
    c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8] = (dt, v_init, v_rest, v_reset, v_thresh, cm, tau_m,
                                                            i_offset, tau_refrac)
    c[10], c[11], c[12], c[13], c[14], c[15], c[16], c[17] = tau_syn_E, tau_syn_I, C_E, C_I, v_tmp, CE, CI, Cm
    c[21], c[22], c[23], c[24], c[25] = (f32(0.951229453), f32(0.967216074), f32(0.0799999982),
                                         f32(0.0599999987), f32(0.90483743))
    return c


def life_const1(dt=1e-4):
    """ConstGen::gen_life_const1 (constgen.h:135-138): the Vogels-Abbott constants."""
    return gen_life_const(-0.06, -0.06, -0.06, 0, 0, 0.005, 1, 1, -0.05, 0, dt=dt)


def lifb_consts(v_thresh=0.020):
    """gen_life_const(v_init 0, v_rest 0, v_reset 10 mV, cm 0, tau_m 0, t_ref 2 ms, 1, 1,
    V_th, 0) at dt = 0.1 ms; LIFB reads DT, V_INIT, V_RESET, V_THRESH and TAU_REFRAC only."""
    return gen_life_const(0.0, 0.0, 0.010, 0.0, 0.0, 0.002, 1.0, 1.0, v_thresh, 0.0, dt=DT)


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


def vogels_abbott(scale=1.0, p=0.02, seed=0, g_exc=0.04, g_inh=0.51):
    """64k E + 16k I conductance-based LIFE neurons (x scale), Bernoulli p, delay 0.8 ms,
    V0 uniform on [-60, -50] mV.  g_exc / g_inh are the conductance increments in the LIFE
    update's units, multiples of the leak conductance (the update advances V by
    dt/tau_m * (E_L - V + g_e (E_e - V) + g_i (E_i - V)), mgcore.cu:197-233): 0.4 nS and
    5.1 nS over g_L = 10 nS (C_m 200 pF, tau_m 20 ms) give 0.04 / 0.51.  The reference's
    make_vogel uses 0.4 / 5.1 scaled by 16e6 / N^2 (maker.h:34-35); unscaled at 80k neurons
    that drive saturates every neuron near its refractory limit (~190 Hz)."""
    nE, nI = int(64000 * scale), int(16000 * scale)
    vlo, vhi = float(f32(-0.060)), float(f32(-0.050))
    d = float(f32(0.0008))
    pops = [dict(num=nE, type=N.LIFE, v_min=vlo, v_max=vhi), dict(num=nI, type=N.LIFE, v_min=vlo, v_max=vhi)]
    we, wi = float(f32(g_exc)), float(f32(-g_inh))
    projs = [dict(src=s, tar=t, kind=N.CONN_BERNOULLI, p=p, w=w, d=d, index=k, seed=seed)
             for k, (s, t, w) in enumerate(((0, 0, we), (0, 1, we), (1, 0, wi), (1, 1, wi)))]
    desc = (f"Vogels-Abbott COBA: {nE} E + {nI} I LIFE, Bernoulli p={p}, g_exc {g_exc}, g_inh {g_inh} "
            f"(x g_leak), delay 0.8 ms, "
            f"V0 ~ U[-60, -50] mV, dt=0.1 ms")
    return pops, projs, desc, nE + nI


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
    (80 % E / 20 % I), g = SWEEP_G[b % 4], Poisson rate SWEEP_NU_HZ[(b // 4) % 4],
    seed b."""
    rng = np.random.default_rng(size_seed)
    sizes = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), n_nets))
    out = []
    for b in range(n_nets):
        n = int(round(sizes[b]))
        out.append(dict(n=n, g=SWEEP_G[b % 4], nu=SWEEP_NU_HZ[(b // 4) % 4], seed=b))
    return out


SWEEP_K_EXT = 1000          # external Poisson synapses per LIF neuron


def brunel_sweep_net(n, g, nu, seed, p=0.1, k_cap=5000, v_thresh=0.020, k_ext=SWEEP_K_EXT):
    """One network of the sweep in generator form (the keyword arguments of
    native.Simulator.generated).  Recurrent connectivity: Bernoulli with p_eff =
    min(p, k_cap / N) between the LIF neurons, so the expected recurrent in-degree is
    at most k_cap.  External drive: every LIF neuron receives k_ext synapses from a
    pool of max(2000, N / 10) Poisson neurons firing at nu (fixed in-degree), so that
    networks of every size are driven (with the pool's in-degree set by p instead,
    networks below ~20k neurons never reach threshold)."""
    nE = int(round(0.8 * n))
    nI = n - nE
    nP = max(2000, int(round(0.1 * n)))
    p_eff = min(p, k_cap / float(n))
    J = float(f32(1e-4))
    wi = float(f32(-g * 1e-4))
    dmin, dmax = float(f32(0.001)), float(f32(0.002))
    pops = [dict(num=nP, type=N.POISSON, rate=nu * DT), dict(num=nE, type=N.LIFB), dict(num=nI, type=N.LIFB)]
    projs = []
    for s, t in ((0, 1), (0, 2)):
        projs.append(dict(src=s, tar=t, kind=N.CONN_FIXED_INDEGREE, k=k_ext, wmin=J, wmax=J, dmin=dmin, dmax=dmax,
                          index=len(projs), seed=seed))
    for s, t, w in ((1, 1, J), (1, 2, J), (2, 1, wi), (2, 2, wi)):
        projs.append(dict(src=s, tar=t, kind=N.CONN_BERNOULLI, p=p_eff, wmin=w, wmax=w, dmin=dmin, dmax=dmax,
                          index=len(projs), seed=seed))
    return dict(pops=pops, projs=projs, lifconst=lifb_consts(v_thresh), step=DT, dt=DT, seed=seed,
                net_seed=seed)


def sweep_nets(n_nets=64, n_lo=2000, n_hi=50000, size_seed=2024):
    params = sweep_params(n_nets, n_lo, n_hi, size_seed)
    return params, [brunel_sweep_net(q["n"], q["g"], q["nu"], q["seed"]) for q in params]


def net_size(net):
    return sum(p["num"] for p in net["pops"])
