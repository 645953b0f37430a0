"""Model descriptions used by the tests (pure Python, no GPU).

brunel_ref / vogel_ref restate ModelMaker::make_brunel / make_vogel
(model/maker.h:10-42) for the reference connectors; config1..config5 are the
workloads of the benchmark, described with the new connectors (Bernoulli,
fixed in-degree) that the device generator and the oracle share.
"""
from __future__ import annotations

import math

import numpy as np

RAW, LIF0, POISSON, LIF, LIFE, LIFB = range(6)
BERN, FIXK = 1, 2
f32 = np.float32


from mgbrain_amd.workloads import gen_life_const  # noqa: E402  (shared with bench.py)


def life_const1(dt=1e-4):
    return gen_life_const(-0.06, -0.06, -0.06, 0, 0, 0.005, 1, 1, -0.05, 0, dt=dt)


def life_const2(dt=1e-4):
    return gen_life_const(0.0, 0.0, 0.0, 0.0, 0.0, 0.002, 1.0, 1.0, 0.02, 0.0, dt=dt)


def default_consts(dt=1e-4):
    """copy_consts_gpu(max_delay, dt) defaults (mgcore.cu:685-733): V_THRESH = 15."""
    c = np.zeros(30, dtype=f32)
    c[0], c[4], c[5], c[6], c[8], c[10], c[11] = dt, 15.0, 0.25, 10.0, 0.0004, 1.0, 1.0
    c[21], c[22], c[23], c[24], c[25] = 0.951229453, 0.967216074, 0.0799999982, 0.0599999987, 0.90483743
    return c


STDP_DEFAULT = np.array([0.1, -0.01, 17, 34, 40, 0], dtype=f32)


class Net:
    """Populations + projections + constants of one network."""

    def __init__(self, pops, projs, lifconst, name=""):
        self.pops = pops      # list of (num, type, source)
        self.projs = projs    # list of dicts (oracle Proj fields + kind)
        self.lifconst = lifconst
        self.name = name
        self.rates = [0.0] * len(pops)
        self.vrange = [None] * len(pops)

    @property
    def n(self):
        return sum(p[0] for p in self.pops)

    def first(self):
        out, acc = [], 0
        for p in self.pops:
            out.append(acc)
            acc += p[0]
        return out

    def neuron_arrays(self):
        types = np.concatenate([np.full(p[0], p[1], dtype=np.int32) for p in self.pops])
        rate = np.concatenate([np.full(p[0], r, dtype=f32) for p, r in zip(self.pops, self.rates)])
        return types, rate

    def gen_pops(self, stream=0):
        out = []
        v1 = float(self.lifconst[1])
        for p, r, vr in zip(self.pops, self.rates, self.vrange):
            lo, hi = (v1, v1) if vr is None else vr
            out.append(dict(num=p[0], type=p[1], rate=r, v_min=lo, v_max=hi, cset=0, stream=stream))
        return out


def brunel_ref(nsyn, rate=0.002):
    """ModelMaker::make_brunel (maker.h:10-27): quota connector, p = 0.1."""
    cp = f32(0.1)
    pop = int(math.sqrt(float(f32(nsyn)) / (float(cp) * 0.5)))
    pops = [(pop * 5 // 10, POISSON, 1), (pop * 4 // 10, LIFB, 0), (pop * 1 // 10, LIFB, 0)]
    w1 = float(f32(0.0001 * 20000 / pop))
    w2 = float(f32(-0.0005 * 20000 / pop))
    d1 = float(f32(0.0016))
    pr = [dict(src=s, tar=t, w=w, d=d1, ctype=float(cp), kind=0)
          for s, t, w in ((0, 1, w1), (0, 2, w1), (1, 1, w1), (1, 2, w1), (2, 2, w2), (2, 1, w2))]
    net = Net(pops, pr, life_const2(), name=f"brunel_ref_{nsyn}")
    net.rates = [rate, 0.0, 0.0]
    return net


def vogel_ref(nsyn):
    """ModelMaker::make_vogel (maker.h:28-42)."""
    cp = f32(0.02)
    pop = int(math.sqrt(float(f32(nsyn)) / float(cp)))
    pops = [(pop * 8 // 10, LIFE, 0), (pop * 2 // 10, LIFE, 0)]
    w1 = float(f32(0.4 * 16000000 / pop / pop))
    w2 = float(f32(-5.1 * 16000000 / pop / pop))
    d1 = float(f32(0.0008))
    pr = [dict(src=s, tar=t, w=w, d=d1, ctype=float(cp), kind=0)
          for s, t, w in ((0, 0, w1), (0, 1, w1), (1, 0, w2), (1, 1, w2))]
    return Net(pops, pr, life_const1(), name=f"vogel_ref_{nsyn}")


def config1(scale=1.0, p=0.1, poisson_rate=0.002):
    """Brunel 8000 E + 2000 I LIFB + 1000 Poisson, Bernoulli p, J = 0.1 mV, g = 5, 1.4 ms."""
    nP, nE, nI = int(1000 * scale), int(8000 * scale), int(2000 * scale)
    J, g, d = float(f32(1e-4)), 5.0, float(f32(0.0014))
    wi = float(f32(-g * 1e-4))
    pops = [(nP, POISSON, 1), (nE, LIFB, 0), (nI, LIFB, 0)]
    pr = [dict(src=s, tar=t, w=w, d=d, kind=BERN, p=p)
          for s, t, w in ((0, 1, J), (0, 2, J), (1, 1, J), (1, 2, J), (2, 1, wi), (2, 2, wi))]
    lc = gen_life_const(0.0, 0.0, 0.010, 0.0, 0.0, 0.002, 1.0, 1.0, 0.020, 0.0)
    net = Net(pops, pr, lc, name=f"config1_x{scale}")
    net.rates = [poisson_rate, 0.0, 0.0]
    return net


def config2(scale=1.0, K=10000, poisson_rate=0.0015, stdp=False):
    """Brunel 80k E + 20k I LIFB + 10k Poisson, fixed in-degree K split by source size,
    delays on the 1.0-2.0 ms grid (11 values)."""
    nP, nE, nI = int(10000 * scale), int(80000 * scale), int(20000 * scale)
    ntot = nP + nE + nI
    Ks = [int(round(K * n / ntot)) for n in (nP, nE, nI)]
    J = float(f32(1e-4))
    wi = float(f32(-5e-4))
    dmin, dmax = float(f32(0.001)), float(f32(0.002))
    pops = [(nP, POISSON, 1), (nE, LIFB, 0), (nI, LIFB, 0)]
    pr = []
    for tgt in (1, 2):
        for src, k, w in ((0, Ks[0], J), (1, Ks[1], J), (2, Ks[2], wi)):
            pr.append(dict(src=src, tar=tgt, wmin=w, wmax=w, dmin=dmin, dmax=dmax, kind=FIXK, k=k,
                           plastic=1 if (stdp and src == 1 and tgt == 1) else 0))
    lc = gen_life_const(0.0, 0.0, 0.010, 0.0, 0.0, 0.002, 1.0, 1.0, 0.020, 0.0)
    net = Net(pops, pr, lc, name=f"config2_x{scale}")
    net.rates = [poisson_rate, 0.0, 0.0]
    return net


def config3(scale=1.0, p=0.02, g_exc=0.04, g_inh=0.51):
    """Vogels-Abbott COBA: 64k E + 16k I LIFE, Bernoulli p = 0.02, delay 0.8 ms, V0 ~ U[-60, -50] mV;
    g_exc 0.4 nS / g_inh 5.1 nS over g_L = 10 nS (see mgbrain_amd.workloads.vogels_abbott)."""
    nE, nI = int(64000 * scale), int(16000 * scale)
    d = float(f32(0.0008))
    pops = [(nE, LIFE, 0), (nI, LIFE, 0)]
    we, wi = float(f32(g_exc)), float(f32(-g_inh))
    pr = [dict(src=s, tar=t, w=w, d=d, kind=BERN, p=p) for s, t, w in ((0, 0, we), (0, 1, we), (1, 0, wi), (1, 1, wi))]
    net = Net(pops, pr, life_const1(), name=f"config3_x{scale}")
    net.vrange = [(float(f32(-0.060)), float(f32(-0.050)))] * 2
    return net


def stdp_consts_config4(J=1e-4):
    """A+ = 0.01, A- = 0.0105 in units of W_max = 2J; tau = 20 ms = 200 steps; W in [0, 2J]."""
    wmax = 2 * J
    return np.array([0.01 * wmax, -0.0105 * wmax, 200.0, 200.0, wmax, 0.0], dtype=f32)


def lif_net(J=0.008, g=5.0, poisson_rate=0.02):
    """Current-based LIF with exponential synaptic currents (LIFECONSTS from
    gen_life_const with cm 0.25, tau_m 20 ms, tau_syn 5 ms; mgcore.cu:132-163):
    400 Poisson + 1600 E + 400 I, Bernoulli p = 0.1, delay 1.2 ms (~22 Hz)."""
    lc = gen_life_const(0.0, 0.0, 0.010, 0.25, 0.020, 0.002, 0.005, 0.005, 0.020, 0.0)
    pops = [(400, POISSON, 1), (1600, LIF, 0), (400, LIF, 0)]
    w, wi, d = float(f32(J)), float(f32(-g * J)), float(f32(0.0012))
    pr = [dict(src=s, tar=t, w=ww, d=d, kind=BERN, p=0.1)
          for s, t, ww in ((0, 1, w), (0, 2, w), (1, 1, w), (1, 2, w), (2, 1, wi), (2, 2, wi))]
    net = Net(pops, pr, lc, name="lif")
    net.rates = [poisson_rate, 0.0, 0.0]
    return net
