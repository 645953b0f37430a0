"""Reduced-precision weight storage (BF16, FP8-E4M3 with per-row scales) and
the fixed-point FP32 format.

1. The device's stored weights equal an independent numpy rounding of the
   list weights (BF16 round-to-nearest-even of w / scale, scale = row max |w|;
   E4M3 of w / scale, scale = row max |w| / 448), and with those weights the
   simulation is bit-exact against the oracle's fixed-point mode; likewise for
   FP8 with the caller's scales, one per source population.
2. Against the FP64 oracle on the Vogels-Abbott COBA network (80,000 LIFE
   neurons, 1.28e8 synapses, 2 s at ~1.2 Hz): population rates within 2 % for every format;
   the membrane-potential error up to the first spike divergence is reported.
"""
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def bf16_round(w):
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


E4M3 = np.array([(1 + (b & 7) / 8.0) * 2.0 ** (((b >> 3) & 15) - 7) if (b >> 3) & 15 else (b & 7) * 2.0 ** -9
                 for b in range(127)], np.float64)      # non-negative E4M3 values, codes 0..126


def e4m3_round(x):
    """Nearest E4M3 value (ties to even code), saturating at 448."""
    a = np.minimum(np.abs(x.astype(np.float64)), 448.0)
    idx = np.clip(np.searchsorted(E4M3, a), 1, 126)
    lo, hi = E4M3[idx - 1], E4M3[idx]
    pick_hi = (hi - a < a - lo) | ((hi - a == a - lo) & ((idx % 2) == 0))
    v = np.where(pick_hi, hi, lo)
    return np.sign(x) * v


def ranged_net():
    # reference connectors with weight ranges -> many distinct weights per target
    pops = [(400, M.POISSON, 1), (300, M.LIFB, 0), (100, M.LIFB, 0)]
    pr = [dict(src=0, tar=1, wmin=0.001, wmax=0.004, dmin=0.001, dmax=0.002, ctype=0.2, kind=0),
          dict(src=0, tar=2, wmin=0.002, wmax=0.006, dmin=0.001, dmax=0.0015, ctype=0.2, kind=0),
          dict(src=1, tar=1, wmin=0.0005, wmax=0.003, dmin=0.0012, dmax=0.002, ctype=0.1, kind=0),
          dict(src=2, tar=1, wmin=-0.01, wmax=-0.003, dmin=0.001, dmax=0.001, ctype=0.1, kind=0),
          dict(src=1, tar=2, wmin=0.001, wmax=0.002, dmin=0.001, dmax=0.002, ctype=0.1, kind=0),
          dict(src=2, tar=2, wmin=-0.008, wmax=-0.008, dmin=0.001, dmax=0.001, ctype=0.1, kind=0)]
    net = M.Net(pops, pr, M.life_const2(), name="ranged")
    net.rates = [0.05, 0.0, 0.0]
    return net


@pytest.mark.parametrize("prop", [N.PROP_DIRECT, N.PROP_TILED])
@pytest.mark.parametrize("wf", [N.WF_BF16, N.WF_FP8, N.WF_F32])
def test_stored_weights_and_bit_exact_dynamics(wf, prop):
    _say(f"case: ranged-weight net, weight_format={wf}, propagation={prop}")
    net = ranged_net()
    syn = H.host_synapses(net, mt_seed=77, rand_seed=5, dw=0.0001)
    assert len(np.unique(syn.weight)) > 20
    g = H.gpu_from_list(net, syn, seed=9, weight_format=wf, record_capacity=1_000_000, propagation=prop)
    assert g.info()["propagation"] == prop
    rp, col, dl, w = g.csr()
    c = po.build_csr(net.n, syn)
    assert np.array_equal(rp, c.row_ptr) and np.array_equal(col, c.col) and np.array_equal(dl, c.dly)
    if wf == N.WF_F32:
        want = c.w.copy()
    else:
        want = np.zeros_like(c.w)
        top = np.float32(448.0) if wf == N.WF_FP8 else np.float32(1.0)
        for i in range(net.n):
            a, b = c.row_ptr[i], c.row_ptr[i + 1]
            if a == b:
                continue
            scale = np.float32(np.max(np.abs(c.w[a:b])) / top)
            x = (c.w[a:b] / scale).astype(np.float32)
            q = e4m3_round(x).astype(np.float32) if wf == N.WF_FP8 else bf16_round(x)
            want[a:b] = (q * scale).astype(np.float32)
    assert np.array_equal(w.view(np.uint32), want.view(np.uint32))
    # oracle (fixed point, same scale) on the stored weights
    info = g.info()
    q = po.SynapseList(syn.src, syn.tar, syn.weight.copy(), syn.delay, syn.plastic)
    q.weight[c.perm] = w
    _, o = H.oracle(net, q, seed=9, acc=po.ACC_FIXED, scale=(info["fix_scale_exc"], info["fix_scale_inh"]))
    g.run(1500)
    st, ids = g.spikes()
    k = H.first_mismatch(H.split_steps(st, ids, 1500), o.run(1500))
    assert k < 0, k
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    g.close()
    o.close()


@pytest.mark.parametrize("prop", [N.PROP_DIRECT, N.PROP_TILED])
def test_fp8_per_population_scales(prop):
    """FP8 with the caller's scales, one per source population (the largest |w| of the
    population's outgoing synapses / 448) instead of one per source row: stored weights
    equal numpy's E4M3 rounding of w / s_pop times s_pop, and the run is bit-exact
    against the oracle's fixed point on them."""
    _say(f"case: ranged-weight net, FP8 with per-population scales, propagation={prop}")
    net = ranged_net()
    syn = H.host_synapses(net, mt_seed=78, rand_seed=6, dw=0.0001)
    c = po.build_csr(net.n, syn)
    bounds = np.cumsum([0] + [p[0] for p in net.pops])
    sc = np.ones(net.n, np.float32)
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        a, b = c.row_ptr[lo], c.row_ptr[hi]
        if a < b:
            sc[lo:hi] = np.float32(np.max(np.abs(c.w[a:b])) / np.float32(448.0))
    assert len(np.unique(sc)) == 3
    g = H.gpu_from_list(net, syn, seed=9, weight_format=N.WF_FP8, fp8_scale=sc, record_capacity=1_000_000,
                        propagation=prop)
    rp, col, dl, w = g.csr()
    want, per_row = np.zeros_like(c.w), np.zeros_like(c.w)
    for i in range(net.n):
        a, b = c.row_ptr[i], c.row_ptr[i + 1]
        if a == b:
            continue
        x = (c.w[a:b] / sc[i]).astype(np.float32)
        want[a:b] = (e4m3_round(x).astype(np.float32) * sc[i]).astype(np.float32)
        rs = np.float32(np.max(np.abs(c.w[a:b])) / np.float32(448.0))
        per_row[a:b] = (e4m3_round((c.w[a:b] / rs).astype(np.float32)).astype(np.float32) * rs).astype(np.float32)
    assert np.array_equal(w.view(np.uint32), want.view(np.uint32))
    # the default per-row scales would have stored other values
    assert np.sum(per_row != want) > 100
    info = g.info()
    q = po.SynapseList(syn.src, syn.tar, syn.weight.copy(), syn.delay, syn.plastic)
    q.weight[c.perm] = w
    _, o = H.oracle(net, q, seed=9, acc=po.ACC_FIXED, scale=(info["fix_scale_exc"], info["fix_scale_inh"]))
    g.run(1500)
    st, ids = g.spikes()
    k = H.first_mismatch(H.split_steps(st, ids, 1500), o.run(1500))
    assert k < 0, k
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    g.close()
    o.close()


def test_e4m3_grid_known_values():
    assert E4M3[0x38] == 1.0 and E4M3[0x7E] == 448.0 and E4M3[0x08] == 2.0 ** -6 and E4M3[0x01] == 2.0 ** -9
    assert e4m3_round(np.array([0.4 / (5.1 / 448)]))[0] == 36.0   # 35.14 -> 36 on the 2^2 grid


def test_coba_formats_against_fp64():
    _say("case: Vogels-Abbott 80k LIFE, 1.28e8 synapses, 50000 steps (5 s), formats vs FP64 oracle")
    net = M.config3(scale=1.0)
    nsteps, every = 50000, 50
    sims = {name: H.gpu_generated(net, seed=4, weight_format=wf, record_capacity=2_000_000)
            for name, wf in (("f32", N.WF_CLASS), ("f32_fixed", N.WF_F32), ("bf16", N.WF_BF16),
                             ("fp8", N.WF_FP8))}
    ref = sims["f32"]
    rp, col, dl, w = ref.csr()
    n = net.n
    types, rate = net.neuron_arrays()
    cs = po.CsrArrays(n, rp, col, dl, w, np.zeros(1, np.uint8), np.zeros(1, np.int64), np.zeros(n + 1, np.int64),
                      np.zeros(1, np.int64), ref.info()["ring_len"], ref.info()["min_delay"])
    o = po.OracleSim(cs, types, rate, H.initial_v(net, 4), net.lifconst, seed=4, acc=po.ACC_F64)
    vmax = {k: 0.0 for k in sims}
    diverged = {k: None for k in sims}
    want = []
    t0 = 0
    # spike lists and V against FP64 every 50 steps until every format has diverged
    while t0 < nsteps and any(v is None for v in diverged.values()):
        want += o.run(every, cap=n * every)
        v64 = o.state(po.F_V)
        for name, g in sims.items():
            g.run(every)
            if diverged[name] is None:
                st, ids = g.spikes()
                k = H.first_mismatch(H.split_steps(st, ids, t0 + every), want)
                if k >= 0:
                    diverged[name] = k
                else:
                    vmax[name] = max(vmax[name], float(np.max(np.abs(g.state(N.F_V) - v64))))
        t0 += every
    # then rates over the whole 5 s
    o.run(nsteps - t0, cap=4_000_000)
    for g in sims.values():
        g.run(nsteps - t0)
    c64 = o.state(po.F_CNT)
    nE = net.pops[0][0]
    errs = {}
    for name, g in sims.items():
        c = g.fire_counts()
        errE = abs(c[:nE].sum() - c64[:nE].sum()) / c64[:nE].sum()
        errI = abs(c[nE:].sum() - c64[nE:].sum()) / c64[nE:].sum()
        errs[name] = (errE, errI)
        _say(f"  {name}: rate error E {errE * 100:.3f} % I {errI * 100:.3f} %, first spike divergence at step "
             f"{diverged[name]}, max |V - V_fp64| before it {vmax[name]:.3e} V")
        g.close()
    o.close()
    for name, (errE, errI) in errs.items():
        assert errE < 0.02 and errI < 0.02, name
