"""Batched networks (mgb_create_batch_generated).

Bar: every network of a batch evolves exactly as it does alone -- bit-identical
per-step spike lists, fire counts and synaptic events -- while all of them
advance in the same kernel launches (one fused-kernel workgroup per tile of
every network; each network is its own segment with its own spike lists).
Standalone runs are pinned to the oracle by test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from mgbrain_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 1500


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _nets():
    params, nets = W.sweep_nets(n_nets=6, n_lo=1500, n_hi=6000, size_seed=7)
    # one network with its own threshold: constant sets are rebased per network
    nets[2]["lifconst"] = W.lifb_consts(0.018)
    return params, nets


@pytest.mark.parametrize("prop,graph", [(0, 0), (0, 1), (1, 0)])
def test_batch_matches_separate_runs(prop, graph):
    params, nets = _nets()
    _say(f"case: batch of {len(nets)} sweep networks {[q['n'] for q in params]}, prop={prop}, graph={graph}, "
         f"{STEPS} steps")
    b = N.Simulator.batch_generated(nets, record_capacity=2_000_000, use_graph=graph, propagation=prop)
    info = b.info()
    if info["propagation"] == N.PROP_TILED:
        assert info["segments"] == len(nets)
    b.run(STEPS)
    bs, bi = b.spikes()
    cnt = b.fire_counts()
    se_b = b.info()["syn_events"]
    b.close()
    base, se_sum, total = 0, 0, 0
    for k, net in enumerate(nets):
        n = W.net_size(net)
        s = N.Simulator.generated(**net, record_capacity=2_000_000, propagation=prop)
        assert s.info()["segments"] == 1
        s.run(STEPS)
        ss, si = s.spikes()
        m = (bi >= base) & (bi < base + n)
        assert np.array_equal(ss, bs[m]), f"network {k}: spike steps differ"
        assert np.array_equal(si, bi[m] - base), f"network {k}: spike ids differ"
        assert np.array_equal(s.fire_counts(), cnt[base:base + n])
        se_sum += s.info()["syn_events"]
        total += len(si)
        s.close()
        base += n
    assert se_sum == se_b
    assert total > 1000          # the networks are active


def _two_nets_arrays(cross):
    """Two disjoint random LIFB networks with Poisson drive on streams 0 / 1; with
    cross=True one synapse joins them (then they cannot be separate segments)."""
    rng = np.random.default_rng(11)
    sizes = (700, 1100)
    types, rate, stream, src, tar, w = [], [], [], [], [], []
    base = 0
    for k, n in enumerate(sizes):
        npois = n // 5
        types += [N.POISSON] * npois + [N.LIFB] * (n - npois)
        rate += [0.004] * npois + [0.0] * (n - npois)
        stream += [k] * n
        m = n * 60
        s = rng.integers(0, n, m) + base
        t = rng.integers(npois, n, m) + base
        src.append(s)
        tar.append(t)
        w.append(np.where(s - base < int(0.8 * n), np.float32(2e-4), np.float32(-6e-4)))
        base += n
    src, tar, w = np.concatenate(src), np.concatenate(tar), np.concatenate(w).astype(np.float32)
    if cross:
        src = np.append(src, 5)
        tar = np.append(tar, sizes[0] + 300)
        w = np.append(w, np.float32(2e-4))
    d = np.full(len(src), np.float32(0.0012))
    return dict(types=np.array(types), rate=np.array(rate, np.float32), stream=np.array(stream, np.uint16),
                stream_base=np.array([0, sizes[0]], np.int32), stream_seed=np.array([5, 9], np.uint64),
                src=src, tar=tar, weight=w, delay_s=d)


@pytest.mark.parametrize("cross", [False, True])
def test_segments_from_streams(cross):
    _say(f"case: two stream networks from arrays, cross synapse={cross}")
    a = _two_nets_arrays(cross)
    lc = W.lifb_consts()
    out = {}
    for prop in (N.PROP_TILED, N.PROP_DIRECT):
        s = N.Simulator.from_arrays(a["types"], a["rate"], lc, a["src"], a["tar"], a["weight"], a["delay_s"],
                                    stream=a["stream"], stream_base=a["stream_base"],
                                    stream_seed=a["stream_seed"], record_capacity=1_000_000, propagation=prop)
        if prop == N.PROP_TILED:
            assert s.info()["segments"] == (1 if cross else 2)
        s.run(1200)
        out[prop] = (s.spikes(), s.fire_counts(), s.info()["syn_events"])
        s.close()
    (ts, ti), tc, tse = out[N.PROP_TILED]
    (ds, di), dc, dse = out[N.PROP_DIRECT]
    assert len(ti) > 100
    assert np.array_equal(ts, ds) and np.array_equal(ti, di)
    assert np.array_equal(tc, dc) and tse == dse


def test_batch_stdp_fused_equals_separate_runs():
    """Two networks with E->E STDP batched on the fused path: per-network spikes, fire
    counts and weights (deferred updates applied at the read) equal separate runs."""
    print("case: batch of two sweep networks with E->E STDP (fused, FP32)", flush=True)
    sys.stdout.flush()
    sc = W.stdp_consts()
    nets = []
    # Poisson drive strong enough for the LIF neurons to fire (~0.15 mV per step of input)
    for n, g, nu, seed in ((3000, 4.0, 500.0, 11), (5000, 5.0, 600.0, 12)):
        net = W.brunel_sweep_net(n, g, nu, seed)
        for q in net["projs"]:
            if q["src"] == 1 and q["tar"] == 1:
                q["plastic"] = 1
        net.update(stdp=True, stdpconst=sc)
        nets.append(net)
    opts = dict(weight_format=N.WF_F32, propagation=N.PROP_TILED, record_capacity=1_000_000)
    b = N.Simulator.batch_generated(nets, **opts)
    assert b.info()["segments"] == 2 and b.info()["window"] == 10
    b.run(1500)
    singles = [N.Simulator.generated(**net, **opts) for net in nets]
    for s in singles:
        s.run(1500)
    cb, wb = b.fire_counts(), b.weights()
    w0 = N.Simulator.batch_generated(nets, **opts)
    changed = int(np.sum(w0.weights() != wb))
    w0.close()
    off, woff = 0, 0
    for s in singles:
        c, w = s.fire_counts(), s.weights()
        assert np.array_equal(cb[off:off + len(c)], c)
        assert np.array_equal(wb[woff:woff + len(w)].view(np.uint32), w.view(np.uint32))
        off += len(c)
        woff += len(w)
        s.close()
    print(f"spikes {cb.sum()}, weights changed {changed}", flush=True)
    assert off == len(cb) and woff == len(wb) and cb.sum() > 20000 and changed > 10000
    b.close()


@pytest.mark.parametrize("wf,prop", [(N.WF_BF16, N.PROP_TILED), (N.WF_F32, N.PROP_DIRECT), (N.WF_FP8, N.PROP_DIRECT)])
def test_batch_fixed_point_formats_equal_separate_runs(wf, prop):
    """Fixed-point formats: every network of a batch accumulates with its own scale (the
    one it has alone), so per-network fire counts equal separate runs bit for bit."""
    print(f"case: batch of two driven sweep networks, weight_format={wf}, propagation={prop}", flush=True)
    sys.stdout.flush()
    nets = [W.brunel_sweep_net(n, g, nu, seed) for n, g, nu, seed in ((3000, 4.0, 500.0, 21), (6000, 5.0, 600.0, 22))]
    opts = dict(weight_format=wf, propagation=prop)
    b = N.Simulator.batch_generated(nets, **opts)
    assert b.info()["segments"] == 2
    b.run(1200)
    cb = b.fire_counts()
    off = 0
    for net in nets:
        s = N.Simulator.generated(**net, **opts)
        s.run(1200)
        c = s.fire_counts()
        assert np.array_equal(cb[off:off + len(c)], c)
        off += len(c)
        s.close()
    assert off == len(cb) and cb.sum() > 20000
    b.close()
