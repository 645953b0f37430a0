"""Persistent launches of the fused kernel (MGB_PERSIST_AUTO, _REGIONS, _LISTS).

Bar: a run whose windows execute in one persistent launch (ring slices kept in
LDS and neuron state in registers from window to window, a grid-wide wait
between windows) equals the same run with one launch per window bit for bit:
per-step spike lists, every state field and synaptic events, and the state
each saves (rings and pending window lists included) continues alike when
restored, for every synapse word of the fused kernel, batched segments and
graph replay, with the spike lists passed through per-tile regions (16-bit
CLASS words) and without.  Runs mix lengths that give persistent launches,
single windows and partial windows.  The per-window path is pinned to the
oracle by test_gpu_parity.py; a persistent run is also compared with the
oracle directly.  A wait that exceeds its bound fails the run with
MGB_ERR_HIP, with every workgroup leaving the kernel.
"""
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from mgbrain_amd import workloads as W
from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu

RUNS = (37, 250, 10, 5, 401)      # whole windows, one window, partial windows


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _state(g):
    return [g.state(f).view(np.uint32).copy() for f in (N.F_V, N.F_IEXC, N.F_IINH, N.F_REFR, N.F_LAST, N.F_CNT)]


def _compare(make, mode=N.PERSIST_AUTO, regions=None):
    a = make(mode)
    b = make(N.PERSIST_OFF)
    ia, ib = a.info(), b.info()
    assert ia["propagation"] == N.PROP_TILED and ib["propagation"] == N.PROP_TILED
    assert ia["persistent"] in (1, 2) and ib["persistent"] == 0
    if regions is not None:
        assert (ia["persistent"] == 2) == regions
    for n in RUNS:
        a.run(n)
        b.run(n)
    sa, xa = a.spikes()
    sb, xb = b.spikes()
    assert len(xa) > 200
    assert np.array_equal(sa, sb) and np.array_equal(xa, xb)
    for fa, fb in zip(_state(a), _state(b)):
        assert np.array_equal(fa, fb)
    ia, ib = a.info(), b.info()
    assert ia["syn_events"] == ib["syn_events"] and ia["step"] == ib["step"] == sum(RUNS)
    assert ia["launches"] < ib["launches"]
    # the dynamic state each saves (rings, pending window lists) continues alike:
    # both checkpoints restored into per-window simulators run on identically
    ca, cb = a.checkpoint(), b.checkpoint()
    a.close()
    b.close()
    ra, rb = make(N.PERSIST_OFF), make(N.PERSIST_OFF)
    ra.restore(ca)
    rb.restore(cb)
    ra.run(333)
    rb.run(333)
    sa, xa = ra.spikes()
    sb, xb = rb.spikes()
    assert len(xa) > 100 and np.array_equal(sa, sb) and np.array_equal(xa, xb)
    for fa, fb in zip(_state(ra), _state(rb)):
        assert np.array_equal(fa, fb)
    ra.close()
    rb.close()


# (16-bit CLASS words: the default at this size has no list regions; REGIONS passes
# every window's list but a launch's last in them, LISTS none)
@pytest.mark.parametrize("wf,mode,regions", [(N.WF_CLASS, N.PERSIST_AUTO, False),
                                             (N.WF_CLASS, N.PERSIST_REGIONS, True),
                                             (N.WF_CLASS, N.PERSIST_LISTS, False),
                                             (N.WF_BF16, N.PERSIST_AUTO, False),
                                             (N.WF_BF16, N.PERSIST_REGIONS, False),
                                             (N.WF_FP8, N.PERSIST_AUTO, False),
                                             (N.WF_F32, N.PERSIST_AUTO, False)])
def test_brunel_bernoulli_every_word(wf, mode, regions):
    _say(f"case: config1 x0.3 generated, weight_format={wf}, persistent={mode}")
    net = M.config1(scale=0.3, poisson_rate=0.01)
    _compare(lambda p: H.gpu_generated(net, seed=3, weight_format=wf, propagation=N.PROP_TILED, persistent=p,
                                       record_capacity=2_000_000), mode, regions)


def test_conductance_two_rings():
    _say("case: config3 x0.05 (LIFE, two fixed-point rings), F32")
    net = M.config3(scale=0.05)
    syn = H.host_synapses(net, seed=4)
    _compare(lambda p: H.gpu_from_list(net, syn, seed=4, weight_format=N.WF_F32, propagation=N.PROP_TILED,
                                       persistent=p, record_capacity=2_000_000))


@pytest.mark.parametrize("mode,regions", [(N.PERSIST_AUTO, False), (N.PERSIST_REGIONS, True)])
def test_fixed_indegree_wide_tiles(mode, regions):
    _say(f"case: config2 x0.05, K=1000 (CLASS16, mixed Poisson and LIFB tiles), persistent={mode}")
    net = M.config2(scale=0.05, K=1000, poisson_rate=0.05)
    _compare(lambda p: H.gpu_generated(net, seed=5, propagation=N.PROP_TILED, persistent=p,
                                       record_capacity=2_000_000), mode, regions)


@pytest.mark.parametrize("graph,mode", [(0, N.PERSIST_AUTO), (1, N.PERSIST_AUTO), (0, N.PERSIST_REGIONS),
                                        (1, N.PERSIST_REGIONS)])
def test_batched_segments(graph, mode):
    # (with list regions, each segment's workgroups scan the counts of its own tiles)
    _params, nets = W.sweep_nets(n_nets=5, n_lo=1500, n_hi=5000, size_seed=11)
    _say(f"case: batch of {len(nets)} sweep networks, graph={graph}, persistent={mode}")
    _compare(lambda p: N.Simulator.batch_generated(nets, record_capacity=2_000_000, use_graph=graph,
                                                   propagation=N.PROP_TILED, persistent=p),
             mode, mode == N.PERSIST_REGIONS)


@pytest.mark.parametrize("mode", [N.PERSIST_AUTO, N.PERSIST_REGIONS])
def test_graph_replay_blocks(mode):
    _say(f"case: config1 x0.3, graph replay of persistent blocks (1200 steps), persistent={mode}")
    net = M.config1(scale=0.3, poisson_rate=0.01)
    a = H.gpu_generated(net, seed=8, use_graph=1, record_capacity=2_000_000, persistent=mode)
    b = H.gpu_generated(net, seed=8, persistent=N.PERSIST_OFF, record_capacity=2_000_000)
    assert a.info()["persistent"] == (2 if mode == N.PERSIST_REGIONS else 1)
    for n in (1200, 3):
        a.run(n)
        b.run(n)
    sa, xa = a.spikes()
    sb, xb = b.spikes()
    assert len(xa) > 200 and np.array_equal(sa, sb) and np.array_equal(xa, xb)
    for fa, fb in zip(_state(a), _state(b)):
        assert np.array_equal(fa, fb)
    a.close()
    b.close()


@pytest.mark.parametrize("mode", [N.PERSIST_AUTO, N.PERSIST_REGIONS, N.PERSIST_LISTS])
def test_persistent_run_against_the_oracle(mode):
    _say(f"case: make_brunel(2e5), persistent={mode}, 600 steps against the oracle")
    net = M.brunel_ref(200_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=5)
    g = H.gpu_from_list(net, syn, seed=9, record_capacity=2_000_000, persistent=mode)
    assert g.info()["persistent"] in (1, 2)
    g.run(600)
    st, ids = g.spikes()
    got = H.split_steps(st, ids, 600)
    _, o = H.oracle(net, syn, seed=9)
    want = o.run(600)
    assert H.first_mismatch(got, want) < 0
    assert sum(len(x) for x in want) > 1000
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    g.close()
    o.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_wait_beyond_its_bound_fails_the_run(graph):
    # 8192 Poisson neurons firing every step, all projecting to 512 LIFB neurons: each
    # LIFB tile adds millions of synapses per window, the Poisson tiles none, so the
    # Poisson tiles wait far longer than the 1 us bound at the first window's end
    # (graph: the run's persistent launch and the kernel that reports the error word
    # replayed from a captured graph)
    _say(f"case: unequal tiles, spin_limit_us=1, graph={graph}")
    pops = [dict(num=8192, type=N.POISSON, rate=1.0), dict(num=512, type=N.LIFB, rate=0.0)]
    projs = [dict(src=0, tar=1, kind=N.CONN_BERNOULLI, p=1.0, wmin=1e-6, wmax=1e-6, dmin=1e-3, dmax=1e-3,
                  index=0, seed=1)]
    g = N.Simulator.generated(pops, projs, W.lifb_consts(), seed=1, propagation=N.PROP_TILED, tile_size=512,
                              spin_limit_us=1, use_graph=graph)
    info = g.info()
    assert info["persistent"] in (1, 2) and info["tiles"] >= 2
    g.run(600 if graph else 100)
    with pytest.raises(N.MGBError) as e:
        g.sync()
    assert e.value.status == N.ERR_HIP and "bound" in str(e.value)
    with pytest.raises(N.MGBError):
        g.run(10)
    g.close()
    # the device is fine afterwards
    h = N.Simulator.generated(pops, projs, W.lifb_consts(), seed=1, propagation=N.PROP_TILED, tile_size=512)
    h.run(30)
    assert h.info()["spikes"] > 0
    h.close()
