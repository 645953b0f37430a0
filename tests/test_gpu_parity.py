"""HIP path vs the CPU oracle on identical seeded networks (through the C ABI).

Bar: bit-exact CSR (row_ptr, target, delay, weight in the reference's axon
order), bit-exact per-step spike index lists, bit-exact final state.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare_run(net, syn, nsteps, seed=0, gpu_kw=None, generated=False, acc=po.ACC_F32, stdp=False,
                 stdpconst=None):
    gpu_kw = dict(gpu_kw or {})
    gpu_kw.setdefault("record_capacity", 4_000_000)
    if generated:
        g = H.gpu_generated(net, seed=seed, stdp=stdp, stdpconst=stdpconst, **gpu_kw)
    else:
        g = H.gpu_from_list(net, syn, seed=seed, stdp=stdp, stdpconst=stdpconst, **gpu_kw)
    info = g.info()
    scale = (info["fix_scale_exc"], info["fix_scale_inh"])
    csr, o = H.oracle(net, syn, seed=seed, acc=acc, stdp=stdp, stdpconst=stdpconst, scale=scale)
    rp, col, dl, w = g.csr()
    assert np.array_equal(rp, csr.row_ptr)
    assert np.array_equal(col, csr.col)
    assert np.array_equal(dl, csr.dly)
    assert np.array_equal(w.view(np.uint32), csr.w.view(np.uint32))
    assert info["ring_len"] == csr.D and info["min_delay"] == csr.min_delay
    g.run(nsteps)
    steps, ids = g.spikes()
    got = H.split_steps(steps, ids, nsteps)
    want = o.run(nsteps)
    k = H.first_mismatch(got, want)
    assert k < 0, f"spike lists differ first at step {k}: gpu {got[k][:10]} oracle {want[k][:10]}"
    v_g = g.state(0)
    assert np.array_equal(v_g.view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert np.array_equal(g.state(5), o.state(po.F_CNT).astype(np.int32))
    assert np.array_equal(g.state(4), o.state(po.F_LAST).astype(np.int32))
    assert np.array_equal(g.state(3), o.state(po.F_REFR).astype(np.int32))
    assert g.info()["syn_events"] == o.syn_events()
    n_spk = sum(len(x) for x in want)
    g.close()
    o.close()
    return n_spk


@pytest.mark.parametrize("prop", [1, 2])
def test_brunel_reference_generator_class_format(prop):
    _say(f"case: make_brunel(2e5) quota generator, Poisson p=0.01, 1000 steps, CLASS format, prop={prop}")
    net = M.brunel_ref(200_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=4242)
    n = _compare_run(net, syn, 1000, seed=7, gpu_kw=dict(weight_format=1, propagation=prop))
    assert n > 1000


@pytest.mark.parametrize("tile,chunk,block", [(64, 4, 256), (100, 8, 512), (333, 16, 1024), (2000, 8, 64),
                                              (150, 1, 512), (77, 2, 1024)])
def test_tiled_shapes(tile, chunk, block):
    _say(f"case: config1 x0.2 tiled tile={tile} chunk={chunk} block={block}")
    net = M.config1(scale=0.2, poisson_rate=0.2)
    syn = H.host_synapses(net, seed=12)
    _compare_run(net, syn, 700, seed=12, gpu_kw=dict(propagation=2, tile_size=tile, chunk=chunk,
                                                      prop_block=block))


def test_tiled_split_runs():
    _say("case: tiled path advanced in uneven mgb_run calls")
    net = M.config1(scale=0.2, poisson_rate=0.2)
    syn = H.host_synapses(net, seed=13)
    g = H.gpu_from_list(net, syn, seed=13, record_capacity=2_000_000, propagation=2)
    for k in (1, 5, 13, 100, 7, 374):
        g.run(k)
    steps, ids = g.spikes()
    got = H.split_steps(steps, ids, 500)
    _, o = H.oracle(net, syn, seed=13)
    want = o.run(500)
    assert H.first_mismatch(got, want) < 0
    assert np.array_equal(g.state(0).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("prop", [1, 2])
def test_brunel_reference_generator_fixed_format(prop):
    _say(f"case: make_brunel(2e5), F32 fixed-point format vs oracle fixed mode, prop={prop}")
    net = M.brunel_ref(200_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=99)
    n = _compare_run(net, syn, 1000, seed=3, gpu_kw=dict(weight_format=2, propagation=prop), acc=po.ACC_FIXED)
    assert n > 1000


@pytest.mark.parametrize("window", [1, 3, 0])
def test_window_sizes_identical(window):
    _say(f"case: config1 x0.2 Bernoulli, window={window}")
    net = M.config1(scale=0.2, poisson_rate=0.2)
    syn = H.host_synapses(net, seed=11)
    _compare_run(net, syn, 600, seed=5, gpu_kw=dict(window=window))
    _compare_run(net, syn, 600, seed=5, gpu_kw=dict(window=window, propagation=1))


def test_generated_bernoulli_matches_host_list():
    _say("case: config1 x0.2 device-generated Bernoulli vs host list")
    net = M.config1(scale=0.2, poisson_rate=0.2)
    syn = H.host_synapses(net, seed=21)
    _compare_run(net, syn, 500, seed=21, generated=True)


def test_generated_fixed_indegree_delays():
    _say("case: config2 x0.02 device-generated fixed in-degree, 1.0-2.0 ms delays")
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3)
    syn = H.host_synapses(net, seed=8)
    _compare_run(net, syn, 500, seed=8, generated=True)


def test_vogels_abbott_life():
    _say("case: config3 x0.05 LIFE (COBA) with V0 ~ U[-60,-50] mV")
    net = M.config3(scale=0.05, p=0.1, g_exc=0.4, g_inh=5.1)   # make_vogel's units at 4k neurons: active
    syn = H.host_synapses(net, seed=2)
    _compare_run(net, syn, 500, seed=2)


@pytest.mark.parametrize("prop,wf", [(1, 1), (2, 1), (2, 2)])
def test_lif_exponential_currents(prop, wf):
    _say(f"case: current-based LIF network (exponential synaptic currents), prop={prop}, weight_format={wf}")
    net = M.lif_net()
    syn = H.host_synapses(net, seed=3)
    n = _compare_run(net, syn, 1500, seed=3, gpu_kw=dict(propagation=prop, weight_format=wf),
                     acc=po.ACC_FIXED if wf == 2 else po.ACC_F32)
    assert n > 3000


@pytest.mark.parametrize("prop", [1, 2])
def test_mixed_lif_lifb_fixed_point(prop):
    """LIF and LIFB neurons in one fixed-point network: two rings, the LIF neurons
    read both sums, the LIFB neurons their sum converted once."""
    _say(f"case: LIF + LIFB network, F32 fixed point, prop={prop}")
    net = M.lif_net()
    net.pops[2] = (400, M.LIFB, 0)
    syn = H.host_synapses(net, seed=4)
    n = _compare_run(net, syn, 1500, seed=4, gpu_kw=dict(propagation=prop, weight_format=2), acc=po.ACC_FIXED)
    assert n > 3000


def test_graph_replay_identical():
    _say("case: config1 x0.2, hipGraph replay vs oracle")
    net = M.config1(scale=0.2, poisson_rate=0.2)
    syn = H.host_synapses(net, seed=31)
    _compare_run(net, syn, 1200, seed=31, gpu_kw=dict(use_graph=1))
    _compare_run(net, syn, 1200, seed=31, gpu_kw=dict(use_graph=1, propagation=1))
