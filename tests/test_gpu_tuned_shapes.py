"""Every launch shape of mgbrain_amd/tuned_shapes.json, at reduced network size.

Bar: each workload's tuned (tile_size, block, chunk), on its fused path
(persistent launches; fused STDP one launch per window), gives results
bit-exact to the oracle -- per-step spike lists, V_m, spike
counters, last-spike and refractory state, synaptic events, and the weights
with STDP -- over several windows; a batched entry equals the same batch built
with the default shape one launch per window (which test_gpu_batch.py pins to
separate runs and test_gpu_parity.py to the oracle).  The shapes table is what
bench.py launches the benchmark workloads with, so a shape that breaks a
workload is caught before it is timed.
"""
import json
import os
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from mgbrain_amd import workloads as W
from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(__file__), "..", "mgbrain_amd", "tuned_shapes.json")
SHAPES = {k: v for k, v in json.load(open(TABLE)).items() if not k.startswith("_")}
STEPS = 600


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _shape(name):
    s = SHAPES[name]
    return dict(tile_size=s["tile_size"], prop_block=s["block"], chunk=s["chunk"])


def _against_oracle(g, net, syn, seed, acc=po.ACC_F32, stdp=False, stdpconst=None):
    info = g.info()
    # the fused STDP kernel runs one launch per window; the others are persistent
    assert info["propagation"] == N.PROP_TILED and (info["persistent"] == 0 if stdp else info["persistent"] >= 1)
    _, o = H.oracle(net, syn, seed=seed, acc=acc, stdp=stdp, stdpconst=stdpconst,
                    scale=(info["fix_scale_exc"], info["fix_scale_inh"]))
    w0 = g.weights() if stdp else None
    for n in (STEPS // 2, STEPS // 2):
        g.run(n)
    st, ids = g.spikes()
    got = H.split_steps(st, ids, STEPS)
    want = o.run(STEPS)
    k = H.first_mismatch(got, want)
    assert k < 0, f"spike lists differ first at step {k}"
    assert sum(len(x) for x in want) > 500
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert np.array_equal(g.state(N.F_CNT), o.state(po.F_CNT).astype(np.int32))
    assert np.array_equal(g.state(N.F_LAST), o.state(po.F_LAST).astype(np.int32))
    assert np.array_equal(g.state(N.F_REFR), o.state(po.F_REFR).astype(np.int32))
    assert g.info()["syn_events"] == o.syn_events()
    if stdp:
        wg = g.weights()
        assert np.array_equal(wg.view(np.uint32), o.weights().astype(np.float32).view(np.uint32))
        assert int(np.sum(wg != w0)) > 100
    g.close()
    o.close()


@pytest.mark.parametrize("name", ["brunel10k_bernoulli_p0.1", "brunel100k_fixed_indegree_K10000"])
def test_brunel_shapes(name):
    if name not in SHAPES:
        pytest.skip(f"{name} is not in the shapes table")
    _say(f"case: {name} shape {SHAPES[name]}")
    if name.startswith("brunel10k"):
        net = M.config1(scale=0.5, poisson_rate=0.01)
    else:
        net = M.config2(scale=0.05, K=1000, poisson_rate=0.05)
    syn = H.host_synapses(net, seed=21)
    g = H.gpu_from_list(net, syn, seed=21, propagation=N.PROP_TILED, record_capacity=4_000_000, **_shape(name))
    _against_oracle(g, net, syn, seed=21)


@pytest.mark.parametrize("wf", [N.WF_F32, N.WF_BF16])
def test_stdp_shape(wf):
    name = "brunel100k_K10000_stdp_ee"
    if name not in SHAPES:
        pytest.skip(f"{name} is not in the shapes table")
    _say(f"case: {name} shape {SHAPES[name]}, E->E STDP, weight_format={wf}")
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=6)     # (rows of equal static weights: exact in BF16)
    sc = M.stdp_consts_config4(J=1e-4)
    g = H.gpu_from_list(net, syn, seed=6, stdp=True, stdpconst=sc, weight_format=wf,
                        propagation=N.PROP_TILED, record_capacity=4_000_000, **_shape(name))
    assert g.info()["window"] == 10
    _against_oracle(g, net, syn, seed=6, acc=po.ACC_FIXED, stdp=True, stdpconst=sc)


def test_conductance_shape():
    name = "vogels_abbott80k_p0.02"
    if name not in SHAPES:
        pytest.skip(f"{name} is not in the shapes table")
    _say(f"case: {name} shape {SHAPES[name]}, config3 x0.05, CLASS words")
    net = M.config3(scale=0.05)
    syn = H.host_synapses(net, seed=4)
    g = H.gpu_from_list(net, syn, seed=4, weight_format=N.WF_CLASS, propagation=N.PROP_TILED,
                        record_capacity=4_000_000, **_shape(name))
    _against_oracle(g, net, syn, seed=4)


@pytest.mark.parametrize("name,wf", [("brunel_sweep64", N.WF_CLASS), ("brunel_sweep64_bf16", N.WF_BF16)])
def test_batched_shapes(name, wf):
    if name not in SHAPES:
        pytest.skip(f"{name} is not in the shapes table")
    _params, nets = W.sweep_nets(n_nets=6, n_lo=1500, n_hi=6000, size_seed=13)
    _say(f"case: {name} shape {SHAPES[name]}, batch of {len(nets)} sweep networks, weight_format={wf}")
    s = SHAPES[name]
    a = N.Simulator.batch_generated(nets, weight_format=wf, record_capacity=4_000_000, propagation=N.PROP_TILED,
                                    tile_size=s["tile_size"])
    a.set_launch(s["block"], s["chunk"])
    b = N.Simulator.batch_generated(nets, weight_format=wf, record_capacity=4_000_000, propagation=N.PROP_TILED,
                                    persistent=N.PERSIST_OFF)
    assert a.info()["persistent"] >= 1
    for n in (STEPS // 2, STEPS // 2):
        a.run(n)
        b.run(n)
    sa, xa = a.spikes()
    sb, xb = b.spikes()
    assert len(xa) > 1000 and np.array_equal(sa, sb) and np.array_equal(xa, xb)
    for f in (N.F_V, N.F_IEXC, N.F_IINH, N.F_REFR, N.F_LAST, N.F_CNT):
        assert np.array_equal(a.state(f).view(np.uint32), b.state(f).view(np.uint32))
    assert a.info()["syn_events"] == b.info()["syn_events"]
    a.close()
    b.close()
