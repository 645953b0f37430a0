"""Checkpoint / resume (mgb_checkpoint_save / mgb_checkpoint_load).

Bar: a simulator restored from a checkpoint and run on continues bit for bit
like the original run on — per-step spike lists (recorded spikes included),
fire counts, membrane potentials, plastic weights — for the fused and direct
paths, the wide formats, STDP and batched networks.  The checkpoint is taken
after a partial window so that pending spikes and ring contents matter.  A run
split by a checkpoint into two simulators (the first closed before the second
restores it) also equals the oracle's uninterrupted run bit for bit.
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

FIRST, SECOND = 703, 900


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _check_resume(make):
    a = make()
    a.run(FIRST)
    blob = a.checkpoint()
    a.run(SECOND)
    b = make()
    b.restore(blob)
    b.run(SECOND)
    sa, ia = a.spikes()
    sb, ib = b.spikes()
    assert len(ia) > 100
    assert np.array_equal(sa, sb) and np.array_equal(ia, ib)
    assert np.array_equal(a.fire_counts(), b.fire_counts())
    assert np.array_equal(a.state(N.F_V).view(np.uint32), b.state(N.F_V).view(np.uint32))
    assert np.array_equal(a.weights().view(np.uint32), b.weights().view(np.uint32))
    ia_, ib_ = a.info(), b.info()
    assert ia_["step"] == ib_["step"] == FIRST + SECOND and ia_["syn_events"] == ib_["syn_events"]
    a.close()
    b.close()


@pytest.mark.parametrize("prop,wf,graph", [(N.PROP_TILED, N.WF_CLASS, 0), (N.PROP_DIRECT, N.WF_CLASS, 0),
                                           (N.PROP_TILED, N.WF_BF16, 1), (N.PROP_DIRECT, N.WF_F32, 1)])
def test_resume_bit_identical(prop, wf, graph):
    _say(f"case: config1 x0.3 generated, propagation={prop}, weight_format={wf}, graph={graph}")
    net = M.config1(scale=0.3, poisson_rate=0.01)
    _check_resume(lambda: H.gpu_generated(net, seed=21, propagation=prop, weight_format=wf, use_graph=graph,
                                          record_capacity=1_000_000))


@pytest.mark.parametrize("prop", [N.PROP_DIRECT, N.PROP_TILED])
def test_resume_stdp(prop):
    _say(f"case: fixed in-degree net with E->E STDP, checkpoint mid-run, propagation={prop}")
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=6)
    sc = M.stdp_consts_config4(J=1e-4)

    def make():
        return H.gpu_from_list(net, syn, seed=6, stdp=True, stdpconst=sc, weight_format=N.WF_F32,
                               record_capacity=1_000_000, propagation=prop)
    _check_resume(make)


def test_resume_batch():
    _say("case: batch of 4 sweep networks, checkpoint mid-run")
    _, nets = W.sweep_nets(n_nets=4, n_lo=1500, n_hi=4000, size_seed=3)
    _check_resume(lambda: N.Simulator.batch_generated(nets, record_capacity=1_000_000))


def test_restore_rejects_other_layout():
    _say("case: checkpoint loaded into a simulator with another tiling")
    net = M.config1(scale=0.2)
    a = H.gpu_generated(net, seed=2, propagation=N.PROP_TILED, tile_size=64)
    a.run(50)
    blob = a.checkpoint()
    b = H.gpu_generated(net, seed=2, propagation=N.PROP_TILED, tile_size=96)
    with pytest.raises(N.MGBError):
        b.restore(blob)
    with pytest.raises(N.MGBError):
        a.restore(blob[:200])           # truncated
    a.restore(blob)                     # the whole blob is accepted
    a.close()
    b.close()


@pytest.mark.parametrize("kind", ["tiled_class", "direct_f32", "tiled_stdp", "tiled_stdp_bf16"])
def test_resumed_run_against_the_oracle(kind):
    _say(f"case: checkpointed run against the oracle, {kind}")
    sc, stdp, acc = None, False, po.ACC_F32
    if kind.startswith("tiled_stdp"):
        # (tiled_stdp_bf16: BF16 static weights; this network's rows of equal static
        # weights are stored exactly)
        net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
        syn = H.host_synapses(net, seed=6)
        sc, stdp, acc = M.stdp_consts_config4(J=1e-4), True, po.ACC_FIXED
        opts = dict(weight_format=N.WF_BF16 if kind.endswith("bf16") else N.WF_F32, propagation=N.PROP_TILED)
    else:
        net = M.config1(scale=0.3, poisson_rate=0.01)
        syn = H.host_synapses(net, seed=31)
        opts = dict(weight_format=N.WF_CLASS if kind == "tiled_class" else N.WF_F32,
                    propagation=N.PROP_TILED if kind == "tiled_class" else N.PROP_DIRECT)
        acc = po.ACC_F32 if kind == "tiled_class" else po.ACC_FIXED    # FP32 words sum in fixed point
    seed = 6 if stdp else 31

    def make():
        return H.gpu_from_list(net, syn, seed=seed, stdp=stdp, stdpconst=sc, record_capacity=2_000_000, **opts)
    a = make()
    a.run(FIRST)
    blob = a.checkpoint()
    info = a.info()
    a.close()
    b = make()
    b.restore(blob)
    b.run(SECOND)
    st, ids = b.spikes()
    got = H.split_steps(st, ids, FIRST + SECOND)
    _, o = H.oracle(net, syn, seed=seed, acc=acc, stdp=stdp, stdpconst=sc,
                    scale=(info["fix_scale_exc"], info["fix_scale_inh"]))
    want = o.run(FIRST + SECOND)
    k = H.first_mismatch(got, want)
    assert k < 0, f"spike lists differ first at step {k}"
    assert sum(len(x) for x in want[FIRST:]) > 100
    assert np.array_equal(b.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert np.array_equal(b.fire_counts(), o.state(po.F_CNT).astype(np.int32))
    assert b.info()["syn_events"] == o.syn_events()
    if stdp:
        assert np.array_equal(b.weights().view(np.uint32), o.weights().astype(np.float32).view(np.uint32))
    b.close()
    o.close()
