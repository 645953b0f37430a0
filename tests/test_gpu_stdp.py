"""Pair-based STDP on the GPU against the oracle (mgcore.cu:346-412).

Bit-exact spikes and weights over the run, for (a) the reference's global
STDP (every synapse plastic, default STDPCONSTS, inhibitory weights clipped to
W_MIN = 0 at their first update) and (b) plasticity on E->E only with the
benchmark's constants (A+ = 0.01 W_max, A- = -0.0105 W_max, tau = 200 steps,
W in [0, 2J]).  The oracle processes spikers sequentially in id order; the GPU
updates all spikers' rows in parallel: each synapse receives at most one
non-zero update per step, so both give the same weights.

Both GPU paths are checked: the direct kernels (one step per window) and the
fused tiled kernel, whose windows span up to min(min_delay, shortest
inter-spike interval of the plastic synapses' ends) steps — one step when
Poisson neurons have plastic synapses (a), ten steps for E->E plasticity with
a 2 ms refractory period and 1-2 ms delays (b).
"""
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu


def _run(net, syn, sc, nsteps, seed, wf, prop=N.PROP_DIRECT, window=1, **kw):
    g = H.gpu_from_list(net, syn, seed=seed, stdp=True, stdpconst=sc, weight_format=wf, record_capacity=4_000_000,
                        propagation=prop, **kw)
    info = g.info()
    assert info["window"] == window
    assert info["propagation"] == (N.PROP_DIRECT if prop == N.PROP_DIRECT else N.PROP_TILED)
    _, o = H.oracle(net, syn, seed=seed, acc=po.ACC_FIXED, stdp=True, stdpconst=sc,
                    scale=(info["fix_scale_exc"], info["fix_scale_inh"]))
    w0 = g.weights()
    g.run(nsteps)
    st, ids = g.spikes()
    got = H.split_steps(st, ids, nsteps)
    want = o.run(nsteps)
    k = H.first_mismatch(got, want)
    assert k < 0, f"spikes differ at step {k}"
    wg = g.weights()
    wo = o.weights().astype(np.float32)
    assert np.array_equal(wg.view(np.uint32), wo.view(np.uint32))
    changed = int(np.sum(wg != w0))
    g.close()
    o.close()
    return changed, sum(len(x) for x in want)


@pytest.mark.parametrize("prop", [N.PROP_DIRECT, N.PROP_TILED])
def test_reference_global_stdp(prop):
    print(f"case: make_brunel(2e5) global STDP, default STDPCONSTS, propagation={prop}", flush=True)
    sys.stdout.flush()
    net = M.brunel_ref(200_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=31)
    syn.plastic[:] = 1
    changed, n = _run(net, syn, M.STDP_DEFAULT, 600, 3, N.WF_F32, prop=prop, window=1)
    assert changed > 1000 and n > 1000


@pytest.mark.parametrize("wf,prop", [(N.WF_F32, N.PROP_DIRECT), (N.WF_BF16, N.PROP_DIRECT),
                                     (N.WF_F32, N.PROP_TILED), (N.WF_BF16, N.PROP_TILED),
                                     (N.WF_F32, N.PROP_AUTO)])
def test_ee_stdp_benchmark_constants(wf, prop):
    print(f"case: fixed in-degree net, E->E STDP, weight_format={wf}, propagation={prop}", flush=True)
    sys.stdout.flush()
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=6)
    assert 0 < syn.plastic.sum() < syn.n
    sc = M.stdp_consts_config4(J=1e-4)
    if wf == N.WF_BF16:
        _bf16_static(net, syn)
    window = 1 if prop == N.PROP_DIRECT else 10     # min_delay 10, refractory interval 20
    changed, n = _run(net, syn, sc, 800, 6, wf, prop=prop, window=window)
    assert changed > 100 and n > 1000


def _bf16_static(net, syn):
    """Static weights are stored as BF16 (round to nearest even) of w / s, s the largest
    |w| of the source row (plastic synapses included), and read as bf16 * s: gives the
    oracle the same values."""
    stat = syn.plastic == 0
    smax = np.zeros(net.n, np.float32)
    np.maximum.at(smax, syn.src, np.abs(syn.weight))
    sc_row = np.where(smax > 0, smax, np.float32(1.0)).astype(np.float32)[syn.src]
    x = (syn.weight / sc_row).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    syn.weight[stat] = (r * sc_row).astype(np.float32)[stat]
    return stat


@pytest.mark.parametrize("prop,chunk", [(N.PROP_DIRECT, 0), (N.PROP_TILED, 1), (N.PROP_TILED, 2),
                                        (N.PROP_TILED, 16)])
def test_stdp_bf16_static_weights_spread(prop, chunk):
    """BF16 static weights beside FP32 plastic ones (the storage of the STDP benchmark):
    static weights spread over [0.5, 1.5) x their value, so that the BF16 codes and the
    per-row scales matter (rows of equal weights are stored exactly).  On the fused path
    the word holds the BF16 code of a static synapse and marks a plastic one, whose
    weight is read beside it; chunks 1, 2 and 16 take the scalar, 2-word and 4-word
    vector loads."""
    print(f"case: fixed in-degree net, E->E STDP, BF16 static weights spread, propagation={prop}, "
          f"chunk={chunk}", flush=True)
    sys.stdout.flush()
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=6)
    stat = syn.plastic == 0
    rng = np.random.default_rng(5)
    syn.weight[stat] = (syn.weight[stat] * rng.uniform(0.5, 1.5, int(stat.sum()))).astype(np.float32)
    before = syn.weight.copy()
    _bf16_static(net, syn)
    assert np.count_nonzero(syn.weight != before) > stat.sum() // 2     # the rounding changed most of them
    sc = M.stdp_consts_config4(J=1e-4)
    kw = {} if chunk == 0 else {"chunk": chunk}
    window = 1 if prop == N.PROP_DIRECT else 10
    changed, n = _run(net, syn, sc, 800, 6, N.WF_BF16, prop=prop, window=window, **kw)
    assert changed > 100 and n > 1000


@pytest.mark.parametrize("chunk", [1, 2, 16])
def test_fused_stdp_chunks(chunk):
    """Fused path with work items of 1, 2 and 16 synapses (the default is 8): items
    are aligned vector loads of the padded tile order; a 16-synapse item reads past
    its sub-row's padding into the next sub-row, whose words must be ignored."""
    print(f"case: fixed in-degree net, E->E STDP, fused, chunk={chunk}", flush=True)
    sys.stdout.flush()
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=6)
    sc = M.stdp_consts_config4(J=1e-4)
    changed, n = _run(net, syn, sc, 800, 6, N.WF_F32, prop=N.PROP_TILED, window=10, chunk=chunk)
    assert changed > 100 and n > 1000


@pytest.mark.parametrize("tau_ref,window", [(0.002, 10), (0.0006, 6)])
def test_fused_stdp_weights_between_runs(tau_ref, window):
    """Fused path: weights read between runs (the deferred updates applied to a copy)
    equal the oracle's at every read, and reading them changes nothing; a 0.6 ms
    refractory period (inter-spike interval >= 6 steps) caps the window below min_delay.
    The runs are long enough for neurons to fire well over 16 times, so the flushes of
    the 16-entry spike history (deferred dendrite updates) are exercised."""
    print(f"case: E->E STDP, tau_ref={tau_ref}, runs of 457 / 600 / 443 steps", flush=True)
    sys.stdout.flush()
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    net.lifconst = M.gen_life_const(0.0, 0.0, 0.010, 0.0, 0.0, tau_ref, 1.0, 1.0, 0.020, 0.0)
    syn = H.host_synapses(net, seed=9)
    sc = M.stdp_consts_config4(J=1e-4)
    g = H.gpu_from_list(net, syn, seed=9, stdp=True, stdpconst=sc, weight_format=N.WF_F32,
                        record_capacity=4_000_000, propagation=N.PROP_TILED)
    info = g.info()
    assert info["window"] == window
    _, o = H.oracle(net, syn, seed=9, acc=po.ACC_FIXED, stdp=True, stdpconst=sc,
                    scale=(info["fix_scale_exc"], info["fix_scale_inh"]))
    w0 = g.weights()
    done, want = 0, []
    for n in (457, 600, 443):
        g.run(n)
        want += o.run(n)
        done += n
        wo = o.weights().astype(np.float32)
        assert np.array_equal(g.weights().view(np.uint32), wo.view(np.uint32)), f"weights differ after {done}"
        assert np.array_equal(g.weights().view(np.uint32), wo.view(np.uint32))   # a second read agrees
    st, ids = g.spikes()
    assert H.first_mismatch(H.split_steps(st, ids, done), want) < 0
    assert np.array_equal(g.fire_counts(), o.state(po.F_CNT).astype(np.int32))
    assert np.array_equal(g.state(N.F_LAST), o.state(po.F_LAST).astype(np.int32))
    assert int(np.sum(g.weights() != w0)) > 100
    counts = g.fire_counts()
    print(f"fire counts: max {counts.max()}, neurons above 31: {int(np.sum(counts > 31))}", flush=True)
    assert int(np.sum(counts > 31)) > 50        # each of these was flushed at least twice
    g.close()
    o.close()


def test_fused_stdp_conductance_neurons():
    """Fused STDP on E->E of a LIFE (conductance-based) network against the oracle: the
    window is min(min_delay 8, refractory interval 50) = 8 steps."""
    print("case: config3 x0.05 LIFE with E->E STDP, fused", flush=True)
    sys.stdout.flush()
    net = M.config3(scale=0.05, p=0.1, g_exc=0.4, g_inh=5.1)
    for q in net.projs:
        if q["src"] == 0 and q["tar"] == 0:
            q["plastic"] = 1
    syn = H.host_synapses(net, seed=4)
    assert 0 < syn.plastic.sum() < syn.n
    sc = M.stdp_consts_config4(J=0.4)
    changed, n = _run(net, syn, sc, 800, 4, N.WF_F32, prop=N.PROP_TILED, window=8)
    assert changed > 100 and n > 500


def test_fused_stdp_graph_replay_identical():
    """Fused STDP replayed as hipGraphs (fused kernel + flush kernel per window) gives the
    eager run's spikes and weights bit for bit, over runs that split windows."""
    print("case: E->E STDP fused, graph replay vs eager", flush=True)
    sys.stdout.flush()
    net = M.config2(scale=0.02, K=200, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=12)
    sc = M.stdp_consts_config4(J=1e-4)
    out = []
    for graph in (0, 1):
        g = H.gpu_from_list(net, syn, seed=12, stdp=True, stdpconst=sc, weight_format=N.WF_F32,
                            record_capacity=4_000_000, propagation=N.PROP_TILED, use_graph=graph)
        for n in (500, 503, 997):
            g.run(n)
        st, ids = g.spikes()
        out.append((st, ids, g.weights(), g.fire_counts()))
        g.close()
    (s0, i0, w0, c0), (s1, i1, w1, c1) = out
    assert np.array_equal(s0, s1) and np.array_equal(i0, i1)
    assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)) and np.array_equal(c0, c1)
    assert int(np.sum(c0 > 31)) > 50                          # flushes happened
