"""Edge cases of the HIP path against the oracle (bit-exact spike lists and state):
no synapses, a silent network, delays beyond the 5-bit field (direct path),
one-neuron tiles, rows of about 2^16 synapses, a one-neuron network with a self-synapse, a batch with an
empty and a Poisson-only network, and ragged tile shapes; the tiles' invariants (mgb_read_tiles)."""
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from mgbrain_amd import workloads as W
from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _arrays_vs_oracle(types, rate, lc, src, tar, w, d, nsteps, seed=1, v0=None, **opts):
    """Runs an explicit network on the GPU and the oracle; returns the spike count."""
    n = len(types)
    types = np.asarray(types, np.int32)
    rate = np.asarray(rate, np.float32)
    v0 = np.zeros(n, np.float32) if v0 is None else np.asarray(v0, np.float32)
    g = N.Simulator.from_arrays(types, rate, lc, src, tar, w, d, v_init=v0, seed=seed, record_capacity=1_000_000,
                                **opts)
    syn = po.SynapseList(np.asarray(src, np.int32), np.asarray(tar, np.int32), np.asarray(w, np.float32),
                         np.asarray(d, np.float32), np.zeros(len(src), np.uint8))
    csr = po.build_csr(n, syn)
    info = g.info()
    fixed = info["weight_format"] != N.WF_CLASS        # the wide formats accumulate in fixed point
    o = po.OracleSim(csr, types, rate, v0, lc, seed=seed, acc=po.ACC_FIXED if fixed else po.ACC_F32,
                     scale_exc=info["fix_scale_exc"], scale_inh=info["fix_scale_inh"])
    g.run(nsteps)
    st, ids = g.spikes()
    got = H.split_steps(st, ids, nsteps)
    want = o.run(nsteps)
    k = H.first_mismatch(got, want)
    assert k < 0, f"spikes differ at step {k}"
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert np.array_equal(g.fire_counts(), o.state(po.F_CNT).astype(np.int32))
    info = g.info()
    g.close()
    o.close()
    return sum(len(x) for x in want), info


def test_no_synapses():
    _say("case: 300 Poisson + 200 LIFB, no synapses")
    types = [M.POISSON] * 300 + [M.LIFB] * 200
    rate = [0.05] * 300 + [0.0] * 200
    e = np.zeros(0, np.int32)
    n, info = _arrays_vs_oracle(types, rate, W.lifb_consts(), e, e, np.zeros(0, np.float32), np.zeros(0, np.float32),
                                500)
    assert n > 1000 and info["n_synapses"] == 0


def test_silent_network():
    _say("case: LIFB network without drive (no spikes at all)")
    rng = np.random.default_rng(3)
    n = 400
    src, tar = rng.integers(0, n, 4000), rng.integers(0, n, 4000)
    w = np.full(4000, np.float32(1e-4))
    d = np.full(4000, np.float32(0.0015))
    cnt, info = _arrays_vs_oracle([M.LIFB] * n, [0.0] * n, W.lifb_consts(), src, tar, w, d, 300)
    assert cnt == 0 and info["syn_events"] == 0


@pytest.mark.parametrize("prop", [0, N.PROP_DIRECT])
def test_delays_beyond_class_field(prop):
    _say(f"case: delays up to 9 ms (90 steps), prop={prop}")
    rng = np.random.default_rng(5)
    nP, nL = 200, 800
    n = nP + nL
    s_ = rng.integers(0, n, 30000)
    t_ = rng.integers(nP, n, 30000)
    w = np.where(s_ < nP + 640, np.float32(2e-3), np.float32(-6e-3)).astype(np.float32)
    d = (rng.integers(10, 91, 30000) * 1e-4).astype(np.float32)
    cnt, info = _arrays_vs_oracle([M.POISSON] * nP + [M.LIFB] * nL, [0.03] * nP + [0.0] * nL, W.lifb_consts(),
                                  s_, t_, w, d, 800, propagation=prop)
    assert info["ring_len"] > 33 and info["propagation"] == N.PROP_DIRECT and cnt > 2000


@pytest.mark.parametrize("tile", [1, 7, 1000])
def test_ragged_tiles(tile):
    _say(f"case: config1 x0.1 with tile_size={tile}")
    net = M.config1(scale=0.1, poisson_rate=0.02)
    syn = H.host_synapses(net, seed=8)
    g = H.gpu_from_list(net, syn, seed=8, propagation=N.PROP_TILED, tile_size=tile, record_capacity=1_000_000)
    info = g.info()
    # about n / tile tiles (at least one per neuron type), none mixing types
    assert 2 <= info["tiles"] <= max(-(-net.n // tile), 2) and info["tile_size"] <= max(tile, 1024)
    csr, o = H.oracle(net, syn, seed=8)
    g.run(600)
    st, ids = g.spikes()
    want = o.run(600)
    assert H.first_mismatch(H.split_steps(st, ids, 600), want) < 0
    assert sum(len(x) for x in want) > 500
    g.close()
    o.close()


@pytest.mark.parametrize("n_tar", [65_530, 70_000])
def test_long_rows(n_tar):
    """Rows of about 2^16 synapses: the padded 16-bit tile-order copy needs 32-bit
    row offsets (70,000: already unpadded; 65,530: only once padded)."""
    _say(f"case: 4 Poisson sources x {n_tar} LIFB targets, tiled CLASS16")
    n_src = 4
    types = [M.POISSON] * n_src + [M.LIFB] * n_tar
    rate = [0.2] * n_src + [0.0] * n_tar
    rng = np.random.default_rng(5)
    src = np.repeat(np.arange(n_src, dtype=np.int32), n_tar)
    tar = np.tile(np.arange(n_src, n_src + n_tar, dtype=np.int32), n_src)
    w = np.full(src.size, 4e-3, np.float32)
    d = (rng.integers(10, 21, src.size) * 1e-4).astype(np.float32)
    cnt, info = _arrays_vs_oracle(types, rate, W.lifb_consts(), src, tar, w, d, 300, propagation=N.PROP_TILED,
                                  tile_size=1000)
    assert info["word_bytes"] == 2 and cnt > n_tar


def test_single_neuron_self_synapse():
    _say("case: one LIFB neuron above threshold with an excitatory self-synapse")
    cnt, info = _arrays_vs_oracle([M.LIFB], [0.0], W.lifb_consts(), [0], [0], [np.float32(3e-2)],
                                  [np.float32(0.0025)], 400, v0=[0.025])
    assert cnt > 10


def test_batch_with_empty_and_poisson_only_networks():
    _say("case: batch of a sweep network, a network with no neurons and a Poisson-only network")
    _, nets = W.sweep_nets(n_nets=1, n_lo=2000, n_hi=2000, size_seed=1)
    empty = dict(pops=[dict(num=0, type=N.LIFB)], projs=[], lifconst=W.lifb_consts(), step=W.DT, dt=W.DT, seed=5)
    pois = dict(pops=[dict(num=300, type=N.POISSON, rate=0.02)], projs=[], lifconst=W.lifb_consts(), step=W.DT,
                dt=W.DT, seed=6)
    b = N.Simulator.batch_generated([nets[0], empty, pois], record_capacity=1_000_000)
    b.run(900)
    cb = b.fire_counts()
    n0 = W.net_size(nets[0])
    s1 = N.Simulator.generated(**nets[0])
    s1.run(900)
    s3 = N.Simulator.generated(**pois)
    s3.run(900)
    assert np.array_equal(cb[:n0], s1.fire_counts()) and np.array_equal(cb[n0:], s3.fire_counts())
    assert cb[n0:].sum() > 1000
    for x in (b, s1, s3):
        x.close()


@pytest.mark.parametrize("wf", [N.WF_CLASS, N.WF_BF16])
def test_tiles_partition_the_neurons(wf):
    """mgb_read_tiles: the tiles cover every neuron once, in order, none wider than
    the widest tile reported, none straddling two networks of a batch or mixing
    neuron types; the direct path has none."""
    _params, nets = W.sweep_nets(n_nets=5, n_lo=1500, n_hi=5000, size_seed=19)
    _say(f"case: tiles of a batch of {len(nets)} sweep networks, weight_format={wf}")
    b = N.Simulator.batch_generated(nets, weight_format=wf, propagation=N.PROP_TILED, tile_size=200)
    info = b.info()
    t = b.tiles()
    assert len(t) == info["tiles"] + 1 and t[0] == 0 and t[-1] == info["n_neurons"]
    widths = np.diff(t)
    assert widths.min() >= 1 and widths.max() == info["tile_size"]
    bounds = np.cumsum([0] + [W.net_size(x) for x in nets])
    types = np.concatenate([np.full(p["num"], p["type"], np.int32) for x in nets for p in x["pops"]])
    for lo, hi in zip(t[:-1], t[1:]):
        net = np.searchsorted(bounds, lo, side="right") - 1
        assert hi <= bounds[net + 1]                       # one network
        assert np.all(types[lo:hi] == types[lo])          # one neuron type
    b.close()
    d = H.gpu_generated(M.config1(scale=0.1), seed=2, propagation=N.PROP_DIRECT)
    with pytest.raises(N.MGBError) as e:
        d.tiles()
    assert e.value.status == N.ERR_UNSUPPORTED
    d.close()
