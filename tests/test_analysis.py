"""Spike-record file format and statistics (CPU): round trip through .npz and
closed forms for rates and CV of the inter-spike intervals."""
import numpy as np

from mgbrain_amd import analysis as A


def test_spike_record_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    steps = rng.integers(0, 1000, 5000).astype(np.int32)
    ids = rng.integers(0, 300, 5000).astype(np.int32)
    p = tmp_path / "rec.npz"
    A.save_spikes(p, steps, ids, 300, 1e-4, pop_bounds=[(0, 100), (100, 300)])
    r = A.load_spikes(p)
    order = np.lexsort((ids, steps))
    assert np.array_equal(r["steps"], steps[order]) and np.array_equal(r["ids"], ids[order])
    assert r["n_neurons"] == 300 and r["dt"] == 1e-4
    assert np.array_equal(r["pop_bounds"], [[0, 100], [100, 300]])


def test_rates_and_cv_closed_forms():
    # neuron 0: regular every 10 steps (CV 0); neuron 1: intervals 5, 15 alternating
    s0 = np.arange(0, 1000, 10)
    s1 = np.r_[0, np.cumsum(np.tile([5, 15], 50))]   # 50 intervals of 5 and 50 of 15
    steps = np.r_[s0, s1]
    ids = np.r_[np.zeros_like(s0), np.ones_like(s1)]
    (r0, r1) = A.population_rates(steps, ids, [(0, 1), (1, 2)], 0.1)
    assert r0 == len(s0) / 0.1 and r1 == len(s1) / 0.1
    (c0, n0), (c1, n1) = A.cv_isi(steps, ids, [(0, 1), (1, 2)])
    assert c0 == 0.0 and n0 == 1
    assert abs(c1 - 0.5) < 1e-12 and n1 == 1          # std 5 / mean 10
    (cb, nb), = A.cv_isi(steps, ids, [(0, 2)])
    assert abs(cb - 0.25) < 1e-12 and nb == 2


def test_weight_histogram():
    c, e = A.weight_histogram(np.array([0.0, 0.5, 1.0, 1.0]), bins=2, lo=0.0, hi=1.0)
    assert list(c) == [1, 3] and list(e) == [0.0, 0.5, 1.0]
