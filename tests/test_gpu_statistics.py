"""Whole-network firing rates of the HIP path against the reference's
published single-GPU runs (benchdata, make_brunel, 1 s simulated).

The reference's runs were clock-seeded (network and Poisson streams), so only
statistics can be compared: mean rate over all neurons within 2 %.
Sources: benchdata/single/mgbrain_brunel_10000000_nvprof.txt:28 (27.1845 Hz),
benchdata/single2/mgbrain_single_brunel_0.01_20000000_nvprof.txt:11 (156.988),
..._0.05_20000000_nvprof.txt:11 (450.556), ..._0.1_20000000_nvprof.txt:11
(718.904), ..._0.01_100000000_nvprof.txt:11 (163.416).

Beyond the bit-exact window, per-population rates and the CV of the
inter-spike intervals over 1 s are compared with the FP64 oracle (2 %).
"""
import sys

import numpy as np
import pytest

from mgbrain_amd import analysis as A
from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu

CASES = [(10_000_000, 0.002, 27.1845), (20_000_000, 0.01, 156.988), (20_000_000, 0.05, 450.556),
         (20_000_000, 0.1, 718.904), (100_000_000, 0.01, 163.416)]


@pytest.mark.parametrize("nsyn,p,ref", CASES)
def test_brunel_rate_against_benchdata(nsyn, p, ref):
    print(f"case: make_brunel({nsyn}) Poisson p={p}, 10000 steps", flush=True)
    sys.stdout.flush()
    net = M.brunel_ref(nsyn, rate=p)
    syn = H.host_synapses(net, mt_seed=12345)
    g = H.gpu_from_list(net, syn, seed=3)
    g.run(10000)
    cnt = g.fire_counts()
    g.close()
    rate = cnt.sum() / net.n / 1.0
    P = net.pops[0][0]
    assert abs(cnt[:P].mean() / (p / 1e-4) - 1.0) < 0.02       # Poisson inputs at p / dt
    assert abs(rate - ref) / ref < 0.02, (rate, ref)


def _rates_cv(steps, ids, bounds, seconds):
    rates = A.population_rates(steps, ids, bounds, seconds)
    return [(r, c, k) for r, (c, k) in zip(rates, A.cv_isi(steps, ids, bounds))]


def test_population_rates_and_cv_isi_against_fp64():
    """After the bit-exact window: per-population rates and CV of ISI of the FP32
    HIP path over 1 s against the FP64 oracle on the same network, within 2 %."""
    print("case: make_brunel(1e7) Poisson p=0.002, 10000 steps, FP32 GPU vs FP64 oracle", flush=True)
    net = M.brunel_ref(10_000_000, rate=0.002)
    syn = H.host_synapses(net, mt_seed=4242)
    g = H.gpu_from_list(net, syn, seed=5, record_capacity=4_000_000)
    g.run(10000)
    gs, gi = g.spikes()
    g.close()
    _, o = H.oracle(net, syn, seed=5, acc=po.ACC_F64)
    want = o.run(10000, cap=4_000_000)
    o.close()
    os_ = np.concatenate([np.full(len(x), k, np.int32) for k, x in enumerate(want)])
    oi = np.concatenate(want).astype(np.int32)
    first = net.first()
    bounds = [(first[k], first[k] + p[0]) for k, p in enumerate(net.pops)]
    got, ref = _rates_cv(gs, gi, bounds, 1.0), _rates_cv(os_, oi, bounds, 1.0)
    for (r1, c1, n1), (r2, c2, n2), (num, typ, _src) in zip(got, ref, net.pops):
        print(f"  pop n={num} type={typ}: rate {r1:.3f} vs {r2:.3f} Hz, CV-ISI {c1:.4f} vs {c2:.4f}", flush=True)
        assert abs(r1 - r2) <= 0.02 * r2, (r1, r2)
        assert abs(c1 - c2) <= 0.02 * c2, (c1, c2)
