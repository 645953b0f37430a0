"""Whole-network firing rates of the HIP path against the reference's
published single-GPU runs (benchdata, make_brunel, 1 s simulated).

The reference's runs were clock-seeded (network and Poisson streams), so only
statistics can be compared: mean rate over all neurons within 2 %.
Sources: benchdata/single/mgbrain_brunel_10000000_nvprof.txt:28 (27.1845 Hz),
benchdata/single2/mgbrain_single_brunel_0.01_20000000_nvprof.txt:11 (156.988),
..._0.05_20000000_nvprof.txt:11 (450.556), ..._0.1_20000000_nvprof.txt:11
(718.904), ..._0.01_100000000_nvprof.txt:11 (163.416).
"""
import sys

import pytest

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
