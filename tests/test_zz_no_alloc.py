"""mgb_read_weights_noalloc against mgb_read_weights, through the C ABI.

The no-allocation read-back works in one fixed device array per GPU (2^24
floats).  Calls in a row on one GPU, each on a smaller network than the one
before (so every call reuses what the previous one left in the array), cover
each way the weights are produced: fused STDP (pending updates applied in the
array), BF16 static + FP32 plastic and FP8 (codes decoded in the array), the
CLASS format (class tables decoded in the array) and FP32 (no array use).
Each result must equal mgb_read_weights bit for bit.  A network larger than
the array is refused with MGB_ERR_UNSUPPORTED.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from mgbrain_amd import native as N
from tests import harness as H
from tests import netmodels as M

pytestmark = pytest.mark.gpu


def _both(g):
    """Weights from both entry points, called directly through the C ABI."""
    s = g.info()["n_synapses"]
    a = np.zeros(max(s, 1), np.float32)
    b = np.full(max(s, 1), np.nan, np.float32)
    L = N.lib()
    st_b = L.mgb_read_weights_noalloc(g.h, b.ctypes.data_as(C.c_void_p))
    assert st_b == N.OK, L.mgb_last_error().decode()
    st_a = L.mgb_read_weights(g.h, a.ctypes.data_as(C.c_void_p))
    assert st_a == N.OK, L.mgb_last_error().decode()
    return a[:s], b[:s]


def _same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _stdp_net(scale, K):
    net = M.config2(scale=scale, K=K, poisson_rate=0.3, stdp=True)
    syn = H.host_synapses(net, seed=6)
    return net, syn, M.stdp_consts_config4(J=1e-4)


def test_noalloc_weights_equal_read_weights_on_calls_in_a_row():
    sizes = []
    # 1. fused STDP (FP32 weights in tile order; the read applies the pending window)
    print("case: fused STDP, config2(scale=0.03, K=300)", flush=True)
    net, syn, sc = _stdp_net(0.03, 300)
    g = H.gpu_from_list(net, syn, seed=2, stdp=True, stdpconst=sc, weight_format=N.WF_F32,
                        propagation=N.PROP_TILED)
    assert g.info()["propagation"] == N.PROP_TILED and g.info()["window"] > 1
    w0 = g.weights()
    g.run(400)
    a, b = _both(g)
    assert np.sum(a != w0) > 100            # STDP changed weights
    _same(a, b)
    g.run(250)                              # the read left the run's state alone
    a2, b2 = _both(g)
    _same(a2, b2)
    assert not np.array_equal(a2, a)
    sizes.append(syn.n)
    g.close()

    # 2. BF16 static + FP32 plastic (direct path), smaller
    print("case: BF16 static + plastic, config2(scale=0.02, K=200)", flush=True)
    net, syn, sc = _stdp_net(0.02, 200)
    g = H.gpu_from_list(net, syn, seed=3, stdp=True, stdpconst=sc, weight_format=N.WF_BF16,
                        propagation=N.PROP_DIRECT)
    g.run(200)
    a, b = _both(g)
    _same(a, b)
    assert syn.n < sizes[-1]
    sizes.append(syn.n)
    g.close()

    # 3. FP8, smaller
    print("case: FP8, config3(scale=0.05)", flush=True)
    net = M.config3(scale=0.05)
    syn = H.host_synapses(net, seed=4)
    g = H.gpu_from_list(net, syn, seed=4, weight_format=N.WF_FP8)
    g.run(100)
    a, b = _both(g)
    _same(a, b)
    assert np.any(a < 0) and np.any(a > 0)
    assert syn.n < sizes[-1]
    sizes.append(syn.n)
    g.close()

    # 4. CLASS format, smaller
    print("case: CLASS, brunel_ref(20000)", flush=True)
    net = M.brunel_ref(20_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=7)
    g = H.gpu_from_list(net, syn, seed=5, weight_format=N.WF_CLASS)
    g.run(50)
    a, b = _both(g)
    _same(a, b)
    _same(b, syn_weights_in_csr_order(g))
    assert syn.n < sizes[-1]
    sizes.append(syn.n)
    g.close()

    # 5. FP32 (read straight from the CSR weights), smaller
    print("case: FP32, brunel_ref(5000)", flush=True)
    net = M.brunel_ref(5_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=8)
    g = H.gpu_from_list(net, syn, seed=6, weight_format=N.WF_F32)
    a, b = _both(g)
    _same(a, b)
    assert syn.n < sizes[-1]
    g.close()
    sys.stdout.flush()


def syn_weights_in_csr_order(g):
    _rp, _col, _dl, w = g.csr()
    return w


def test_noalloc_refuses_networks_beyond_its_array():
    print("case: BF16 network of 2.2e7 synapses (above the 2^24-float array)", flush=True)
    net = M.config2(scale=0.2, K=1000)
    g = H.gpu_generated(net, seed=1, weight_format=N.WF_BF16)
    s = g.info()["n_synapses"]
    assert s > (1 << 24)
    out = np.zeros(s, np.float32)
    L = N.lib()
    st = L.mgb_read_weights_noalloc(g.h, out.ctypes.data_as(C.c_void_p))
    assert st == N.ERR_UNSUPPORTED
    assert b"scratch" in L.mgb_last_error()
    # the handle still works, and a smaller network afterwards still reads correctly
    g.run(20)
    g.close()
    net = M.brunel_ref(20_000, rate=0.01)
    syn = H.host_synapses(net, mt_seed=9)
    g = H.gpu_from_list(net, syn, seed=7, weight_format=N.WF_BF16)
    a, b = _both(g)
    _same(a, b)
    g.close()
