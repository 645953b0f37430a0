"""The C++ facade end to end (Model -> NetGen -> MGSimulator -> C ABI -> HIP)
through mgbrain_amd/lib/mgb_run_model, against the oracle."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import harness as H
from tests import netmodels as M
from tests import refspec

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "mgbrain_amd", "lib", "mgb_run_model")


def run_model(spec):
    with tempfile.TemporaryDirectory() as d:
        sp, out = os.path.join(d, "spec.txt"), os.path.join(d, "out.bin")
        open(sp, "w").write("\n".join(x.replace("{dir}", d) for x in spec) + "\n")
        r = subprocess.run([PROG, sp, out], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(out, "rb").read()
    hdr = np.frombuffer(raw, np.int64, 4)
    n, s, ns, se = (int(x) for x in hdr)
    o = 32

    def take(dt, c):
        nonlocal o
        a = np.frombuffer(raw, dt, c, o).copy()
        o += a.nbytes
        return a

    res = dict(n=n, se=se, src=take(np.int32, s), tar=take(np.int32, s), w=take(np.float32, s),
               d=take(np.float32, s), pl=take(np.uint8, s), steps=take(np.int32, ns), ids=take(np.int32, ns),
               v=take(np.float32, n), cnt=take(np.int32, n))
    res["v0"] = take(np.float32, len(res["v"]) if s else 0)
    res["rate"] = take(np.float32, len(res["v"]) if s else 0)
    return res


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def test_facade_brunel_reference_generator_vs_oracle():
    _say("case: facade make_brunel(2e5) host NetGen, 1000 steps")
    spec = ["timestep 0.0001", "clock 4242", "srand 1", "seed 7", "brunel 200000", "rate 0 0.01", "steps 1000",
            "record 2000000"]
    r = run_model(spec)
    syn, t = refspec.oracle_synapses(["timestep 0.0001", "clock 4242", "srand 1", "brunel 200000"])
    assert np.array_equal(r["src"], syn.src) and np.array_equal(r["tar"], syn.tar)
    assert np.array_equal(r["w"], syn.weight) and np.array_equal(r["d"], syn.delay)
    net = M.brunel_ref(200_000, rate=0.01)
    _, o = H.oracle(net, syn, seed=7)
    want = o.run(1000)
    got = H.split_steps(r["steps"], r["ids"], 1000)
    assert H.first_mismatch(got, want) < 0
    assert np.array_equal(r["v"].view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert r["se"] == o.syn_events()


def test_facade_device_generation_equals_host_generation():
    _say("case: facade Bernoulli + fixed in-degree, device-generated vs host NetGen")
    base = ["timestep 0.0001", "clock 99", "seed 5", "lifconst 0 0 0.01 0 0 0.002 1 1 0.02 0",
            "pop 300 2 1", "pop 800 5 0", "pop 200 5 0", "rate 0 0.2",
            "bern 0 1 0.0001 0.0001 0.0014 0.0014 0.1", "bern 0 2 0.0001 0.0001 0.0014 0.0014 0.1",
            "fixk 1 1 0.0001 0.0001 0.001 0.002 80", "fixk 1 2 0.0001 0.0001 0.001 0.002 80",
            "fixk 2 1 -0.0005 -0.0005 0.001 0.002 20", "bern 2 2 -0.0005 -0.0005 0.0014 0.0014 0.1",
            "vinit 1 0.0 0.015", "steps 800", "record 2000000"]
    host = run_model(base + ["device_gen 0"])
    dev = run_model(base + ["device_gen 1"])
    assert np.array_equal(host["steps"], dev["steps"]) and np.array_equal(host["ids"], dev["ids"])
    assert np.array_equal(host["v"].view(np.uint32), dev["v"].view(np.uint32))
    assert len(host["ids"]) > 1000
    # and against the oracle
    n = 1300
    L = po.SynapseList(host["src"], host["tar"], host["w"], host["d"], host["pl"])
    csr = po.build_csr(n, L)
    types = np.array([2] * 300 + [5] * 1000, np.int32)
    lc = M.gen_life_const(0, 0, 0.01, 0, 0, 0.002, 1, 1, 0.02, 0)
    o = po.OracleSim(csr, types, host["rate"], host["v0"], lc, seed=5)
    assert H.first_mismatch(H.split_steps(host["steps"], host["ids"], 800), o.run(800)) < 0


def test_facade_per_projection_stdp():
    _say("case: facade STDP on E->E only (Config::enable_stdp_per_projection)")
    spec = ["timestep 0.0001", "clock 3", "seed 2", "srand 1", "lifconst2", "stdp", "stdp_per_projection",
            "stdpconst 2e-6 -2.1e-6 200 200 2e-4 0", "pop 200 2 1", "pop 400 5 0", "pop 100 5 0", "rate 0 0.3",
            "bern 0 1 1e-4 1e-4 0.001 0.001 0.2", "bern 1 1 1e-4 1e-4 0.0012 0.0012 0.1 1",
            "bern 2 1 -5e-4 -5e-4 0.001 0.001 0.2", "bern 1 2 1e-4 1e-4 0.001 0.001 0.1",
            "steps 600", "record 2000000"]
    r = run_model(spec)
    assert r["pl"].sum() > 0 and r["pl"].sum() < len(r["pl"])
    n = 700
    L = po.SynapseList(r["src"], r["tar"], r["w"], r["d"], r["pl"])
    csr = po.build_csr(n, L)
    types = np.array([2] * 200 + [5] * 500, np.int32)
    sc = np.array([2e-6, -2.1e-6, 200, 200, 2e-4, 0], np.float32)
    # the facade chooses the fixed-point scale from the network; recompute it the same way
    from mgbrain_amd import native as N
    g = N.Simulator.from_arrays(types, r["rate"], M.life_const2(), r["src"], r["tar"], r["w"], r["d"],
                                stdp=True, plastic=r["pl"], stdpconst=sc, seed=2)
    info = g.info()
    g.close()
    o = po.OracleSim(csr, types, r["rate"], r["v0"], M.life_const2(), stdpconst=sc, stdp=True, seed=2,
                     acc=po.ACC_FIXED, scale_exc=info["fix_scale_exc"], scale_inh=info["fix_scale_inh"])
    assert H.first_mismatch(H.split_steps(r["steps"], r["ids"], 600), o.run(600)) < 0
    assert np.array_equal(r["v"].view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("device_gen", [0, 1])
def test_facade_checkpoint_resume(device_gen):
    _say(f"case: facade save after 333 steps, finish in a new MGSimulator loaded from the file, device_gen={device_gen}")
    base = ["timestep 0.0001", "clock 7", "seed 3", "lifconst 0 0 0.01 0 0 0.002 1 1 0.02 0",
            "pop 300 2 1", "pop 800 5 0", "pop 200 5 0", "rate 0 0.2",
            "bern 0 1 0.0001 0.0001 0.0014 0.0014 0.1", "bern 0 2 0.0001 0.0001 0.0014 0.0014 0.1",
            "fixk 1 1 0.0001 0.0001 0.001 0.002 80", "fixk 1 2 0.0001 0.0001 0.001 0.002 80",
            "fixk 2 1 -0.0005 -0.0005 0.001 0.002 20", "bern 2 2 -0.0005 -0.0005 0.0014 0.0014 0.1",
            "steps 900", "record 2000000", f"device_gen {device_gen}"]
    whole = run_model(base)
    resumed = run_model(base + ["checkpoint 333 {dir}/ck.bin"])
    assert len(whole["ids"]) > 1000
    assert np.array_equal(whole["steps"], resumed["steps"]) and np.array_equal(whole["ids"], resumed["ids"])
    assert np.array_equal(whole["v"].view(np.uint32), resumed["v"].view(np.uint32))
    assert whole["se"] == resumed["se"]
