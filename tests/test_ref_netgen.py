"""The oracle's network generator against the reference's own NetGen.

Fixtures (tests/golden/ref_netgen.json) were written by
tests/golden/make_ref_fixtures.py from oracle/_ref/ref_netgen, which is the
reference's model layer (macro.cpp, model/model.cpp, model/netgen.cpp)
compiled from its sources.  When that binary is present the comparison also
runs live on freshly drawn specs.
"""
import hashlib
import json
import os
import random
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import netmodels as M
from tests import refspec

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_netgen.json")))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_generator_matches_reference_fixture(name):
    fx = GOLDEN[name]
    syn, t = refspec.oracle_synapses(fx["spec"])
    assert sum(p[0] for p in t["pops"]) == fx["n"]
    assert syn.n == fx["n_syn"]
    assert _digest(syn.src) == fx["src"]
    assert _digest(syn.tar) == fx["tar"]
    assert _digest(syn.weight) == fx["weight"]
    assert _digest(syn.delay) == fx["delay"]
    head = [[int(a), int(b), float(c), float(e)] for a, b, c, e in
            zip(syn.src[:8], syn.tar[:8], syn.weight[:8], syn.delay[:8])]
    assert head == fx["head"]
    # neuron out/in degrees (Neuron::nxt / pre sizes)
    n = fx["n"]
    assert _digest(np.bincount(syn.src, minlength=n).astype(np.int32)) == fx["nxt"]
    assert _digest(np.bincount(syn.tar, minlength=n).astype(np.int32)) == fx["pre"]


@pytest.mark.parametrize("name", ["brunel_200k_clock123456789012", "quota_tail_single", "quota_tail_chain_clock27",
                                  "quota_tail_chain_clock33"])
def test_quota_generator_counts_closed_form(name):
    """netgen.cpp:169-226: every quota projection emits exactly gen_num = round(sn*tn*p)
    synapses.  When the 1.1p acceptance runs out of pairs first, the zero entries left in
    genlist (:172) become synapses from the projection's first source to its first target,
    so that pair then occurs 1 + (gen_num - accepted) times (or gen_num - accepted times
    when the pair itself was not accepted)."""
    fx = GOLDEN[name]
    syn, t = refspec.oracle_synapses(fx["spec"])
    tails = 0
    for pr in t["projs"]:
        sn, tn = t["pops"][pr["src"]][0], t["pops"][pr["tar"]][0]
        gen_num = int(np.round(np.float32(sn * tn) * np.float32(pr["ctype"])))
        first = sum(p[0] for p in t["pops"][:pr["src"]])
        tfirst = sum(p[0] for p in t["pops"][:pr["tar"]])
        mask = (syn.src >= first) & (syn.src < first + sn) & (syn.tar >= tfirst) & (syn.tar < tfirst + tn)
        assert int(np.sum(mask)) == gen_num
        pairs = syn.src[mask].astype(np.int64) * (1 << 32) + syn.tar[mask]
        uniq, cnt = np.unique(pairs, return_counts=True)
        # only the (first source, first target) pair may repeat
        rep = uniq[cnt > 1]
        assert all(int(x) == first * (1 << 32) + tfirst for x in rep)
        tails += int(cnt[cnt > 1].sum()) if rep.size else 0
    if name.startswith("quota_tail"):
        assert tails > 0


def test_constgen_matches_reference():
    fx = GOLDEN["brunel_20k_clock777"]
    assert np.array_equal(M.life_const1(1e-4).view(np.uint32), np.array(fx["life1"], np.float32).view(np.uint32))
    assert np.array_equal(M.life_const2(1e-4).view(np.uint32), np.array(fx["life2"], np.float32).view(np.uint32))
    assert np.array_equal(np.array(fx["lifconst"], np.float32).view(np.uint32),
                          M.life_const2(1e-4).view(np.uint32))
    c3 = M.gen_life_const(0.0, 0.0, 0.0, 0.25, 10.0, 0.0004, 1.0, 1.0, 15.0, 0.0, dt=1e-4)
    assert np.array_equal(c3.view(np.uint32), np.array(fx["life3"], np.float32).view(np.uint32))
    fx2 = GOLDEN["timestep_1ms"]
    assert np.array_equal(M.life_const2(1e-3).view(np.uint32), np.array(fx2["life2"], np.float32).view(np.uint32))


def _compare_live(spec):
    with tempfile.TemporaryDirectory() as d:
        r = po.run_ref_netgen(spec, d)
    syn, _ = refspec.oracle_synapses(spec)
    assert r["src"].size == syn.n
    assert np.array_equal(r["src"], syn.src)
    assert np.array_equal(r["tar"], syn.tar)
    assert np.array_equal(r["weight"].view(np.uint32), syn.weight.view(np.uint32))
    assert np.array_equal(r["delay"].view(np.uint32), syn.delay.view(np.uint32))


@pytest.mark.skipif(not po.ref_netgen_available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_live_random_specs(seed):
    rng = random.Random(seed)
    spec = [f"timestep {rng.choice(['0.0001', '0.0002', '0.001'])}", "dw 0.005", f"clock {rng.randrange(1 << 40)}",
            f"srand {rng.randrange(1, 1000)}"]
    npop = rng.randrange(2, 5)
    sizes = [rng.randrange(20, 120) for _ in range(npop)]
    for s in sizes:
        spec.append(f"pop {s} {rng.choice([2, 5, 4])} 0")
    for _ in range(rng.randrange(2, 7)):
        s, t = rng.randrange(npop), rng.randrange(npop)
        w0 = rng.choice([0.001, -0.02, 0.05])
        w1 = w0 + rng.choice([0, 0.01, 0.03])
        d0 = rng.choice([0.0005, 0.001, 0.0016])
        d1 = d0 + rng.choice([0, 0.0005, 0.001])
        ct = rng.choice([0.05, 0.3, 0.7, 1.0, 0.0])
        spec.append(f"proj {s} {t} {w0} {w1} {d0} {d1} {ct}")
    _compare_live(spec)


@pytest.mark.skipif(not po.ref_netgen_available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("seed", range(24))
def test_live_small_quota_projections(seed):
    """Small populations at low p: the acceptance often runs out before gen_num, so the
    generator's zero-entry tail (netgen.cpp:172,210) and the rand() draws it takes from
    every later projection are exercised."""
    rng = random.Random(1000 + seed)
    spec = ["timestep 0.0001", f"dw {rng.choice(['0.001', '0.002', '0.005'])}", f"clock {rng.randrange(1, 1 << 32)}",
            f"srand {rng.randrange(1, 100)}"]
    npop = rng.randrange(2, 4)
    for _ in range(npop):
        spec.append(f"pop {rng.randrange(3, 16)} {rng.choice([2, 5, 4])} 0")
    for _ in range(rng.randrange(2, 5)):
        s, t = rng.randrange(npop), rng.randrange(npop)
        w0 = rng.choice([0.001, -0.02, 0.004])
        w1 = w0 + rng.choice([0, 0.004, 0.01])
        d0 = rng.choice([0.0005, 0.001])
        d1 = d0 + rng.choice([0, 0.0005, 0.001])
        spec.append(f"proj {s} {t} {w0} {w1} {d0} {d1} {rng.choice([0.05, 0.1, 0.2, 0.3])}")
    _compare_live(spec)


@pytest.mark.skipif(not po.ref_netgen_available(), reason="oracle/_ref not built (needs the reference sources)")
def test_live_tail_clock_sweep():
    """pop 10 / pop 10 at p = 0.2 over NetGen clocks 1-40 (8 of them run short of gen_num)."""
    for clock in range(1, 41):
        _compare_live(["timestep 0.0001", "dw 0.01", f"clock {clock}", "srand 1", "pop 10 2 1", "pop 10 5 0",
                       "proj 0 1 0.001 0.001 0.001 0.001 0.2"])
