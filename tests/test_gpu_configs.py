"""The benchmark workloads at their stated sizes against the oracle.

Bar (per workload):
  * the device-built CSR, delay indices and weights equal the oracle's
    (bit for bit; for the 1e9-synapse network: every structural invariant of
    the whole CSR, and the exact in-rows of sampled targets, which the fixed
    in-degree generator draws target by target);
  * per-step spike index lists bit-exact for the first 1000 steps (100 ms),
    and per-population firing rates and CV of the inter-spike intervals within
    2 % over 1 s (10,000 steps) -- with the exact formats the whole second is
    bit-exact, so they are equal;
  * with STDP on the E->E synapses (the 1e9-synapse network, fused path):
    spikes and every weight bit-exact for 1000 steps;
  * the batched BF16 sweep against the FP64 oracle: per-population rates of
    every network within 2 % over 1 s, and the V_m error up to the first spike
    divergence;
  * configs 1, 2 and 4 against the FP64 oracle likewise (their exact run equals
    the GPU's bit for bit): the V_m error up to the first spike divergence below
    1 uV and per-population rates within 2 % (or within the counting noise of two
    realisations, where that is larger) over 1 s (config 4: over its 1000 steps).
Oracle runs without STDP use its multithreaded step loop (bit-identical to the
sequential one, tests/test_oracle_semantics.py).
"""
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

STEPS = 10_000          # 1 s at 0.1 ms
EXACT = 1_000           # 100 ms
THREADS = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pop_stats(steps, ids, bounds, nsteps, dt=1e-4):
    """Per population (id ranges bounds[k]..bounds[k+1]): mean rate (Hz) and the mean
    CV of the inter-spike intervals over neurons with at least three spikes."""
    out = []
    order = np.lexsort((steps, ids))
    sid, sst = ids[order], steps[order]
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        m = (sid >= lo) & (sid < hi)
        a, t = sid[m], sst[m]
        rate = len(a) / (hi - lo) / (nsteps * dt)
        cvs = []
        if len(a):
            cuts = np.flatnonzero(np.diff(a)) + 1
            for seg in np.split(t, cuts):
                if len(seg) >= 3:
                    isi = np.diff(seg).astype(np.float64)
                    cvs.append(isi.std() / isi.mean())
        out.append((rate, float(np.mean(cvs)) if cvs else 0.0))
    return out


def _assert_stats_close(got, want, tol=0.02):
    for (rg, cg), (rw, cw) in zip(got, want):
        assert abs(rg - rw) <= tol * max(rw, 1e-9), (rg, rw)
        if cw > 0:
            assert abs(cg - cw) <= tol * cw, (cg, cw)


def _flat(lists):
    st = np.concatenate([np.full(len(x), k, np.int32) for k, x in enumerate(lists)])
    ids = np.concatenate(lists).astype(np.int32)
    return st, ids


def _spikes_and_stats(g, o_lists, bounds):
    steps, ids = g.spikes()
    got = H.split_steps(steps, ids, STEPS)
    k = H.first_mismatch(got[:EXACT], o_lists[:EXACT])
    assert k < 0, f"spike lists differ first at step {k}"
    exact_all = H.first_mismatch(got, o_lists) < 0
    os_, oi = _flat(o_lists)
    sg, so = pop_stats(steps, ids, bounds, STEPS), pop_stats(os_, oi, bounds, STEPS)
    _say(f"  per population (rate Hz, CV-ISI): gpu {np.round(sg, 4).tolist()} oracle {np.round(so, 4).tolist()}; "
         f"all {STEPS} steps bit-exact: {exact_all}")
    _assert_stats_close(sg, so)
    return exact_all


def _fp64_report(label, make, bounds, nsteps=STEPS):
    """The exact run -- equal to the GPU's bit for bit, as the caller has checked --
    against the FP64 oracle from the same initial state: the V_m error up to the first
    spike divergence (10-step resolution) and each integrating population's rate over
    nsteps.  Asserted: V_m error below 1 uV (FP32 rounding only), rates within 2 % or
    within three standard deviations of the difference of two independent spike
    counts, whichever is larger (past the divergence the two runs are two realisations
    of the same chaotic network)."""
    o32, o64 = make(po.ACC_F32), make(po.ACC_F64)
    done, diverged, verr = 0, None, 0.0
    while done < nsteps and diverged is None:
        a = o32.run_mt(10, THREADS, record=True)
        b = o64.run_mt(10, THREADS, record=True)
        if H.first_mismatch(a, b) >= 0:
            diverged = done
        else:
            verr = max(verr, float(np.max(np.abs(o32.state(po.F_V).astype(np.float64) - o64.state(po.F_V)))))
        done += 10
    if nsteps > done:
        o32.run_mt(nsteps - done, THREADS)
        o64.run_mt(nsteps - done, THREADS)
    c32, c64 = o32.state(po.F_CNT), o64.state(po.F_CNT)
    rows = []
    for lo, hi in zip(bounds[1:-1], bounds[2:]):        # the integrating populations (after the Poisson one)
        n32, n64 = float(c32[lo:hi].sum()), float(c64[lo:hi].sum())
        err = abs(n32 - n64) / max(n64, 1.0)
        tol = max(0.02, 3.0 * np.sqrt(2.0 * max(n64, 1.0)) / max(n64, 1.0))
        rows.append((n64 / (hi - lo) / (nsteps * 1e-4), err, tol))
    _say(f"  {label} FP32 vs FP64: first spike divergence at step {diverged}, max |V32 - V64| before it "
         f"{verr * 1e3:.3g} mV; population rates (FP64 Hz, rate error, bound) "
         f"{[(round(r, 3), round(e * 100, 2), round(t * 100, 2)) for r, e, t in rows]}")
    assert verr < 1e-6
    for _r, e, t in rows:
        assert e <= t
    o32.close()
    o64.close()


def test_config1_brunel_10k_bernoulli_full_size():
    """8000 E + 2000 I LIFB + 1000 Poisson at p = 0.002 per step, Bernoulli p = 0.1,
    J = 0.1 mV, g = 5, 1.4 ms delays: ~1.1e7 synapses, generated on the device."""
    _say("case: config 1 at full size (Bernoulli p=0.1, 1.1e7 synapses), 10000 steps")
    net = M.config1(scale=1.0, p=0.1, poisson_rate=0.002)
    syn = H.host_synapses(net, seed=0)
    g = H.gpu_generated(net, seed=0, record_capacity=4_000_000)
    csr, o = H.oracle(net, syn, seed=0)
    rp, col, dl, w = g.csr()
    assert rp[-1] == syn.n and 1.05e7 < syn.n < 1.15e7
    assert np.array_equal(rp, csr.row_ptr) and np.array_equal(col, csr.col) and np.array_equal(dl, csr.dly)
    assert np.array_equal(w.view(np.uint32), csr.w.view(np.uint32))
    assert g.info()["ring_len"] == csr.D == 15 and g.info()["min_delay"] == 14
    del rp, col, dl, w
    g.run(STEPS)
    want = o.run_mt(STEPS, THREADS, record=True, cap=4_000_000)
    bounds = np.cumsum([0] + [p[0] for p in net.pops])
    assert _spikes_and_stats(g, want, bounds)
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert g.info()["syn_events"] == o.syn_events()
    g.close()
    o.close()
    _fp64_report("config 1", lambda acc: H.oracle(net, syn, seed=0, acc=acc)[1], bounds)


def test_config2_brunel_100k_fixed_indegree_full_size():
    """80k E + 20k I LIFB + 10k Poisson at 15 Hz, fixed in-degree K = 10,000, delays on
    the 1.0-2.0 ms grid: 1e9 synapses generated on the device.  The whole CSR is checked
    by its invariants and sampled targets' in-rows exactly; the oracle then runs on the
    read-back CSR."""
    _say("case: config 2 at full size (K=10000, 1e9 synapses), 10000 steps")
    net = M.config2(scale=1.0)
    g = H.gpu_generated(net, seed=0, record_capacity=4_000_000)
    info = g.info()
    n = info["n_neurons"]
    bounds = np.cumsum([0] + [p[0] for p in net.pops])
    nP, nE = net.pops[0][0], net.pops[1][0]
    assert n == 110_000 and info["n_synapses"] == 1_000_000_000
    assert info["ring_len"] == 21 and info["min_delay"] == 10
    rp, col, dl, w = g.csr()
    _say("  CSR read back")
    # invariants of the whole CSR
    assert rp[0] == 0 and rp[-1] == 1_000_000_000 and np.all(np.diff(rp) >= 0)
    indeg = np.bincount(col, minlength=n)
    assert np.all(indeg[:nP] == 0) and np.all(indeg[nP:] == 10_000)
    assert dl.min() == 10 and dl.max() == 20
    J, wi = np.float32(1e-4), np.float32(-5e-4)
    rows = np.diff(rp)
    inh_from = rp[nP + nE]                      # rows of I sources start here (rows in id order)
    assert np.all(w[:inh_from] == J) and np.all(w[inh_from:] == wi)
    del indeg
    # sampled targets: their synapses in CSR order equal the generator's draws in list order
    rng = np.random.default_rng(7)
    targets = np.sort(np.concatenate([[nP, nP + nE - 1, nP + nE, n - 1],
                                      rng.choice(np.arange(nP, n), 60, replace=False)])).astype(np.int32)
    mask = np.zeros(n, bool)
    mask[targets] = True
    slots = np.flatnonzero(mask[col])
    src_g = (np.searchsorted(rp, slots, side="right") - 1).astype(np.int32)
    g_seq = np.stack([src_g, col[slots], dl[slots], w[slots].view(np.int32)])
    o_src, o_tar, o_dl, o_w = [], [], [], []
    for j, pr in enumerate(net.projs):
        tlo, thi = bounds[pr["tar"]], bounds[pr["tar"] + 1]
        sel = targets[(targets >= tlo) & (targets < thi)]
        s_, w_, d_ = po.fixk_targets(0, j, net.pops[pr["src"]][0], pr["k"], pr["wmin"], pr["wmax"], pr["dmin"],
                                     pr["dmax"], sel - tlo)
        o_src.append((s_ + bounds[pr["src"]]).ravel())
        o_tar.append(np.repeat(sel, pr["k"]))
        o_dl.append(d_.ravel())
        o_w.append(w_.ravel())
    o_src, o_tar, o_dl, o_w = (np.concatenate(x) for x in (o_src, o_tar, o_dl, o_w))
    order = np.argsort(o_src, kind="stable")       # list order within each source row
    o_seq = np.stack([o_src[order], o_tar[order], o_dl[order], o_w[order].view(np.int32)])
    assert g_seq.shape == o_seq.shape == (4, targets.size * 10_000)
    assert np.array_equal(g_seq, o_seq)
    _say(f"  {targets.size} sampled targets' in-rows equal the generator's draws")
    del mask, slots, src_g, g_seq, o_seq, rows
    # dynamics against the oracle on the read-back CSR
    types, rate = net.neuron_arrays()
    cs = po.CsrArrays(n, rp, col, dl, w, np.zeros(1, np.uint8), np.zeros(1, np.int64), np.zeros(n + 1, np.int64),
                      np.zeros(1, np.int64), info["ring_len"], info["min_delay"])
    o = po.OracleSim(cs, types, rate, H.initial_v(net, 0), net.lifconst, seed=0, acc=po.ACC_F32)
    g.run(STEPS)
    want = o.run_mt(STEPS, THREADS, record=True, cap=4_000_000)
    assert _spikes_and_stats(g, want, bounds)
    assert np.array_equal(g.state(N.F_V).view(np.uint32), o.state(po.F_V).astype(np.float32).view(np.uint32))
    assert g.info()["syn_events"] == o.syn_events()
    g.close()
    o.close()
    v0 = H.initial_v(net, 0)
    _fp64_report("config 2", lambda acc: po.OracleSim(cs, types, rate, v0, net.lifconst, seed=0, acc=acc), bounds)


# the batched sweep's networks at the corners of the g x nu grid: the two smallest of
# the four networks of each corner (sizes from the sweep's fixed size seed)
CORNERS = ((3.0, 10.0), (3.0, 25.0), (6.0, 10.0), (6.0, 25.0))


def _corner_nets():
    params, nets = W.sweep_nets()
    pick = []
    for g_, nu in CORNERS:
        c = sorted((q["n"], b) for b, q in enumerate(params) if q["g"] == g_ and q["nu"] == nu)
        pick += [b for _, b in c[:2]]
    return [params[b] for b in pick], [nets[b] for b in pick]


def test_config5_sweep_corners_batched_bf16_against_fp64():
    params, nets = _corner_nets()
    _say(f"case: config 5 corners {[(q['g'], q['nu'], q['n']) for q in params]} batched BF16 vs FP64, 1 s")
    b = N.Simulator.batch_generated(nets, weight_format=N.WF_BF16, record_capacity=8_000_000)
    assert b.info()["segments"] == len(nets)
    v0_all = b.state(N.F_V)
    rp, col, dl, w = b.csr()
    # V_m error up to the first spike divergence, in 10-step chunks
    base, oracles = 0, []
    for q, net in zip(params, nets):
        sz = W.net_size(net)
        lo, hi = int(rp[base]), int(rp[base + sz])
        cs = po.CsrArrays(sz, rp[base:base + sz + 1] - lo, col[lo:hi] - base, dl[lo:hi], w[lo:hi],
                          np.zeros(1, np.uint8), np.zeros(1, np.int64), np.zeros(sz + 1, np.int64),
                          np.zeros(1, np.int64), int(dl[lo:hi].max()) + 1, int(dl[lo:hi].min()))
        types = np.concatenate([np.full(p["num"], p["type"], np.int32) for p in net["pops"]])
        rate = np.concatenate([np.full(p["num"], p.get("rate", 0.0), np.float32) for p in net["pops"]])
        oracles.append((base, sz, po.OracleSim(cs, types, rate, v0_all[base:base + sz], net["lifconst"],
                                               seed=net["seed"], acc=po.ACC_F64)))
        # the BF16 codes of these equal-weight rows are exact: the oracle runs the true network
        J, wi = np.float32(1e-4), np.float32(-q["g"] * 1e-4)
        assert np.all((w[lo:hi] == J) | (w[lo:hi] == wi))
        base += sz
    diverged = [None] * len(nets)
    verr = [0.0] * len(nets)
    done = 0
    while done < 3000 and any(d is None for d in diverged):
        b.run(10)
        st, ids = b.spikes()
        cur = H.split_steps(st, ids, 10, step0=done)
        vg = b.state(N.F_V)
        for k, (base, sz, o) in enumerate(oracles):
            ol = o.run_mt(10, THREADS, record=True)
            if diverged[k] is not None:
                continue
            gl = [x[(x >= base) & (x < base + sz)] - base for x in cur]
            if H.first_mismatch(gl, ol) >= 0:
                diverged[k] = done
            else:
                verr[k] = max(verr[k], float(np.max(np.abs(vg[base:base + sz].astype(np.float64) -
                                                            o.state(po.F_V)))))
        done += 10
    b.run(STEPS - done)
    cnt = b.fire_counts()
    for k, ((base, sz, o), q, net) in enumerate(zip(oracles, params, nets)):
        o.run_mt(STEPS - done, THREADS)
        oc = o.state(po.F_CNT)
        nP, nE = net["pops"][0]["num"], net["pops"][1]["num"]
        errs = []
        for lo_, hi_ in ((nP, nP + nE), (nP + nE, sz)):
            rg = cnt[base + lo_:base + hi_].sum() / (hi_ - lo_)
            ro = oc[lo_:hi_].sum() / (hi_ - lo_)
            errs.append(abs(rg - ro) / max(ro, 1e-9))
        _say(f"  net g={q['g']} nu={q['nu']} N={q['n']}: E/I rate {cnt[base + nP:base + nP + nE].mean():.2f}/"
             f"{cnt[base + nP + nE:base + sz].mean():.2f} Hz, rate error vs FP64 E {errs[0] * 100:.2f} % "
             f"I {errs[1] * 100:.2f} %, first spike divergence at step {diverged[k]}, "
             f"max |V - V_fp64| before it {verr[k] * 1e3:.3g} mV")
        assert verr[k] < 1e-6                     # FP32 rounding only: below 1 uV
        assert max(errs) <= 0.02, errs
        o.close()
    b.close()


def test_config4_brunel_100k_stdp_full_size():
    """Config 2 with pair-based STDP on the E->E synapses (A+ 0.01, A- 0.0105 x W_max,
    tau 20 ms, W in [0, 2J]) on the fused path with BF16 static weights (E->I, I->*,
    Poisson->*) and FP32 plastic ones: 1000 steps against the oracle in fixed point on
    the read-back CSR -- every spike and all 1e9 weights."""
    _say("case: config 4 at full size (1e9 synapses, 5.8e8 plastic), fused STDP, BF16 static, 1000 steps")
    net = M.config2(scale=1.0, stdp=True)
    sc = M.stdp_consts_config4(J=1e-4)
    g = H.gpu_generated(net, seed=0, stdp=True, stdpconst=sc, weight_format=N.WF_BF16,
                        propagation=N.PROP_TILED, record_capacity=4_000_000)
    info = g.info()
    n = info["n_neurons"]
    # (4 B words: BF16 static weights in the word, plastic ones beside it)
    assert info["propagation"] == N.PROP_TILED and info["window"] == 10 and info["word_bytes"] == 4
    nP, nE = net.pops[0][0], net.pops[1][0]
    rp, col, dl, w = g.csr()
    plastic = np.zeros(col.size, np.uint8)
    e0, e1 = int(rp[nP]), int(rp[nP + nE])
    plastic[e0:e1] = (col[e0:e1] >= nP) & (col[e0:e1] < nP + nE)
    assert 5.5e8 < int(plastic.sum()) < 6.1e8
    cs = po.csr_from_arrays(n, rp, col, dl, w, info["ring_len"], info["min_delay"], plastic=plastic)
    types, rate = net.neuron_arrays()
    o = po.OracleSim(cs, types, rate, H.initial_v(net, 0), net.lifconst, stdpconst=sc, stdp=True, seed=0,
                     acc=po.ACC_FIXED, scale_exc=info["fix_scale_exc"], scale_inh=info["fix_scale_inh"])
    _say("  oracle built")
    g.run(EXACT)
    steps, ids = g.spikes()
    got = H.split_steps(steps, ids, EXACT)
    want = o.run(EXACT)
    k = H.first_mismatch(got, want)
    assert k < 0, f"spike lists differ first at step {k}"
    assert sum(len(x) for x in want) > 20_000
    wg = g.weights()
    wo = o.weights().astype(np.float32)
    assert np.array_equal(wg.view(np.uint32), wo.view(np.uint32))
    changed = int(np.count_nonzero(wg != w))
    _say(f"  {sum(len(x) for x in want)} spikes bit-exact; all weights bit-exact, {changed} changed by STDP")
    assert changed > 1_000_000
    g.close()
    o.close()
    del wg, wo
    # against FP64: the exact run (the fixed-point oracle, equal to the GPU's bit for bit
    # above) and the FP64 oracle from the same state, to the first spike divergence and
    # over the 1000 steps
    ofx = po.OracleSim(cs, types, rate, H.initial_v(net, 0), net.lifconst, stdpconst=sc, stdp=True, seed=0,
                       acc=po.ACC_FIXED, scale_exc=info["fix_scale_exc"], scale_inh=info["fix_scale_inh"])
    o64 = po.OracleSim(cs, types, rate, H.initial_v(net, 0), net.lifconst, stdpconst=sc, stdp=True, seed=0,
                       acc=po.ACC_F64)
    done, diverged, verr = 0, None, 0.0
    while done < EXACT and diverged is None:
        a, b = ofx.run(10), o64.run(10)
        if H.first_mismatch(a, b) >= 0:
            diverged = done
        else:
            verr = max(verr, float(np.max(np.abs(ofx.state(po.F_V).astype(np.float64) - o64.state(po.F_V)))))
        done += 10
    if EXACT > done:
        ofx.run(EXACT - done)
        o64.run(EXACT - done)
    cf, c64 = ofx.state(po.F_CNT), o64.state(po.F_CNT)
    rows = []
    for lo, hi in zip(bounds_of(net)[1:-1], bounds_of(net)[2:]):
        nf, n64 = float(cf[lo:hi].sum()), float(c64[lo:hi].sum())
        rows.append((n64 / (hi - lo) / (EXACT * 1e-4), abs(nf - n64) / max(n64, 1.0),
                     max(0.02, 3.0 * np.sqrt(2.0 * max(n64, 1.0)) / max(n64, 1.0))))
    _say(f"  config 4 exact vs FP64: first spike divergence at step {diverged}, max |V - V64| before it "
         f"{verr * 1e3:.3g} mV; rates over "
         f"{EXACT} steps (FP64 Hz, rate error %, bound %) "
         f"{[(round(r, 3), round(e * 100, 2), round(t * 100, 2)) for r, e, t in rows]}")
    assert verr < 1e-6
    for _r, e, t in rows:
        assert e <= t
    ofx.close()
    o64.close()


def bounds_of(net):
    return np.cumsum([0] + [p[0] for p in net.pops])
