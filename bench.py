#!/usr/bin/env python3
"""Single-GPU throughput of the MGBrain simulation step on MI355X.

Workload (default): Brunel network of 80,000 excitatory + 20,000 inhibitory
LIFB neurons and 10,000 Poisson inputs at 15 Hz, fixed in-degree K = 10,000
(1.0e9 synapses), J = 0.1 mV, g = 5, delays on the 1.0-2.0 ms grid, dt = 0.1 ms,
FP32 weights (exact per-class weights: a 4 B CSR word per synapse and a 2 B
tile-local word in the fused kernel's copy).  The network is generated on the
device from a seed; inputs are resident in HBM before the timed region.

A "step" is one simulation time step (0.1 ms of biological time) of the whole
network.  The timed region runs --steps steps after --settle + --warmup untimed
steps; a --long run beside it gives the steady-state figures and the roofline.
Prints one JSON line: synaptic events/s (value), realtime factor, the HBM
roofline of the dominant kernel (the fused window kernel), the CPU baseline
(the oracle's FP32 restatement of the reference step on the host's cores), and
the config-1, batched-sweep, conductance-network (config 3) and STDP (config 4)
legs, 1 s of simulated time each.  --config 1..5 runs one workload alone;
--tune runs the autotuner.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mgbrain_amd import native as N  # noqa: E402
from mgbrain_amd import tune as T  # noqa: E402
from mgbrain_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
DT = W.DT
lifb_consts = W.lifb_consts
brunel_fixed_k = W.brunel_fixed_k
brunel_bernoulli = W.brunel_bernoulli


def fused_roofline(b, info0, steps, se, spikes, n_lif, n_poisson, kname, lif_state=16.0, lif_layout=32.0):
    """HBM roofline of the fused kernel over the timed region.

    achieved = algorithmic bytes per launch / average launch duration, the bytes being the
    reference step's per-unit figures: per step 16 B of state per LIFB neuron (32 B LIFE,
    lif_state) and 4 B per Poisson neuron, 16 B ring read + clear per receiving neuron; per
    spike 8 B row pointers
    + 4 B list entry; per synaptic event 4 B target + 4 B FP32 weight + 1 B delay.
    layout_bytes_per_launch = what this layout itself must move: word_bytes per event,
    12 B per spike, 4 B per spike and tile of its segment, the ring rows live across
    windows in and out (each tile's slice is as wide as the tile: one column per neuron; none
    with persistent launches, whose slices stay in LDS), 32 B per LIFB (48 B LIFE, lif_layout) /
    20 B per Poisson neuron of state in and out.  With persistent launches a "launch" here is
    one window (the profiled spans count the windows of each launch)."""
    launches = b["prop_launches"]
    if not launches or b["prop_ms"] <= 0:
        return None
    avg_s = b["prop_ms"] / 1e3 / launches
    alg = steps * ((lif_state + 16.0) * n_lif + 4.0 * n_poisson) + 12.0 * spikes + 9.0 * se
    rw = 1 if b["weight_format"] == N.WF_CLASS else 2       # ring words per (slot, target)
    # (a persistent launch keeps the ring slices in LDS across its windows: they move once per
    # launch of tens of windows, not per window)
    live = 0 if b.get("persistent") else max(info0["ring_len"] - info0["window"], 0)
    g_seg = b["tiles"] / max(1, b["segments"])
    lay = (b["word_bytes"] * se + 12.0 * spikes + 4.0 * spikes * g_seg
           + launches * (8.0 * rw * live * info0["n_neurons"] + lif_layout * n_lif + 20.0 * n_poisson))
    ach = alg / launches / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
            "avg_launch_us": round(avg_s * 1e6, 3), "bytes_per_launch": round(alg / launches),
            "layout_bytes_per_launch": round(lay / launches),
            "layout_frac": round(lay / launches / avg_s / 1e9 / HBM_PEAK_GBS, 4),
            "launches": launches, "tile_size": b["tile_size"], "tiles": b["tiles"], "block": b["block"],
            "word_bytes": b["word_bytes"]}


def cpu_baseline(sim, budget_s=12.0, max_steps=4000, seed=0):
    """The oracle (FP32 restatement of the reference step) on the same network (CSR
    read back from the device), timed on this host: one thread (the checker's step
    loop), then all cores (its OpenMP step loop: neurons in parallel, fired rows in
    parallel with atomic ring adds), each on half the time budget from step 0."""
    from oracle import pyoracle as po  # CPU checker / baseline only

    inf = sim.info()
    n = inf["n_neurons"]
    t0 = time.perf_counter()
    rp, col, dl, w = sim.csr()
    types = sim_types = None
    types = sim._bench_types
    rate = sim._bench_rate
    v0 = np.zeros(n, dtype=np.float32)
    cs = po.CsrArrays(n, rp, col, dl, w, np.zeros(1, np.uint8), np.zeros(1, np.int64),
                      np.zeros(n + 1, np.int64), np.zeros(1, np.int64), inf["ring_len"], inf["min_delay"])
    o = po.OracleSim(cs, types, rate, v0, sim._bench_lifconst, seed=seed, acc=po.ACC_F32)
    setup = time.perf_counter() - t0
    del sim_types
    steps = 0
    t1 = time.perf_counter()
    while steps < max_steps and (time.perf_counter() - t1) < budget_s / 2:
        o.run(100, cap=n * 100)
        steps += 100
    el = time.perf_counter() - t1
    se = o.syn_events()
    o.close()
    host = f"host {platform.processor() or platform.machine()}, nproc={os.cpu_count()}"
    single = dict(value=se / el, cores=1, steps=steps, seconds=el, syn_events=se, rtf=el / (steps * DT))
    # all cores (the box's CPU share: OMP_NUM_THREADS when set)
    nt = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    o = po.OracleSim(cs, types, rate, v0, sim._bench_lifconst, seed=seed, acc=po.ACC_F32)
    steps_mt = 0
    t2 = time.perf_counter()
    while steps_mt < 4 * max_steps and (time.perf_counter() - t2) < budget_s / 2:
        o.run_mt(100, nt)
        steps_mt += 100
    el_mt = time.perf_counter() - t2
    se_mt = o.syn_events()
    o.close()
    return dict(value=se_mt / el_mt, unit="syn_events/s", cores=nt, kind="port",
                sample=f"oracle FP32 step loop with OpenMP ({nt} threads), steps 0..{steps_mt} "
                       f"({steps_mt * DT * 1e3:.0f} ms simulated) of the same network, {el_mt:.1f} s timed; "
                       f"one thread: steps 0..{steps}, {el:.1f} s timed (+{setup:.1f} s read-back/setup); {host}",
                steps=steps_mt, seconds=el_mt, syn_events=se_mt, rtf=el_mt / (steps_mt * DT),
                single_thread=single)


def batched_leg(args, shapes=None, log=None):
    """The batched sweep: 64 independent Brunel networks (sizes log-uniform on
    [2k, 50k] LIF neurons, g x Poisson rate grid, seeds 0-63, 1000 external Poisson
    synapses per neuron) advanced together by one simulator (mgb_create_batch_generated;
    every fused-kernel launch covers all networks), against the same 64 networks run one
    after another, each alone; then the same batch with BF16 weights, whose per-population
    rates over --bf16-window seconds are compared with the exact batch's, beside the spread
    between two exact batches that differ only in their Poisson keys."""
    log = log or (lambda m: None)
    params, nets = W.sweep_nets(args.nets)
    sizes = [W.net_size(x) for x in nets]
    shp = dict(tile_size=0, prop_block=0, chunk=0)
    if shapes:
        shp = dict(tile_size=shapes["tile_size"], prop_block=shapes["block"], chunk=shapes["chunk"])
    t0 = time.perf_counter()
    sim = N.Simulator.batch_generated(nets, use_graph=args.graph, profile=0 if args.no_profile else 1, **shp)
    sim.sync()
    setup_s = time.perf_counter() - t0
    info0 = sim.info()
    log(f"# batched: {len(nets)} networks, {info0['n_neurons']} neurons, {info0['n_synapses']} synapses, "
        f"{info0['segments']} segments, {info0['tiles']} tiles of {info0['tile_size']}, persistent "
        f"{info0['persistent']}, setup {setup_s:.2f} s, {info0['device_bytes'] / 2**30:.2f} GiB")
    sim.run(args.settle + args.warmup)
    sim.sync()
    a = sim.info()
    t1 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    wall = time.perf_counter() - t1
    b = sim.info()
    counts = sim.fire_counts()
    se = b["syn_events"] - a["syn_events"]
    spikes = b["spikes"] - a["spikes"]
    roof = None
    n_poisson = sum(p["num"] for x in nets for p in x["pops"] if p["type"] == N.POISSON)
    if b["propagation"] == N.PROP_TILED:
        roof = fused_roofline(b, info0, args.steps, se, spikes, info0["n_neurons"] - n_poisson, n_poisson,
                              "k_tiled (fused, batched segments)")
    log(f"# batched: {args.steps} steps in {wall * 1e3:.1f} ms, {se:.4g} events, {se / wall:.4g} SE/s, "
        f"RTF {wall / (args.steps * DT):.4f}")
    # the same networks one after another, each in its own simulator
    seq_wall, seq_se, identical = 0.0, 0, True
    base = 0
    for k, net in enumerate([] if args.no_sequential else nets):
        s1 = N.Simulator.generated(**net, use_graph=args.graph, profile=0)
        s1.run(args.settle + args.warmup)
        s1.sync()
        a1 = s1.info()
        t = time.perf_counter()
        s1.run(args.steps)
        s1.sync()
        seq_wall += time.perf_counter() - t
        seq_se += s1.info()["syn_events"] - a1["syn_events"]
        c1 = s1.fire_counts()
        s1.close()
        identical = identical and bool(np.array_equal(c1, counts[base:base + sizes[k]]))
        base += sizes[k]
    if args.no_sequential:
        seq_wall, seq_se, identical = None, None, None
    else:
        log(f"# sequential: {seq_se:.4g} events in {seq_wall * 1e3:.1f} ms, {seq_se / seq_wall:.4g} SE/s; "
            f"per-network fire counts identical: {identical}")
    bf16 = None
    if not args.no_batch_bf16:
        # the same batch with BF16 weights (per-source scales, fixed-point sums)
        table = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
        tb = json.load(open(table)).get(f"brunel_sweep{len(nets)}_bf16") if os.path.exists(table) else None
        shb = dict(tile_size=tb["tile_size"], prop_block=tb["block"], chunk=tb["chunk"]) if tb else {}
        sb = N.Simulator.batch_generated(nets, use_graph=args.graph, weight_format=N.WF_BF16, **shb)
        ib = sb.info()
        sb.run(args.settle + args.warmup)
        sb.sync()
        a2 = sb.info()
        t = time.perf_counter()
        sb.run(args.steps)
        sb.sync()
        wall2 = time.perf_counter() - t
        se2 = sb.info()["syn_events"] - a2["syn_events"]
        # rate error: every network's E and I rates over the next args.bf16_window seconds,
        # BF16 against exact; beside it the same error between the exact batch and an exact
        # batch whose Poisson inputs draw from other keys (same networks, same initial V):
        # the spread of a population's rate over input realisations, which bounds what a
        # comparison of two chaotic runs can resolve
        acc_steps = int(round(args.bf16_window / DT))
        alt_nets = [dict(x, seed=x["seed"] + 1000) for x in nets]
        sa = N.Simulator.batch_generated(alt_nets, use_graph=args.graph, **shp)
        sa.run(args.settle + args.warmup + args.steps)
        c_ex0, c_bf0, c_al0 = sim.fire_counts(), sb.fire_counts(), sa.fire_counts()
        sim.run(acc_steps)
        sb.run(acc_steps)
        sa.run(acc_steps)
        d_ex, d_bf, d_al = sim.fire_counts() - c_ex0, sb.fire_counts() - c_bf0, sa.fire_counts() - c_al0
        sb.close()
        sa.close()
        errs, floor, rates, base = [], [], [], 0
        for k, net in enumerate(nets):
            nP, nE = net["pops"][0]["num"], net["pops"][1]["num"]
            for lo, hi in ((nP, nP + nE), (nP + nE, sizes[k])):
                r_ex = float(d_ex[base + lo:base + hi].sum()) / (hi - lo)
                r_bf = float(d_bf[base + lo:base + hi].sum()) / (hi - lo)
                r_al = float(d_al[base + lo:base + hi].sum()) / (hi - lo)
                rates.append(r_ex / (acc_steps * DT))
                if r_ex > 0:
                    errs.append(abs(r_bf - r_ex) / r_ex)
                    floor.append(abs(r_al - r_ex) / r_ex)
            base += sizes[k]
        bf16 = {"value": se2 / wall2, "rtf": wall2 / (args.steps * DT), "tiles": ib["tiles"],
                "tile_size": ib["tile_size"], "word_bytes": ib["word_bytes"],
                "speedup_vs_sequential_exact": (se2 / wall2) / (seq_se / seq_wall) if seq_wall else None,
                "rate_error_window_s": acc_steps * DT,
                "max_pop_rate_err_vs_exact": max(errs) if errs else 0.0,
                "mean_pop_rate_err_vs_exact": float(np.mean(errs)) if errs else 0.0,
                "pops_within_2pct": [int(sum(e <= 0.02 for e in errs)), len(errs)],
                "realisation_spread": {"max": max(floor) if floor else 0.0,
                                       "mean": float(np.mean(floor)) if floor else 0.0,
                                       "what": "exact batch vs exact batch with other Poisson keys"},
                "pop_rates_hz_exact": [round(min(rates), 2), round(float(np.median(rates)), 2), round(max(rates), 2)]}
        log(f"# batched BF16: {se2 / wall2:.4g} SE/s, RTF {wall2 / (args.steps * DT):.4f}; population rates over "
            f"{acc_steps * DT:g} s: {min(rates):.1f}..{max(rates):.1f} Hz, error vs exact max "
            f"{bf16['max_pop_rate_err_vs_exact'] * 100:.2f} % mean {bf16['mean_pop_rate_err_vs_exact'] * 100:.3f} % "
            f"({bf16['pops_within_2pct'][0]}/{len(errs)} within 2 %); exact vs other Poisson keys max "
            f"{bf16['realisation_spread']['max'] * 100:.2f} % mean {bf16['realisation_spread']['mean'] * 100:.3f} %")
    sim.close()
    return {
        "workload": f"brunel_sweep{len(nets)}",
        "description": (f"{len(nets)} independent Brunel networks: N log-uniform on [2000, 50000] LIFB "
                        f"(80/20 E/I), recurrent Bernoulli p = min(0.1, 5000/N), {W.SWEEP_K_EXT} external "
                        f"Poisson synapses per neuron from max(2000, N/10) Poisson neurons, g in "
                        f"{list(W.SWEEP_G)}, Poisson rate in {list(W.SWEEP_NU_HZ)} Hz, J = 0.1 mV, delays "
                        f"U{{1.0..2.0}} ms, seeds 0-{len(nets) - 1}"),
        "value": se / wall, "unit": "syn_events/s", "rtf": wall / (args.steps * DT),
        "neurons": info0["n_neurons"], "synapses": info0["n_synapses"], "segments": info0["segments"],
        "weight_format": f"f32 per-class ({info0['word_bytes']} B synapse word; exact)", "graph": args.graph,
        "setup_s": round(setup_s, 2), "syn_events": se, "settle_steps": args.settle,
        "sequential_value": seq_se / seq_wall if seq_wall else None,
        "sequential_rtf": seq_wall / (args.steps * DT) if seq_wall else None,
        "speedup_vs_sequential": (se / wall) / (seq_se / seq_wall) if seq_wall else None,
        "fire_counts_identical_to_sequential": identical,
        "roofline": roof, "shapes": shapes, "bf16": bf16,
    }


def stdp_leg(args, log):
    """Brunel 100k (K = 10,000) with pair-based STDP on the E->E synapses, static weights
    (E->I, I->*, Poisson->*) in the format chosen by --precision (BF16 by default), plastic
    weights FP32.  --stdp-path fused (default): k_tiled<CH, kTStdpB> (kTStdp with
    --precision f32; "k_tiled_stdp" below) + k_stdp_flush, one launch each per window of
    min(min_delay, shortest inter-spike interval of E neurons) = 10 steps.  direct: a window
    of one step (k_neurons + k_prop + k_stdp per step)."""
    pops, projs, desc, ntot = W.brunel_fixed_k(args.scale, seed=args.seed, plastic_ee=True)
    sc, lc = W.stdp_consts(), W.lifb_consts()
    fused = args.stdp_path == "fused"
    wf = N.WF_BF16 if args.precision == "bf16" else N.WF_F32
    prop = N.PROP_TILED if fused else N.PROP_DIRECT
    t0 = time.perf_counter()
    shape = {}
    if fused:
        shape = dict(tile_size=args.tile, prop_block=args.block, chunk=args.chunk)
        table = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
        tuned = json.load(open(table)).get("brunel100k_K10000_stdp_ee") if os.path.exists(table) else None
        if tuned and args.scale == 1.0 and not (args.tile or args.block or args.chunk):
            shape = dict(tile_size=tuned["tile_size"], prop_block=tuned["block"], chunk=tuned["chunk"])
    sim = N.Simulator.generated(pops, projs, lc, step=DT, dt=DT, seed=args.seed, stdp=True, stdpconst=sc,
                                weight_format=wf, use_graph=args.graph, profile=0 if args.no_profile else 1,
                                propagation=prop, **shape)
    sim.sync()
    setup_s = time.perf_counter() - t0
    info0 = sim.info()
    log(f"# stdp: setup {setup_s:.2f} s, {info0['n_synapses']} synapses, window {info0['window']}, "
        f"tile {info0['tile_size']} block {info0['block']}, "
        f"{info0['device_bytes'] / 2**30:.2f} GiB")
    sim.run(args.warmup)
    sim.sync()
    a = sim.info()
    t1 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    wall = time.perf_counter() - t1
    b = sim.info()
    se = b["syn_events"] - a["syn_events"]
    spikes = b["spikes"] - a["spikes"]
    log(f"# stdp: {args.steps} steps in {wall:.2f} s, {se:.4g} events, {se / wall:.4g} SE/s, RTF "
        f"{wall / (args.steps * DT):.3f}")
    rp, col, _dl, w = sim.csr()
    sim.close()
    nP, nE = pops[0]["num"], pops[1]["num"]
    lo, hi = int(rp[nP]), int(rp[nP + nE])
    c, ww = col[lo:hi], w[lo:hi]
    wee = ww[(c >= nP) & (c < nP + nE)]
    del rp, col, w, c, ww
    hist, edges = np.histogram(wee, bins=20, range=(0.0, float(sc[4])))
    # per-kernel durations: a profiled pass (no graph replay) on a second simulator
    roof = None
    if not args.no_profile:
        prof = N.Simulator.generated(pops, projs, lc, step=DT, dt=DT, seed=args.seed, stdp=True, stdpconst=sc,
                                     weight_format=wf, use_graph=0, profile=1, propagation=prop, **shape)
        prof.run(500)
        pa = prof.info()
        prof.run(1000)
        pb = prof.info()
        prof.close()
        pse = pb["syn_events"] - pa["syn_events"]
        psp = pb["spikes"] - pa["spikes"]
        n_poisson = int(sum(p["num"] for p in pops if p["type"] == N.POISSON))
        n_lif = ntot - n_poisson
        plastic_share = len(wee) / max(1, info0["n_synapses"])
        # reference step's per-unit bytes (9 B per event with FP32 weights; STDP: 8 B per plastic
        # out-synapse of a spike, 16 B per plastic in-synapse, taken at the plastic share of
        # the spike's out-degree for both)
        k_bytes = {"k_neurons": 1000 * ((16.0 + 16.0) * n_lif + 4.0 * n_poisson) + 4.0 * psp,
                   "k_prop": 12.0 * psp + 9.0 * pse,
                   "k_stdp": (8.0 + 16.0) * plastic_share * pse}
        k_ms = {"k_neurons": (pb["neuron_ms"], pb["neuron_launches"]), "k_prop": (pb["prop_ms"], pb["prop_launches"]),
                "k_stdp": (pb["stdp_ms"], pb["stdp_launches"])}
        if fused:
            # the fused kernel and the flush of deferred dendrite updates after it do the
            # three kernels' work: the same per-unit bytes, summed, over both kernels' time
            k_bytes = {"k_tiled_stdp+k_stdp_flush": sum(k_bytes.values())}
            k_ms = {"k_tiled_stdp+k_stdp_flush": (pb["prop_ms"] + pb["stdp_ms"], pb["prop_launches"])}
            log(f"# stdp kernels: k_tiled_stdp {pb['prop_ms'] * 1e3 / max(1, pb['prop_launches']):.1f} us, "
                f"k_stdp_flush {pb['stdp_ms'] * 1e3 / max(1, pb['prop_launches']):.1f} us per window "
                f"({pb['stdp_launches']} launches in {pb['prop_launches']} windows)")
        kernels = {}
        for k, (ms, nl) in k_ms.items():
            if nl and ms > 0:
                ach = k_bytes[k] / (ms / 1e3) / 1e9
                kernels[k] = {"avg_launch_us": round(ms * 1e3 / nl, 3), "achieved": round(ach, 1),
                              "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_launch": round(k_bytes[k] / nl)}
        if kernels:
            dom = max(kernels, key=lambda k: k_ms[k][0])
            roof = dict(bound="hbm", achieved=kernels[dom]["achieved"], peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=kernels[dom]["frac"], traffic=None, kernel=dom, kernels=kernels,
                        avg_launch_us=kernels[dom]["avg_launch_us"])
            log(f"# stdp profile (1000 steps, no graph): " + ", ".join(
                f"{k} {v['avg_launch_us']:.1f} us {v['achieved']:.0f} GB/s" for k, v in kernels.items()))
    return {
        "workload": "brunel100k_K10000_stdp_ee", "description": desc + "; STDP on E->E (A+ 0.01, A- 0.0105 "
        "x W_max, tau 20 ms, W in [0, 0.2 mV])", "value": se / wall, "unit": "syn_events/s",
        "rtf": wall / (args.steps * DT), "simulated_s": args.steps * DT,
        "path": args.stdp_path, "window": info0["window"],
        "static_weights": "f32" if wf == N.WF_F32 else "bf16", "plastic_weights": "f32",
        "mean_rate_hz": spikes / ntot / (args.steps * DT), "setup_s": round(setup_s, 2),
        "ee_weights": {"n": int(wee.size), "mean": float(wee.mean()) if wee.size else 0.0,
                       "hist_edges": [float(x) for x in edges], "hist": [int(x) for x in hist]},
        "roofline": roof,
    }


def coba_leg(args, log):
    """Vogels-Abbott COBA (80k LIFE, p = 0.02, 1.28e8 synapses) in each weight storage:
    FP32 per-class (exact), FP32 fixed-point, BF16 and FP8-E4M3 with per-source scales.
    Population rates of each format are compared with the exact FP32 run of the same
    network; the error against the FP64 oracle is the test suite's
    (tests/test_gpu_precision.py)."""
    pops, projs, desc, ntot = W.vogels_abbott(args.scale, seed=args.seed)
    lc = W.life_const1()
    nE = pops[0]["num"]
    table = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
    tc = json.load(open(table)).get("vogels_abbott80k_p0.02") if os.path.exists(table) else None
    shc = dict(tile_size=tc["tile_size"], prop_block=tc["block"], chunk=tc["chunk"]) if tc and args.scale == 1.0 else {}
    if args.tile or args.block or args.chunk:          # explicit shapes win over the table
        shc = dict(tile_size=args.tile, prop_block=args.block, chunk=args.chunk)
    out, ref = {}, None
    for name, wf in (("f32_class", N.WF_CLASS), ("f32_fixed", N.WF_F32), ("bf16", N.WF_BF16), ("fp8", N.WF_FP8)):
        t0 = time.perf_counter()
        sim = N.Simulator.generated(pops, projs, lc, step=DT, dt=DT, seed=args.seed, weight_format=wf,
                                    use_graph=args.graph, profile=0 if args.no_profile else 1, **shc)
        sim.sync()
        setup_s = time.perf_counter() - t0
        info0 = sim.info()
        sim.run(args.warmup)
        sim.sync()
        a = sim.info()
        t1 = time.perf_counter()
        sim.run(args.steps)
        sim.sync()
        wall = time.perf_counter() - t1
        b = sim.info()
        cnt = sim.fire_counts()
        sim.close()
        se = b["syn_events"] - a["syn_events"]
        spikes = b["spikes"] - a["spikes"]
        secs = (args.warmup + args.steps) * DT
        rates = [float(cnt[:nE].sum()) / nE / secs, float(cnt[nE:].sum()) / (ntot - nE) / secs]
        if ref is None:
            ref = rates
        roof = None
        if b["propagation"] == N.PROP_TILED:
            roof = fused_roofline(b, info0, args.steps, se, spikes, ntot, 0, "k_tiled (" + name + ")",
                                  lif_state=32.0, lif_layout=48.0)
        out[name] = {"value": se / wall, "rtf": wall / (args.steps * DT), "word_bytes": b["word_bytes"],
                     "propagation": b["propagation"], "rate_e_hz": rates[0], "rate_i_hz": rates[1],
                     "rate_err_vs_f32": max(abs(x - y) / y for x, y in zip(rates, ref) if y > 0),
                     "device_gib": info0["device_bytes"] / 2**30, "setup_s": round(setup_s, 2), "roofline": roof}
        log(f"# coba {name}: {se / wall:.4g} SE/s, RTF {wall / (args.steps * DT):.4f}, rates E {rates[0]:.3f} "
            f"I {rates[1]:.3f} Hz, {b['word_bytes']} B words")
    return {"workload": "vogels_abbott80k_p0.02", "description": desc, "formats": out}


def with_incumbent(cands, workload):
    """Tile candidates plus the shapes table's current tile for the workload, so a
    re-tune is measured against what it would replace."""
    src = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
    cur = (json.load(open(src)) if os.path.exists(src) else {}).get(workload)
    out = list(cands)
    if cur and cur.get("tile_size", 0) > 0 and cur["tile_size"] not in out:
        out.append(cur["tile_size"])
    return out


def tune_batched(args, log, wf=N.WF_AUTO):
    """Autotunes the batched sweep's launch shapes (each tile size builds the batch anew)
    for the exact (CLASS) or the BF16 weight format."""
    _, nets = W.sweep_nets(args.nets)
    probe = N.Simulator.batch_generated(nets, weight_format=wf)
    pinfo = probe.info()
    probe.close()
    if pinfo["propagation"] != N.PROP_TILED:
        return None
    mw = 1024 if pinfo["word_bytes"] == 2 else 2048
    cands = sorted(set(T.tile_candidates(max(W.net_size(x) for x in nets), pinfo["ring_len"], max_width=mw)[:2]
                       + T.tile_candidates(pinfo["n_neurons"], pinfo["ring_len"], max_width=mw)[:3]))
    cands = with_incumbent(cands, f"brunel_sweep{args.nets}" + ("_bf16" if wf == N.WF_BF16 else ""))
    best, res = T.autotune(lambda b: N.Simulator.batch_generated(nets, tile_size=b, weight_format=wf), cands,
                           blocks=(512, 1024), chunks=(4, 8, 16), steps=2000, warmup=500,
                           log=lambda m: log("# " + m))
    return dict(best, candidates=len(res)) if best else None


def tune_coba(args, log):
    """Autotunes the Vogels-Abbott network's launch shapes (exact CLASS format; the
    other formats use the same shape)."""
    pops, projs, _desc, _ntot = W.vogels_abbott(args.scale, seed=args.seed)
    lc = W.life_const1()

    def make(tile):
        return N.Simulator.generated(pops, projs, lc, step=DT, dt=DT, seed=args.seed, weight_format=N.WF_CLASS,
                                     tile_size=tile)
    probe = make(0)
    pinfo = probe.info()
    probe.close()
    if pinfo["propagation"] != N.PROP_TILED:
        return None
    cands = with_incumbent(T.tile_candidates(pinfo["n_neurons"], pinfo["ring_len"], rw=2, max_width=1024),
                           "vogels_abbott80k_p0.02")
    best, res = T.autotune(make, cands, blocks=(256, 512, 1024), chunks=(2, 4, 8, 16), steps=2000, warmup=500,
                           log=lambda m: log("# " + m))
    return dict(best, candidates=len(res)) if best else None


def tune_stdp(args, log):
    """Autotunes the fused STDP path's launch shapes (static weights as --precision)."""
    pops, projs, _desc, _ntot = W.brunel_fixed_k(args.scale, seed=args.seed, plastic_ee=True)
    sc, lc = W.stdp_consts(), W.lifb_consts()

    def make(tile):
        return N.Simulator.generated(pops, projs, lc, step=DT, dt=DT, seed=args.seed, stdp=True, stdpconst=sc,
                                     weight_format=N.WF_BF16 if args.precision == "bf16" else N.WF_F32,
                                     propagation=N.PROP_TILED, tile_size=tile)
    probe = make(0)
    pinfo = probe.info()
    probe.close()
    cands = with_incumbent(T.tile_candidates(pinfo["n_neurons"], pinfo["ring_len"], max_width=2048),
                           "brunel100k_K10000_stdp_ee")
    best, res = T.autotune(make, cands, blocks=(512, 1024), chunks=(1, 2, 4, 8), steps=2000, warmup=500,
                           log=lambda m: log("# " + m))
    return dict(best, candidates=len(res)) if best else None


def write_tuned(path, entries):
    """Merges autotuned shapes {workload: best} into a copy of the shapes table."""
    src = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
    table = json.load(open(src)) if os.path.exists(src) else {}
    for k, v in entries.items():
        if v:
            table[k] = {"tile_size": v["tile_size"], "block": v["block"], "chunk": v["chunk"]}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=2)


def copy_probe(nbytes=1 << 30):
    """A device-to-device copy of exactly nbytes read + nbytes written per launch (torch's
    copy kernel), run under rocprofv3 --pmc beside the workload to scale its counters."""
    import torch  # plumbing only: device buffers and a copy kernel

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {nbytes} bytes x 3", flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default 10,000; 50,000 for --config 4)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", type=int, default=2, choices=(1, 2, 3, 4, 5),
                    help="1: Brunel 10k Bernoulli; 2: Brunel 100k K=10000 (default, plus the batched leg); "
                         "3: Vogels-Abbott COBA in four weight formats; 4: config 2 with STDP on E->E; "
                         "5: batched sweep only")
    ap.add_argument("--precision", default="bf16", choices=("bf16", "f32"),
                    help="static weight storage of --config 4")
    ap.add_argument("--nets", type=int, default=64, help="networks in the batched sweep")
    ap.add_argument("--stdp-path", default="fused", choices=("fused", "direct"),
                    help="--config 4: fused window kernel or the direct kernels (static weights as "
                         "--precision, plastic weights FP32 on both)")
    ap.add_argument("--settle", type=int, default=2000,
                    help="untimed steps run first, so that the timed steps see the network past its start-up "
                         "transient (default 2000 = 200 ms)")
    ap.add_argument("--long", type=int, default=10000,
                    help="steady-state run after the timed steps (reported beside them; the roofline comes "
                         "from it); 0 = none")
    ap.add_argument("--no-config1", action="store_true", help="skip the config-1 leg of the default run")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg of the default run")
    ap.add_argument("--no-config3", action="store_true", help="skip the conductance-network leg of the default run")
    ap.add_argument("--no-config4", action="store_true", help="skip the STDP leg of the default run")
    ap.add_argument("--no-batch-bf16", action="store_true", help="skip the BF16 run of the batched leg")
    ap.add_argument("--no-sequential", action="store_true",
                    help="skip the batched leg's one-network-at-a-time comparison (profiling runs)")
    ap.add_argument("--bf16-window", type=float, default=5.0,
                    help="simulated seconds over which the batched BF16 run's population rates are compared")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--graph", type=int, default=1, help="replay the step loop as hipGraphs (default 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--tune", action="store_true", help="run the autotuner instead of the tuned-shapes table")
    ap.add_argument("--tuned-out", default="", help="with --tune: write the shapes table with the tuned entries here")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--propagation", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device three times: the known-size reference that "
                         "converts rocprofv3 FETCH_SIZE / WRITE_SIZE counters into HBM bytes")
    args = ap.parse_args(argv)
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU only; multi-GPU sharding is a design note, "
                                   "not an implemented path", "n_gpus": args.gpus}))
        return 2

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if not args.steps:
        args.steps = 50000 if args.config == 4 else 10000
    if args.config == 3:
        if args.tune:
            best = tune_coba(args, log)
            log(f"# tuned: {best}")
            if args.tuned_out:
                write_tuned(args.tuned_out, {"vogels_abbott80k_p0.02": best})
            return 0
        r = coba_leg(args, log)
        best = r["formats"]["f32_class"]
        out = {
            "metric": "syn. events/s & realtime factor, 1 MI355X; plain vs batched-BF16 @<=2% rate err",
            "value": best["value"], "unit": "syn_events/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": best["rtf"] * DT * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (network generated on the device from seed %d)" % args.seed,
            "config": {"workload": r["workload"], "description": r["description"], "simulated_s": args.steps * DT},
            "rtf": best["rtf"], "coba": r, "roofline": best["roofline"], "cpu_baseline": None,
        }
        print(json.dumps(out))
        return 0
    if args.config == 4:
        if args.tune:
            best = tune_stdp(args, log)
            log(f"# tuned: {best}")
            if args.tuned_out:
                write_tuned(args.tuned_out, {"brunel100k_K10000_stdp_ee": best})
            return 0
        r = stdp_leg(args, log)
        out = {
            "metric": "syn. events/s & realtime factor, 1 MI355X; plain vs batched-BF16 @<=2% rate err",
            "value": r["value"], "unit": "syn_events/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": r["rtf"] * DT * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (network generated on the device from seed %d)" % args.seed,
            "config": {"workload": r["workload"], "description": r["description"], "simulated_s": r["simulated_s"]},
            "rtf": r["rtf"], "stdp": r, "roofline": r["roofline"], "cpu_baseline": None,
        }
        print(json.dumps(out))
        return 0
    table = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
    tuned_table = json.load(open(table)) if os.path.exists(table) else {}
    batch_shapes = None if args.tune else tuned_table.get(f"brunel_sweep{args.nets}")
    if args.config == 5 and (args.tile or args.block or args.chunk):   # explicit shapes win over the table
        batch_shapes = {"tile_size": args.tile, "block": args.block, "chunk": args.chunk}
    if args.config == 5:
        if args.tune:
            batch_shapes = tune_batched(args, log)
            bf = tune_batched(args, log, wf=N.WF_BF16)
            log(f"# tuned: exact {batch_shapes}, bf16 {bf}")
            if args.tuned_out:
                write_tuned(args.tuned_out, {f"brunel_sweep{args.nets}": batch_shapes,
                                             f"brunel_sweep{args.nets}_bf16": bf})
            return 0
        r = batched_leg(args, batch_shapes, log)
        out = {
            "metric": "syn. events/s & realtime factor, 1 MI355X; plain vs batched-BF16 @<=2% rate err",
            "value": r["value"], "unit": "syn_events/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": r["rtf"] * DT * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (networks generated on the device from seeds 0-%d)" % (args.nets - 1),
            "config": {"workload": r["workload"], "description": r["description"], "neurons": r["neurons"],
                       "synapses": r["synapses"], "simulated_s": args.steps * DT},
            "rtf": r["rtf"], "batched": r, "roofline": r["roofline"], "cpu_baseline": None,
        }
        print(json.dumps(out))
        return 0
    main_cfg = 2 if args.config == 2 else 1
    r = brunel_leg(args, main_cfg, tuned_table, log, cpu=not args.no_cpu)
    if args.tune:
        log(f"# tuned: {r['autotune']}")
        if args.tuned_out:
            write_tuned(args.tuned_out, {r["config"]["workload"]: r["autotune"]["best"] if r["autotune"] else None})
        return 0
    side = None
    if args.config == 2 and not args.no_config1:
        side = brunel_leg(args, 1, tuned_table, log, cpu=False)
    batched = None
    if args.config == 2 and not args.no_batch:
        batched = batched_leg(args, batch_shapes, log)
    # the conductance-based network (config 3) and STDP (config 4), 1 s each
    coba = stdp = None
    if args.config == 2 and not args.no_config3:
        a3 = argparse.Namespace(**vars(args))
        a3.steps, a3.warmup = 10000, 100
        coba = coba_leg(a3, log)
    if args.config == 2 and not args.no_config4:
        a4 = argparse.Namespace(**vars(args))
        a4.steps, a4.warmup = 10000, 1000
        stdp = stdp_leg(a4, log)

    out = {
        "metric": "syn. events/s & realtime factor, 1 MI355X; plain vs batched-BF16 @<=2% rate err",
        "value": r["value"],
        "unit": "syn_events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (network generated on the device from seed %d)" % args.seed,
        "config": r["config"],
        "rtf": r["rtf"],
        "syn_events": r["syn_events"],
        "settle_steps": args.settle,
        "steady": r["steady"],
        "roofline": r["roofline"],
        "autotune": r["autotune"],
        "cpu_baseline": r["cpu_baseline"],
        "config1": side,
        "batched": batched,
        "config3": coba,
        "config4": stdp,
    }
    print(json.dumps(out))
    return 0


def brunel_leg(args, cfg, tuned_table, log, cpu=False):
    """Brunel config 2 (100k LIFB, fixed in-degree K = 10,000, 1e9 synapses) or config 1
    (10k LIFB, Bernoulli p = 0.1): settle, warm up, time args.steps steps, then a long
    untimed-by-contract run (args.long steps) that gives the steady-state figures and the
    roofline of the fused kernel."""
    if cfg == 2:
        pops, projs, desc, ntot = brunel_fixed_k(args.scale, seed=args.seed)
        workload = "brunel100k_fixed_indegree_K10000"
    else:
        pops, projs, desc, ntot = brunel_bernoulli(args.scale, seed=args.seed)
        workload = "brunel10k_bernoulli_p0.1"
    lc = lifb_consts()

    def make(tile, block=0, chunk=0, profile=0):
        return N.Simulator.generated(pops, projs, lc, step=DT, dt=DT, seed=args.seed, use_graph=args.graph,
                                     profile=profile, propagation=args.propagation, tile_size=tile,
                                     prop_block=block, chunk=chunk)

    tuned = None
    explicit = cfg == args.config and (args.tile or args.block or args.chunk)
    tile, block, chunk = (args.tile, args.block, args.chunk) if explicit else (0, 0, 0)
    if not args.tune and not explicit and args.scale == 1.0:
        shapes = tuned_table.get(workload)
        if shapes:
            tile, block, chunk = shapes["tile_size"], shapes["block"], shapes["chunk"]
            tuned = dict(best=shapes, source="mgbrain_amd/tuned_shapes.json")
    if args.tune and args.propagation in (0, 2) and not (tile and block and chunk):
        probe = make(0)
        pinfo = probe.info()
        probe.close()
        if pinfo["propagation"] == N.PROP_TILED:
            cands = [tile] if tile else with_incumbent(
                T.tile_candidates(pinfo["n_neurons"], pinfo["ring_len"], max_width=1024), workload)
            best, res = T.autotune(lambda b: make(b), cands, steps=2000, warmup=500,
                                   log=lambda m: print("# " + m, file=sys.stderr, flush=True))
            if best:
                tile, block, chunk = best["tile_size"], best["block"], best["chunk"]
                tuned = dict(best=best, candidates=len(res))
    t0 = time.perf_counter()
    sim = make(tile, block, chunk, profile=0 if args.no_profile else 1)
    tuned_chunk = chunk if chunk else 8
    sim.sync()
    setup_s = time.perf_counter() - t0
    info0 = sim.info()
    log(f"# config {cfg} setup {setup_s:.2f} s: {info0['n_neurons']} neurons, {info0['n_synapses']} synapses, "
        f"D={info0['ring_len']}, window={info0['window']}, format={info0['weight_format']}, "
        f"persistent={info0['persistent']}, {info0['device_bytes'] / 2**30:.2f} GiB on device")
    n_poisson = int(sum(p["num"] for p in pops if p["type"] == N.POISSON))
    n_lif = info0["n_neurons"] - n_poisson

    def timed(nsteps):
        a = sim.info()
        t1 = time.perf_counter()
        sim.run(nsteps)
        sim.sync()
        wall = time.perf_counter() - t1
        b = sim.info()
        se = b["syn_events"] - a["syn_events"]
        spikes = b["spikes"] - a["spikes"]
        return wall, se, spikes, b

    # settle (stated, untimed): the network leaves its start-up transient (first spikes of
    # the integrating neurons arrive after the longest delay) before anything is timed
    sim.run(args.settle)
    sim.run(args.warmup)
    sim.sync()
    wall, se, spikes, b = timed(args.steps)
    sim_s = args.steps * DT
    value = se / wall
    rates_hz = spikes / ntot / sim_s
    log(f"# config {cfg} timed {args.steps} steps (after {args.settle} settle + {args.warmup} warm-up) in "
        f"{wall * 1e3:.2f} ms: {se:.4g} events, {spikes} spikes ({rates_hz:.2f} Hz mean), {value:.4g} SE/s, "
        f"RTF {wall / sim_s:.4f}")
    roof_src = (b, args.steps, se, spikes, f"timed run (steps {args.settle + args.warmup}.."
                                               f"{args.settle + args.warmup + args.steps})")
    steady = None
    if args.long:
        lw, lse, lsp, lb = timed(args.long)
        s0 = args.settle + args.warmup + args.steps
        steady = {"steps": args.long, "first_step": s0, "value": lse / lw, "rtf": lw / (args.long * DT),
                  "ms_per_step": lw * 1e3 / args.long, "mean_rate_hz": lsp / ntot / (args.long * DT)}
        log(f"# config {cfg} steady: {args.long} steps from step {s0} in {lw * 1e3:.1f} ms, {lse / lw:.4g} SE/s, "
            f"RTF {lw / (args.long * DT):.4f}, {steady['mean_rate_hz']:.2f} Hz")
        roof_src = (lb, args.long, lse, lsp, f"steady run (steps {s0}..{s0 + args.long})")
    roof = None
    rb, rsteps, rse, rsp, rname = roof_src
    if rb["prop_launches"] > 0 and rb["propagation"] == N.PROP_TILED:
        roof = fused_roofline(rb, info0, rsteps, rse, rsp, n_lif, n_poisson,
                              "k_tiled (fused neuron update + CLASS propagation)")
        if roof:
            roof["run"] = rname
            roof["persistent"] = info0["persistent"]
            # measured HBM bytes per window: FETCH_SIZE / WRITE_SIZE counter passes of this
            # workload in the same regime and launch shape, scaled by a copy of known size
            tfile = os.path.join(ROOT, "profiles", "traffic_" + workload.replace("_fixed_indegree", "") + ".json")
            if os.path.exists(tfile):
                t = json.load(open(tfile))
                if (t.get("tile_size"), t.get("block"), t.get("chunk"), t.get("word_bytes"), t.get("persistent")) == \
                        (rb["tile_size"], rb["block"], tuned_chunk, rb["word_bytes"], info0["persistent"]):
                    roof["traffic"] = round(t["traffic_bytes_per_window"])
                    roof["traffic_source"] = t["source"]
            log(f"# config {cfg} {roof['kernel']}: {roof['launches']} windows ({roof['run']}), avg "
                f"{roof['avg_launch_us']:.2f} us per window, {roof['bytes_per_launch'] / 1e6:.3f} MB algorithmic -> "
                f"{roof['achieved']:.1f} GB/s ({roof['frac']:.3f} of peak)")
    elif rb["prop_launches"] > 0:
        avg_s = rb["prop_ms"] / 1e3 / rb["prop_launches"]
        alg = 12.0 * rsp + 9.0 * rse
        ach = alg / rb["prop_launches"] / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_prop (direct)",
                "avg_launch_us": round(avg_s * 1e6, 3), "bytes_per_launch": round(alg / rb["prop_launches"]),
                "run": rname}

    cpu_res = None
    if cpu:
        log("# cpu baseline (oracle) ...")
        sim._bench_types = np.concatenate([np.full(p["num"], p["type"], np.int32) for p in pops])
        sim._bench_rate = np.concatenate([np.full(p["num"], p.get("rate", 0.0), np.float32) for p in pops])
        sim._bench_lifconst = lc
        cpu_res = cpu_baseline(sim, budget_s=args.cpu_budget, seed=args.seed)
        log(f"# cpu: {cpu_res['value']:.4g} SE/s with {cpu_res['cores']} threads over {cpu_res['steps']} steps in "
            f"{cpu_res['seconds']:.1f} s; one thread {cpu_res['single_thread']['value']:.4g} SE/s")
    sim.close()
    return {
        "value": value, "rtf": wall / sim_s, "ms_per_step": wall * 1e3 / args.steps, "syn_events": se,
        "config": {"workload": workload, "description": desc, "neurons": info0["n_neurons"],
                   "synapses": info0["n_synapses"], "simulated_s": sim_s, "weight_format": "f32 per-class",
                   "mean_rate_hz": round(rates_hz, 3), "setup_s": round(setup_s, 2),
                   "persistent": info0["persistent"], "window_steps": info0["window"]},
        "steady": steady, "roofline": roof, "autotune": tuned, "cpu_baseline": cpu_res,
    }

if __name__ == "__main__":
    sys.exit(main())
