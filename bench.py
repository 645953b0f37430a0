#!/usr/bin/env python3
"""Single-GPU throughput of the MGBrain simulation step on MI355X.

Workload (default): Brunel network of 80,000 excitatory + 20,000 inhibitory
LIFB neurons and 10,000 Poisson inputs at 15 Hz, fixed in-degree K = 10,000
(1.0e9 synapses), J = 0.1 mV, g = 5, delays on the 1.0-2.0 ms grid, dt = 0.1 ms,
FP32 weights (exact per-class weights, 4 B per synapse on the device).  The
network is generated on the device from a seed; inputs are resident in HBM
before the timed region.

A "step" is one simulation time step (0.1 ms of biological time) of the whole
network.  The timed region runs --steps steps after --warmup untimed steps.
Prints one JSON line: synaptic events/s (value), realtime factor, the HBM
roofline of the dominant kernel (spike propagation) and the CPU baseline
(the oracle's FP32 restatement of the reference step on the host's cores).
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

HBM_PEAK_GBS = 8000.0
DT = 1e-4
f32 = np.float32


def lifb_consts():
    """gen_life_const(v_init 0, v_rest 0, v_reset 10 mV, cm 0, tau_m 0, t_ref 2 ms, 1, 1,
    V_th 20 mV, 0) at dt = 0.1 ms (constgen.h:9-107); LIFB reads DT, V_INIT, V_RESET,
    V_THRESH and TAU_REFRAC only."""
    c = np.zeros(30, dtype=f32)
    c[0], c[3], c[4], c[8] = DT, 0.010, 0.020, 0.002
    c[10], c[11] = 1.0, 1.0
    e = np.exp(f32(-DT), dtype=f32)
    c[12], c[13], c[15], c[16] = e, e, e, e
    c[21], c[22], c[23], c[24], c[25] = 0.951229453, 0.967216074, 0.0799999982, 0.0599999987, 0.90483743
    return c


def brunel_fixed_k(scale=1.0, K=10000, rate=0.0015, seed=0):
    nP, nE, nI = int(10000 * scale), int(80000 * scale), int(20000 * scale)
    ntot = nP + nE + nI
    Ks = [int(round(K * n / ntot)) for n in (nP, nE, nI)]
    J, wi = float(f32(1e-4)), float(f32(-5e-4))
    dmin, dmax = float(f32(0.001)), float(f32(0.002))
    pops = [dict(num=nP, type=N.POISSON, rate=rate), dict(num=nE, type=N.LIFB), dict(num=nI, type=N.LIFB)]
    projs = []
    for tgt in (1, 2):
        for src, k, w in ((0, Ks[0], J), (1, Ks[1], J), (2, Ks[2], wi)):
            projs.append(dict(src=src, tar=tgt, kind=N.CONN_FIXED_INDEGREE, k=k, wmin=w, wmax=w, dmin=dmin,
                              dmax=dmax, index=len(projs), seed=seed))
    desc = (f"Brunel fixed in-degree: {nE} E + {nI} I LIFB + {nP} Poisson @ {rate / DT:.0f} Hz, "
            f"K={K}, J=0.1 mV, g=5, delay U{{1.0..2.0}} ms, dt=0.1 ms")
    return pops, projs, desc, ntot


def brunel_bernoulli(scale=1.0, p=0.1, rate=0.002, seed=0):
    nP, nE, nI = int(1000 * scale), int(8000 * scale), int(2000 * scale)
    J, wi, d = float(f32(1e-4)), float(f32(-5e-4)), float(f32(0.0014))
    pops = [dict(num=nP, type=N.POISSON, rate=rate), dict(num=nE, type=N.LIFB), dict(num=nI, type=N.LIFB)]
    projs = []
    for s, t, w in ((0, 1, J), (0, 2, J), (1, 1, J), (1, 2, J), (2, 1, wi), (2, 2, wi)):
        projs.append(dict(src=s, tar=t, kind=N.CONN_BERNOULLI, p=p, w=w, d=d, index=len(projs), seed=seed))
    desc = (f"Brunel Bernoulli p={p}: {nE} E + {nI} I LIFB + {nP} Poisson @ {rate / DT:.0f} Hz, "
            f"J=0.1 mV, g=5, delay 1.4 ms, dt=0.1 ms")
    return pops, projs, desc, nP + nE + nI


def cpu_baseline(sim, budget_s=12.0, max_steps=4000, seed=0):
    """The oracle (FP32 restatement of the reference step, one thread) on the same
    network (CSR read back from the device), timed on this host."""
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
    while steps < max_steps and (time.perf_counter() - t1) < budget_s:
        o.run(100, cap=n * 100)
        steps += 100
    el = time.perf_counter() - t1
    se = o.syn_events()
    o.close()
    return dict(value=se / el, unit="syn_events/s", cores=1, kind="port",
                sample=f"oracle FP32 step loop, steps 0..{steps} ({steps * DT * 1e3:.0f} ms simulated) of the "
                       f"same network, 1 thread, {el:.1f} s timed (+{setup:.1f} s read-back/setup); "
                       f"host {platform.processor() or platform.machine()}, nproc={os.cpu_count()}",
                steps=steps, seconds=el, syn_events=se, rtf=el / (steps * DT))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--config", type=int, default=2, choices=(1, 2))
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--tune", action="store_true", help="run the autotuner instead of the tuned-shapes table")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--propagation", type=int, default=0)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design note, "
                                   "not an implemented path", "n_gpus": args.gpus}))
        return 2

    if args.config == 2:
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
    tile, block, chunk = args.tile, args.block, args.chunk
    table = os.path.join(ROOT, "mgbrain_amd", "tuned_shapes.json")
    if not args.tune and not (tile or block or chunk) and args.scale == 1.0 and os.path.exists(table):
        shapes = json.load(open(table)).get(workload)
        if shapes:
            tile, block, chunk = shapes["tile_size"], shapes["block"], shapes["chunk"]
            tuned = dict(best=shapes, source="mgbrain_amd/tuned_shapes.json")
    if args.tune and args.propagation in (0, 2) and not (tile and block and chunk):
        probe = make(0)
        pinfo = probe.info()
        probe.close()
        if pinfo["propagation"] == N.PROP_TILED:
            cands = [tile] if tile else T.tile_candidates(pinfo["n_neurons"], pinfo["ring_len"])
            best, res = T.autotune(lambda b: make(b), cands,
                                   log=lambda m: print("# " + m, file=sys.stderr, flush=True))
            if best:
                tile, block, chunk = best["tile_size"], best["block"], best["chunk"]
                tuned = dict(best=best, candidates=len(res))
    t0 = time.perf_counter()
    sim = make(tile, block, chunk, profile=0 if args.no_profile else 1)
    tuned_chunk = chunk if chunk else 8
    sim.sync()
    setup_s = time.perf_counter() - t0
    sim._bench_types = np.concatenate([np.full(p["num"], p["type"], np.int32) for p in pops])
    sim._bench_rate = np.concatenate([np.full(p["num"], p.get("rate", 0.0), np.float32) for p in pops])
    sim._bench_lifconst = lc
    info0 = sim.info()
    print(f"# setup {setup_s:.2f} s: {info0['n_neurons']} neurons, {info0['n_synapses']} synapses, "
          f"D={info0['ring_len']}, window={info0['window']}, format={info0['weight_format']}, "
          f"{info0['device_bytes'] / 2**30:.2f} GiB on device", file=sys.stderr, flush=True)

    sim.run(args.warmup)
    sim.sync()
    a = sim.info()
    t1 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    t2 = time.perf_counter()
    b = sim.info()
    wall = t2 - t1
    se = b["syn_events"] - a["syn_events"]
    spikes = b["spikes"] - a["spikes"]
    sim_s = args.steps * DT
    value = se / wall
    rates_hz = spikes / ntot / sim_s
    print(f"# timed {args.steps} steps in {wall * 1e3:.1f} ms (device {b['last_run_ms']:.1f} ms): "
          f"{se:.4g} events, {spikes} spikes ({rates_hz:.2f} Hz mean), {value:.4g} SE/s, "
          f"RTF {wall / sim_s:.4f}", file=sys.stderr, flush=True)

    roof = None
    if b["prop_launches"] > 0:
        n = info0["n_neurons"]
        D = info0["ring_len"]
        n_poisson = int(sum(p["num"] for p in pops if p["type"] == N.POISSON))
        n_lif = n - n_poisson
        launches = b["prop_launches"]
        if b["propagation"] == N.PROP_TILED:
            # fused window kernel, bytes each launch must move through HBM (each datum once):
            # synapse words 4 B/event; spike entry 4 B written + row pointers 8 B per spike;
            # (source, tile) offset pairs 4 B per spike and tile; ring slice in + out 8 B per
            # (slot, target); LIFB state V, refr, last, cnt read + write 32 B, Poisson rate +
            # last + cnt 20 B per neuron
            G = b["tiles"]
            alg = (4.0 * se + 12.0 * spikes + 4.0 * spikes * G
                   + launches * (8.0 * D * G * b["tile_size"] + 32.0 * n_lif + 20.0 * n_poisson))
            kname = "k_tiled (fused neuron update + CLASS propagation)"
        else:
            alg = 12.0 * spikes + 4.0 * se
            kname = "k_prop (CLASS, DIRECT)"
        per_launch = alg / launches
        avg_s = b["prop_ms"] / 1e3 / launches
        ach = per_launch / avg_s / 1e9
        # measured HBM bytes per launch: FETCH_SIZE / WRITE_SIZE counter passes of this
        # workload scaled by a copy kernel of known size (profiles/traffic_*.json)
        traffic = None
        tfile = os.path.join(ROOT, "profiles", "traffic_" + workload.replace("_fixed_indegree", "") + ".json")
        if os.path.exists(tfile):
            t = json.load(open(tfile))
            if (t.get("tile_size"), t.get("block"), t.get("chunk")) == (b["tile_size"], b["block"], tuned_chunk):
                traffic = round(t["traffic_bytes_per_launch"])
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                "avg_launch_us": round(avg_s * 1e6, 3), "bytes_per_launch": round(per_launch),
                "launches": launches, "tile_size": b["tile_size"], "tiles": b["tiles"], "block": b["block"]}
        print(f"# {kname}: {launches} launches, avg {avg_s * 1e6:.2f} us, {per_launch / 1e6:.3f} MB "
              f"algorithmic -> {ach:.1f} GB/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sim, budget_s=args.cpu_budget, seed=args.seed)
        print(f"# cpu: {cpu['value']:.4g} SE/s over {cpu['steps']} steps in {cpu['seconds']:.1f} s",
              file=sys.stderr, flush=True)
    sim.close()

    out = {
        "metric": "syn. events/s & realtime factor, 1 MI355X; plain vs batched-BF16 @<=2% rate err",
        "value": value,
        "unit": "syn_events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (network generated on the device from seed %d)" % args.seed,
        "config": {"workload": workload, "description": desc, "neurons": info0["n_neurons"],
                   "synapses": info0["n_synapses"], "simulated_s": sim_s, "weight_format": "f32 per-class",
                   "mean_rate_hz": round(rates_hz, 3), "setup_s": round(setup_s, 2)},
        "rtf": wall / sim_s,
        "syn_events": se,
        "roofline": roof,
        "autotune": tuned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
