"""Autotuner of the step kernels' launch shapes.

The fused tiled kernel has three shape parameters:
  tile_size  target neurons owned by one workgroup (its LDS ring slice is
             D x tile_size words; it sets the number of workgroups and the
             length of the per-(spike, tile) CSR sub-rows),
  block      workgroup size,
  chunk      synapses per propagation work item.
tile_size is a data-layout choice (the tile-sorted CSR copy depends on it), so
every candidate tile size builds its own simulator; block and chunk are
changed on a live simulator with mgb_set_launch.  Each candidate is timed with
the library's hipEvents over a fixed number of steps after a warm-up.
"""
from __future__ import annotations

import math
import time


def tile_candidates(n, ring_len, rw=1, max_width=1024, lds_bytes=160 * 1024, batch_words=3 * 1024 + 4096,
                    static_bytes=1024, aux_words=0):
    """Mean tile widths to try for n target neurons: about n/128 ... n/2048 wide,
    each below the widest tile the fused kernel can hold -- its LDS ring slice
    (rw words per slot and target, ring_len slots) beside the staging area
    (batch_words, aux_words) -- and below the synapse word's local-target range
    (max_width: 1024 for 16-bit CLASS words, 2048 for the 32-bit words)."""
    room = (lds_bytes - static_bytes) // 4 - batch_words - aux_words
    bmax = max(1, min(max_width, room // max(1, ring_len * rw)))
    out = []
    for g in (128, 256, 512, 1024, 2048):
        b = min(bmax, max(32, math.ceil(n / g)))
        if b not in out:
            out.append(b)
    return out


def autotune(make_sim, tiles, blocks=(256, 512, 1024), chunks=(1, 2, 4, 8, 16), steps=400, warmup=200,
             budget_s=90.0, log=None):
    """make_sim(tile_size) -> Simulator.  Returns (best, results) where each
    result is a dict(tile_size, block, chunk, ms_per_step, widest): tile_size is
    the mean tile width asked for (what to pass back to build the same tiling),
    widest the simulator's widest tile."""
    t_start = time.perf_counter()
    results = []
    for tile in tiles:
        if time.perf_counter() - t_start > budget_s:
            break
        sim = make_sim(tile)
        try:
            sim.run(warmup)
            for block in blocks:
                for chunk in chunks:
                    sim.set_launch(block, chunk)
                    sim.run(steps // 4)
                    sim.run(steps)
                    ms = sim.info()["last_run_ms"]
                    r = dict(tile_size=tile, block=block, chunk=chunk, ms_per_step=ms / steps,
                             widest=sim.info()["tile_size"])
                    results.append(r)
                    if log:
                        log(f"tune tile={tile} (widest {r['widest']}) block={block} chunk={chunk}: "
                            f"{r['ms_per_step'] * 1e3:.3f} us/step")
        finally:
            sim.close()
    best = min(results, key=lambda r: r["ms_per_step"]) if results else None
    return best, results
