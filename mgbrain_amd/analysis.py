"""Spike records on disk and their statistics.

The reference only prints fire counts and rates (mgsim.cpp:596-615); here a
run's recorded spikes can be written to and read from an ``.npz`` file (arrays
``steps`` int32, ``ids`` int32, sorted by (step, id); scalars ``dt`` and
``n_neurons``; optional ``pop_bounds`` [k, 2] of id ranges) and reduced to
per-population rates, CV of the inter-spike intervals and weight histograms.
"""
from __future__ import annotations

import numpy as np

FORMAT_VERSION = 1


def save_spikes(path, steps, ids, n_neurons, dt, pop_bounds=None):
    """Writes a spike record (steps / ids as returned by Simulator.spikes())."""
    steps = np.ascontiguousarray(steps, dtype=np.int32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    if steps.shape != ids.shape:
        raise ValueError("steps and ids differ in length")
    order = np.lexsort((ids, steps))
    extra = {}
    if pop_bounds is not None:
        extra["pop_bounds"] = np.asarray(pop_bounds, dtype=np.int64).reshape(-1, 2)
    np.savez(path, format_version=np.int32(FORMAT_VERSION), steps=steps[order], ids=ids[order],
             n_neurons=np.int64(n_neurons), dt=np.float64(dt), **extra)


def load_spikes(path):
    """Returns a dict with steps, ids, n_neurons, dt and (if stored) pop_bounds."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["format_version"]) != FORMAT_VERSION:
            raise ValueError(f"spike record format {int(z['format_version'])} is not {FORMAT_VERSION}")
        out = {"steps": z["steps"], "ids": z["ids"], "n_neurons": int(z["n_neurons"]), "dt": float(z["dt"])}
        if "pop_bounds" in z:
            out["pop_bounds"] = z["pop_bounds"]
    return out


def population_rates(steps, ids, bounds, seconds):
    """Mean rate (Hz) of each id range [lo, hi) over `seconds`."""
    ids = np.asarray(ids)
    return [float(np.count_nonzero((ids >= lo) & (ids < hi))) / max(hi - lo, 1) / seconds for lo, hi in bounds]


def cv_isi(steps, ids, bounds, min_spikes=3):
    """Per id range: (mean CV of the inter-spike intervals over neurons with at
    least min_spikes spikes, number of such neurons)."""
    steps = np.asarray(steps, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.int64)
    order = np.lexsort((steps, ids))
    s, i = steps[order], ids[order]
    out = []
    for lo, hi in bounds:
        m = (i >= lo) & (i < hi)
        si, ii = s[m], i[m]
        cvs = []
        if len(si):
            cut = np.flatnonzero(np.diff(ii)) + 1
            for a, b in zip(np.r_[0, cut], np.r_[cut, len(si)]):
                if b - a >= min_spikes:
                    isi = np.diff(si[a:b]).astype(np.float64)
                    cvs.append(isi.std() / isi.mean())
        out.append((float(np.mean(cvs)) if cvs else float("nan"), len(cvs)))
    return out


def weight_histogram(weights, bins=20, lo=None, hi=None):
    """(counts, edges) of a weight array (e.g. the plastic E->E weights)."""
    w = np.asarray(weights, dtype=np.float64)
    lo = float(w.min()) if lo is None else lo
    hi = float(w.max()) if hi is None else hi
    counts, edges = np.histogram(w, bins=bins, range=(lo, hi))
    return counts, edges
