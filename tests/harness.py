"""Builds the same network for the oracle and for the HIP path (test helper)."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po
from tests.netmodels import Net


def host_synapses(net: Net, seed=0, mt_seed=12345, rand_seed=1, dw=0.01, dt=1e-4):
    return po.gen_net(net.pops, net.projs, mt_seed=mt_seed, rand_seed=rand_seed, philox_seed=seed, dw=dw, dt=dt)


def initial_v(net: Net, seed=0):
    v = []
    base = 0
    for (num, _t, _s), vr in zip(net.pops, net.vrange):
        if vr is None:
            v.append(np.full(num, net.lifconst[1], dtype=np.float32))
        else:
            v.append(po.vinit(seed, range(base, base + num), vr[0], vr[1]))
        base += num
    return np.concatenate(v)


def oracle(net: Net, syn, seed=0, acc=po.ACC_F32, stdp=False, stdpconst=None, scale=(2.0 ** 30, 2.0 ** 30),
           step=1e-4):
    csr = po.build_csr(net.n, syn, step=step)
    types, rate = net.neuron_arrays()
    sim = po.OracleSim(csr, types, rate, initial_v(net, seed), net.lifconst, stdpconst=stdpconst, stdp=stdp,
                       seed=seed, acc=acc, scale_exc=scale[0], scale_inh=scale[1])
    return csr, sim


def gpu_from_list(net: Net, syn, seed=0, stdp=False, stdpconst=None, step=1e-4, **opts):
    from mgbrain_amd.native import Simulator
    types, rate = net.neuron_arrays()
    return Simulator.from_arrays(types, rate, net.lifconst, syn.src, syn.tar, syn.weight, syn.delay, step=step,
                                 v_init=initial_v(net, seed), stdpconst=stdpconst, stdp=stdp,
                                 plastic=syn.plastic if stdp else None, seed=seed, **opts)


def gpu_generated(net: Net, seed=0, stdp=False, stdpconst=None, step=1e-4, **opts):
    from mgbrain_amd.native import Simulator
    projs = []
    for j, pr in enumerate(net.projs):
        q = dict(pr)
        q.setdefault("wmin", pr.get("w", 0.0))
        q.setdefault("wmax", pr.get("w", 0.0))
        q.setdefault("dmin", pr.get("d", 0.0))
        q.setdefault("dmax", pr.get("d", 0.0))
        q["index"] = j
        q["seed"] = seed
        projs.append(q)
    return Simulator.generated(net.gen_pops(), projs, net.lifconst, step=step, stdp=stdp, stdpconst=stdpconst,
                               seed=seed, **opts)


def split_steps(steps, ids, nsteps, step0=0):
    """(steps, ids) sorted records -> list of per-step id arrays."""
    out = []
    bounds = np.searchsorted(steps, np.arange(step0, step0 + nsteps + 1))
    for k in range(nsteps):
        out.append(ids[bounds[k]:bounds[k + 1]])
    return out


def first_mismatch(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        if not np.array_equal(x, y):
            return k
    return -1
