"""Translates the reference-harness spec lines into oracle populations and
projections, applying the reference Model's own rules (model/model.cpp:20-57):
one-to-one projections between populations of different size are rejected,
and a repeated connect overwrites the earlier projection only when its index
is > 0 (so the first projection can be duplicated)."""
from __future__ import annotations

from oracle import pyoracle as po
from tests import netmodels as M


def translate(spec):
    dt, dw, mt_seed, rand_seed, shuffle = 1e-3, 0.01, 0, 1, False
    pops, projs = [], []
    for line in spec:
        k, *a = line.split()
        if k == "timestep":
            dt = float(a[0])
        elif k == "dw":
            dw = float(a[0])
        elif k == "clock":
            mt_seed = int(a[0]) & 0xFFFFFFFF
        elif k == "srand":
            rand_seed = int(a[0])
        elif k == "shuffle":
            shuffle = a[0] == "1"
        elif k == "pop":
            pops.append((int(a[0]), int(a[1]), int(a[2])))
        elif k == "proj":
            s, t = int(a[0]), int(a[1])
            wmin, wmax, dmin, dmax, ctype = (float(M.f32(x)) for x in a[2:7])
            if ctype == 0.0 and pops[s][0] != pops[t][0]:
                continue
            pr = dict(src=s, tar=t, wmin=wmin, wmax=wmax, dmin=dmin, dmax=dmax, ctype=ctype, kind=po.CONN_REF)
            idx = -1
            for i, q in enumerate(projs):
                if q["src"] == s and q["tar"] == t:
                    idx = i
            if idx > 0:
                projs[idx].update(wmin=wmin, wmax=wmax, dmin=dmin, dmax=dmax, ctype=ctype)
            else:
                projs.append(pr)
        elif k in ("brunel", "vogel"):
            net = M.brunel_ref(int(a[0])) if k == "brunel" else M.vogel_ref(int(a[0]))
            pops = list(net.pops)
            projs = [dict(p, wmin=p["w"], wmax=p["w"], dmin=p["d"], dmax=p["d"]) for p in net.projs]
    if shuffle:
        for q in projs:
            if 0.0 < q["ctype"] < 1.0:
                q["kind"] = po.CONN_REF_SHUFFLE
    return dict(pops=pops, projs=projs, dt=dt, dw=dw, mt_seed=mt_seed, rand_seed=rand_seed)


def oracle_synapses(spec):
    t = translate(spec)
    return po.gen_net(t["pops"], t["projs"], mt_seed=t["mt_seed"], rand_seed=t["rand_seed"], philox_seed=0,
                      dw=t["dw"], dt=t["dt"]), t
