"""ctypes binding of the C ABI in include/mgbrain_hip.h (libmgbrain_hip.so).

This is plumbing for tests and bench.py: the product is the shared library and
the C++ facade.  Loading fails loudly when the library is missing; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libmgbrain_hip.so")

OK, ERR_INVALID, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_CAPACITY = range(6)
RAW, LIF0, POISSON, LIF, LIFE, LIFB = range(6)
WF_AUTO, WF_CLASS, WF_F32, WF_BF16, WF_FP8 = range(5)
PROP_AUTO, PROP_DIRECT, PROP_TILED = range(3)
PERSIST_AUTO, PERSIST_OFF, PERSIST_REGIONS, PERSIST_LISTS = range(4)
CONN_BERNOULLI, CONN_FIXED_INDEGREE = 1, 2
F_V, F_IEXC, F_IINH, F_REFR, F_LAST, F_CNT = range(6)

EXPORTS = ["mgb_abi_version", "mgb_last_error", "mgb_create", "mgb_create_generated",
           "mgb_create_batch_generated", "mgb_run",
           "mgb_sync", "mgb_set_launch", "mgb_get_info", "mgb_read_spikes", "mgb_clear_spikes",
           "mgb_read_state", "mgb_fire_counts", "mgb_read_tiles", "mgb_read_csr", "mgb_read_weights", "mgb_read_weights_noalloc",
           "mgb_checkpoint_size",
           "mgb_checkpoint_save", "mgb_checkpoint_load", "mgb_destroy"]

_p = C.c_void_p


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("weight_format", C.c_int32), ("window", C.c_int32),
                ("use_graph", C.c_int32), ("record_capacity", C.c_int64), ("prop_block", C.c_int32),
                ("prop_grid", C.c_int32), ("profile", C.c_int32), ("propagation", C.c_int32),
                ("tile_size", C.c_int32), ("chunk", C.c_int32), ("persistent", C.c_int32),
                ("spin_limit_us", C.c_int32), ("reserved", C.c_int32 * 2)]


class NetDesc(C.Structure):
    _fields_ = [("n_neurons", C.c_int32), ("type", _p), ("rate", _p), ("v_init", _p), ("cset", _p),
                ("stream", _p), ("n_csets", C.c_int32), ("lifconst", _p), ("stdpconst", _p),
                ("n_streams", C.c_int32), ("stream_base", _p), ("stream_seed", _p),
                ("n_synapses", C.c_int64), ("syn_src", _p), ("syn_tar", _p), ("syn_weight", _p),
                ("syn_delay", _p), ("syn_plastic", _p), ("fp8_scale", _p), ("step", C.c_float),
                ("stdp", C.c_int32), ("seed", C.c_uint64)]


class GenPop(C.Structure):
    _fields_ = [("num", C.c_int32), ("type", C.c_int32), ("rate", C.c_float), ("v_min", C.c_float),
                ("v_max", C.c_float), ("cset", C.c_int32), ("stream", C.c_int32)]


class GenProj(C.Structure):
    _fields_ = [("src", C.c_int32), ("tar", C.c_int32), ("kind", C.c_int32), ("p", C.c_float),
                ("k", C.c_int32), ("wmin", C.c_float), ("wmax", C.c_float), ("dmin", C.c_float),
                ("dmax", C.c_float), ("plastic", C.c_int32), ("index", C.c_int32), ("seed", C.c_uint64)]


class GenDesc(C.Structure):
    _fields_ = [("n_pops", C.c_int32), ("pops", _p), ("n_projs", C.c_int32), ("projs", _p),
                ("n_csets", C.c_int32), ("lifconst", _p), ("stdpconst", _p), ("n_streams", C.c_int32),
                ("stream_seed", _p), ("step", C.c_float), ("dt", C.c_float), ("dw", C.c_float),
                ("stdp", C.c_int32), ("seed", C.c_uint64), ("net_seed", C.c_uint64)]


class Info(C.Structure):
    _fields_ = [("n_neurons", C.c_int32), ("ring_len", C.c_int32), ("min_delay", C.c_int32),
                ("window", C.c_int32), ("n_synapses", C.c_int64), ("step", C.c_int64),
                ("syn_events", C.c_int64), ("spikes_recorded", C.c_int64), ("weight_format", C.c_int32),
                ("n_classes", C.c_int32), ("fix_scale_exc", C.c_float), ("fix_scale_inh", C.c_float),
                ("device_bytes", C.c_int64), ("last_run_ms", C.c_double), ("launches", C.c_int64),
                ("spikes", C.c_int64), ("neuron_ms", C.c_double), ("prop_ms", C.c_double),
                ("stdp_ms", C.c_double), ("neuron_launches", C.c_int64), ("prop_launches", C.c_int64),
                ("stdp_launches", C.c_int64), ("propagation", C.c_int32), ("tile_size", C.c_int32),
                ("tiles", C.c_int32), ("block", C.c_int32), ("segments", C.c_int32),
                ("word_bytes", C.c_int32), ("persistent", C.c_int32)]


_lib = None


class MGBError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mgb status {status}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"HIP library not built: {HIP_LIB} (run __graft_entry__.build())")
        L = C.CDLL(HIP_LIB)
        L.mgb_abi_version.restype = C.c_int32
        L.mgb_last_error.restype = C.c_char_p
        for name in EXPORTS:
            if name not in ("mgb_abi_version", "mgb_last_error"):
                getattr(L, name).restype = C.c_int
        L.mgb_create.argtypes = [C.POINTER(NetDesc), C.POINTER(Opts), C.POINTER(_p)]
        L.mgb_create_generated.argtypes = [C.POINTER(GenDesc), C.POINTER(Opts), C.POINTER(_p)]
        L.mgb_create_batch_generated.argtypes = [C.POINTER(GenDesc), C.c_int32, C.POINTER(Opts), C.POINTER(_p)]
        L.mgb_run.argtypes = [_p, C.c_int32]
        L.mgb_sync.argtypes = [_p]
        L.mgb_set_launch.argtypes = [_p, C.c_int32, C.c_int32]
        L.mgb_read_tiles.argtypes = [_p, _p]
        L.mgb_get_info.argtypes = [_p, C.POINTER(Info)]
        L.mgb_read_spikes.argtypes = [_p, C.c_int64, _p, _p, C.POINTER(C.c_int64)]
        L.mgb_clear_spikes.argtypes = [_p]
        L.mgb_read_state.argtypes = [_p, C.c_int32, _p]
        L.mgb_fire_counts.argtypes = [_p, _p]
        L.mgb_read_csr.argtypes = [_p, _p, _p, _p, _p]
        L.mgb_read_weights.argtypes = [_p, _p]
        L.mgb_read_weights_noalloc.argtypes = [_p, _p]
        L.mgb_checkpoint_size.argtypes = [_p, C.POINTER(C.c_int64)]
        L.mgb_checkpoint_save.argtypes = [_p, _p, C.c_int64]
        L.mgb_checkpoint_load.argtypes = [_p, _p, C.c_int64]
        L.mgb_destroy.argtypes = [_p]
        _lib = L
    return _lib


def _check(status):
    if status != OK:
        raise MGBError(status, lib().mgb_last_error().decode(errors="replace"))


def _addr(a):
    return None if a is None else a.ctypes.data


def make_opts(weight_format=WF_AUTO, window=0, use_graph=0, record_capacity=0, device=0, prop_block=0,
              prop_grid=0, profile=0, propagation=0, tile_size=0, chunk=0, persistent=PERSIST_AUTO,
              spin_limit_us=0):
    o = Opts()
    o.persistent = persistent
    o.spin_limit_us = spin_limit_us
    o.device = device
    o.weight_format = weight_format
    o.window = window
    o.use_graph = use_graph
    o.record_capacity = record_capacity
    o.prop_block = prop_block
    o.prop_grid = prop_grid
    o.profile = profile
    o.propagation = propagation
    o.tile_size = tile_size
    o.chunk = chunk
    return o


class Simulator:
    """One mgb_sim handle."""

    def __init__(self, handle, keep=None):
        self.h = handle
        self._keep = keep

    # -- construction ---------------------------------------------------------
    @classmethod
    def from_arrays(cls, types, rate, lifconst, src, tar, weight, delay_s, step=1e-4, v_init=None,
                    stdpconst=None, stdp=False, plastic=None, seed=0, cset=None, stream=None,
                    stream_base=None, stream_seed=None, fp8_scale=None, **opts):
        n = len(types)
        arrs = dict(
            type=np.ascontiguousarray(types, dtype=np.int32),
            rate=np.ascontiguousarray(rate, dtype=np.float32),
            lifconst=np.ascontiguousarray(lifconst, dtype=np.float32).reshape(-1),
            src=np.ascontiguousarray(src, dtype=np.int32),
            tar=np.ascontiguousarray(tar, dtype=np.int32),
            w=np.ascontiguousarray(weight, dtype=np.float32),
            d=np.ascontiguousarray(delay_s, dtype=np.float32),
        )
        if v_init is not None:
            arrs["v"] = np.ascontiguousarray(v_init, dtype=np.float32)
        if stdpconst is not None:
            arrs["sc"] = np.ascontiguousarray(stdpconst, dtype=np.float32)
        if plastic is not None:
            arrs["pl"] = np.ascontiguousarray(plastic, dtype=np.uint8)
        if cset is not None:
            arrs["cs"] = np.ascontiguousarray(cset, dtype=np.uint8)
        if stream is not None:
            arrs["st"] = np.ascontiguousarray(stream, dtype=np.uint16)
        if stream_base is not None:
            arrs["sb"] = np.ascontiguousarray(stream_base, dtype=np.int32)
            arrs["ss"] = np.ascontiguousarray(stream_seed, dtype=np.uint64)
        if fp8_scale is not None:
            arrs["f8"] = np.ascontiguousarray(fp8_scale, dtype=np.float32)
        d = NetDesc()
        d.n_neurons = n
        d.type = _addr(arrs["type"])
        d.rate = _addr(arrs["rate"])
        d.v_init = _addr(arrs.get("v"))
        d.cset = _addr(arrs.get("cs"))
        d.stream = _addr(arrs.get("st"))
        d.n_csets = arrs["lifconst"].size // 30
        d.lifconst = _addr(arrs["lifconst"])
        d.stdpconst = _addr(arrs.get("sc"))
        d.n_streams = len(arrs["sb"]) if "sb" in arrs else 1
        d.stream_base = _addr(arrs.get("sb"))
        d.stream_seed = _addr(arrs.get("ss"))
        d.n_synapses = arrs["src"].size
        d.syn_src = _addr(arrs["src"])
        d.syn_tar = _addr(arrs["tar"])
        d.syn_weight = _addr(arrs["w"])
        d.syn_delay = _addr(arrs["d"])
        d.syn_plastic = _addr(arrs.get("pl"))
        d.fp8_scale = _addr(arrs.get("f8"))
        d.step = step
        d.stdp = int(bool(stdp))
        d.seed = seed
        o = make_opts(**opts)
        h = _p()
        _check(lib().mgb_create(C.byref(d), C.byref(o), C.byref(h)))
        return cls(h)

    @staticmethod
    def _gen_desc(g, keep, pops, projs, lifconst, step=1e-4, dt=1e-4, dw=0.01, stdpconst=None, stdp=False,
                  seed=0, stream_seed=None, net_seed=None):
        """Fills GenDesc g; the arrays it points to are appended to keep."""
        P = (GenPop * len(pops))()
        for i, q in enumerate(pops):
            P[i] = GenPop(q["num"], q["type"], q.get("rate", 0.0), q.get("v_min", 0.0), q.get("v_max", 0.0),
                          q.get("cset", 0), q.get("stream", 0))
        J = (GenProj * max(1, len(projs)))()
        for i, q in enumerate(projs):
            J[i] = GenProj(q["src"], q["tar"], q["kind"], q.get("p", 0.0), q.get("k", 0),
                           q.get("wmin", q.get("w", 0.0)), q.get("wmax", q.get("w", 0.0)),
                           q.get("dmin", q.get("d", 0.0)), q.get("dmax", q.get("d", 0.0)),
                           q.get("plastic", 0), q.get("index", i), q.get("seed", seed))
        lc = np.ascontiguousarray(lifconst, dtype=np.float32).reshape(-1)
        sc = np.ascontiguousarray(stdpconst if stdpconst is not None else [0.1, -0.01, 17, 34, 40, 0],
                                  dtype=np.float32)
        ss = None if stream_seed is None else np.ascontiguousarray(stream_seed, dtype=np.uint64)
        keep.extend([P, J, lc, sc, ss])
        g.n_pops = len(pops)
        g.pops = C.cast(P, _p)
        g.n_projs = len(projs)
        g.projs = C.cast(J, _p)
        g.n_csets = lc.size // 30
        g.lifconst = _addr(lc)
        g.stdpconst = _addr(sc)
        g.n_streams = 1 if ss is None else len(ss)
        g.stream_seed = _addr(ss)
        g.step = step
        g.dt = dt
        g.dw = dw
        g.stdp = int(bool(stdp))
        g.seed = seed
        g.net_seed = seed if net_seed is None else net_seed

    @classmethod
    def generated(cls, pops, projs, lifconst, step=1e-4, dt=1e-4, dw=0.01, stdpconst=None, stdp=False,
                  seed=0, stream_seed=None, net_seed=None, **opts):
        """pops: list of dicts (num, type, rate, v_min, v_max, cset, stream);
        projs: list of dicts (src, tar, kind, p, k, wmin, wmax, dmin, dmax, plastic, index, seed)."""
        g, keep = GenDesc(), []
        cls._gen_desc(g, keep, pops, projs, lifconst, step, dt, dw, stdpconst, stdp, seed, stream_seed,
                      net_seed)
        o = make_opts(**opts)
        h = _p()
        _check(lib().mgb_create_generated(C.byref(g), C.byref(o), C.byref(h)))
        return cls(h)

    @classmethod
    def batch_generated(cls, nets, **opts):
        """nets: list of dicts with the keyword arguments of generated() (pops, projs,
        lifconst, step, dt, dw, stdpconst, stdp, seed, stream_seed, net_seed).  Neuron
        ids of network b follow those of network b-1."""
        G = (GenDesc * len(nets))()
        keep = []
        for b, net in enumerate(nets):
            cls._gen_desc(G[b], keep, **net)
        o = make_opts(**opts)
        h = _p()
        _check(lib().mgb_create_batch_generated(G, len(nets), C.byref(o), C.byref(h)))
        return cls(h)

    # -- stepping -------------------------------------------------------------
    def run(self, nsteps):
        _check(lib().mgb_run(self.h, int(nsteps)))

    def sync(self):
        _check(lib().mgb_sync(self.h))

    def set_launch(self, block=0, chunk=0):
        _check(lib().mgb_set_launch(self.h, int(block), int(chunk)))

    def tiles(self):
        """Tile boundaries of the fused path: tile g holds neurons [t[g], t[g + 1])."""
        out = np.zeros(self.info()["tiles"] + 1, dtype=np.int32)
        _check(lib().mgb_read_tiles(self.h, out.ctypes.data))
        return out

    def info(self):
        i = Info()
        _check(lib().mgb_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in Info._fields_}

    # -- read-back ------------------------------------------------------------
    def spikes(self, cap=None):
        """(steps, ids) sorted by (step, id)."""
        if cap is None:
            cap = max(1, self.info()["spikes_recorded"])
        steps = np.zeros(cap, dtype=np.int32)
        ids = np.zeros(cap, dtype=np.int32)
        n = C.c_int64()
        _check(lib().mgb_read_spikes(self.h, cap, _addr(steps), _addr(ids), C.byref(n)))
        return steps[: n.value], ids[: n.value]

    def clear_spikes(self):
        _check(lib().mgb_clear_spikes(self.h))

    def state(self, field):
        n = self.info()["n_neurons"]
        dt = np.float32 if field in (F_V, F_IEXC, F_IINH) else np.int32
        out = np.zeros(n, dtype=dt)
        _check(lib().mgb_read_state(self.h, field, _addr(out)))
        return out

    def fire_counts(self):
        return self.state(F_CNT)

    def csr(self):
        inf = self.info()
        n, s = inf["n_neurons"], inf["n_synapses"]
        rp = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(max(s, 1), dtype=np.int32)
        dl = np.zeros(max(s, 1), dtype=np.int32)
        w = np.zeros(max(s, 1), dtype=np.float32)
        _check(lib().mgb_read_csr(self.h, _addr(rp), _addr(col), _addr(dl), _addr(w)))
        return rp, col[:s], dl[:s], w[:s]

    def weights(self):
        s = self.info()["n_synapses"]
        w = np.zeros(max(s, 1), dtype=np.float32)
        _check(lib().mgb_read_weights(self.h, _addr(w)))
        return w[:s]

    def weights_noalloc(self):
        """mgb_read_weights_noalloc: the same weights, read through the library's fixed
        device scratch (no device allocation)."""
        s = self.info()["n_synapses"]
        w = np.zeros(max(s, 1), dtype=np.float32)
        _check(lib().mgb_read_weights_noalloc(self.h, _addr(w)))
        return w[:s]

    # -- checkpoint -------------------------------------------------------------
    def checkpoint(self):
        """bytes of the simulator's mutable state (mgb_checkpoint_save)."""
        n = C.c_int64()
        _check(lib().mgb_checkpoint_size(self.h, C.byref(n)))
        buf = np.zeros(max(1, n.value), dtype=np.uint8)
        _check(lib().mgb_checkpoint_save(self.h, _addr(buf), n.value))
        return buf[:n.value].tobytes()

    def restore(self, blob):
        """Loads a checkpoint taken from a simulator of the same network and options."""
        buf = np.frombuffer(blob, dtype=np.uint8).copy()
        _check(lib().mgb_checkpoint_load(self.h, _addr(buf), buf.size))

    def save(self, path):
        with open(path, "wb") as f:
            f.write(self.checkpoint())

    def load(self, path):
        with open(path, "rb") as f:
            self.restore(f.read())

    def close(self):
        if self.h:
            lib().mgb_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
