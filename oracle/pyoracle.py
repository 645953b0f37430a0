"""ctypes binding of the CPU oracle (oracle/mgb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path (mgbrain_amd/).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
REF_NETGEN = os.path.join(HERE, "_ref", "ref_netgen")

ACC_F32, ACC_FIXED, ACC_F64 = 0, 1, 2
CONN_REF, CONN_BERNOULLI, CONN_FIXED_INDEGREE, CONN_REF_SHUFFLE = 0, 1, 2, 3
RAW, LIF0, POISSON, LIF, LIFE, LIFB = 0, 1, 2, 3, 4, 5
F_V, F_IEXC, F_IINH, F_REFR, F_LAST, F_CNT = range(6)

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)


class Pop(C.Structure):
    _fields_ = [("num", C.c_int32), ("type", C.c_int32), ("source", C.c_int32)]


class Proj(C.Structure):
    _fields_ = [("src", C.c_int32), ("tar", C.c_int32), ("wmin", C.c_float), ("wmax", C.c_float),
                ("dmin", C.c_float), ("dmax", C.c_float), ("ctype", C.c_float), ("kind", C.c_int32),
                ("p", C.c_float), ("k", C.c_int32), ("plastic", C.c_int32)]


class SynList(C.Structure):
    _fields_ = [("n_syn", C.c_int64), ("src", _i32p), ("tar", _i32p), ("weight", _f32p),
                ("delay", _f32p), ("plastic", _u8p)]


class Csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("s", C.c_int64), ("row_ptr", _i64p), ("col", _i32p),
                ("dly", _i32p), ("w", _f32p), ("plastic", _u8p), ("perm", _i64p),
                ("in_ptr", _i64p), ("in_pos", _i64p), ("max_delay", C.c_int32),
                ("min_delay", C.c_int32)]


class Desc(C.Structure):
    _fields_ = [("n", C.c_int32), ("type", _i32p), ("rate", _f32p), ("v0", _f32p),
                ("cset", _i32p), ("n_cset", C.c_int32), ("lifconst", _f32p),
                ("stdpconst", _f32p), ("stdp", C.c_int32), ("csr", C.POINTER(Csr)),
                ("seed", C.c_uint64), ("acc", C.c_int32), ("scale_exc", C.c_float),
                ("scale_inh", C.c_float)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


class GlibcRand(C.Structure):
    _fields_ = [("r", C.c_uint32 * 31), ("idx", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
        L.orc_mt_next.argtypes = [C.POINTER(MT)]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_glibc_srand.argtypes = [C.POINTER(GlibcRand), C.c_uint32]
        L.orc_glibc_rand_next.argtypes = [C.POINTER(GlibcRand)]
        L.orc_glibc_rand_next.restype = C.c_int32
        L.orc_gen_net.argtypes = [C.POINTER(Pop), C.c_int32, C.POINTER(Proj), C.c_int32, C.c_uint32,
                                  C.c_uint32, C.c_uint64, C.c_float, C.c_float, C.POINTER(SynList)]
        L.orc_gen_net.restype = C.c_int
        L.orc_free_synlist.argtypes = [C.POINTER(SynList)]
        L.orc_build_csr.argtypes = [C.c_int32, C.c_int64, _i32p, _i32p, _f32p, _f32p, _u8p, C.c_float,
                                    C.POINTER(Csr)]
        L.orc_build_csr.restype = C.c_int
        L.orc_free_csr.argtypes = [C.POINTER(Csr)]
        L.orc_create.argtypes = [C.POINTER(Desc)]
        L.orc_create.restype = C.c_void_p
        L.orc_fixk_targets.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                       C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _i32p, C.c_int32,
                                       _i32p, _f32p, _i32p]
        L.orc_in_csr.argtypes = [C.c_int32, C.c_int64, _i32p, _u8p, _i64p, _i64p]
        L.orc_in_csr.restype = C.c_int64
        L.orc_run.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int64, _i64p]
        L.orc_run.restype = C.c_int64
        L.orc_run_mt.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i32p, C.c_int64, _i64p]
        L.orc_run_mt.restype = C.c_int64
        L.orc_get_state.argtypes = [C.c_void_p, C.c_int, _f64p]
        L.orc_get_weights.argtypes = [C.c_void_p, _f64p]
        L.orc_syn_events.argtypes = [C.c_void_p]
        L.orc_syn_events.restype = C.c_int64
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_poisson_u.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_poisson_u.restype = C.c_float
        L.orc_vinit.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.c_float]
        L.orc_vinit.restype = C.c_float
        L.orc_stdp_dw.argtypes = [_f32p, C.c_int32]
        L.orc_stdp_dw.restype = C.c_float
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


# ---------------------------------------------------------------------------
# random number generators
# ---------------------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def mt19937(seed, count):
    g = MT()
    lib().orc_mt_seed(C.byref(g), seed & 0xFFFFFFFF)
    return np.array([lib().orc_mt_next(C.byref(g)) for _ in range(count)], dtype=np.uint64)


def glibc_rand(seed, count):
    g = GlibcRand()
    lib().orc_glibc_srand(C.byref(g), seed & 0xFFFFFFFF)
    return np.array([lib().orc_glibc_rand_next(C.byref(g)) for _ in range(count)], dtype=np.int64)


def poisson_u(seed, nid, step):
    return float(lib().orc_poisson_u(seed, nid, step))


def vinit(seed, local_ids, lo, hi):
    f = lib().orc_vinit
    return np.array([f(seed, int(i), lo, hi) for i in local_ids], dtype=np.float32)


def stdp_dw(stdpconst, dt):
    sc = np.asarray(stdpconst, dtype=np.float32)
    return float(lib().orc_stdp_dw(_ptr(sc, _f32p), int(dt)))


# ---------------------------------------------------------------------------
# network generation and CSR
# ---------------------------------------------------------------------------
@dataclass
class SynapseList:
    src: np.ndarray
    tar: np.ndarray
    weight: np.ndarray
    delay: np.ndarray
    plastic: np.ndarray

    @property
    def n(self):
        return int(self.src.shape[0])


def gen_net(pops, projs, mt_seed=0, rand_seed=1, philox_seed=0, dw=0.01, dt=1e-4) -> SynapseList:
    """pops: list of (num, type, source); projs: list of dicts with Proj fields."""
    P = (Pop * len(pops))(*[Pop(int(a), int(b), int(c)) for a, b, c in pops])
    J = (Proj * max(1, len(projs)))()
    for i, pr in enumerate(projs):
        J[i] = Proj(pr["src"], pr["tar"], pr.get("wmin", pr.get("w", 0.0)), pr.get("wmax", pr.get("w", 0.0)),
                    pr.get("dmin", pr.get("d", 0.0)), pr.get("dmax", pr.get("d", 0.0)), pr.get("ctype", 0.5),
                    pr.get("kind", CONN_REF), pr.get("p", 0.0), pr.get("k", 0), pr.get("plastic", 0))
    out = SynList()
    rc = lib().orc_gen_net(P, len(pops), J, len(projs), mt_seed & 0xFFFFFFFF, rand_seed & 0xFFFFFFFF,
                           philox_seed, dw, dt, C.byref(out))
    if rc != 0:
        raise RuntimeError("orc_gen_net failed")
    n = out.n_syn

    def arr(p, dt_):
        if n == 0:
            return np.zeros(0, dtype=dt_)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    res = SynapseList(arr(out.src, np.int32), arr(out.tar, np.int32), arr(out.weight, np.float32),
                      arr(out.delay, np.float32), arr(out.plastic, np.uint8))
    lib().orc_free_synlist(C.byref(out))
    return res


def fixk_targets(philox_seed, j, sn, k, wmin, wmax, dmin, dmax, tloc, dw=0.01, dt=1e-4, step=1e-4):
    """Draws of the fixed in-degree connector for the local targets tloc of projection j:
    (local source, weight, delay steps) arrays of shape (len(tloc), k), in list order."""
    tl = np.ascontiguousarray(tloc, dtype=np.int32)
    src = np.zeros((tl.size, k), np.int32)
    w = np.zeros((tl.size, k), np.float32)
    d = np.zeros((tl.size, k), np.int32)
    lib().orc_fixk_targets(philox_seed, j, sn, k, wmin, wmax, dmin, dmax, dw, dt, step, _ptr(tl, _i32p), tl.size,
                           _ptr(src, _i32p), _ptr(w, _f32p), _ptr(d, _i32p))
    return src, w, d


@dataclass
class CsrArrays:
    n: int
    row_ptr: np.ndarray
    col: np.ndarray
    dly: np.ndarray
    w: np.ndarray
    plastic: np.ndarray
    perm: np.ndarray
    in_ptr: np.ndarray
    in_pos: np.ndarray
    D: int
    min_delay: int
    _c: object = field(default=None, repr=False)
    _keep: list = field(default_factory=list, repr=False)


def build_csr(n, syn: SynapseList, step=1e-4) -> CsrArrays:
    src = np.ascontiguousarray(syn.src, dtype=np.int32)
    tar = np.ascontiguousarray(syn.tar, dtype=np.int32)
    w = np.ascontiguousarray(syn.weight, dtype=np.float32)
    d = np.ascontiguousarray(syn.delay, dtype=np.float32)
    pl = np.ascontiguousarray(syn.plastic, dtype=np.uint8)
    c = Csr()
    rc = lib().orc_build_csr(n, syn.n, _ptr(src, _i32p), _ptr(tar, _i32p), _ptr(w, _f32p), _ptr(d, _f32p),
                             _ptr(pl, _u8p), step, C.byref(c))
    if rc != 0:
        raise RuntimeError(f"orc_build_csr failed ({rc})")
    s = syn.n

    def arr(p, m, dt_):
        if m == 0:
            return np.zeros(0, dtype=dt_)
        return np.ctypeslib.as_array(p, shape=(m,)).copy()

    res = CsrArrays(n, arr(c.row_ptr, n + 1, np.int64), arr(c.col, s, np.int32), arr(c.dly, s, np.int32),
                    arr(c.w, s, np.float32), arr(c.plastic, s, np.uint8), arr(c.perm, s, np.int64),
                    arr(c.in_ptr, n + 1, np.int64), arr(c.in_pos, s, np.int64), c.max_delay, c.min_delay)
    lib().orc_free_csr(C.byref(c))
    return res


def csr_from_arrays(n, row_ptr, col, dly, w, D, min_delay, plastic=None):
    """CsrArrays over a CSR given as arrays (e.g. read back from the device); with
    plastic flags per slot, also the incoming lists the STDP step needs."""
    col = np.ascontiguousarray(col, dtype=np.int32)
    if plastic is None:
        return CsrArrays(n, row_ptr, col, dly, w, np.zeros(1, np.uint8), np.zeros(1, np.int64),
                         np.zeros(n + 1, np.int64), np.zeros(1, np.int64), D, min_delay)
    pl = np.ascontiguousarray(plastic, dtype=np.uint8)
    in_ptr = np.zeros(n + 1, np.int64)
    in_pos = np.zeros(max(1, int(np.count_nonzero(pl))), np.int64)
    got = lib().orc_in_csr(n, col.size, _ptr(col, _i32p), _ptr(pl, _u8p), _ptr(in_ptr, _i64p), _ptr(in_pos, _i64p))
    if got < 0:
        raise RuntimeError("orc_in_csr failed")
    return CsrArrays(n, row_ptr, col, dly, w, pl, np.zeros(1, np.int64), in_ptr, in_pos, D, min_delay)


def csr_struct(cs: CsrArrays):
    """ctypes Csr view over numpy arrays (arrays kept alive on cs)."""
    c = Csr()
    c.n = cs.n
    c.s = int(cs.col.shape[0])
    keep = []
    for name, t in (("row_ptr", _i64p), ("col", _i32p), ("dly", _i32p), ("w", _f32p), ("plastic", _u8p),
                    ("perm", _i64p), ("in_ptr", _i64p), ("in_pos", _i64p)):
        a = np.ascontiguousarray(getattr(cs, name))
        if a.size == 0:
            a = np.zeros(1, dtype=a.dtype)
        keep.append(a)
        setattr(c, name, a.ctypes.data_as(t))
    c.max_delay = cs.D
    c.min_delay = cs.min_delay
    cs._keep = keep
    cs._c = c
    return c


# ---------------------------------------------------------------------------
# simulation
# ---------------------------------------------------------------------------
class OracleSim:
    def __init__(self, csr: CsrArrays, types, rate, v0, lifconst, stdpconst=None, stdp=False, seed=0,
                 acc=ACC_F32, scale_exc=2.0 ** 30, scale_inh=2.0 ** 30, cset=None):
        self.csr = csr
        self.n = csr.n
        self._types = np.ascontiguousarray(types, dtype=np.int32)
        self._rate = np.ascontiguousarray(rate, dtype=np.float32)
        self._v0 = np.ascontiguousarray(v0, dtype=np.float32)
        lc = np.ascontiguousarray(lifconst, dtype=np.float32).reshape(-1)
        self._lc = lc
        self._sc = np.ascontiguousarray(stdpconst if stdpconst is not None else [0.1, -0.01, 17, 34, 40, 0],
                                        dtype=np.float32)
        self._cset = None if cset is None else np.ascontiguousarray(cset, dtype=np.int32)
        self._c = csr_struct(csr)
        d = Desc()
        d.n = self.n
        d.type = _ptr(self._types, _i32p)
        d.rate = _ptr(self._rate, _f32p)
        d.v0 = _ptr(self._v0, _f32p)
        d.cset = _ptr(self._cset, _i32p)
        d.n_cset = lc.size // 30
        d.lifconst = _ptr(lc, _f32p)
        d.stdpconst = _ptr(self._sc, _f32p)
        d.stdp = int(bool(stdp))
        d.csr = C.pointer(self._c)
        d.seed = seed
        d.acc = acc
        d.scale_exc = scale_exc
        d.scale_inh = scale_inh
        self._d = d
        self.h = lib().orc_create(C.byref(d))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.acc = acc

    def run(self, nsteps, cap=None):
        """Returns a list (one per step) of ascending spike-id arrays."""
        if cap is None:
            cap = max(1024, self.n * nsteps)
        ids = np.zeros(cap, dtype=np.int32)
        off = np.zeros(nsteps + 1, dtype=np.int64)
        w = lib().orc_run(self.h, nsteps, _ptr(ids, _i32p), cap, _ptr(off, _i64p))
        if w < 0:
            raise RuntimeError("spike capacity exceeded")
        return [ids[off[k]:off[k + 1]].copy() for k in range(nsteps)]

    def run_mt(self, nsteps, nthreads, record=False, cap=None):
        """Multithreaded steps of a static network (every mode; the same results as
        run()).  record: also return the per-step ascending spike-id arrays."""
        if not record:
            n = lib().orc_run_mt(self.h, nsteps, nthreads, None, 0, None)
            if n < 0:
                raise RuntimeError("run_mt needs an oracle without STDP")
            return int(n)
        if cap is None:
            cap = max(1024, self.n * nsteps // 8)
        ids = np.zeros(cap, dtype=np.int32)
        off = np.zeros(nsteps + 1, dtype=np.int64)
        n = lib().orc_run_mt(self.h, nsteps, nthreads, _ptr(ids, _i32p), cap, _ptr(off, _i64p))
        if n == -2:
            raise RuntimeError("spike capacity exceeded")
        if n < 0:
            raise RuntimeError("run_mt needs an oracle without STDP")
        return [ids[off[k]:off[k + 1]].copy() for k in range(nsteps)]

    def state(self, field_):
        out = np.zeros(self.n, dtype=np.float64)
        lib().orc_get_state(self.h, field_, _ptr(out, _f64p))
        return out

    def weights(self):
        out = np.zeros(max(1, self.csr.col.size), dtype=np.float64)
        lib().orc_get_weights(self.h, _ptr(out, _f64p))
        return out[: self.csr.col.size]

    def syn_events(self):
        return int(lib().orc_syn_events(self.h))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# reference netgen (oracle/_ref/ref_netgen, built from the reference sources)
# ---------------------------------------------------------------------------
def ref_netgen_available():
    return os.path.exists(REF_NETGEN)


def run_ref_netgen(spec_lines, workdir):
    spec = os.path.join(workdir, "spec.txt")
    out = os.path.join(workdir, "out.bin")
    with open(spec, "w") as f:
        f.write("\n".join(spec_lines) + "\n")
    subprocess.run([REF_NETGEN, spec, out], check=True, stdout=subprocess.DEVNULL)
    raw = open(out, "rb").read()
    o = 0

    def take(dt_, cnt):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt_, count=cnt, offset=o).copy()
        o += a.nbytes
        return a

    nn, ns = (int(x) for x in take(np.int64, 2))
    res = {"n": nn, "types": take(np.int32, nn), "nxt": take(np.int32, nn), "pre": take(np.int32, nn),
           "src": take(np.int32, ns), "tar": take(np.int32, ns), "weight": take(np.float32, ns),
           "delay": take(np.float32, ns)}
    res["nlifconst"] = int(take(np.int32, 1)[0])
    res["lifconst"] = take(np.float32, 30)
    res["life1"] = take(np.float32, 30)
    res["life2"] = take(np.float32, 30)
    res["life3"] = take(np.float32, 30)
    res["stdp1"] = take(np.float32, 6)
    return res
