"""ctypes front end of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker or as the
timed CPU baseline.  The product (libcdx.so) never calls it.

The oracle restates the reference's OpenMP path:
  aabb          AABBGenerator.cu:84-130
  broadphase    Broadphase.cu:344-545
  narrowphase   Narrowphase.cu:316-595
  mpr_steps     the narrowphase's support steps per pair (MPR work histogram)
  mpr_flops     the narrowphase's FP64 operations per pair (FLOP census)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# Reference default of setParallelConfiguration (Broadphase.cu:40).
DEFAULT_FLAGS = (1, 0, 1, 0, 0, 0, 0, 0, 1, 0)
ALL_PARALLEL = (1,) * 10


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


class _BP(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("m", C.c_uint64), ("a", C.c_uint64), ("pdup", C.c_uint64), ("p", C.c_uint64),
        ("bin_counts", C.POINTER(C.c_uint32)), ("bins_intersected", C.POINTER(C.c_uint32)),
        ("sorted_keys", C.POINTER(C.c_uint32)), ("sorted_bodies", C.POINTER(C.c_uint32)),
        ("active_keys", C.POINTER(C.c_uint32)), ("active_counts", C.POINTER(C.c_uint32)),
        ("bin_end", C.POINTER(C.c_uint32)), ("num_contactd", C.POINTER(C.c_uint32)),
        ("pairs", C.POINTER(C.c_int64)),
        ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
        ("origin", C.c_double * 3), ("inv", C.c_double * 3),
    ]


class _NP(C.Structure):
    _fields_ = [
        ("p", C.c_uint64), ("c", C.c_uint64),
        ("norm", C.c_void_p), ("cpta", C.c_void_p), ("cptb", C.c_void_p), ("dpth", C.c_void_p),
        ("bids", C.POINTER(C.c_int32)), ("pairs_permuted", C.POINTER(C.c_int64)),
        ("hit", C.POINTER(C.c_uint8)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def planted(k: int):
    """The oracle with planted slip k in the MPR contact point (ORACLE_PLANT
    in oracle.cpp), for tests that check the pins catch it."""
    path = os.path.join(HERE, "_build", f"liboracle_plant{k}.so")
    subprocess.check_call(["make", "-s", "-C", HERE, "planted"])  # rebuilt when oracle.cpp changed
    return _configure(C.CDLL(path))


def _configure(L):
    if True:
        vp = C.c_void_p
        for suf in ("f64", "f32"):
            getattr(L, f"oracle_aabb_{suf}").argtypes = [C.c_uint64] + [vp] * 9
            getattr(L, f"oracle_aabb_{suf}").restype = C.c_int
            getattr(L, f"oracle_broadphase_{suf}").argtypes = [vp, C.c_uint64, C.c_int, C.c_int, C.c_int, vp]
            getattr(L, f"oracle_broadphase_{suf}").restype = C.POINTER(_BP)
            rt = C.c_double if suf == "f64" else C.c_float
            getattr(L, f"oracle_narrowphase_{suf}").argtypes = [C.c_uint64] + [vp] * 10 + [rt]
            getattr(L, f"oracle_narrowphase_{suf}").restype = C.POINTER(_NP)
        L.oracle_bp_free.argtypes = [C.POINTER(_BP)]
        L.oracle_np_free.argtypes = [C.POINTER(_NP)]
        L.oracle_mpr_f64.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.oracle_mpr_f64.restype = C.c_int
        L.oracle_support_f64.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        L.oracle_quat_rotate_f64.argtypes = [vp, vp, vp]
        L.oracle_quat_mult_f64.argtypes = [vp, vp, vp]
        L.oracle_mpr_steps_f64.argtypes = [C.c_uint64] + [vp] * 11
        L.oracle_max_threads.restype = C.c_int
        L.oracle_mpr_flops_f64.argtypes = [C.c_uint64] + [vp] * 10 + [C.c_double, vp]
        L.oracle_mpr_portal_f64.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_mpr_portal_f64.restype = C.c_int
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def aabb(scene) -> np.ndarray:
    """AABBs of every shape as the reference lays them out: (2n, 3), mins then maxes."""
    s = scene
    n = len(s.T)
    out = np.zeros((2 * n, 3), dtype=s.dtype)
    getattr(lib(), f"oracle_aabb_{_suffix(s.dtype)}")(
        n, _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID), _p(s.pos), _p(s.rot), _p(out))
    return out


@dataclass
class BPResult:
    aabb_shifted: np.ndarray
    bin_counts: np.ndarray
    bins_intersected: np.ndarray
    sorted_keys: np.ndarray
    sorted_bodies: np.ndarray
    active_keys: np.ndarray
    active_counts: np.ndarray
    bin_end: np.ndarray
    num_contactd: np.ndarray
    pairs: np.ndarray
    pdup: int
    bmin: np.ndarray
    bmax: np.ndarray
    origin: np.ndarray
    inv: np.ndarray


def broadphase(aabb_data: np.ndarray, grid, flags=DEFAULT_FLAGS) -> BPResult:
    """detectPossibleCollisions (Broadphase.cu:344-545) on a copy of aabb_data."""
    a = np.ascontiguousarray(aabb_data).copy()
    n = a.shape[0] // 2
    fl = np.asarray(flags, dtype=np.int32)
    L = lib()
    r = getattr(L, f"oracle_broadphase_{_suffix(a.dtype)}")(_p(a), n, int(grid[0]), int(grid[1]), int(grid[2]), _p(fl))
    try:
        c = r.contents
        u = np.uint32
        res = BPResult(
            aabb_shifted=a,
            bin_counts=_arr(c.bin_counts, c.n, u),
            bins_intersected=_arr(c.bins_intersected, c.n, u),
            sorted_keys=_arr(c.sorted_keys, c.m, u),
            sorted_bodies=_arr(c.sorted_bodies, c.m, u),
            active_keys=_arr(c.active_keys, c.a, u),
            active_counts=_arr(c.active_counts, c.a, u),
            bin_end=_arr(c.bin_end, c.a, u),
            num_contactd=_arr(c.num_contactd, c.a, u),
            pairs=_arr(c.pairs, c.p, np.int64),
            pdup=int(c.pdup),
            bmin=np.array(list(c.bmin), dtype=np.float32),
            bmax=np.array(list(c.bmax), dtype=np.float32),
            origin=np.array(list(c.origin)),
            inv=np.array(list(c.inv)),
        )
    finally:
        L.oracle_bp_free(r)
    return res


@dataclass
class NPResult:
    norm: np.ndarray
    cpta: np.ndarray
    cptb: np.ndarray
    dpth: np.ndarray
    bids: np.ndarray
    pairs_permuted: np.ndarray
    hit: np.ndarray
    count: int


def narrowphase(scene, pairs: np.ndarray, envelope: float = 0.0) -> NPResult:
    """DoNarrowphase (Narrowphase.cu:511-595)."""
    s = scene
    pr = np.ascontiguousarray(pairs, dtype=np.int64)
    L = lib()
    suf = _suffix(s.dtype)
    r = getattr(L, f"oracle_narrowphase_{suf}")(
        len(pr), _p(pr), _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID), _p(s.active), _p(s.pos),
        _p(s.rot), envelope)
    try:
        c = r.contents
        dt = s.dtype
        fp = C.POINTER(C.c_double if dt == np.float64 else C.c_float)
        k = int(c.c)
        res = NPResult(
            norm=_arr(C.cast(c.norm, fp), 3 * k, dt).reshape(k, 3),
            cpta=_arr(C.cast(c.cpta, fp), 3 * k, dt).reshape(k, 3),
            cptb=_arr(C.cast(c.cptb, fp), 3 * k, dt).reshape(k, 3),
            dpth=_arr(C.cast(c.dpth, fp), k, dt),
            bids=_arr(c.bids, 2 * k, np.int32).reshape(k, 2),
            pairs_permuted=_arr(c.pairs_permuted, int(c.p), np.int64),
            hit=_arr(c.hit, int(c.p), np.uint8),
            count=k,
        )
    finally:
        L.oracle_np_free(r)
    return res


def mpr_steps(scene, pairs: np.ndarray) -> np.ndarray:
    """Support steps (pairs of support-map calls) the reference narrowphase
    spends on each candidate pair, contact points included; 0 where it ends
    before the MPR.  FP64 scenes only."""
    s = scene
    assert s.dtype == np.float64
    pr = np.ascontiguousarray(pairs, dtype=np.int64)
    out = np.zeros(len(pr), dtype=np.int32)
    lib().oracle_mpr_steps_f64(len(pr), _p(pr), _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID),
                               _p(s.active), _p(s.pos), _p(s.rot), _p(out))
    return out


def mpr_flops(scene, pairs: np.ndarray, envelope: float = 0.0) -> np.ndarray:
    """FP64 operations (+, -, *, /, sqrt) function_MPR_Store performs per
    candidate pair once the two shapes are placed in the world (MPR and
    contact points; per-shape placement excluded); 0 where it ends earlier.
    FP64 scenes only."""
    s = scene
    assert s.dtype == np.float64
    pr = np.ascontiguousarray(pairs, dtype=np.int64)
    out = np.zeros(len(pr), dtype=np.uint64)
    lib().oracle_mpr_flops_f64(len(pr), _p(pr), _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID),
                               _p(s.active), _p(s.pos), _p(s.rot), float(envelope), _p(out))
    return out


def mpr(tA, AX, AY, AR, tB, BX, BY, BR):
    """CollideAndFindPoint (Narrowphase.cu:122-314) in FP64 on world-placed shapes."""
    f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    n = np.zeros(3)
    p = np.zeros(3)
    d = np.zeros(1)
    hit = lib().oracle_mpr_f64(int(tA), _p(f(AX)), _p(f(AY)), _p(f(AR)), int(tB), _p(f(BX)), _p(f(BY)), _p(f(BR)),
                               _p(n), _p(p), _p(d))
    return bool(hit), n, p, float(d[0])


@dataclass
class Portal:
    """Final MPR portal of a hit that reached the barycentric contact point
    (Narrowphase.cu:265-289): vertices v (4x3, Minkowski-difference points,
    v[0] the centre difference), the support points of A and B making each
    vertex (a, b: 4x3), the weights w as computed, the portal normal n, and
    whether the sum <= 0 fallback was taken.  filled is False when the MPR
    ended elsewhere (miss, degenerate early hit)."""
    v: np.ndarray
    a: np.ndarray
    b: np.ndarray
    w: np.ndarray
    n: np.ndarray
    fallback: bool
    filled: bool


def mpr_portal(tA, AX, AY, AR, tB, BX, BY, BR, L=None):
    """mpr() plus the final portal; L selects another build of the oracle
    (planted())."""
    f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    n = np.zeros(3)
    p = np.zeros(3)
    d = np.zeros(1)
    pt = np.zeros(45)
    L = L or lib()
    hit = L.oracle_mpr_portal_f64(int(tA), _p(f(AX)), _p(f(AY)), _p(f(AR)), int(tB), _p(f(BX)), _p(f(BY)), _p(f(BR)),
                                  _p(n), _p(p), _p(d), _p(pt))
    portal = Portal(v=pt[0:12].reshape(4, 3), a=pt[12:24].reshape(4, 3), b=pt[24:36].reshape(4, 3), w=pt[36:40].copy(),
                    n=pt[40:43].copy(), fallback=bool(pt[43]), filled=bool(pt[44]))
    return bool(hit), n, p, float(d[0]), portal


def support(t, X, Y, Rq, direction):
    f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(3)
    lib().oracle_support_f64(int(t), _p(f(X)), _p(f(Y)), _p(f(Rq)), _p(f(direction)), _p(out))
    return out


def quat_rotate(v, q):
    out = np.zeros(3)
    lib().oracle_quat_rotate_f64(_p(np.ascontiguousarray(v, dtype=np.float64)),
                                 _p(np.ascontiguousarray(q, dtype=np.float64)), _p(out))
    return out


def quat_mult(a, b):
    out = np.zeros(4)
    lib().oracle_quat_mult_f64(_p(np.ascontiguousarray(a, dtype=np.float64)),
                               _p(np.ascontiguousarray(b, dtype=np.float64)), _p(out))
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())
