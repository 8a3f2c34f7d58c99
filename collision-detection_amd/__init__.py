"""Python front end of libcdx.so (include/cdx.h) for tests and benchmarks.

The product is the C ABI; this module only binds it with ctypes and moves
numpy arrays across.  It never falls back to a CPU path: if libcdx.so is
missing or no GPU is visible, every call raises.

Load it from the repository root with::

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "collision_detection_amd", "collision-detection_amd/__init__.py",
        submodule_search_locations=["collision-detection_amd"])
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcdx.so")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

REAL_F64, REAL_F32 = 0, 1
NP_NATIVE, NP_MIXED, NP_FAST = 0, 1, 2
PAIRS_AUTO, PAIRS_BINS, PAIRS_CELLS = 0, 1, 2
FLAG_CELL_SEARCH, FLAG_NO_ACTIVE_BINS = 0x100, 0x200
STAGES = ["aabb", "bp_bounds", "bp_count", "bp_scan", "bp_emit", "bp_sort", "bp_runs", "bp_pairs",
          "bp_pairsort", "np_main", "np_resolve", "np_compact", "bp_pairscan", "bp_pairemit", "np_mpr"]

# Every extern "C" symbol include/cdx.h declares.
EXPORTS = [
    "cdx_version", "cdx_status_string", "cdx_device_count", "cdx_ctx_create", "cdx_ctx_destroy",
    "cdx_ctx_last_error", "cdx_generate_aabb", "cdx_detect_possible_collisions", "cdx_read_possible_collisions",
    "cdx_detect_possible_collisions_noalloc",
    "cdx_do_narrowphase", "cdx_read_contacts", "cdx_mpr_store", "cdx_scene_create", "cdx_scene_destroy", "cdx_scene_last_error",
    "cdx_scene_set_shapes", "cdx_scene_set_bodies", "cdx_scene_update_bodies", "cdx_scene_set_aabbs",
    "cdx_scene_set_pairs", "cdx_scene_set_bins", "cdx_scene_set_envelope", "cdx_scene_set_narrowphase_mode",
    "cdx_scene_set_profiling", "cdx_scene_set_pair_search", "cdx_scene_aabb", "cdx_scene_broadphase", "cdx_scene_narrowphase",
    "cdx_scene_step", "cdx_scene_sync", "cdx_scene_counts", "cdx_scene_read_aabb", "cdx_scene_read_bins",
    "cdx_scene_read_pairs", "cdx_scene_read_pairs_permuted", "cdx_scene_read_mpr_queue", "cdx_scene_read_contacts",
    "cdx_scene_stage_ms",
    "cdx_scene_device_bytes", "cdx_batch_create", "cdx_batch_destroy", "cdx_batch_last_error",
    "cdx_batch_add_scene", "cdx_batch_scene_bins", "cdx_batch_scene_count", "cdx_batch_update_bodies",
    "cdx_batch_set_envelope", "cdx_batch_set_narrowphase_mode", "cdx_batch_select",
    "cdx_batch_step",
    "cdx_batch_finish", "cdx_batch_scene_counts", "cdx_batch_read_pairs", "cdx_batch_read_contacts",
    "cdx_batch_step_ms", "cdx_batch_device_bytes", "cdx_batch_scene_bytes", "cdx_server_create",
    "cdx_server_destroy", "cdx_server_last_error", "cdx_server_add_scene", "cdx_server_scene_info",
    "cdx_server_submit", "cdx_server_poll", "cdx_server_wait", "cdx_server_result", "cdx_server_read_contacts",
    "cdx_server_release", "cdx_server_stats", "cdx_server_memory", "cdx_server_pose_buffer",
    "cdx_server_submit_buffer", "cdx_server_drop_buffer",
]


class ServerConfig(C.Structure):
    _fields_ = [
        ("real_kind", C.c_int), ("np_mode", C.c_int), ("tau", C.c_double), ("envelope", C.c_double),
        ("hbm_budget", C.c_uint64), ("max_batch_shapes", C.c_uint64), ("max_pending", C.c_uint64),
        ("use_graph", C.c_int),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64), ("steps", C.c_uint64),
        ("stepped_requests", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
        ("max_ms", C.c_double), ("busy_ms", C.c_double), ("window_s", C.c_double), ("host_ms", C.c_double * 6),
        ("class_completed", C.c_uint64 * 4), ("class_steps", C.c_uint64 * 4), ("class_p50_ms", C.c_double * 4),
        ("class_p99_ms", C.c_double * 4), ("class_busy_ms", C.c_double * 4),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "host_ms" and not k.startswith("class_")}
        d["host_ms"] = dict(zip(["upload", "select", "launch", "wait", "block", "readback"], list(self.host_ms)))
        d["classes"] = [{"completed": self.class_completed[c], "steps": self.class_steps[c],
                         "p50_ms": self.class_p50_ms[c], "p99_ms": self.class_p99_ms[c],
                         "busy_ms": self.class_busy_ms[c]} for c in range(4)]
        return d


class CdxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"cdx status {status}: {msg}")
        self.status = status


class Counts(C.Structure):
    _fields_ = [
        ("n_shapes", C.c_uint64), ("bin_entries", C.c_uint64), ("active_bins", C.c_uint64),
        ("pairs", C.c_uint64), ("contacts", C.c_uint64), ("resolved", C.c_uint64),
        ("key_max", C.c_uint32), ("flags", C.c_uint32),
        ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
    ]


def build(arch: str = "gfx950", force: bool = False) -> str:
    """Compile libcdx.so in-tree with hipcc (cross-compiles without a GPU);
    force recompiles every source even when the library is up to date."""
    subprocess.check_call(["make", "-s", "-C", CSRC, f"ARCH={arch}", "-j8"] + (["-B"] if force else []))
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CdxError(-1, f"{LIB_PATH} is not built: run build() first")
    L = C.CDLL(LIB_PATH)
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    sig = {
        "cdx_version": ([], i32),
        "cdx_status_string": ([i32], C.c_char_p),
        "cdx_device_count": ([vp], i32),
        "cdx_ctx_create": ([i32, vp], i32),
        "cdx_ctx_destroy": ([vp], i32),
        "cdx_ctx_last_error": ([vp], C.c_char_p),
        "cdx_generate_aabb": ([vp, i32, u64, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp], i32),
        "cdx_detect_possible_collisions": ([vp, i32, vp, u64, i32, i32, i32, vp], i32),
        "cdx_read_possible_collisions": ([vp, vp, u64], i32),
        "cdx_detect_possible_collisions_noalloc": ([vp, i32, vp, u64, i32, i32, i32, vp, u64, vp, vp], i32),
        "cdx_do_narrowphase": ([vp, i32, i32, C.c_double, u64, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, u64, vp],
                               i32),
        "cdx_read_contacts": ([vp, vp, vp, vp, vp, vp, u64], i32),
        "cdx_mpr_store": ([vp, i32, i32, C.c_double, u64, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, u64, vp, vp,
                           vp, vp, vp, vp], i32),
        "cdx_scene_create": ([vp, i32, vp], i32),
        "cdx_scene_destroy": ([vp], i32),
        "cdx_scene_last_error": ([vp], C.c_char_p),
        "cdx_scene_set_shapes": ([vp, u64, vp, vp, vp, vp, vp, vp], i32),
        "cdx_scene_set_bodies": ([vp, u64, vp, vp, vp], i32),
        "cdx_scene_update_bodies": ([vp, u64, vp, vp, vp], i32),
        "cdx_scene_set_aabbs": ([vp, u64, vp], i32),
        "cdx_scene_set_pairs": ([vp, u64, vp], i32),
        "cdx_scene_set_bins": ([vp, i32, i32, i32], i32),
        "cdx_scene_set_envelope": ([vp, C.c_double], i32),
        "cdx_scene_set_narrowphase_mode": ([vp, i32, C.c_double], i32),
        "cdx_scene_set_profiling": ([vp, i32], i32),
        "cdx_scene_set_pair_search": ([vp, i32], i32),
        "cdx_scene_aabb": ([vp], i32),
        "cdx_scene_broadphase": ([vp], i32),
        "cdx_scene_narrowphase": ([vp], i32),
        "cdx_scene_step": ([vp], i32),
        "cdx_scene_sync": ([vp], i32),
        "cdx_scene_counts": ([vp, vp], i32),
        "cdx_scene_read_aabb": ([vp, i32, vp], i32),
        "cdx_scene_read_bins": ([vp, vp, vp, vp, vp, vp], i32),
        "cdx_scene_read_pairs": ([vp, vp], i32),
        "cdx_scene_read_pairs_permuted": ([vp, vp], i32),
        "cdx_scene_read_mpr_queue": ([vp, vp, u64, vp], i32),
        "cdx_scene_read_contacts": ([vp, vp, vp, vp, vp, vp], i32),
        "cdx_scene_stage_ms": ([vp, vp], i32),
        "cdx_scene_device_bytes": ([vp, vp], i32),
        "cdx_batch_create": ([vp, i32, vp], i32),
        "cdx_batch_destroy": ([vp], i32),
        "cdx_batch_last_error": ([vp], C.c_char_p),
        "cdx_batch_add_scene": ([vp, u64, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, i32, i32, i32, vp], i32),
        "cdx_batch_scene_bins": ([vp, C.c_uint32, vp], i32),
        "cdx_batch_scene_count": ([vp, vp], i32),
        "cdx_batch_update_bodies": ([vp, C.c_uint32, u64, vp, vp, vp], i32),
        "cdx_batch_set_envelope": ([vp, C.c_double], i32),
        "cdx_batch_set_narrowphase_mode": ([vp, i32, C.c_double], i32),
        "cdx_batch_select": ([vp, vp, u64], i32),
        "cdx_batch_step": ([vp, i32], i32),
        "cdx_batch_finish": ([vp, vp], i32),
        "cdx_batch_scene_counts": ([vp, u64, vp], i32),
        "cdx_batch_read_pairs": ([vp, u64, vp, u64], i32),
        "cdx_batch_read_contacts": ([vp, u64, vp, vp, vp, vp, vp, u64], i32),
        "cdx_batch_step_ms": ([vp, vp], i32),
        "cdx_batch_device_bytes": ([vp, vp], i32),
        "cdx_batch_scene_bytes": ([vp, u64, u64, vp], i32),
        "cdx_server_create": ([vp, vp, vp], i32),
        "cdx_server_destroy": ([vp], i32),
        "cdx_server_last_error": ([vp], C.c_char_p),
        "cdx_server_add_scene": ([vp, u64, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, i32, i32, i32, vp], i32),
        "cdx_server_scene_info": ([vp, C.c_uint32, vp, vp, vp, vp], i32),
        "cdx_server_submit": ([vp, C.c_uint32, u64, vp, vp, vp, vp], i32),
        "cdx_server_poll": ([vp, u64, vp], i32),
        "cdx_server_wait": ([vp, u64, C.c_double, vp], i32),
        "cdx_server_result": ([vp, u64, vp, vp], i32),
        "cdx_server_read_contacts": ([vp, u64, vp, vp, vp, vp, vp, u64], i32),
        "cdx_server_release": ([vp, u64], i32),
        "cdx_server_stats": ([vp, vp, i32], i32),
        "cdx_server_memory": ([vp, vp, vp, vp], i32),
        "cdx_server_pose_buffer": ([vp, C.c_uint32, u64, vp, vp, vp], i32),
        "cdx_server_submit_buffer": ([vp, u64, vp], i32),
        "cdx_server_drop_buffer": ([vp, u64], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def _kind(dtype):
    return REAL_F64 if np.dtype(dtype) == np.float64 else REAL_F32


def device_count() -> int:
    c = C.c_int(0)
    lib().cdx_device_count(C.byref(c))
    return c.value


class Context:
    """One cdx_ctx on a device; also hosts the drop-in one-shot calls."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        st = lib().cdx_ctx_create(device, C.byref(self.h))
        if st:
            raise CdxError(st, lib().cdx_status_string(st).decode())

    def close(self):
        if self.h:
            lib().cdx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _chk(self, st):
        if st:
            raise CdxError(st, lib().cdx_ctx_last_error(self.h).decode())

    # -- drop-in calls ------------------------------------------------------
    def generate_aabb(self, s) -> np.ndarray:
        out = np.zeros((2 * s.n, 3), dtype=s.dtype)
        self._chk(lib().cdx_generate_aabb(self.h, _kind(s.dtype), s.n, _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R),
                                          _p(s.ID), s.nb, _p(s.pos), _p(s.rot), _p(out)))
        return out

    def detect_possible_collisions(self, aabb: np.ndarray, bins) -> np.ndarray:
        """Runs on aabb in place (shifted like the reference) and returns the pair list."""
        assert aabb.flags.c_contiguous
        npairs = C.c_uint64(0)
        self._chk(lib().cdx_detect_possible_collisions(self.h, _kind(aabb.dtype), _p(aabb), aabb.shape[0] // 2,
                                                       int(bins[0]), int(bins[1]), int(bins[2]), C.byref(npairs)))
        out = np.zeros(npairs.value, dtype=np.int64)
        self._chk(lib().cdx_read_possible_collisions(self.h, _p(out), npairs.value))
        return out

    def do_narrowphase(self, s, pairs: np.ndarray, envelope=0.0, mode=NP_NATIVE):
        """Returns (norm, cpta, cptb, dpth, bids); pairs is permuted in place."""
        assert pairs.dtype == np.int64 and pairs.flags.c_contiguous
        nc = C.c_uint64(0)
        self._chk(lib().cdx_do_narrowphase(self.h, _kind(s.dtype), mode, float(envelope), s.n, _p(s.T), _p(s.A),
                                           _p(s.B), _p(s.C), _p(s.R), _p(s.ID), s.nb, _p(s.active), _p(s.pos),
                                           _p(s.rot), _p(pairs), len(pairs), C.byref(nc)))
        k = nc.value
        dt = s.dtype
        norm, pa, pb = (np.zeros((k, 3), dt) for _ in range(3))
        d = np.zeros(k, dt)
        ids = np.zeros((k, 2), np.int32)
        self._chk(lib().cdx_read_contacts(self.h, _p(norm), _p(pa), _p(pb), _p(d), _p(ids), k))
        return norm, pa, pb, d, ids

    def mpr_store(self, s, pairs: np.ndarray, envelope=0.0, mode=NP_NATIVE, fill=0.0, active_fill=1):
        """Narrowphase::host_MPR_Store: per-pair arrays at the pair index,
        written only for contacts (others keep `fill` / `active_fill`).
        Returns (contact_active, norm, cpta, cptb, dpth, bids)."""
        pairs = np.ascontiguousarray(pairs, dtype=np.int64)
        P = len(pairs)
        dt = s.dtype
        act = np.full(P, active_fill, np.uint32)
        norm, pa, pb = (np.full((P, 3), fill, dt) for _ in range(3))
        d = np.full(P, fill, dt)
        ids = np.full((P, 2), -1, np.int32)
        self._chk(lib().cdx_mpr_store(self.h, _kind(dt), mode, float(envelope), s.n, _p(s.T), _p(s.A), _p(s.B),
                                      _p(s.C), _p(s.R), _p(s.ID), s.nb, _p(s.active), _p(s.pos), _p(s.rot), _p(pairs),
                                      P, _p(act), _p(norm), _p(pa), _p(pb), _p(d), _p(ids)))
        return act, norm, pa, pb, d, ids


class Scene:
    """A device-resident scene (cdx_scene)."""

    def __init__(self, ctx: Context, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.h = C.c_void_p()
        ctx._chk(lib().cdx_scene_create(ctx.h, _kind(dtype), C.byref(self.h)))
        self.n = 0

    def close(self):
        if self.h:
            lib().cdx_scene_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _chk(self, st):
        if st:
            raise CdxError(st, lib().cdx_scene_last_error(self.h).decode())

    def load(self, s, bins=None):
        assert np.dtype(s.dtype) == self.dtype
        self._keep = s
        self.n = s.n
        self._chk(lib().cdx_scene_set_shapes(self.h, s.n, _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID)))
        self._chk(lib().cdx_scene_set_bodies(self.h, s.nb, _p(s.pos), _p(s.rot), _p(s.active)))
        g = bins if bins is not None else s.grid
        self.set_bins(*g)
        return self

    def set_bins(self, gx, gy, gz):
        self._chk(lib().cdx_scene_set_bins(self.h, int(gx), int(gy), int(gz)))

    def set_aabbs(self, aabb):
        aabb = np.ascontiguousarray(aabb, dtype=self.dtype)
        self.n = aabb.shape[0] // 2
        self._chk(lib().cdx_scene_set_aabbs(self.h, self.n, _p(aabb)))

    def set_pairs(self, pairs):
        pairs = np.ascontiguousarray(pairs, dtype=np.int64)
        self._chk(lib().cdx_scene_set_pairs(self.h, len(pairs), _p(pairs)))

    def update_bodies(self, idx, pos, rot):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        pos = np.ascontiguousarray(pos, dtype=self.dtype)
        rot = np.ascontiguousarray(rot, dtype=self.dtype)
        self._chk(lib().cdx_scene_update_bodies(self.h, len(idx), _p(idx), _p(pos), _p(rot)))

    def set_envelope(self, e):
        self._chk(lib().cdx_scene_set_envelope(self.h, float(e)))

    def set_mode(self, mode, tau=1e-2):
        self._chk(lib().cdx_scene_set_narrowphase_mode(self.h, int(mode), float(tau)))

    def set_profiling(self, on=True):
        self._chk(lib().cdx_scene_set_profiling(self.h, 1 if on else 0))

    def set_pair_search(self, mode):
        """PAIRS_AUTO, PAIRS_BINS or PAIRS_CELLS (cdx_scene_set_pair_search)."""
        self._chk(lib().cdx_scene_set_pair_search(self.h, int(mode)))

    def aabb(self):
        self._chk(lib().cdx_scene_aabb(self.h))

    def broadphase(self):
        self._chk(lib().cdx_scene_broadphase(self.h))

    def narrowphase(self):
        self._chk(lib().cdx_scene_narrowphase(self.h))

    def step(self):
        self._chk(lib().cdx_scene_step(self.h))

    def sync(self):
        self._chk(lib().cdx_scene_sync(self.h))

    def counts(self) -> Counts:
        c = Counts()
        self._chk(lib().cdx_scene_counts(self.h, C.byref(c)))
        return c

    def read_aabb(self, shifted=False):
        out = np.zeros((2 * self.n, 3), dtype=self.dtype)
        self._chk(lib().cdx_scene_read_aabb(self.h, 1 if shifted else 0, _p(out)))
        return out

    def read_bins(self):
        c = self.counts()
        bc = np.zeros(self.n, np.uint32)
        sk = np.zeros(c.bin_entries, np.uint32)
        sb = np.zeros(c.bin_entries, np.uint32)
        ak = np.zeros(c.active_bins, np.uint32)
        ac = np.zeros(c.active_bins, np.uint32)
        self._chk(lib().cdx_scene_read_bins(self.h, _p(bc), _p(sk), _p(sb), _p(ak), _p(ac)))
        return bc, sk, sb, ak, ac

    def read_pairs(self):
        out = np.zeros(self.counts().pairs, np.int64)
        self._chk(lib().cdx_scene_read_pairs(self.h, _p(out)))
        return out

    def read_pairs_permuted(self):
        out = np.zeros(self.counts().pairs, np.int64)
        self._chk(lib().cdx_scene_read_pairs_permuted(self.h, _p(out)))
        return out

    def read_mpr_queue(self):
        """Indices (into read_pairs()) of the pairs the last narrowphase's
        reference-order MPR launch ran on."""
        n = C.c_uint64(0)
        self._chk(lib().cdx_scene_read_mpr_queue(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint32)
        self._chk(lib().cdx_scene_read_mpr_queue(self.h, _p(out), n.value, C.byref(n)))
        return out

    def read_contacts(self):
        k = self.counts().contacts
        dt = self.dtype
        norm, pa, pb = (np.zeros((k, 3), dt) for _ in range(3))
        d = np.zeros(k, dt)
        ids = np.zeros((k, 2), np.int32)
        self._chk(lib().cdx_scene_read_contacts(self.h, _p(norm), _p(pa), _p(pb), _p(d), _p(ids)))
        return norm, pa, pb, d, ids

    def stage_ms(self) -> dict:
        out = np.zeros(len(STAGES), np.float64)
        self._chk(lib().cdx_scene_stage_ms(self.h, _p(out)))
        return dict(zip(STAGES, out.tolist()))

    def device_bytes(self) -> int:
        b = C.c_uint64(0)
        self._chk(lib().cdx_scene_device_bytes(self.h, C.byref(b)))
        return b.value


class Batch:
    """Many device-resident scenes stepped together (cdx_batch)."""

    def __init__(self, ctx: Context, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.h = C.c_void_p()
        ctx._chk(lib().cdx_batch_create(ctx.h, _kind(dtype), C.byref(self.h)))
        self._scenes = []
        self._selection = None

    def close(self):
        if self.h:
            lib().cdx_batch_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _chk(self, st):
        if st:
            raise CdxError(st, lib().cdx_batch_last_error(self.h).decode())

    def add_scene(self, s, bins=None) -> int:
        """Makes scene `s` resident; bins=None uses s.grid, bins=(0,0,0) autotunes."""
        assert np.dtype(s.dtype) == self.dtype
        g = s.grid if bins is None else bins
        out = C.c_uint32(0)
        self._chk(lib().cdx_batch_add_scene(self.h, s.n, _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID),
                                            s.nb, _p(s.pos), _p(s.rot), _p(s.active), int(g[0]), int(g[1]),
                                            int(g[2]), C.byref(out)))
        self._scenes.append(s.n)
        return out.value

    def scene_bins(self, scene):
        g = np.zeros(3, np.int32)
        self._chk(lib().cdx_batch_scene_bins(self.h, int(scene), _p(g)))
        return tuple(int(v) for v in g)

    def scene_count(self) -> int:
        c = C.c_uint64(0)
        self._chk(lib().cdx_batch_scene_count(self.h, C.byref(c)))
        return c.value

    def update_bodies(self, scene, idx, pos, rot):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        pos = np.ascontiguousarray(pos, dtype=self.dtype)
        rot = np.ascontiguousarray(rot, dtype=self.dtype)
        self._chk(lib().cdx_batch_update_bodies(self.h, int(scene), len(idx), _p(idx), _p(pos), _p(rot)))

    def set_envelope(self, e):
        self._chk(lib().cdx_batch_set_envelope(self.h, float(e)))

    def set_mode(self, mode, tau=1e-2):
        self._chk(lib().cdx_batch_set_narrowphase_mode(self.h, int(mode), float(tau)))

    def select(self, scenes=None):
        if scenes is None:
            self._chk(lib().cdx_batch_select(self.h, None, 0))
            self._selection = None
        else:
            a = np.ascontiguousarray(scenes, dtype=np.uint32)
            self._chk(lib().cdx_batch_select(self.h, _p(a), len(a)))
            self._selection = a

    def step(self, use_graph=True):
        self._chk(lib().cdx_batch_step(self.h, 1 if use_graph else 0))

    def finish(self) -> Counts:
        c = Counts()
        self._chk(lib().cdx_batch_finish(self.h, C.byref(c)))
        return c

    def scene_counts(self, j) -> Counts:
        c = Counts()
        self._chk(lib().cdx_batch_scene_counts(self.h, int(j), C.byref(c)))
        return c

    def read_pairs(self, j):
        out = np.zeros(self.scene_counts(j).pairs, np.int64)
        self._chk(lib().cdx_batch_read_pairs(self.h, int(j), _p(out), len(out)))
        return out

    def read_contacts(self, j):
        k = self.scene_counts(j).contacts
        dt = self.dtype
        norm, pa, pb = (np.zeros((k, 3), dt) for _ in range(3))
        d = np.zeros(k, dt)
        ids = np.zeros((k, 2), np.int32)
        self._chk(lib().cdx_batch_read_contacts(self.h, int(j), _p(norm), _p(pa), _p(pb), _p(d), _p(ids), k))
        return norm, pa, pb, d, ids

    def step_ms(self) -> float:
        v = C.c_double(0)
        self._chk(lib().cdx_batch_step_ms(self.h, C.byref(v)))
        return v.value

    def device_bytes(self) -> int:
        b = C.c_uint64(0)
        self._chk(lib().cdx_batch_device_bytes(self.h, C.byref(b)))
        return b.value

    def scene_bytes(self, n, nb) -> int:
        b = C.c_uint64(0)
        self._chk(lib().cdx_batch_scene_bytes(self.h, int(n), int(nb), C.byref(b)))
        return b.value


class Server:
    """Served request loop over device-resident scenes (cdx_server)."""

    def __init__(self, ctx: Context, dtype=np.float64, mode=NP_MIXED, tau=1e-2, envelope=0.0, hbm_budget=0,
                 max_batch_shapes=0, max_pending=0, use_graph=True):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        cfg = ServerConfig(_kind(dtype), int(mode), float(tau), float(envelope), int(hbm_budget),
                           int(max_batch_shapes), int(max_pending), 1 if use_graph else 0)
        self.h = C.c_void_p()
        ctx._chk(lib().cdx_server_create(ctx.h, C.byref(cfg), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().cdx_server_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _chk(self, st):
        if st:
            raise CdxError(st, lib().cdx_server_last_error(self.h).decode())

    def add_scene(self, s, bins=None) -> int:
        assert np.dtype(s.dtype) == self.dtype
        g = s.grid if bins is None else bins
        out = C.c_uint32(0)
        self._chk(lib().cdx_server_add_scene(self.h, s.n, _p(s.T), _p(s.A), _p(s.B), _p(s.C), _p(s.R), _p(s.ID),
                                             s.nb, _p(s.pos), _p(s.rot), _p(s.active), int(g[0]), int(g[1]),
                                             int(g[2]), C.byref(out)))
        return out.value

    def scene_info(self, scene):
        n, nb, cls = C.c_uint64(0), C.c_uint64(0), C.c_int(0)
        g = np.zeros(3, np.int32)
        self._chk(lib().cdx_server_scene_info(self.h, int(scene), C.byref(n), C.byref(nb), C.byref(cls), _p(g)))
        return {"n": n.value, "nb": nb.value, "size_class": cls.value, "bins": tuple(int(v) for v in g)}

    def submit(self, scene, pos, rot, idx=None) -> int:
        pos = np.ascontiguousarray(pos, dtype=self.dtype)
        rot = np.ascontiguousarray(rot, dtype=self.dtype)
        if idx is not None:
            idx = np.ascontiguousarray(idx, dtype=np.uint32)
        out = C.c_uint64(0)
        self._chk(lib().cdx_server_submit(self.h, int(scene), len(pos), _p(idx), _p(pos), _p(rot), C.byref(out)))
        return out.value

    def pose_buffer(self, scene, k):
        """Pinned (positions (k,3), rotations (k,4)) arrays for bodies 0..k-1, filled in place, and their handle."""
        p, r, h = C.c_void_p(), C.c_void_p(), C.c_uint64(0)
        self._chk(lib().cdx_server_pose_buffer(self.h, int(scene), int(k), C.byref(p), C.byref(r), C.byref(h)))
        ct = C.c_double if self.dtype == np.float64 else C.c_float
        pos = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(int(k), 3))
        rot = np.ctypeslib.as_array(C.cast(r, C.POINTER(ct)), shape=(int(k), 4))
        return h.value, pos, rot

    def submit_buffer(self, handle) -> int:
        out = C.c_uint64(0)
        self._chk(lib().cdx_server_submit_buffer(self.h, int(handle), C.byref(out)))
        return out.value

    def drop_buffer(self, handle):
        self._chk(lib().cdx_server_drop_buffer(self.h, int(handle)))

    def poll(self, req) -> bool:
        d = C.c_int(0)
        self._chk(lib().cdx_server_poll(self.h, int(req), C.byref(d)))
        return bool(d.value)

    def wait(self, req, timeout_ms=-1.0) -> bool:
        d = C.c_int(0)
        self._chk(lib().cdx_server_wait(self.h, int(req), float(timeout_ms), C.byref(d)))
        return bool(d.value)

    def result(self, req):
        n, lat = C.c_uint64(0), C.c_double(0)
        self._chk(lib().cdx_server_result(self.h, int(req), C.byref(n), C.byref(lat)))
        return n.value, lat.value

    def read_contacts(self, req):
        k, _ = self.result(req)
        dt = self.dtype
        norm, pa, pb = (np.zeros((k, 3), dt) for _ in range(3))
        d = np.zeros(k, dt)
        ids = np.zeros((k, 2), np.int32)
        self._chk(lib().cdx_server_read_contacts(self.h, int(req), _p(norm), _p(pa), _p(pb), _p(d), _p(ids), k))
        return norm, pa, pb, d, ids

    def release(self, req):
        self._chk(lib().cdx_server_release(self.h, int(req)))

    def stats(self, reset=False) -> dict:
        s = ServerStats()
        self._chk(lib().cdx_server_stats(self.h, C.byref(s), 1 if reset else 0))
        return s.as_dict()

    def memory(self):
        b, c, d = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        self._chk(lib().cdx_server_memory(self.h, C.byref(b), C.byref(c), C.byref(d)))
        return {"hbm_budget": b.value, "hbm_committed": c.value, "device_bytes": d.value}
