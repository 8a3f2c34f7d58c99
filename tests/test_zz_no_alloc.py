"""cdx_detect_possible_collisions_noalloc against cdx_detect_possible_collisions.

The allocation-free entry point runs the same broadphase on device buffers,
with every intermediate in the library's fixed scratch array.  Its pairs and
in-place shifted AABBs must equal the allocating entry point's bit for bit,
on two calls in a row (the second smaller, so it runs on the first call's
leftover scratch contents), and it must refuse inputs beyond its limits.
Device buffers come from hipMalloc through ctypes (the test allocates; the
entry point does not).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 256)) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(cdx, ctx, hip, aabb, bins, pair_cap):
    d_aabb = DevArray(hip, aabb.nbytes)
    d_pairs = DevArray(hip, pair_cap * 8)
    try:
        d_aabb.put(aabb)
        n = C.c_uint64(0)
        st = cdx.lib().cdx_detect_possible_collisions_noalloc(
            ctx.h, cdx._kind(aabb.dtype), d_aabb.p, aabb.shape[0] // 2, int(bins[0]), int(bins[1]), int(bins[2]),
            d_pairs.p, pair_cap, C.byref(n), None)
        if st:
            raise cdx.CdxError(st, cdx.lib().cdx_ctx_last_error(ctx.h).decode())
        pairs = d_pairs.get(np.zeros(n.value, np.int64)) if n.value else np.zeros(0, np.int64)
        shifted = d_aabb.get(np.zeros_like(aabb))
        return pairs, shifted
    finally:
        d_aabb.free()
        d_pairs.free()


@pytest.mark.parametrize("first,second", [
    (dict(n=20_000, seed=51, dtype=np.float64), dict(n=3_000, seed=52, dtype=np.float64)),
    (dict(n=65_536, seed=53, dtype=np.float32), dict(n=700, seed=54, dtype=np.float32, grid=(3, 3, 3))),
])
def test_noalloc_equals_allocating_entry_point(cdx, ctx, oracle, scenes, hip, first, second):
    for spec in (first, second):
        s = scenes.mixed_bodies(**spec)
        aabb = oracle.aabb(s)
        print(f"noalloc n={s.n} grid={s.grid} dtype={np.dtype(s.dtype).name}", flush=True)
        ref_aabb = aabb.copy()
        ref_pairs = ctx.detect_possible_collisions(ref_aabb, s.grid)
        pairs, shifted = _noalloc(cdx, ctx, hip, aabb, s.grid, len(ref_pairs) + 17)
        np.testing.assert_array_equal(pairs, ref_pairs)
        np.testing.assert_array_equal(shifted, ref_aabb)


def test_noalloc_limits(cdx, ctx, oracle, scenes, hip):
    s = scenes.mixed_bodies(n=5_000, seed=55)
    aabb = oracle.aabb(s)
    ref = ctx.detect_possible_collisions(aabb.copy(), s.grid)
    assert len(ref) > 10
    with pytest.raises(cdx.CdxError) as e:  # output buffer one pair short
        _noalloc(cdx, ctx, hip, aabb, s.grid, len(ref) - 1)
    assert e.value.status == 4
    big = scenes.mixed_bodies(n=65_537, seed=56, dtype=np.float32)
    with pytest.raises(cdx.CdxError) as e:  # more AABBs than the scratch array is laid out for
        _noalloc(cdx, ctx, hip, oracle.aabb(big), big.grid, 1 << 20)
    assert e.value.status == 4
    # still usable after the refusals
    pairs, _ = _noalloc(cdx, ctx, hip, aabb, s.grid, len(ref))
    np.testing.assert_array_equal(pairs, ref)
