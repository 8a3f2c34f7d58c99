"""The oracle against the reference's own known-answer tests (test/*.cu).

These run on the CPU only; they pin the oracle's broadphase before it is
trusted as the checker of the HIP library.
"""
import ctypes

import numpy as np
import pytest


def test_single_bin_all_pairs(oracle):
    """test/test1.cu:9-33: N identical AABBs [0,1]^3, 1x1x1 grid -> N(N-1)/2 pairs.

    The 1^3 grid makes every AABB span cells 0..1 on each axis, so Hash_Index
    aliases (Broadphase.cu:140-147) and bodies repeat inside one key.
    """
    for n in (2, 10, 57):
        ab = np.zeros((2 * n, 3))
        ab[n:] = 1.0
        r = oracle.broadphase(ab, (1, 1, 1))
        assert len(r.pairs) == n * (n - 1) // 2
        a = r.pairs >> 32
        b = r.pairs & 0xffffffff
        assert np.all(a < b)
        # every body spans 2^3 cells whose keys alias onto 0..3
        np.testing.assert_array_equal(r.bin_counts, np.full(n, 8, np.uint32))


def test_duplicated_lattice(oracle):
    """test/binTest_Broadphase.cu:13-89: two spheres per lattice site -> exactly n^3 pairs (2i, 2i+1)."""
    for n, bpa in ((10, 10), (6, 4)):
        r_ = 0.25
        k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        c = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
        c = np.repeat(c, 2, axis=0)
        ab = np.concatenate([c - r_, c + r_])
        r = oracle.broadphase(ab, (bpa, bpa, bpa))
        assert len(r.pairs) == n ** 3
        exp = (np.arange(0, 2 * n ** 3, 2, dtype=np.int64) << 32) | np.arange(1, 2 * n ** 3, 2, dtype=np.int64)
        np.testing.assert_array_equal(r.pairs, exp)


def _glibc_rand_spheres(n):
    """The inputs of test/nSquared_Broadphase.cu:30-51 (glibc rand(), never seeded => seed 1)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    RAND_MAX = 2147483647
    vals = np.array([libc.rand() for _ in range(4 * n)], dtype=np.float64) / RAND_MAX
    vals = vals.reshape(n, 4)
    pos = -30 + vals[:, :3] * 60.0
    r = 0.1 + vals[:, 3] * 0.9
    return pos, r


def test_nsquared_superset(oracle):
    """test/nSquared_Broadphase.cu: every strict sphere contact (dist^2 < (rA+rB)^2) is a candidate pair.

    Same 300,000 glibc-rand() spheres in [-30,30]^3, r in [0.1, 1], 30^3 grid.
    The true contacts come from scipy's KD-tree instead of the O(N^2) loop.
    """
    from scipy.spatial import cKDTree
    n = 300_000
    pos, r = _glibc_rand_spheres(n)
    ab = np.concatenate([pos - r[:, None], pos + r[:, None]])
    res = oracle.broadphase(ab, (30, 30, 30))
    tree = cKDTree(pos)
    cand = tree.query_pairs(2.0, output_type="ndarray")
    d2 = np.sum((pos[cand[:, 0]] - pos[cand[:, 1]]) ** 2, axis=1)
    true = cand[d2 < (r[cand[:, 0]] + r[cand[:, 1]]) ** 2]
    true = np.sort(true, axis=1)
    keys = (true[:, 0].astype(np.int64) << 32) | true[:, 1].astype(np.int64)
    found = np.isin(keys, res.pairs)
    assert found.all(), f"{(~found).sum()} true contacts missing"
    assert len(res.pairs) >= len(keys)
