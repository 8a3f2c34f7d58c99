"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit (guards the oracle
against drift).  GPU: libcdx.so reproduces the fixtures: AABBs, bin counts,
sorted entries, active bins, pair list, shifted AABBs, contact ids and the
permuted pair list bit-exact; contact normal / points / depth within
1e-12 x scene length (FP64) or 1e-5 x scene length (FP32) for NATIVE, and
MIXED equal to NATIVE bit for bit (FP64 fixtures; with FP32 input MIXED
re-solves hits in FP64 and is more accurate than the FP32 reference path,
so only NATIVE is held to the FP32 fixtures).
"""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FILES]


def _load(scenes, name):
    z = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    s = scenes.Scene(z["T"], z["A"], z["B"], z["C"], z["R"], z["ID"], z["pos"].copy(), z["rot"].copy(),
                     z["active"], tuple(int(v) for v in z["grid"]), name)
    return s, z


def _tol(s):
    L = float(np.abs(s.pos).max()) + 1.0
    return (1e-12 if s.dtype == np.float64 else 1e-5) * L


def test_fixtures_present():
    assert len(FILES) >= 5


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(oracle, scenes, name):
    s, z = _load(scenes, name)
    ab = oracle.aabb(s)
    np.testing.assert_array_equal(ab, z["aabb"])
    bp = oracle.broadphase(ab, s.grid)
    for k in ("aabb_shifted", "bin_counts", "sorted_keys", "sorted_bodies", "active_keys", "active_counts", "pairs"):
        np.testing.assert_array_equal(getattr(bp, k), z[k], err_msg=k)
    npr = oracle.narrowphase(s, bp.pairs, float(z["envelope"]))
    for k in ("norm", "cpta", "cptb", "dpth", "bids", "pairs_permuted"):
        np.testing.assert_array_equal(getattr(npr, k), z[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_fixture(cdx, ctx, scenes, name):
    s, z = _load(scenes, name)
    print(f"fixture {name} n={s.n}", flush=True)
    got = {}
    modes = (cdx.NP_NATIVE, cdx.NP_MIXED) if s.dtype == np.float64 else (cdx.NP_NATIVE,)
    for mode in modes:
        sc = cdx.Scene(ctx, s.dtype).load(s)
        sc.set_envelope(float(z["envelope"]))
        sc.set_mode(mode, 1e-2)
        sc.aabb()
        sc.sync()
        np.testing.assert_array_equal(sc.read_aabb(), z["aabb"])
        sc.broadphase()
        sc.narrowphase()
        sc.sync()
        bc, sk, sb, ak, ac = sc.read_bins()
        for k, v in (("bin_counts", bc), ("sorted_keys", sk), ("sorted_bodies", sb), ("active_keys", ak),
                     ("active_counts", ac)):
            np.testing.assert_array_equal(v, z[k], err_msg=k)
        np.testing.assert_array_equal(sc.read_pairs(), z["pairs"])
        np.testing.assert_array_equal(sc.read_aabb(shifted=True), z["aabb_shifted"])
        np.testing.assert_array_equal(sc.read_pairs_permuted(), z["pairs_permuted"])
        norm, pa, pb, d, ids = sc.read_contacts()
        np.testing.assert_array_equal(ids, z["bids"])
        t = _tol(s)
        for k, v in (("norm", norm), ("cpta", pa), ("cptb", pb), ("dpth", d)):
            np.testing.assert_allclose(v, z[k], rtol=0, atol=t, err_msg=k)
        got[mode] = (norm, pa, pb, d)
        sc.close()
    if s.dtype == np.float64:
        for a, b in zip(got[cdx.NP_NATIVE], got[cdx.NP_MIXED]):
            np.testing.assert_array_equal(a, b)
