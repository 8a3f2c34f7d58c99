"""GPU parity: libcdx.so (through the C ABI) against the CPU oracle.

Bit-exact: AABBs, per-AABB bin counts, sorted (bin, body) entries, active
bins and their occupancy, the pair list, the shifted AABBs.
Contacts (NATIVE mode): the contact set, body ids and the permuted pair list
exactly; normal, points and depth within 1e-12 x scene length (FP64) or
1e-5 x scene length (FP32).  MIXED mode: the contact set and ids exactly;
normal and depth within the tolerances stated in test_gpu_mixed_precision.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene_cases(scenes):
    return {
        "spheres8k_f64": scenes.small_spheres(),
        "mixed20k_f64": scenes.mixed_bodies(n=20_000, seed=11),
        "mixed20k_f32": scenes.mixed_bodies(n=20_000, seed=12, dtype=np.float32),
        "lattice16_f32": scenes.lattice_spheres(side=16, seed=13),
        # every reference shape type plus the capsule (rectangles, cones and
        # triangles take the fixed AABB paths and the extended MPR launch)
        "alltypes4k_f64": scenes.all_shape_types(n=4000, seed=14),
        "alltypes4k_f32": scenes.all_shape_types(n=4000, seed=15, dtype=np.float32),
    }


@pytest.fixture(scope="module")
def cases(scenes):
    return _scene_cases(scenes)


CASE_NAMES = ["spheres8k_f64", "mixed20k_f64", "mixed20k_f32", "lattice16_f32", "alltypes4k_f64", "alltypes4k_f32"]


@pytest.mark.parametrize("name", CASE_NAMES)
def test_aabb_bit_exact(cdx, ctx, oracle, cases, name):
    s = cases[name]
    print(f"case {name} n={s.n}", flush=True)
    ref = oracle.aabb(s)
    got = ctx.generate_aabb(s)
    assert got.dtype == ref.dtype
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("name", CASE_NAMES)
def test_broadphase_bit_exact(cdx, ctx, oracle, cases, name):
    s = cases[name]
    print(f"case {name} n={s.n} grid={s.grid}", flush=True)
    _check_broadphase(cdx, ctx, oracle, s)


LONG_RUN_CASES = {
    # runs of several hundred entries per bin: the pair search's look-ahead
    # window, its continuation past the window and hit masks with partners
    # 63 or more places on
    "mixed3k_grid2_f64": dict(n=3000, seed=21, grid=(2, 2, 2)),
    "mixed3k_grid2_f32": dict(n=3000, seed=22, dtype=np.float32, grid=(2, 2, 2)),
    # ~60 AABB partners per shape: partner groups longer than the register sorts
    "dense600_grid3_f64": dict(n=600, seed=23, neighbours=60.0, grid=(3, 3, 3)),
}


@pytest.mark.parametrize("name", sorted(LONG_RUN_CASES))
def test_broadphase_long_runs(cdx, ctx, oracle, scenes, name):
    s = scenes.mixed_bodies(**LONG_RUN_CASES[name])
    print(f"case {name} n={s.n} grid={s.grid}", flush=True)
    _check_broadphase(cdx, ctx, oracle, s)


def _check_broadphase(cdx, ctx, oracle, s):
    ab = oracle.aabb(s)
    ref = oracle.broadphase(ab, s.grid)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.aabb()
    sc.broadphase()
    sc.sync()
    c = sc.counts()
    assert c.bin_entries == len(ref.sorted_keys)
    assert c.active_bins == len(ref.active_keys)
    assert c.pairs == len(ref.pairs)
    np.testing.assert_array_equal(np.array(c.bmin), ref.bmin)
    np.testing.assert_array_equal(np.array(c.bmax), ref.bmax)
    bc, sk, sb, ak, ac = sc.read_bins()
    np.testing.assert_array_equal(bc, ref.bin_counts)
    np.testing.assert_array_equal(sk, ref.sorted_keys)
    np.testing.assert_array_equal(sb, ref.sorted_bodies)
    np.testing.assert_array_equal(ak, ref.active_keys)
    np.testing.assert_array_equal(ac, ref.active_counts)
    np.testing.assert_array_equal(sc.read_pairs(), ref.pairs)
    np.testing.assert_array_equal(sc.read_aabb(shifted=True), ref.aabb_shifted)
    np.testing.assert_array_equal(sc.read_aabb(shifted=False), ab)


def test_dropin_broadphase_shifts_in_place(cdx, ctx, oracle, cases):
    s = cases["spheres8k_f64"]
    ab = oracle.aabb(s)
    ref = oracle.broadphase(ab, s.grid)
    mine = ab.copy()
    pairs = ctx.detect_possible_collisions(mine, s.grid)
    np.testing.assert_array_equal(pairs, ref.pairs)
    np.testing.assert_array_equal(mine, ref.aabb_shifted)


def test_dropin_error_does_not_stick(cdx, ctx, oracle, cases):
    """An AABB with max < min is reported (no bins, CDX_ERR_INVALID); the next
    drop-in broadphase on the same context starts with clear error flags."""
    bad = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.5, 0.5, 0.5], [0.2, 0.2, 0.2]])
    with pytest.raises(cdx.CdxError):
        ctx.detect_possible_collisions(bad, (4, 4, 4))
    s = cases["spheres8k_f64"]
    ab = oracle.aabb(s)
    ref = oracle.broadphase(ab, s.grid)
    np.testing.assert_array_equal(ctx.detect_possible_collisions(ab.copy(), s.grid), ref.pairs)


def _tol(s):
    L = float(np.ptp(s.pos, axis=0).max()) + 1.0
    return (1e-12 if s.dtype == np.float64 else 1e-5) * L


@pytest.mark.parametrize("name", CASE_NAMES)
def test_narrowphase_native(cdx, ctx, oracle, cases, name):
    s = cases[name]
    print(f"case {name}", flush=True)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    ref = oracle.narrowphase(s, pairs)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(cdx.NP_NATIVE)
    sc.step()
    sc.sync()
    c = sc.counts()
    assert c.contacts == ref.count
    norm, pa, pb, d, ids = sc.read_contacts()
    np.testing.assert_array_equal(ids, ref.bids)
    np.testing.assert_array_equal(sc.read_pairs_permuted(), ref.pairs_permuted)
    t = _tol(s)
    np.testing.assert_allclose(norm, ref.norm, rtol=0, atol=t)
    np.testing.assert_allclose(pa, ref.cpta, rtol=0, atol=t)
    np.testing.assert_allclose(pb, ref.cptb, rtol=0, atol=t)
    np.testing.assert_allclose(d, ref.dpth, rtol=0, atol=t)


def test_dropin_narrowphase_envelope_and_inactive(cdx, ctx, oracle, scenes):
    s = scenes.mixed_bodies(n=8_000, seed=21)
    rng = np.random.default_rng(5)
    s.active = (rng.random(s.nb) < 0.7).astype(np.uint8)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    ref = oracle.narrowphase(s, pairs, envelope=0.001)
    mine = pairs.copy()
    norm, pa, pb, d, ids = ctx.do_narrowphase(s, mine, envelope=0.001)
    np.testing.assert_array_equal(mine, ref.pairs_permuted)
    np.testing.assert_array_equal(ids, ref.bids)
    t = _tol(s)
    np.testing.assert_allclose(norm, ref.norm, atol=t, rtol=0)
    np.testing.assert_allclose(d, ref.dpth, atol=t, rtol=0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_mpr_store_per_pair(cdx, ctx, oracle, scenes, dtype):
    """host_MPR_Store (Narrowphase.cu:464-504) through cdx_mpr_store: contacts
    at their pair index, contact_active cleared there, other entries kept."""
    s = scenes.mixed_bodies(n=8_000, seed=25, dtype=dtype)
    rng = np.random.default_rng(6)
    s.active = (rng.random(s.nb) < 0.7).astype(np.uint8)
    pairs = oracle.broadphase(oracle.aabb(s), s.grid).pairs
    print(f"mpr_store n={s.n} pairs={len(pairs)} dtype={np.dtype(dtype).name}", flush=True)
    ref = oracle.narrowphase(s, pairs, envelope=0.001)
    hit = ref.hit.astype(bool)
    act, norm, pa, pb, d, ids = ctx.mpr_store(s, pairs, envelope=0.001, fill=np.nan, active_fill=1)
    np.testing.assert_array_equal(act, np.where(hit, 0, 1))
    np.testing.assert_array_equal(ids[hit], ref.bids)
    assert (ids[~hit] == -1).all() and np.isnan(norm[~hit]).all() and np.isnan(d[~hit]).all()
    t = _tol(s)
    for got, want in ((norm, ref.norm), (pa, ref.cpta), (pb, ref.cptb)):
        np.testing.assert_allclose(got[hit], want, atol=t, rtol=0)
    np.testing.assert_allclose(d[hit], ref.dpth, atol=t, rtol=0)


@pytest.mark.parametrize("mode", ["mixed", "native"])
def test_mixed_equals_native_fp64(cdx, ctx, oracle, scenes, mode):
    """MIXED (FP32 filter + FP64 reference path for hits) equals the FP64 oracle."""
    s = scenes.mixed_bodies(n=50_000, seed=31)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    ref = oracle.narrowphase(s, pairs)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(cdx.NP_MIXED if mode == "mixed" else cdx.NP_NATIVE, 1e-2)
    sc.step()
    sc.sync()
    c = sc.counts()
    print(f"mode={mode} pairs={c.pairs} contacts={c.contacts} resolved_fp64={c.resolved}", flush=True)
    norm, pa, pb, d, ids = sc.read_contacts()
    assert c.contacts == ref.count
    np.testing.assert_array_equal(ids, ref.bids)
    np.testing.assert_array_equal(sc.read_pairs_permuted(), ref.pairs_permuted)
    t = _tol(s)
    np.testing.assert_allclose(norm, ref.norm, rtol=0, atol=t)
    np.testing.assert_allclose(pa, ref.cpta, rtol=0, atol=t)
    np.testing.assert_allclose(pb, ref.cptb, rtol=0, atol=t)
    np.testing.assert_allclose(d, ref.dpth, rtol=0, atol=t)


def test_fast_mode_error_report(cdx, ctx, oracle, scenes):
    """FAST keeps FP32 contacts for clear hits.

    Stated tolerance: the contact set and ids equal the FP64 oracle's; the
    median normal difference is below 1e-3 and the median depth difference
    below 1e-4 x (rA + rB).  (FP32 resolves curved MPR portals to about
    sqrt(FLT_EPSILON), and the reference's FP64 MPR stops at its 34-iteration
    cap, so neither side is converged beyond ~1e-4.)
    """
    s = scenes.mixed_bodies(n=50_000, seed=31)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    ref = oracle.narrowphase(s, pairs)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(cdx.NP_FAST, 1e-2)
    sc.step()
    sc.sync()
    c = sc.counts()
    norm, pa, pb, d, ids = sc.read_contacts()
    assert c.contacts == ref.count
    np.testing.assert_array_equal(ids, ref.bids)
    perm = sc.read_pairs_permuted()
    np.testing.assert_array_equal(perm, ref.pairs_permuted)
    hits = perm[: c.contacts]
    a = (hits >> 32).astype(np.int64)
    b = (hits & 0xffffffff).astype(np.int64)
    rad = np.linalg.norm(s.B, axis=1)
    scale = rad[a] + rad[b]
    en = np.linalg.norm(norm - ref.norm, axis=1)
    ed = np.abs(d - ref.dpth) / scale
    print(f"FAST resolved_fp64={c.resolved}/{c.pairs}; normal diff p50={np.median(en):.2e} "
          f"p99={np.quantile(en, 0.99):.2e} max={en.max():.2e}; depth diff/scale p50={np.median(ed):.2e} "
          f"p99={np.quantile(ed, 0.99):.2e} max={ed.max():.2e}", flush=True)
    assert np.median(en) < 1e-3
    assert np.median(ed) < 1e-4


def test_kat_single_bin(cdx, ctx):
    """test/test1.cu: N identical AABBs [0,1]^3 on a 1^3 grid -> N(N-1)/2 pairs."""
    for n in (10, 100):
        ab = np.zeros((2 * n, 3))
        ab[n:] = 1.0
        pairs = ctx.detect_possible_collisions(ab, (1, 1, 1))
        assert len(pairs) == n * (n - 1) // 2


def test_kat_duplicated_lattice(cdx, ctx):
    """test/binTest_Broadphase.cu: two spheres per site of an n^3 lattice -> n^3 pairs (2i, 2i+1)."""
    n, bpa, r = 10, 10, 0.25
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    c = np.repeat(c, 2, axis=0)
    ab = np.concatenate([c - r, c + r])
    pairs = ctx.detect_possible_collisions(ab, (bpa, bpa, bpa))
    assert len(pairs) == n ** 3
    exp = (np.arange(0, 2 * n ** 3, 2, dtype=np.int64) << 32) | np.arange(1, 2 * n ** 3, 2, dtype=np.int64)
    np.testing.assert_array_equal(pairs, exp)


@pytest.mark.parametrize("axis", [0, 2])
def test_broadphase_large_grid(cdx, ctx, oracle, axis):
    """2^24 bins along one axis: the owner-bin hash takes its 32-bit path
    (the 24-bit one needs every hash factor below 2^24); pairs vs the oracle."""
    rng = np.random.default_rng(31)
    n2 = 1000
    c = np.stack([rng.uniform(0, 1e-3, n2), rng.uniform(0, 1e-3, n2), rng.uniform(0.0, 100.0, n2)], axis=1)
    c = np.repeat(c, 2, axis=0)
    c[1::2, 2] += 1e-5  # each twin overlaps its partner: pairs (2i, 2i+1)
    if axis == 0:
        c = c[:, [2, 1, 0]].copy()
    ab = np.concatenate([c - 2e-5, c + 2e-5])
    grid = [1, 1, 1]
    grid[axis] = 1 << 24
    print(f"large grid {grid}", flush=True)
    ref = oracle.broadphase(ab, grid)
    pairs = ctx.detect_possible_collisions(ab, tuple(grid))
    assert len(ref.pairs) >= n2
    np.testing.assert_array_equal(pairs, ref.pairs)


def test_empty_inputs(cdx, ctx):
    ab = np.zeros((0, 3))
    assert len(ctx.detect_possible_collisions(ab, (4, 4, 4))) == 0


def test_mpr_queue_is_the_mpr_work(cdx, ctx, oracle, scenes):
    """The reference-order MPR launch runs on exactly the pairs the reference
    sends into CollideAndFindPoint (NATIVE: every pair with MPR steps in the
    oracle), and on the FP64 re-solves in MIXED."""
    s = scenes.mixed_bodies(n=20_000, seed=33)
    pairs = oracle.broadphase(oracle.aabb(s), s.grid).pairs
    want = np.nonzero(oracle.mpr_steps(s, pairs) > 0)[0]
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(cdx.NP_NATIVE)
    sc.set_profiling(True)
    sc.step()
    sc.sync()
    q = sc.read_mpr_queue()
    np.testing.assert_array_equal(np.sort(q), want)
    assert sc.stage_ms()["np_mpr"] > 0
    sc.set_mode(cdx.NP_MIXED, 1e-2)
    sc.step()
    sc.sync()
    q = sc.read_mpr_queue()
    assert len(q) == sc.counts().resolved and len(np.unique(q)) == len(q)
    assert np.isin(q, want).all()
    sc.close()


@pytest.mark.parametrize("mode", ["mixed", "fast"])
def test_all_shape_types_with_envelope(cdx, ctx, oracle, scenes, mode):
    """Rectangles, cones and triangles skip the FP32 filter and are resolved
    on the reference path in every mode; triangle pairs ignore the envelope
    (Narrowphase.cu:364, :373).  Contacts equal the FP64 oracle's."""
    s = scenes.all_shape_types(n=6000, seed=16)
    pairs = oracle.broadphase(oracle.aabb(s), s.grid).pairs
    ref = oracle.narrowphase(s, pairs, envelope=0.0005)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(cdx.NP_MIXED if mode == "mixed" else cdx.NP_FAST, 1e-2)
    sc.set_envelope(0.0005)
    sc.step()
    sc.sync()
    norm, pa, pb, d, ids = sc.read_contacts()
    print(f"all types mode={mode}: {len(d)} contacts, oracle {ref.count}", flush=True)
    assert len(d) == ref.count
    np.testing.assert_array_equal(ids, ref.bids)
    T = s.T
    ext = np.isin(T[ids[:, 0]], [4, 5, 6]) | np.isin(T[ids[:, 1]], [4, 5, 6])
    assert ext.sum() > 100
    t = _tol(s)
    sel = slice(None) if mode == "mixed" else ext  # FAST keeps FP32 contacts for the other types
    for g, r in ((norm, ref.norm), (pa, ref.cpta), (pb, ref.cptb)):
        np.testing.assert_allclose(g[sel], r[sel], rtol=0, atol=t)
    np.testing.assert_allclose(d[sel], ref.dpth[sel], rtol=0, atol=t)
    sc.close()
