"""The cell search against the oracle's bin broadphase (Broadphase.cu:344-545).

The pair list does not depend on how the candidates are found, so the cell
search must give the oracle's list bit for bit; so must the bin entries,
active bins and occupancy it builds when a caller reads them.  Cases: the
parity scenes with the search forced each way, and the shapes of input the
cell grid has to survive: zero-width AABBs, a flat scene, coincident shapes,
one box spanning the scene (a partner group longer than the in-register
sorts), one or two shapes, negative coordinates far from the origin.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spheres(scenes, pos, r, dtype, grid):
    n = len(pos)
    B = np.zeros((n, 3))
    B[:, 0] = r
    rot = np.zeros((n, 4))
    rot[:, 0] = 1
    return scenes._one_shape_per_body(np.zeros(n, np.int32), B, pos, rot, dtype, grid, "cell-edge", {"n": n})


def _edge_cases(scenes):
    rng = np.random.default_rng(0x5EED0C11)
    out = {}
    # zero-width AABBs, with exact duplicates (closed boxes touch)
    p = np.round(rng.uniform(0, 4, (1500, 3)), 1)
    out["points_f64"] = _spheres(scenes, p, 0.0, np.float64, (8, 8, 8))
    # a flat scene: every centre at z = 0
    p = rng.uniform(-3, 3, (3000, 3))
    p[:, 2] = 0
    out["flat_f32"] = _spheres(scenes, p, rng.uniform(0.02, 0.06, 3000), np.float32, (16, 16, 1))
    # coincident shapes: 40 spheres at each of 50 places
    p = np.repeat(rng.uniform(0, 10, (50, 3)), 40, axis=0)
    out["stacked_f64"] = _spheres(scenes, p, 0.3, np.float64, (4, 4, 4))
    # shape 0 a box spanning the scene among small spheres: its group holds
    # every other shape (longer than the in-register sorts)
    n = 3000
    T = np.zeros(n + 1, np.int32)
    T[0] = 2  # BOX
    B = np.zeros((n + 1, 3))
    B[1:, 0] = 0.05
    B[0] = (6.0, 6.0, 6.0)
    pos = np.vstack([[[5.0, 5.0, 5.0]], rng.uniform(0, 10, (n, 3))])
    rot = np.zeros((n + 1, 4))
    rot[:, 0] = 1
    out["bigbox_f64"] = scenes._one_shape_per_body(T, B, pos, rot, np.float64, (8, 8, 8), "cell-edge", {"n": n + 1})
    # far from the origin, all coordinates negative
    p = rng.uniform(-1.0e4 - 5, -1.0e4, (4000, 3))
    out["negative_f32"] = _spheres(scenes, p, rng.uniform(0.05, 0.15, 4000), np.float32, (10, 10, 10))
    out["one_f64"] = _spheres(scenes, np.array([[1.0, 2.0, 3.0]]), 0.5, np.float64, (2, 2, 2))
    out["two_f32"] = _spheres(scenes, np.array([[0.0, 0.0, 0.0], [0.9, 0.0, 0.0]]), 0.5, np.float32, (2, 2, 2))
    return out


@pytest.fixture(scope="module")
def edge_cases(scenes):
    return _edge_cases(scenes)


PARITY = ["spheres8k_f64", "mixed20k_f64", "mixed20k_f32", "lattice16_f32", "alltypes4k_f64", "alltypes4k_f32"]
EDGES = ["points_f64", "flat_f32", "stacked_f64", "bigbox_f64", "negative_f32", "one_f64", "two_f32"]


def _parity_scene(scenes, name):
    from test_gpu_parity import _scene_cases
    return _scene_cases(scenes)[name]


def _run(cdx, ctx, s, mode):
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_pair_search(mode)
    sc.aabb()
    sc.broadphase()
    sc.sync()
    return sc


def _check(cdx, ctx, oracle, s, mode, expect_cells):
    ab = oracle.aabb(s)
    ref = oracle.broadphase(ab, s.grid)
    sc = _run(cdx, ctx, s, mode)
    pairs = sc.read_pairs()
    c = sc.counts()
    assert bool(c.flags & cdx.FLAG_CELL_SEARCH) == expect_cells
    assert not (c.flags & cdx.FLAG_NO_ACTIVE_BINS)
    np.testing.assert_array_equal(pairs, ref.pairs)
    assert c.pairs == len(ref.pairs)
    assert c.bin_entries == len(ref.sorted_keys)
    assert c.active_bins == len(ref.active_keys)
    np.testing.assert_array_equal(np.array(c.bmin), ref.bmin)
    np.testing.assert_array_equal(np.array(c.bmax), ref.bmax)
    bc, sk, sb, ak, ac = sc.read_bins()
    np.testing.assert_array_equal(bc, ref.bin_counts)
    np.testing.assert_array_equal(sk, ref.sorted_keys)
    np.testing.assert_array_equal(sb, ref.sorted_bodies)
    np.testing.assert_array_equal(ak, ref.active_keys)
    np.testing.assert_array_equal(ac, ref.active_counts)
    np.testing.assert_array_equal(sc.read_aabb(shifted=True), ref.aabb_shifted)
    sc.close()


@pytest.mark.parametrize("name", PARITY)
@pytest.mark.parametrize("mode", ["bins", "cells"])
def test_pair_search_modes_match_oracle(cdx, ctx, oracle, scenes, name, mode):
    s = _parity_scene(scenes, name)
    print(f"case {name} n={s.n} mode={mode}", flush=True)
    m = cdx.PAIRS_CELLS if mode == "cells" else cdx.PAIRS_BINS
    _check(cdx, ctx, oracle, s, m, mode == "cells")


@pytest.mark.parametrize("name", EDGES)
def test_cell_search_edge_cases(cdx, ctx, oracle, edge_cases, name):
    s = edge_cases[name]
    print(f"case {name} n={s.n}", flush=True)
    # one shape has no pairs and no cell grid: the bins run
    _check(cdx, ctx, oracle, s, cdx.PAIRS_CELLS, s.n > 1)


def test_auto_takes_cells_for_uniform_spheres(cdx, ctx, oracle, scenes):
    s = scenes.lattice_spheres(side=40, seed=21)
    print(f"lattice-40 n={s.n}", flush=True)
    _check(cdx, ctx, oracle, s, cdx.PAIRS_AUTO, True)


def test_repeated_cell_steps_count_bins_on_request(cdx, ctx, oracle, scenes):
    """After a cell-search broadphase the next one decides from the bounds
    alone and counts the reference grid's bins only when they are read; full
    steps (a narrowphase between the broadphase and the read) leave the
    totals and bins the oracle's."""
    s = scenes.lattice_spheres(side=30, seed=25)
    print(f"lattice-30 n={s.n}", flush=True)
    ref = oracle.broadphase(oracle.aabb(s), s.grid)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_pair_search(cdx.PAIRS_AUTO)
    for it in range(3):
        print(f"step {it}", flush=True)
        sc.step()
        sc.sync()
        np.testing.assert_array_equal(sc.read_pairs(), ref.pairs)
        c = sc.counts()
        assert c.flags & cdx.FLAG_CELL_SEARCH and not (c.flags & cdx.FLAG_NO_ACTIVE_BINS)
        assert c.pairs == len(ref.pairs)
        assert c.bin_entries == len(ref.sorted_keys)
        assert c.active_bins == len(ref.active_keys)
        assert c.key_max == int(ref.sorted_keys.max())
        bc, sk, sb, ak, ac = sc.read_bins()
        np.testing.assert_array_equal(bc, ref.bin_counts)
        np.testing.assert_array_equal(sk, ref.sorted_keys)
        np.testing.assert_array_equal(sb, ref.sorted_bodies)
        np.testing.assert_array_equal(ak, ref.active_keys)
        np.testing.assert_array_equal(ac, ref.active_counts)
    sc.close()


def test_auto_takes_bins_for_mixed_sizes(cdx, ctx, oracle, scenes):
    s = scenes.mixed_bodies(n=20_000, seed=22)
    _check(cdx, ctx, oracle, s, cdx.PAIRS_AUTO, False)


def test_bins_unavailable_after_the_aabbs_change(cdx, ctx, oracle, scenes):
    s = scenes.lattice_spheres(side=12, seed=23)
    sc = _run(cdx, ctx, s, cdx.PAIRS_CELLS)
    sc.aabb()  # new AABBs before anyone read the cell broadphase's bins
    sc.sync()
    c = sc.counts()
    assert c.flags & cdx.FLAG_CELL_SEARCH and c.flags & cdx.FLAG_NO_ACTIVE_BINS
    with pytest.raises(cdx.CdxError):
        sc.read_bins()
    # the pairs of that broadphase stay readable, and a new one rebuilds everything
    ref = oracle.broadphase(oracle.aabb(s), s.grid)
    np.testing.assert_array_equal(sc.read_pairs(), ref.pairs)
    sc.broadphase()
    sc.sync()
    _, sk, _, ak, _ = sc.read_bins()
    np.testing.assert_array_equal(sk, ref.sorted_keys)
    np.testing.assert_array_equal(ak, ref.active_keys)
    sc.close()


def test_bins_unavailable_after_the_grid_changes(cdx, ctx, oracle, scenes):
    """The bins of a cell-search broadphase are counted on request at the
    grid of that broadphase; after new bins per axis they are unavailable."""
    s = scenes.lattice_spheres(side=12, seed=26)
    sc = _run(cdx, ctx, s, cdx.PAIRS_CELLS)
    sc.set_bins(*[g + 1 for g in s.grid])
    c = sc.counts()
    assert c.flags & cdx.FLAG_CELL_SEARCH and c.flags & cdx.FLAG_NO_ACTIVE_BINS
    assert c.bin_entries == 0 and c.active_bins == 0
    with pytest.raises(cdx.CdxError):
        sc.read_bins()
    ref = oracle.broadphase(oracle.aabb(s), s.grid)
    np.testing.assert_array_equal(sc.read_pairs(), ref.pairs)
    sc.close()


def test_inverted_aabb_reported_after_cell_steps(cdx, ctx, scenes):
    """A scene whose last broadphase took the cell search still reports an
    AABB whose max corner lies below its min corner (the bounds pass flags
    it, so the bin count runs and reports it)."""
    s = scenes.lattice_spheres(side=12, seed=27)
    sc = _run(cdx, ctx, s, cdx.PAIRS_AUTO)
    assert sc.counts().flags & cdx.FLAG_CELL_SEARCH
    ab = sc.read_aabb()
    ab[s.n + 5] = ab[5] - 1.0  # shape 5: max corner one unit below its min corner
    sc.set_aabbs(ab)
    with pytest.raises(cdx.CdxError):
        sc.broadphase()
        sc.sync()
    sc.close()


def test_step_contacts_equal_across_searches(cdx, ctx, oracle, scenes):
    """Full steps (AABB, broadphase, NATIVE narrowphase) through each search."""
    s = scenes.lattice_spheres(side=24, seed=24, dtype=np.float64)
    out = {}
    for mode in (cdx.PAIRS_BINS, cdx.PAIRS_CELLS):
        sc = cdx.Scene(ctx, s.dtype).load(s)
        sc.set_pair_search(mode)
        sc.step()
        sc.sync()
        out[mode] = (sc.read_pairs(), sc.read_contacts())
        sc.close()
    np.testing.assert_array_equal(out[cdx.PAIRS_BINS][0], out[cdx.PAIRS_CELLS][0])
    for a, b in zip(out[cdx.PAIRS_BINS][1], out[cdx.PAIRS_CELLS][1]):
        np.testing.assert_array_equal(a, b)
    ref = oracle.narrowphase(s, out[cdx.PAIRS_CELLS][0])
    assert len(out[cdx.PAIRS_CELLS][1][3]) == ref.count
