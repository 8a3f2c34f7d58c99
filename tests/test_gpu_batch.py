"""Batches of resident scenes (cdx_batch_*) against one-scene calls and the oracle.

Every scene's pair list in a batched step is bit-exact with the one-scene
path on the same scene and grid (and with the oracle for small scenes); its
contacts are equal to the one-scene contacts (same kernels, same order) in
every narrowphase mode.  Recorded graph replays, pose updates between
replays, sub-selections and autotuned grids keep that equality.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [300, 1500, 4000, 700, 12000, 2500, 9000, 1000]


def _scenes(scenes, dtype=np.float64, sizes=SIZES, seed=100):
    out = [scenes.mixed_scene_of_size(n, 4.0, seed + i, dtype) for i, n in enumerate(sizes)]
    if len(out) > 3:
        out[3] = scenes.two_shapes_per_body(out[3])
    return out


def _single(cdx, ctx, s, mode, bins=None):
    sc = cdx.Scene(ctx, s.dtype).load(s, bins)
    sc.set_mode(mode, 1e-2)
    sc.step()
    sc.sync()
    return sc.read_pairs(), sc.read_contacts()


def _single_counts(cdx, ctx, s, mode):
    """Bin entries and active bins of the one-scene path (its broadphase
    searches by bins, as batched steps do)."""
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_pair_search(cdx.PAIRS_BINS)
    sc.set_mode(mode, 1e-2)
    sc.step()
    sc.sync()
    c = sc.counts()
    sc.close()
    return c.bin_entries, c.active_bins


def _check_scene(b, j, ref_pairs, ref_contacts):
    got = b.read_pairs(j)
    np.testing.assert_array_equal(got, ref_pairs)
    gc = b.read_contacts(j)
    for g, r in zip(gc, ref_contacts):
        np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("mode", ["mixed", "native", "fast"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_single(cdx, ctx, scenes, mode, use_graph):
    m = {"mixed": cdx.NP_MIXED, "native": cdx.NP_NATIVE, "fast": cdx.NP_FAST}[mode]
    ss = _scenes(scenes)
    print(f"batch of {len(ss)} scenes mode={mode} graph={use_graph}", flush=True)
    b = cdx.Batch(ctx, np.float64)
    for s in ss:
        b.add_scene(s)
    b.set_mode(m, 1e-2)
    b.step(use_graph)
    tot = b.finish()
    assert tot.n_shapes == sum(s.n for s in ss)
    psum = csum = 0
    for j, s in enumerate(ss):
        rp, rc = _single(cdx, ctx, s, m)
        c = b.scene_counts(j)
        assert c.pairs == len(rp) and c.contacts == len(rc[0])
        psum += c.pairs
        csum += c.contacts
        _check_scene(b, j, rp, rc)
    assert tot.pairs == psum and tot.contacts == csum
    # bin totals of the step: the sums of the one-scene totals (every scene
    # keys its bins from its own key base, so no run spans two scenes)
    ms, as_ = zip(*[_single_counts(cdx, ctx, s, m) for s in ss])
    assert tot.bin_entries == sum(ms) and tot.active_bins == sum(as_)


@pytest.mark.parametrize("sizes", [[300, 700, 1000], [20000, 30000]])
def test_batch_bin_totals(cdx, ctx, scenes, sizes):
    """Steps small enough to fold the active-bin count and the key maximum
    by atomics, and steps that reduce per-workgroup partials instead: the
    same totals as the one-scene path."""
    ss = _scenes(scenes, sizes=sizes, seed=300)
    print(f"batch sizes={sizes}", flush=True)
    b = cdx.Batch(ctx, np.float64)
    for s in ss:
        b.add_scene(s)
    for rep in range(2):
        b.step(True)
        tot = b.finish()
        ms, as_ = zip(*[_single_counts(cdx, ctx, s, cdx.NP_MIXED) for s in ss])
        assert tot.bin_entries == sum(ms) and tot.active_bins == sum(as_)
        for j, s in enumerate(ss):
            assert b.scene_counts(j).bin_entries == ms[j]


def test_batch_small_scenes_match_oracle(cdx, ctx, oracle, scenes):
    ss = _scenes(scenes, sizes=[400, 900, 250, 600], seed=300)
    b = cdx.Batch(ctx, np.float64)
    for s in ss:
        b.add_scene(s)
    b.set_mode(cdx.NP_MIXED, 1e-2)
    b.step(True)
    b.finish()
    for j, s in enumerate(ss):
        print(f"scene {j} n={s.n} nb={s.nb}", flush=True)
        ab = oracle.aabb(s)
        bp = oracle.broadphase(ab, s.grid)
        np.testing.assert_array_equal(b.read_pairs(j), bp.pairs)
        npr = oracle.narrowphase(s, bp.pairs, 0.0)
        norm, pa, pb, d, ids = b.read_contacts(j)
        assert len(d) == npr.count
        np.testing.assert_array_equal(ids, npr.bids)
        np.testing.assert_array_equal(norm, npr.norm)
        np.testing.assert_array_equal(d, npr.dpth)


def test_batch_replays_with_pose_updates(cdx, ctx, scenes):
    ss = _scenes(scenes, sizes=[2000, 5000, 800, 3000], seed=400)
    b = cdx.Batch(ctx, np.float64)
    for s in ss:
        b.add_scene(s)
    b.set_mode(cdx.NP_MIXED, 1e-2)
    rng = np.random.default_rng(7)
    for it in range(4):
        # move a tenth of the bodies of scene it % 4 by up to one shape size
        k = it % len(ss)
        s = ss[k]
        idx = rng.choice(s.nb, s.nb // 10, replace=False).astype(np.uint32)
        s.pos[idx] += rng.uniform(-0.01, 0.01, (len(idx), 3))
        s.rot[idx] = scenes.random_quats(rng, len(idx))
        b.update_bodies(k, idx, s.pos[idx], s.rot[idx])
        b.step(True)
        b.finish()
        print(f"replay {it}: moved {len(idx)} bodies of scene {k}", flush=True)
        for j, sj in enumerate(ss):
            rp, rc = _single(cdx, ctx, sj, cdx.NP_MIXED)
            _check_scene(b, j, rp, rc)


def test_batch_subselection(cdx, ctx, scenes):
    ss = _scenes(scenes, seed=500)
    b = cdx.Batch(ctx, np.float64)
    for s in ss:
        b.add_scene(s)
    b.set_mode(cdx.NP_MIXED, 1e-2)
    refs = [_single(cdx, ctx, s, cdx.NP_MIXED) for s in ss]
    for sel in ([5, 1, 6], [0], [7, 6, 5, 4, 3, 2, 1, 0], [2, 4]):
        print(f"selection {sel}", flush=True)
        b.select(sel)
        b.step(True)
        b.finish()
        for j, k in enumerate(sel):
            _check_scene(b, j, *refs[k])


def test_batch_fp32(cdx, ctx, scenes):
    ss = _scenes(scenes, dtype=np.float32, sizes=[3000, 800, 6000], seed=600)
    b = cdx.Batch(ctx, np.float32)
    for s in ss:
        b.add_scene(s)
    b.set_mode(cdx.NP_NATIVE, 1e-2)
    b.step(True)
    b.finish()
    for j, s in enumerate(ss):
        _check_scene(b, j, *_single(cdx, ctx, s, cdx.NP_NATIVE))


def test_batch_autotuned_grid(cdx, ctx, oracle, scenes):
    ss = _scenes(scenes, sizes=[1200, 1300, 5000, 20000], seed=700)
    b = cdx.Batch(ctx, np.float64)
    for s in ss:
        b.add_scene(s, bins=(0, 0, 0))
    grids = [b.scene_bins(j) for j in range(len(ss))]
    print(f"autotuned grids {grids}", flush=True)
    for g in grids:
        assert all(v >= 1 for v in g)
        assert g[0] * g[1] * g[2] <= max(64, 4 * 20000)
    b.set_mode(cdx.NP_MIXED, 1e-2)
    b.step(True)
    b.finish()
    for j, s in enumerate(ss):
        rp, rc = _single(cdx, ctx, s, cdx.NP_MIXED, bins=grids[j])
        _check_scene(b, j, rp, rc)
    # the pair list does not depend on the grid: the hand-picked grid gives the same pairs
    for j, s in enumerate(ss):
        rp, _ = _single(cdx, ctx, s, cdx.NP_MIXED)
        np.testing.assert_array_equal(b.read_pairs(j), rp)


def test_batch_errors(cdx, ctx, scenes):
    s = scenes.mixed_scene_of_size(500, 4.0, 900)
    b = cdx.Batch(ctx, np.float64)
    with pytest.raises(cdx.CdxError):
        b.step(True)  # no resident scene
    b.add_scene(s)
    with pytest.raises(cdx.CdxError):
        b.select([3])
    with pytest.raises(cdx.CdxError):
        b.update_bodies(0, [s.nb], s.pos[:1], s.rot[:1])
    bad = s.astype(np.float64)
    bad.ID = bad.ID.copy()
    bad.ID[0] = bad.nb
    with pytest.raises(cdx.CdxError):
        b.add_scene(bad)
    assert b.scene_count() == 1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["mixed", "native", "fast"])
def test_graph_step_equals_eager_step_all_types(cdx, ctx, scenes, mode, dtype):
    """A replayed graph step gives the same pairs, contacts and re-solve
    counts as the same step launched eagerly, on every selection in turn
    (the graph is recorded once), every shape type included."""
    m = {"mixed": cdx.NP_MIXED, "native": cdx.NP_NATIVE, "fast": cdx.NP_FAST}[mode]
    ss = [scenes.all_shape_types(n=3000, seed=40, dtype=dtype),
          scenes.mixed_scene_of_size(5000, 6.0, 41, dtype),
          scenes.two_shapes_per_body(scenes.mixed_scene_of_size(1200, 4.0, 42, dtype)),
          scenes.all_shape_types(n=800, seed=43, dtype=dtype)]
    print(f"graph vs eager mode={mode} dtype={np.dtype(dtype).name}", flush=True)
    res = {}
    for use_graph in (True, False):
        b = cdx.Batch(ctx, dtype)
        for s in ss:
            b.add_scene(s)
        b.set_mode(m, 1e-2)
        out = []
        for sel in (None, [2], [3, 0], None):  # all scenes, two sub-selections, all again
            b.select(sel)
            b.step(use_graph)
            tot = b.finish()
            js = range(len(ss)) if sel is None else range(len(sel))
            out.append((tot.pairs, tot.contacts, tot.resolved,
                        [(b.read_pairs(j), b.read_contacts(j)) for j in js]))
        b.close()
        res[use_graph] = out
    ref = res[False]
    for key, got in res.items():
        for (p0, c0, r0, sc0), (p1, c1, r1, sc1) in zip(ref, got):
            assert (p1, c1, r1) == (p0, c0, r0), f"{key}: totals {(p1, c1, r1)} vs {(p0, c0, r0)}"
            for (rp, rc), (gp, gc) in zip(sc0, sc1):
                np.testing.assert_array_equal(gp, rp)
                for g, r in zip(gc, rc):
                    np.testing.assert_array_equal(g, r)
