"""The benchmark workloads themselves against the CPU oracle.

* mixed-1M: 1,048,576 mixed convex bodies, FP64 input, one resident scene
  stepped in MIXED precision: the pair list bit-exact with the oracle's
  detectPossibleCollisions (Broadphase.cu:344-545), the contacts equal to
  the oracle's FP64 DoNarrowphase (Narrowphase.cu:511-595).
* batched-512: 512 scenes (log-uniform 1,024-65,536 bodies, U[2, 8]
  neighbours, autotuned grids) in one recorded step: every scene's pairs
  and contacts against the oracle on that scene and grid.
* served stream: scenes of every size class (256-262,144 bodies), Poisson
  arrivals of full pose updates (Gaussian displacement, small rotation);
  every request's contacts against the oracle on that request's poses.

MIXED re-solves every contact on the FP64 reference path, so contacts are
compared at 1e-12 x scene extent and body ids, counts and order exactly.
"""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _extent(s):
    return float(np.ptp(s.pos, axis=0).max()) + 1.0


def _check_contacts(got, ref, s, what):
    norm, pa, pb, d, ids = got
    assert len(d) == ref.count, f"{what}: {len(d)} contacts, oracle {ref.count}"
    np.testing.assert_array_equal(ids, ref.bids)
    t = 1e-12 * _extent(s)
    for g, r in ((norm, ref.norm), (pa, ref.cpta), (pb, ref.cptb)):
        np.testing.assert_allclose(g, r, atol=t, rtol=0)
    np.testing.assert_allclose(d, ref.dpth, atol=t, rtol=0)


@pytest.mark.timeout(600)
def test_mixed_1m_vs_oracle(cdx, ctx, oracle, scenes):
    s = scenes.mixed_bodies()
    print(f"mixed-1M n={s.n} grid={s.grid}", flush=True)
    ab = oracle.aabb(s)
    bp = oracle.broadphase(ab, s.grid)
    ref = oracle.narrowphase(s, bp.pairs)
    print(f"oracle: {len(bp.pairs)} pairs, {ref.count} contacts", flush=True)
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(cdx.NP_MIXED, 1e-2)
    sc.step()
    sc.sync()
    np.testing.assert_array_equal(sc.read_pairs(), bp.pairs)
    _check_contacts(sc.read_contacts(), ref, s, "mixed-1M")
    sc.close()


@pytest.mark.timeout(900)
def test_batched_512_vs_oracle(cdx, ctx, oracle, scenes):
    ss = scenes.batched_scenes(512)
    print(f"batched-512: {sum(s.n for s in ss)} bodies", flush=True)
    b = cdx.Batch(ctx, np.float64)
    for i, s in enumerate(ss):
        b.add_scene(s, (0, 0, 0))
        if i % 128 == 127:
            print(f"  {i + 1} scenes resident", flush=True)
    b.set_mode(cdx.NP_MIXED, 1e-2)
    b.step(True)
    tot = b.finish()
    print(f"step: {tot.pairs} pairs, {tot.contacts} contacts", flush=True)
    for j, s in enumerate(ss):
        grid = b.scene_bins(j)
        bp = oracle.broadphase(oracle.aabb(s), grid)
        np.testing.assert_array_equal(b.read_pairs(j), bp.pairs, err_msg=f"scene {j} n={s.n} grid={grid}")
        _check_contacts(b.read_contacts(j), oracle.narrowphase(s, bp.pairs), s, f"scene {j}")
        if j % 128 == 127:
            print(f"  {j + 1} scenes match the oracle", flush=True)
    b.close()


# Sizes drawn from the served distribution (log-uniform 256-262,144),
# at least two in each of the four size classes (< 8192, < 49152,
# < 131072, >= 131072 shapes).
SERVED_SIZES = [256, 1500, 7000, 9000, 30000, 45000, 60000, 120000, 140000, 262144]


@pytest.mark.timeout(900)
def test_served_poisson_stream_vs_oracle(cdx, ctx, oracle, scenes):
    rng = np.random.default_rng(0x5EED0004)
    ss = [scenes.fleet_scene(n, 4.0, 4000 + i) for i, n in enumerate(SERVED_SIZES)]
    srv = cdx.Server(ctx, np.float64, mode=cdx.NP_MIXED)
    ids = [srv.add_scene(s, (0, 0, 0)) for s in ss]
    classes = {srv.scene_info(i)["size_class"] for i in ids}
    assert classes == {0, 1, 2, 3}
    # Poisson arrivals at 400 requests/s from one producer; each request a
    # full pose update stepped from the scene's previous request
    nreq = 60
    plan = []
    cur = {j: (s.pos.copy(), s.rot.copy()) for j, s in enumerate(ss)}
    for _ in range(nreq):
        j = int(rng.integers(len(ss)))
        base = ss[j]
        sj = scenes.Scene(base.T, base.A, base.B, base.C, base.R, base.ID, cur[j][0], cur[j][1], base.active,
                          base.grid)
        p, q = scenes.pose_step(sj, rng)
        cur[j] = (p, q)
        plan.append((j, p, q, float(rng.exponential(1.0 / 400.0))))
    reqs = []
    t0 = time.perf_counter()
    at = 0.0
    for j, p, q, gap in plan:
        at += gap
        while time.perf_counter() - t0 < at:
            time.sleep(0.0002)
        reqs.append(srv.submit(ids[j], p, q))
    print(f"served: {nreq} requests over {len(ss)} scenes submitted", flush=True)
    for k, ((j, p, q, _), rid) in enumerate(zip(plan, reqs)):
        assert srv.wait(rid, 60_000)
        base = ss[j]
        sj = scenes.Scene(base.T, base.A, base.B, base.C, base.R, base.ID, p, q, base.active, base.grid)
        bp = oracle.broadphase(oracle.aabb(sj), srv.scene_info(ids[j])["bins"])
        _check_contacts(srv.read_contacts(rid), oracle.narrowphase(sj, bp.pairs), sj, f"request {k} scene {j}")
        srv.release(rid)
        if k % 20 == 19:
            print(f"  {k + 1} requests match the oracle", flush=True)
    st = srv.stats()
    print(f"p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms, {st['steps']} steps", flush=True)
    assert st["failed"] == 0 and st["completed"] == nreq
    srv.close()
