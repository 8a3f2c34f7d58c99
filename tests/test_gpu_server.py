"""Served request loop (cdx_server_*): alone-vs-mix equality, ordering, admission.

A request's contacts must be the same whether its step runs alone or
batched with concurrent requests of other scenes and size classes, and the
same as a one-scene step on the same poses.  The path is deterministic, so
the tolerance is zero (bit-equal arrays).
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# scenes in all four size classes (< 8192, < 49152, < 131072, >= 131072 shapes)
SIZES = [600, 2500, 3900, 5000, 12000, 30000, 60000, 1200, 9000, 140000]


def _single_contacts(cdx, ctx, s, mode):
    sc = cdx.Scene(ctx, s.dtype).load(s)
    sc.set_mode(mode, 1e-2)
    sc.step()
    sc.sync()
    return sc.read_contacts()


def _equal(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.fixture(scope="module")
def served(cdx, ctx, scenes):
    ss = [scenes.mixed_scene_of_size(n, 4.0, 2000 + i) for i, n in enumerate(SIZES)]
    srv = cdx.Server(ctx, np.float64, mode=cdx.NP_MIXED, max_batch_shapes=1 << 17)
    ids = [srv.add_scene(s) for s in ss]
    yield srv, ss, ids
    srv.close()


def test_alone_equals_mix(cdx, ctx, scenes, served):
    srv, ss, ids = served
    rng = np.random.default_rng(11)
    target = 4
    s = ss[target]
    pos, rot = scenes.pose_step(s, rng)
    # alone
    r = srv.submit(ids[target], pos, rot)
    assert srv.wait(r, 30_000)
    alone = srv.read_contacts(r)
    srv.release(r)
    print(f"alone: {len(alone[3])} contacts", flush=True)
    s.pos[:], s.rot[:] = pos, rot
    _equal(alone, _single_contacts(cdx, ctx, s, cdx.NP_MIXED))
    # the same poses again, submitted among requests for every other scene
    updates = {}
    reqs = []
    for j, sj in enumerate(ss):
        if j == target:
            continue
        p, q = scenes.pose_step(sj, rng)
        updates[j] = (p, q)
        reqs.append((j, srv.submit(ids[j], p, q)))
    rt = srv.submit(ids[target], pos, rot)
    for j, sj in enumerate(ss):
        if j != target and j % 2 == 0:
            p, q = scenes.pose_step(sj, rng)
            updates[j] = (p, q)  # the later request for j wins
            reqs.append((j, srv.submit(ids[j], p, q)))
    assert srv.wait(rt, 30_000)
    mixed = srv.read_contacts(rt)
    srv.release(rt)
    _equal(mixed, alone)
    # every other request equals its scene stepped alone on its poses
    last = {}
    for j, rid in reqs:
        assert srv.wait(rid, 30_000)
        last[j] = rid
    for j, rid in last.items():
        sj = ss[j]
        sj.pos[:], sj.rot[:] = updates[j]
        print(f"scene {j} n={sj.n}", flush=True)
        _equal(srv.read_contacts(rid), _single_contacts(cdx, ctx, sj, cdx.NP_MIXED))
    for _, rid in reqs:
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] >= len(reqs) + 2
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_same_scene_requests_run_in_order(cdx, ctx, scenes, served):
    srv, ss, ids = served
    rng = np.random.default_rng(12)
    k = 1
    s = ss[k]
    steps = []
    cur = s
    for _ in range(4):
        p, q = scenes.pose_step(cur, rng)
        cur = scenes.Scene(s.T, s.A, s.B, s.C, s.R, s.ID, p, q, s.active, s.grid)
        steps.append((p, q))
    rids = [srv.submit(ids[k], p, q) for p, q in steps]
    for rid in rids:
        assert srv.wait(rid, 30_000)
    for (p, q), rid in zip(steps, rids):
        s.pos[:], s.rot[:] = p, q
        _equal(srv.read_contacts(rid), _single_contacts(cdx, ctx, s, cdx.NP_MIXED))
        srv.release(rid)


def test_concurrent_submitters(cdx, ctx, scenes, served):
    srv, ss, ids = served
    small = [j for j, s in enumerate(ss) if s.n < 10_000]
    results = {}
    errors = []

    def client(j, seed):
        try:
            rng = np.random.default_rng(seed)
            p, q = scenes.pose_step(ss[j], rng)
            rid = srv.submit(ids[j], p, q)
            srv.wait(rid, 30_000)
            results[j] = (p, q, srv.read_contacts(rid))
            srv.release(rid)
        except Exception as e:  # surfaced below
            errors.append(e)

    th = [threading.Thread(target=client, args=(j, 100 + j)) for j in small]
    for t in th:
        t.start()
    for t in th:
        t.join(60)
    assert not errors, errors
    for j, (p, q, got) in results.items():
        s = ss[j]
        s.pos[:], s.rot[:] = p, q
        _equal(got, _single_contacts(cdx, ctx, s, cdx.NP_MIXED))


def test_partial_pose_updates(cdx, ctx, scenes, served):
    srv, ss, ids = served
    rng = np.random.default_rng(13)
    j = 7
    s = ss[j]
    p, q = scenes.pose_step(s, rng)
    idx = rng.choice(s.nb, s.nb // 3, replace=False).astype(np.uint32)
    rid = srv.submit(ids[j], p[idx], q[idx], idx=idx)
    assert srv.wait(rid, 30_000)
    s.pos[idx], s.rot[idx] = p[idx], q[idx]
    _equal(srv.read_contacts(rid), _single_contacts(cdx, ctx, s, cdx.NP_MIXED))
    srv.release(rid)


def test_admission_by_hbm_budget(cdx, ctx, scenes):
    s = scenes.mixed_scene_of_size(20_000, 4.0, 3000)
    srv = cdx.Server(ctx, np.float64, hbm_budget=64 << 20, max_batch_shapes=1 << 15)
    admitted = 0
    with pytest.raises(cdx.CdxError) as ei:
        for _ in range(64):
            srv.add_scene(s)
            admitted += 1
    assert ei.value.status == 4  # CDX_ERR_CAPACITY
    m = srv.memory()
    print(f"admitted {admitted} scenes, {m}", flush=True)
    assert 1 <= admitted < 64
    assert m["hbm_committed"] <= m["hbm_budget"]
    # the admitted scenes still serve
    rid = srv.submit(0, s.pos, s.rot)
    assert srv.wait(rid, 30_000)
    _equal(srv.read_contacts(rid), _single_contacts(cdx, ctx, s, cdx.NP_MIXED))
    srv.close()


def test_server_errors(cdx, ctx, scenes, served):
    srv, ss, ids = served
    with pytest.raises(cdx.CdxError):
        srv.submit(len(ids) + 5, ss[0].pos, ss[0].rot)
    with pytest.raises(cdx.CdxError):
        srv.submit(ids[0], ss[0].pos[:2], ss[0].rot[:2], idx=np.array([0, ss[0].nb], np.uint32))
    with pytest.raises(cdx.CdxError):
        srv.result(123456789)


def test_pose_buffer_submission(cdx, ctx, scenes, served):
    srv, ss, ids = served
    rng = np.random.default_rng(14)
    j = 2
    s = ss[j]
    p, q = scenes.pose_step(s, rng)
    h, pos, rot = srv.pose_buffer(ids[j], s.nb)
    pos[:] = p
    rot[:] = q
    rid = srv.submit_buffer(h)
    assert srv.wait(rid, 30_000)
    s.pos[:], s.rot[:] = p, q
    _equal(srv.read_contacts(rid), _single_contacts(cdx, ctx, s, cdx.NP_MIXED))
    srv.release(rid)
    h2, _, _ = srv.pose_buffer(ids[j], 10)
    srv.drop_buffer(h2)
    with pytest.raises(cdx.CdxError):
        srv.submit_buffer(h2)
