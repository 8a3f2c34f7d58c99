"""Pins of the oracle by mathematics independent of its code.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a swapped sign, a transposed rotation, a wrong composition order,
an off-by-one cell) fails it.
"""
import numpy as np
import pytest

SPHERE, ELLIPSOID, BOX, CYLINDER, CAPSULE = 0, 1, 2, 3, 7


def quat_to_mat(q):
    """Textbook rotation matrix of unit quaternion (w,x,y,z)."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def hamilton(a, b):
    aw, av = a[0], np.asarray(a[1:])
    bw, bv = b[0], np.asarray(b[1:])
    return np.concatenate([[aw * bw - av @ bv], aw * bv + bw * av + np.cross(av, bv)])


def rand_quat(rng):
    q = rng.standard_normal(4)
    return q / np.linalg.norm(q)


# ------------------------------------------------------------ arithmetic --

def test_rotation_is_q_v_qconj(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rand_quat(rng)
        v = rng.standard_normal(3)
        ref = hamilton(hamilton(q, np.concatenate([[0], v])), q * np.array([1, -1, -1, -1]))[1:]
        np.testing.assert_allclose(oracle.quat_rotate(v, q), ref, atol=1e-14)


def test_mult_is_hamilton_product(oracle):
    rng = np.random.default_rng(2)
    for _ in range(200):
        a, b = rand_quat(rng), rand_quat(rng)
        np.testing.assert_allclose(oracle.quat_mult(a, b), hamilton(a, b), atol=1e-15)
        # composition order: rotating by mult(a,b) == rotating by b then by a
        v = rng.standard_normal(3)
        np.testing.assert_allclose(oracle.quat_rotate(v, oracle.quat_mult(a, b)),
                                   oracle.quat_rotate(oracle.quat_rotate(v, b), a), atol=1e-14)


# ------------------------------------------------------------------ AABB --

def test_sphere_aabb_closed_form(oracle, scenes):
    s = scenes.small_spheres(n=2000)
    ab = oracle.aabb(s)
    r = s.B[:, :1]
    np.testing.assert_array_equal(ab[:s.n], s.pos - r)
    np.testing.assert_array_equal(ab[s.n:], s.pos + r)


def _corner_bounds(center, Rm, half):
    signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=float)
    pts = center + (signs * half) @ Rm.T
    return pts.min(axis=0), pts.max(axis=0)


@pytest.mark.parametrize("t", [BOX, ELLIPSOID, CYLINDER, CAPSULE])
def test_box_like_aabb_vs_corners(oracle, scenes, t):
    """Box-like AABB == bounds of the 8 rotated corners of the bounding box,
    with a non-trivial local position and local rotation on top of the body pose."""
    rng = np.random.default_rng(3 + t)
    n = 300
    s = scenes.mixed_bodies(n=n, seed=4)
    s.T[:] = t
    s.A = rng.uniform(-0.05, 0.05, (n, 3))
    s.R = np.array([rand_quat(rng) for _ in range(n)])
    ab = oracle.aabb(s)
    for i in range(n):
        Rb = quat_to_mat(s.rot[i])
        Rl = quat_to_mat(s.R[i])
        c = Rb @ s.A[i] + s.pos[i]
        half = s.B[i].copy()
        if t == CAPSULE:
            half[1] += half[0]
        lo, hi = _corner_bounds(c, Rb @ Rl, half)
        np.testing.assert_allclose(ab[i], lo, atol=1e-14)
        np.testing.assert_allclose(ab[n + i], hi, atol=1e-14)


# ------------------------------------------------------------ broadphase --

def _brute_pairs(shifted):
    n = len(shifted) // 2
    mn, mx = shifted[:n], shifted[n:]
    out = []
    for i in range(n - 1):
        ov = np.all((mn[i] <= mx[i + 1:]) & (mn[i + 1:] <= mx[i]), axis=1)
        j = np.nonzero(ov)[0] + i + 1
        out.append((np.int64(i) << 32) | j.astype(np.int64))
    return np.sort(np.concatenate(out)) if out else np.zeros(0, np.int64)


def _bin_reference(shifted, inv, grid):
    """Independent cell enumeration: truncation toward zero, keys in unsigned 32-bit."""
    n = len(shifted) // 2
    dt = shifted.dtype
    g0 = (shifted[:n] * inv.astype(dt)).astype(np.int64)
    g1 = (shifted[n:] * inv.astype(dt)).astype(np.int64)
    counts = np.prod(g1 - g0 + 1, axis=1).astype(np.uint32)
    keys, bodies = [], []
    gx, gy = np.uint64(grid[0]), np.uint64(grid[1])
    for i in range(n):
        xs, ys, zs = (np.arange(g0[i, c], g1[i, c] + 1, dtype=np.uint64) for c in range(3))
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        k = ((Z * gy * gx + Y * gx + X) & np.uint64(0xffffffff)).ravel()
        keys.append(k)
        bodies.append(np.full(k.size, i, np.uint64))
    keys = np.concatenate(keys)
    bodies = np.concatenate(bodies)
    order = np.lexsort((bodies, keys))
    return counts, keys[order].astype(np.uint32), bodies[order].astype(np.uint32)


@pytest.mark.parametrize("case", ["spheres_f64", "mixed_f64", "mixed_f32", "lattice_f32", "small_grid"])
def test_broadphase_vs_independent(oracle, scenes, case):
    s = {
        "spheres_f64": lambda: scenes.small_spheres(),
        "mixed_f64": lambda: scenes.mixed_bodies(n=6000, seed=5),
        "mixed_f32": lambda: scenes.mixed_bodies(n=6000, seed=6, dtype=np.float32),
        "lattice_f32": lambda: scenes.lattice_spheres(side=12, seed=7),
        "small_grid": lambda: scenes.mixed_bodies(n=3000, seed=8, grid=(3, 5, 2)),
    }[case]()
    ab = oracle.aabb(s)
    r = oracle.broadphase(ab, s.grid)
    # bounds narrowed to float (Broadphase.cu:368-376) and the in-place shift
    np.testing.assert_array_equal(r.bmin, ab.min(axis=0).astype(np.float32))
    np.testing.assert_array_equal(r.bmax, ab.max(axis=0).astype(np.float32))
    origin = np.abs(r.bmin.astype(s.dtype))
    np.testing.assert_array_equal(r.aabb_shifted, ab + origin)
    ext = np.abs(r.bmax.astype(s.dtype) + np.abs(r.bmin.astype(s.dtype)))
    inv = np.asarray(s.grid, dtype=s.dtype) / ext
    counts, keys, bodies = _bin_reference(r.aabb_shifted, inv, s.grid)
    np.testing.assert_array_equal(r.bin_counts, counts)
    np.testing.assert_array_equal(r.bins_intersected, np.cumsum(counts, dtype=np.uint64).astype(np.uint32))
    np.testing.assert_array_equal(r.sorted_keys, keys)
    np.testing.assert_array_equal(r.sorted_bodies, bodies)
    uk, uc = np.unique(keys, return_counts=True)
    np.testing.assert_array_equal(r.active_keys, uk)
    np.testing.assert_array_equal(r.active_counts, uc.astype(np.uint32))
    np.testing.assert_array_equal(r.pairs, _brute_pairs(r.aabb_shifted))
    assert r.pdup >= len(r.pairs)


def test_broadphase_parallel_flags_agree(oracle, scenes):
    s = scenes.mixed_bodies(n=20000, seed=9)
    ab = oracle.aabb(s)
    a = oracle.broadphase(ab, s.grid, oracle.DEFAULT_FLAGS)
    b = oracle.broadphase(ab, s.grid, oracle.ALL_PARALLEL)
    np.testing.assert_array_equal(a.sorted_keys, b.sorted_keys)
    np.testing.assert_array_equal(a.sorted_bodies, b.sorted_bodies)
    np.testing.assert_array_equal(a.pairs, b.pairs)


# ----------------------------------------------------------- narrowphase --

def test_sphere_sphere_closed_form(oracle, scenes):
    s = scenes.small_spheres()
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    r = oracle.narrowphase(s, pairs)
    a = (pairs >> 32)
    b = pairs & 0xffffffff
    d = s.pos[b] - s.pos[a]
    dist = np.linalg.norm(d, axis=1)
    rs = s.B[a, 0] + s.B[b, 0]
    hit = np.sum(d * d, axis=1) <= rs * rs
    np.testing.assert_array_equal(r.hit.astype(bool), hit)
    a, b, d, dist = a[hit], b[hit], d[hit], dist[hit]
    N = d / dist[:, None]
    np.testing.assert_allclose(r.norm, N, atol=1e-13)
    np.testing.assert_allclose(r.cpta, s.pos[a] + N * s.B[a, :1], atol=1e-13)
    np.testing.assert_allclose(r.cptb, s.pos[b] - N * s.B[b, :1], atol=1e-13)
    np.testing.assert_allclose(r.dpth, dist - s.B[a, 0] - s.B[b, 0], atol=1e-13)
    np.testing.assert_array_equal(r.bids, np.stack([a, b], axis=1))


def _support_brute(t, B, Rm, c, d):
    """Support point by maximising over a dense sample of the shape's surface."""
    if t == BOX:
        signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=float)
        pts = (signs * B) @ Rm.T + c
        return pts[np.argmax(pts @ d)]
    dl = Rm.T @ d
    if t == SPHERE:
        return c + B[0] * d / np.linalg.norm(d)
    if t == ELLIPSOID:
        # argmax over the ellipsoid of <x, dl> = B^2 dl / |B dl|
        x = B * B * dl / np.linalg.norm(B * dl)
        return Rm @ x + c
    if t == CYLINDER:
        w = np.array([dl[0], 0, dl[2]])
        x = np.sign(dl[1]) * B[1] * np.array([0, 1, 0]) + B[0] * w / np.linalg.norm(w)
        return Rm @ x + c
    if t == CAPSULE:
        x = np.sign(dl[1]) * B[1] * np.array([0, 1, 0]) + B[0] * dl / np.linalg.norm(dl)
        return Rm @ x + c


@pytest.mark.parametrize("t", [SPHERE, ELLIPSOID, BOX, CYLINDER, CAPSULE])
def test_support_maps(oracle, t):
    """World support points: maximal along the direction, and equal to the closed form."""
    rng = np.random.default_rng(10 + t)
    for _ in range(100):
        q = rand_quat(rng)
        Rm = quat_to_mat(q)
        B = rng.uniform(0.5, 1.5, 3)
        c = rng.standard_normal(3)
        d = rng.standard_normal(3)
        got = oracle.support(t, c, B, q, d)
        np.testing.assert_allclose(got, _support_brute(t, B, Rm, c, d), atol=1e-12)
        # no sampled surface point lies further along d
        u = rng.standard_normal((400, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        sample = np.array([oracle.support(t, c, B, q, v) for v in u[:40]])
        assert np.all(sample @ d <= got @ d + 1e-12)


IDQ = np.array([1.0, 0, 0, 0])


def _mpr_case(tA, BA, XA, qA, tB, BB, XB, qB):
    import oracle as o
    hit, n, p, depth = o.mpr(tA, XA, BA, qA, tB, XB, BB, qB)
    return hit, n, p, depth


@pytest.mark.parametrize("dx", [0.9, 0.5, 1.1])
def test_mpr_box_box_face(oracle, dx):
    """Axis-aligned unit-half boxes offset along x: overlap depth = 2*0.5 - dx along x."""
    XA = np.zeros(3)
    XB = np.array([dx, 0.013, -0.021])
    B = np.array([0.5, 0.5, 0.5])
    hit, n, p, depth = _mpr_case(BOX, B, XA, IDQ, BOX, B, XB, IDQ)
    if dx > 1.0:
        assert not hit
        return
    assert hit
    np.testing.assert_allclose(np.abs(n), [1, 0, 0], atol=1e-9)
    np.testing.assert_allclose(depth, 1.0 - dx, atol=1e-9)


@pytest.mark.parametrize("t,axis,expect", [
    (CAPSULE, 1, lambda r, h, rs, d: h + r + rs - d),   # along the capsule axis
    (CAPSULE, 0, lambda r, h, rs, d: r + rs - d),       # radial
    (CYLINDER, 1, lambda r, h, rs, d: h + rs - d),      # onto the flat cap
    (CYLINDER, 0, lambda r, h, rs, d: r + rs - d),      # radial
    (ELLIPSOID, 0, lambda r, h, rs, d: r + rs - d),     # equal axes == sphere
])
def test_mpr_closed_forms_vs_sphere(oracle, t, axis, expect):
    """Shape at the origin, sphere of radius rs displaced along one axis (slightly off-axis)."""
    r, h, rs = 0.3, 0.6, 0.25
    B = np.array([r, h, r]) if t != ELLIPSOID else np.array([r, r, r])
    for d in (0.2, 0.45, 0.7):
        # axial cases start from the end of the shape's core (cap plane / segment end)
        dist = d + h if axis == 1 else d
        XB = np.zeros(3)
        XB[axis] = dist
        XB[(axis + 1) % 3] += 1e-4
        hit, n, p, depth = _mpr_case(t, B, np.zeros(3), IDQ, SPHERE, np.array([rs, 0, 0]), XB, IDQ)
        want = expect(r, h, rs, dist)
        if want <= 0:
            continue
        assert hit
        assert abs(n[axis]) > 0.999
        np.testing.assert_allclose(depth, want, atol=2e-3 * rs)


def test_contact_invariants(oracle, scenes):
    """|N| = 1, penetration depth <= 0 and depth = N.(p2 - p1) for every contact."""
    s = scenes.mixed_bodies(n=6000, seed=15)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    r = oracle.narrowphase(s, pairs)
    assert r.count > 100
    np.testing.assert_allclose(np.linalg.norm(r.norm, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.sum(r.norm * (r.cptb - r.cpta), axis=1), r.dpth, atol=1e-13)
    assert np.all(r.dpth <= 1e-12)
    # contacts are the hits in pair order; misses follow, order kept
    hits = pairs[r.hit.astype(bool)]
    np.testing.assert_array_equal(r.pairs_permuted[: r.count], hits)
    np.testing.assert_array_equal(r.pairs_permuted[r.count:], pairs[~r.hit.astype(bool)])


def test_inactive_and_same_body_are_skipped(oracle, scenes):
    s = scenes.small_spheres(n=3000)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    full = oracle.narrowphase(s, pairs)
    s.active[:] = 0
    s.active[::2] = 1
    part = oracle.narrowphase(s, pairs)
    a = pairs >> 32
    b = pairs & 0xffffffff
    keep = (s.active[a] == 1) | (s.active[b] == 1)
    np.testing.assert_array_equal(part.hit.astype(bool), full.hit.astype(bool) & keep)
    s.ID[:] = 0  # every shape on body 0 -> no contacts at all
    assert oracle.narrowphase(s, pairs).count == 0


def test_envelope_adds_twice_to_depth(oracle, scenes):
    s = scenes.small_spheres(n=3000)
    ab = oracle.aabb(s)
    pairs = oracle.broadphase(ab, s.grid).pairs
    r0 = oracle.narrowphase(s, pairs)
    r1 = oracle.narrowphase(s, pairs, envelope=0.002)
    np.testing.assert_allclose(r1.dpth, r0.dpth + 0.004, atol=1e-13)


def _pair_scene(scenes, ta, tb, xb, dims=0.5):
    T = np.array([ta, tb], np.int32)
    z3 = np.zeros((2, 3))
    B = np.full((2, 3), dims)
    R = np.tile([1.0, 0.0, 0.0, 0.0], (2, 1))
    pos = np.array([[0.0, 0.0, 0.0], list(xb)])
    return scenes.Scene(T, z3.copy(), B, z3.copy(), R, np.arange(2, dtype=np.uint32), pos, R.copy(),
                        np.ones(2, np.uint8), (4, 4, 4))


def test_mpr_steps_counts(oracle, scenes):
    """Support steps per pair: none for sphere pairs (closed form), one for a
    pair separated along the centre line (the first portal test ends it),
    two for a hit whose first support lies on the centre line (the
    degenerate early hit, Narrowphase.cu:161-171, plus the contact points),
    and otherwise the discovery steps (2), at least one refinement step and
    the contact points (1), at most the caps (2 + 35 + 35 + 1)."""
    pair = np.array([(0 << 32) | 1], np.int64)
    assert oracle.mpr_steps(_pair_scene(scenes, SPHERE, SPHERE, (0.5, 0, 0)), pair)[0] == 0
    assert oracle.mpr_steps(_pair_scene(scenes, BOX, BOX, (3.0, 0, 0)), pair)[0] == 1
    assert oracle.mpr_steps(_pair_scene(scenes, SPHERE, BOX, (0.7, 0, 0)), pair)[0] == 2
    k = int(oracle.mpr_steps(_pair_scene(scenes, SPHERE, BOX, (0.7, 0.2, 0.1)), pair)[0])
    assert 4 <= k <= 73
    # the count is the work of the narrowphase itself: a hit costs more than a miss
    s = scenes.mixed_bodies(n=4096, seed=3)
    pairs = oracle.broadphase(oracle.aabb(s), s.grid).pairs
    steps = oracle.mpr_steps(s, pairs)
    hit = oracle.narrowphase(s, pairs).hit.astype(bool)
    a, b = pairs >> 32, pairs & 0xffffffff
    mpr = ~((s.T[a] == SPHERE) & (s.T[b] == SPHERE))
    assert (steps[~mpr] == 0).all()
    assert (steps[mpr] >= 1).all() and steps.max() <= 73
    assert steps[mpr & hit].min() >= 2 and np.median(steps[mpr & hit]) >= 4


# ------------------------------------------- MPR contact point (barycentric) --

def _dims(rng, t):
    if t == SPHERE:
        return np.array([rng.uniform(0.2, 0.5), 0.0, 0.0])
    if t in (BOX, ELLIPSOID):
        return rng.uniform(0.2, 0.6, 3)
    r = rng.uniform(0.15, 0.5)
    return np.array([r, rng.uniform(0.2, 0.6), r])  # cylinder / capsule: (r, half length, r)


def _barycentric(portal):
    """Barycentric coordinates of the origin, by a linear solve: in the
    tetrahedron v0..v3, or (fallback) of its projection along the portal
    normal onto the triangle v1 v2 v3."""
    if not portal.fallback:
        M = np.vstack([portal.v.T, np.ones(4)])
        return np.linalg.solve(M, [0.0, 0.0, 0.0, 1.0])
    M = np.zeros((4, 4))
    M[:3, :3] = portal.v[1:].T
    M[:3, 3] = -portal.n
    M[3, :3] = 1.0
    return np.concatenate([[0.0], np.linalg.solve(M, [0.0, 0.0, 0.0, 1.0])[:3]])


def _outside_by(t, B, q, c, p, extra_dirs=()):
    """How far p lies outside the shape: max over directions u of
    u.p - h(u), with h the support function (numpy closed forms)."""
    rng = np.random.default_rng(99)
    u = rng.standard_normal((400, 3))
    u = np.vstack([u, np.eye(3), -np.eye(3)] + [np.atleast_2d(d) for d in extra_dirs])
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    Rm = quat_to_mat(q)
    h = np.array([_support_brute(t, B, Rm, c, d) @ d for d in u])
    return float(np.max(u @ p - h))


_CURVED_TYPES = [SPHERE, ELLIPSOID, BOX, CYLINDER, CAPSULE]


def _random_barycentric_hits(oracle_lib, oracle, count, seed):
    """Penetrating pairs of every shape-type pair (sphere-sphere excluded:
    the narrowphase solves it in closed form) whose MPR ends in the
    barycentric contact point (Narrowphase.cu:265-289)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        tA, tB = rng.choice(_CURVED_TYPES, 2)
        if tA == SPHERE and tB == SPHERE:
            continue
        BA, BB = _dims(rng, tA), _dims(rng, tB)
        qA, qB = rand_quat(rng), rand_quat(rng)
        XA = rng.uniform(-1, 1, 3)
        d = rng.standard_normal(3)
        XB = XA + d * rng.uniform(0.2, 0.9) / np.linalg.norm(d)
        hit, n, p, depth, portal = oracle.mpr_portal(tA, XA, BA, qA, tB, XB, BB, qB, L=oracle_lib)
        if hit and portal.filled:
            out.append(((tA, BA, XA, qA), (tB, BB, XB, qB), n, p, depth, portal))
    return out


def _check_barycentric_point(case, tol=1e-6):
    """The contact point is the mean of the two halves sum(l_i a_i) and
    sum(l_i b_i), l the barycentric coordinates of the origin in the final
    portal; with the origin inside (l >= 0) the halves coincide, so the
    point lies in both shapes."""
    (tA, BA, XA, qA), (tB, BB, XB, qB), n, p, depth, portal = case
    lam = _barycentric(portal)
    assert lam.min() > -1e-9, f"origin outside the final portal: {lam}"
    want = (lam[:, None] * (portal.a + portal.b)).sum(axis=0) / 2
    np.testing.assert_allclose(p, want, rtol=0, atol=tol)
    # the halves of a correct point coincide (sum l v = 0), so it lies in A and in B
    assert _outside_by(tA, BA, qA, XA, p, [n, -n]) < tol
    assert _outside_by(tB, BB, qB, XB, p, [n, -n]) < tol


def test_mpr_contact_point_barycentric(oracle):
    """Pins Narrowphase.cu:265-289 on 400 random penetrating pairs of every
    shape-type pair that reach the barycentric branch: the oracle's point
    equals the point built from a numpy solve of the barycentric system on
    its own final portal, and lies inside both shapes."""
    cases = _random_barycentric_hits(None, oracle, 400, seed=21)
    kinds = {(c[0][0], c[1][0]) for c in cases}
    assert len(kinds) >= 20  # every ordered type pair but sphere-sphere
    for case in cases:
        _check_barycentric_point(case)


@pytest.mark.parametrize("name,tA,BA,qA,tB,BB,XB,qB", [
    ("box-box off-axis", BOX, [0.5, 0.4, 0.3], [0.99, 0.0, 0.0, 0.141], BOX, [0.2, 0.25, 0.15], [0.62, 0.13, -0.07],
     [0.955, 0.296, 0.0, 0.0]),
    ("sphere-box off-centre face", BOX, [0.5, 0.5, 0.5], [1, 0, 0, 0], SPHERE, [0.3, 0, 0], [0.7, 0.2, 0.1],
     [1, 0, 0, 0]),
    ("capsule-box", CAPSULE, [0.2, 0.4, 0.2], [0.92, 0.39, 0.0, 0.0], BOX, [0.3, 0.3, 0.3], [0.1, 0.45, 0.38],
     [0.98, 0.0, 0.2, 0.0]),
    ("ellipsoid (unequal axes)-sphere", ELLIPSOID, [0.5, 0.3, 0.2], [0.9, 0.3, 0.3, 0.1], SPHERE, [0.25, 0, 0],
     [0.35, 0.3, 0.2], [1, 0, 0, 0]),
])
def test_mpr_contact_point_named_cases(oracle, name, tA, BA, qA, tB, BB, XB, qB):
    qA = np.asarray(qA, float) / np.linalg.norm(qA)
    qB = np.asarray(qB, float) / np.linalg.norm(qB)
    XA = np.zeros(3)
    hit, n, p, depth, portal = oracle.mpr_portal(tA, XA, np.asarray(BA, float), qA, tB, np.asarray(XB, float),
                                                 np.asarray(BB, float), qB)
    assert hit and portal.filled, name
    _check_barycentric_point(((tA, np.asarray(BA, float), XA, qA), (tB, np.asarray(BB, float), np.asarray(XB, float),
                                                                      qB), n, p, depth, portal))
    if name.startswith("sphere-box"):
        # sphere centre c over the face x = h of a box at the origin, portal on
        # the face: the halves meet at x = h D / (h + r) on the sphere's axis
        h, r, D = 0.5, 0.3, XB[0]
        np.testing.assert_allclose(p, [h * D / (h + r), XB[1], XB[2]], atol=1e-12)


_NAMED_FLAT = [  # face contacts: the final portal's vertices stay apart, so every term counts
    (BOX, [0.5, 0.4, 0.3], [0.99, 0.0, 0.0, 0.141], BOX, [0.2, 0.25, 0.15], [0.62, 0.13, -0.07], [0.955, 0.296, 0, 0]),
    (BOX, [0.5, 0.5, 0.5], [1, 0, 0, 0], SPHERE, [0.3, 0, 0], [0.7, 0.2, 0.1], [1, 0, 0, 0]),
]


@pytest.mark.parametrize("slip,min_frac", [(1, 0.9), (2, 0.3), (3, 0.3)])
def test_mpr_pins_catch_planted_slips(oracle, slip, min_frac):
    """Oracle builds with a planted slip in the contact point (1: the halving
    dropped, 2: the v2/v3 weights swapped, 3: v1's weight on v2's A point)
    fail the barycentric pin on both face-contact cases and on most random
    cases.  Slips 2 and 3 change nothing where the refinement has shrunk
    the portal onto one point of a curved surface (the swapped support
    points coincide there), hence the lower fraction."""
    L = oracle.planted(slip)
    for tA, BA, qA, tB, BB, XB, qB in _NAMED_FLAT:
        qA = np.asarray(qA, float) / np.linalg.norm(qA)
        qB = np.asarray(qB, float) / np.linalg.norm(qB)
        BA, BB, XB, XA = np.asarray(BA, float), np.asarray(BB, float), np.asarray(XB, float), np.zeros(3)
        hit, n, p, depth, portal = oracle.mpr_portal(tA, XA, BA, qA, tB, XB, BB, qB, L=L)
        with pytest.raises(AssertionError):
            _check_barycentric_point(((tA, BA, XA, qA), (tB, BB, XB, qB), n, p, depth, portal))
    cases = _random_barycentric_hits(L, oracle, 60, seed=22)
    failed = 0
    for case in cases:
        try:
            _check_barycentric_point(case)
        except AssertionError:
            failed += 1
    assert failed >= min_frac * len(cases), f"slip {slip} failed only {failed} of {len(cases)} cases"


def _np_two(scenes, shapes, envelope=0.0):
    """Narrowphase of one pair of world-placed shapes (one body each)."""
    T = np.array([s[0] for s in shapes], np.int32)
    B = np.array([np.asarray(s[1], float) for s in shapes])
    pos = np.array([np.asarray(s[2], float) for s in shapes])
    rot = np.array([np.asarray(s[3], float) for s in shapes])
    z3 = np.zeros((2, 3))
    R = np.tile([1.0, 0.0, 0.0, 0.0], (2, 1))
    sc = scenes.Scene(T, z3.copy(), B, z3.copy(), R, np.arange(2, dtype=np.uint32), pos, rot, np.ones(2, np.uint8),
                      (4, 4, 4))
    return sc


def test_contact_points_on_support_planes(oracle, scenes):
    """function_MPR_Store (Narrowphase.cu:404-415): p1 and p2 lie on the line
    through the MPR point along N, p1 on A's supporting plane of normal N
    and p2 on B's of normal -N (numpy support functions), so depth is the
    negated width of the Minkowski difference along N."""
    cases = _random_barycentric_hits(None, oracle, 150, seed=23)
    pair = np.array([(0 << 32) | 1], np.int64)
    for (tA, BA, XA, qA), (tB, BB, XB, qB), n_mpr, p0, _, _ in cases:
        sc = _np_two(scenes, [(tA, BA, XA, qA), (tB, BB, XB, qB)])
        r = oracle.narrowphase(sc, pair)
        assert r.count == 1
        N, p1, p2 = r.norm[0], r.cpta[0], r.cptb[0]
        np.testing.assert_allclose(N, -n_mpr, atol=1e-15)  # N points from A to B
        hA = _support_brute(tA, BA, quat_to_mat(qA), XA, N) @ N
        hB = _support_brute(tB, BB, quat_to_mat(qB), XB, -N) @ N
        np.testing.assert_allclose(p1 @ N, hA, atol=1e-12)
        np.testing.assert_allclose(p2 @ N, hB, atol=1e-12)
        np.testing.assert_allclose(np.cross(p1 - p0, N), 0, atol=1e-12)
        np.testing.assert_allclose(np.cross(p2 - p0, N), 0, atol=1e-12)
        assert r.dpth[0] <= 0


def test_swap_symmetry(oracle, scenes):
    """(a, b) -> (b, a) flips N, swaps p1 and p2 and keeps the depth.
    Exact to rounding where the contact is between flat faces (box-box,
    where the MPR stops on the face); on curved surfaces the reference's
    34-iteration cap (Narrowphase.cu:263) leaves the portal, and so the
    normal, dependent on the order of exploration, which the swapped pair
    changes: measured up to ~6e-3 in the normal, median ~1e-6."""
    pair = np.array([(0 << 32) | 1], np.int64)
    rng = np.random.default_rng(24)
    flat, curved = [], []
    for _ in range(300):
        tA, tB = rng.choice(_CURVED_TYPES, 2)
        if tA == SPHERE and tB == SPHERE:
            continue
        a = (tA, _dims(rng, tA), np.zeros(3), rand_quat(rng))
        d = rng.standard_normal(3)
        b = (tB, _dims(rng, tB), d * rng.uniform(0.2, 0.9) / np.linalg.norm(d), rand_quat(rng))
        r1 = oracle.narrowphase(_np_two(scenes, [a, b]), pair)
        r2 = oracle.narrowphase(_np_two(scenes, [b, a]), pair)
        if r1.count != 1 or r2.count != 1:
            continue
        err = (np.abs(r1.norm[0] + r2.norm[0]).max(), abs(r1.dpth[0] - r2.dpth[0]),
               np.abs(r1.cpta[0] - r2.cptb[0]).max(), np.abs(r1.cptb[0] - r2.cpta[0]).max())
        (flat if tA == BOX and tB == BOX else curved).append(err)
    flat, curved = np.array(flat), np.array(curved)
    assert len(flat) >= 5 and len(curved) >= 100
    assert flat.max() < 1e-12
    assert np.median(curved[:, 0]) < 1e-5 and curved[:, 0].max() < 2e-2
    assert np.median(curved[:, 1]) < 1e-6 and curved[:, 1].max() < 2e-3
    assert np.median(curved[:, 2:]) < 1e-5 and curved[:, 2:].max() < 2e-2


def test_flop_census_sphere_pair_closed_form(oracle, scenes):
    """The FLOP census counts each +, -, *, / and sqrt of function_MPR_Store
    after placement (a fused multiply-add as two): a touching sphere pair
    (Narrowphase.cu:382-396, then the envelope and depth :412-415) is 52
    operations, a separated one 10, and a pair of one body 0."""
    pair = np.array([(0 << 32) | 1], np.int64)
    hit = _pair_scene(scenes, SPHERE, SPHERE, (0.7, 0.1, 0.0))
    miss = _pair_scene(scenes, SPHERE, SPHERE, (1.7, 0.1, 0.0))
    assert oracle.narrowphase(hit, pair).count == 1 and oracle.narrowphase(miss, pair).count == 0
    assert oracle.mpr_flops(hit, pair)[0] == 52
    assert oracle.mpr_flops(miss, pair)[0] == 10
    same = _pair_scene(scenes, SPHERE, BOX, (0.7, 0.1, 0.0))
    same.ID[:] = 0
    assert oracle.mpr_flops(same, pair)[0] == 0
    # MPR pairs: the census grows with the support steps the pair takes
    s = scenes.mixed_bodies(n=4096, seed=3)
    pairs = oracle.broadphase(oracle.aabb(s), s.grid).pairs
    f = oracle.mpr_flops(s, pairs).astype(np.int64)
    st = oracle.mpr_steps(s, pairs)
    m = st > 0
    assert (f[m] > 0).all()
    per_step = f[m] / st[m]
    assert 150 < np.median(per_step) < 600


# ----------------------------------------- RECT, CONE, TRIANGLEMESH shapes --

RECT, CONE, TRIANGLEMESH = 4, 5, 6


def _cone_points(B, k=3600):
    r, h = B[0], B[1]
    t = np.linspace(0, 2 * np.pi, k, endpoint=False)
    rim = np.stack([r * np.cos(t), np.full(k, -h / 2), r * np.sin(t)], axis=1)
    return np.vstack([[0.0, h / 2, 0.0], rim])


def test_support_maps_rect_cone_triangle(oracle):
    """SupportFunctions.cuh:14-29, :65-99 against brute force: the rectangle's
    four corners at height B.z, the cone's apex and base rim, the triangle's
    vertices (first maximum wins)."""
    rng = np.random.default_rng(41)
    for _ in range(100):
        q = rand_quat(rng)
        Rm = quat_to_mat(q)
        c = rng.standard_normal(3)
        d = rng.standard_normal(3)
        B = np.array([rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), rng.uniform(-0.2, 0.2)])
        corners = np.array([[sx * B[0], sy * B[1], B[2]] for sx in (-1, 1) for sy in (-1, 1)]) @ Rm.T + c
        np.testing.assert_allclose(oracle.support(RECT, c, B, q, d), corners[np.argmax(corners @ d)], atol=1e-12)
        Bc = np.array([rng.uniform(0.2, 0.8), rng.uniform(0.3, 1.5), 0.0])
        pts = _cone_points(Bc) @ Rm.T + c
        got = oracle.support(CONE, c, Bc, q, d)
        assert got @ d >= (pts @ d).max() - 1e-9 * np.linalg.norm(d)
        loc = Rm.T @ (got - c)
        on_apex = np.allclose(loc, [0, Bc[1] / 2, 0], atol=1e-12)
        on_rim = abs(loc[1] + Bc[1] / 2) < 1e-12 and abs(np.hypot(loc[0], loc[2]) - Bc[0]) < 1e-12
        assert on_apex or on_rim
        # triangle: world vertices X, Y and (0, 0, 0) through the test entry point
        Y = rng.standard_normal(3)
        verts = np.array([c, Y, np.zeros(3)])
        dn = d / np.linalg.norm(d)
        np.testing.assert_array_equal(oracle.support(TRIANGLEMESH, c, Y, q, d), verts[np.argmax(verts @ dn)])


def _typed_scene(scenes, T, A, B, C, R, pos, rot, grid=(4, 4, 4)):
    n = len(T)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return scenes.Scene(np.asarray(T, np.int32), f(A), f(B), f(C), f(R), np.arange(n, dtype=np.uint32), f(pos),
                        f(rot), np.ones(n, np.uint8), grid)


def test_aabb_rect_cone_triangle(oracle, scenes):
    """The bounds written for the types the reference leaves unwritten (RECT)
    or sends down the wrong path (CONE, TRIANGLEMESH; AABBGenerator.cu:103-112):
    the rectangle's corners and the cone's box at the narrowphase's placement
    (local position, rotation body x local), and the triangle's world
    vertices R (v + body position) (AABBGenerator.cu:16-23, :103-107)."""
    rng = np.random.default_rng(42)
    n = 60
    T = np.repeat([RECT, CONE, TRIANGLEMESH], n // 3)
    A, B, C = rng.standard_normal((n, 3)), rng.uniform(0.2, 1.0, (n, 3)), rng.standard_normal((n, 3))
    R = np.array([rand_quat(rng) for _ in range(n)])
    pos, rot = rng.standard_normal((n, 3)), np.array([rand_quat(rng) for _ in range(n)])
    s = _typed_scene(scenes, T, A, B, C, R, pos, rot)
    ab = oracle.aabb(s)
    for i in range(n):
        Rm = quat_to_mat(hamilton(rot[i], R[i]))
        if T[i] == RECT:
            pts = np.array([[sx * B[i, 0], sy * B[i, 1], B[i, 2]] for sx in (-1, 1) for sy in (-1, 1)]) @ Rm.T + A[i]
        elif T[i] == CONE:
            box = np.array([B[i, 0], B[i, 1] / 2, B[i, 0]])
            pts = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * box @ Rm.T + A[i]
            inner = _cone_points(B[i], 360) @ Rm.T + A[i]
            assert np.all(inner >= ab[i] - 1e-12) and np.all(inner <= ab[n + i] + 1e-12)
        else:
            pts = np.array([A[i], B[i], C[i]]) + pos[i]
            pts = pts @ Rm.T
        np.testing.assert_allclose(ab[i], pts.min(axis=0), atol=1e-13)
        np.testing.assert_allclose(ab[n + i], pts.max(axis=0), atol=1e-13)


def _pair(oracle, scenes, T, A, B, C, R, pos, rot, envelope=0.0):
    s = _typed_scene(scenes, T, A, B, C, R, pos, rot)
    return oracle.narrowphase(s, np.array([(0 << 32) | 1], np.int64), envelope=envelope)


I4 = [1.0, 0.0, 0.0, 0.0]
Z3 = [0.0, 0.0, 0.0]


@pytest.mark.parametrize("case", ["rect-face", "cone-apex", "cone-base", "triangle-face"])
def test_mpr_new_shapes_vs_sphere_closed_form(oracle, scenes, case):
    """A sphere pressed onto a rectangle, a cone's apex and base, and a
    triangle: depth = -(penetration), normal along the face normal / axis."""
    rs = 0.2
    if case == "rect-face":
        shape = (RECT, Z3, [0.5, 0.4, 0.0], Z3)
        c, want_n, pen = [0.1, -0.05, 0.15], [0, 0, 1], 0.05
    elif case == "cone-apex":  # vertex contact: along the apex-to-centre line
        shape = (CONE, Z3, [0.3, 0.8, 0.0], Z3)
        c = [1e-3, 0.55, -2e-3]
        d = np.array(c) - [0.0, 0.4, 0.0]
        want_n, pen = d / np.linalg.norm(d), rs - np.linalg.norm(d)
    elif case == "cone-base":
        shape = (CONE, Z3, [0.3, 0.8, 0.0], Z3)
        c, want_n, pen = [0.05, -0.55, 0.02], [0, -1, 0], 0.05
    else:
        shape = (TRIANGLEMESH, [-0.5, -0.4, 0.0], [0.5, -0.4, 0.0], [0.0, 0.6, 0.0])
        c, want_n, pen = [0.02, 0.01, 0.15], [0, 0, 1], 0.05
    t, a, b, cc = shape
    r = _pair(oracle, scenes, [t, SPHERE], [a, Z3], [b, [rs, 0, 0]], [cc, Z3], [I4, I4], [Z3, c], [I4, I4])
    assert r.count == 1, case
    # the MPR normal at a vertex contact is the portal's after the 34-step
    # cap (Narrowphase.cu:263): close to the closed form, not equal
    np.testing.assert_allclose(r.norm[0], want_n, atol=5e-3 if case == "cone-apex" else 1e-6)
    np.testing.assert_allclose(r.dpth[0], -pen, atol=1e-4 if case == "cone-apex" else 1e-6)


def test_new_shape_placement_and_envelope_quirks(oracle, scenes):
    """function_MPR_Store (Narrowphase.cu:356-377, :412-415): RECT and CONE are
    placed at their local position (the body position is not applied; the
    body rotation is, through body x local), and a pair with a triangle
    ignores the collision envelope."""
    c = [1e-3, 0.55, -2e-3]
    args = ([CONE, SPHERE], [Z3, Z3], [[0.3, 0.8, 0.0], [0.2, 0, 0]], [Z3, Z3], [I4, I4])
    base = _pair(oracle, scenes, *args, [Z3, c], [I4, I4])
    moved = _pair(oracle, scenes, *args, [[5.0, -3.0, 2.0], c], [I4, I4])
    assert base.count == moved.count == 1
    np.testing.assert_array_equal(base.norm, moved.norm)
    np.testing.assert_array_equal(base.dpth, moved.dpth)
    env = _pair(oracle, scenes, *args, [Z3, c], [I4, I4], envelope=0.01)
    np.testing.assert_allclose(env.dpth, base.dpth + 0.02, atol=1e-14)
    tri = ([TRIANGLEMESH, SPHERE], [[-0.5, -0.4, 0.0], Z3], [[0.5, -0.4, 0.0], [0.2, 0, 0]],
           [[0.0, 0.6, 0.0], Z3], [I4, I4], [Z3, [0.02, 0.01, 0.15]], [I4, I4])
    t0 = _pair(oracle, scenes, *tri)
    t1 = _pair(oracle, scenes, *tri, envelope=0.01)
    np.testing.assert_array_equal(t0.dpth, t1.dpth)


def _segment_closest(p1, q1, p2, q2):
    """Closest points of segments p1-q1 and p2-q2 (the clamped
    parametric solution, as in Ericson, Real-Time Collision Detection 5.1.9)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0.0, 1.0) if den > 1e-14 else 0.0
    t = (b * s + f) / e
    if t < 0.0:
        t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
    elif t > 1.0:
        t, s = 1.0, np.clip((b - c) / a, 0.0, 1.0)
    return p1 + d1 * s, p2 + d2 * t


def test_mpr_capsule_capsule_vs_segment_distance(oracle):
    """Two capsules (local axis y, half-length h, radius r) whose axes cross
    at their midpoints, offset along the common perpendicular w: the axis
    segments' closest points (computed independently) are the centres, so
    the overlap r_a + r_b - |c_b - c_a| lies along the centre line, the ray
    MPR walks (Narrowphase.cu:135-140), and MPR's depth must equal it.  200
    random orientations, sizes and overlaps."""
    rng = np.random.default_rng(0x5EED0C0C)
    checked = 0
    for _ in range(200):
        ra, rb = rng.uniform(0.1, 0.3, 2)
        ha, hb = rng.uniform(0.2, 0.8, 2)
        qa, qb = (q / np.linalg.norm(q) for q in rng.normal(size=(2, 4)))
        ua, ub = quat_to_mat(qa)[:, 1], quat_to_mat(qb)[:, 1]
        w = np.cross(ua, ub)
        if np.linalg.norm(w) < 0.2:  # nearly parallel axes: the closest points are not unique
            continue
        w /= np.linalg.norm(w)
        overlap = rng.uniform(0.02, 0.5) * min(ra, rb)
        XA = rng.normal(size=3)
        XB = XA + w * (ra + rb - overlap)
        ca, cb = _segment_closest(XA - ha * ua, XA + ha * ua, XB - hb * ub, XB + hb * ub)
        want = ra + rb - np.linalg.norm(cb - ca)
        hit, n, p, depth = _mpr_case(CAPSULE, np.array([ra, ha, ra]), XA, qa, CAPSULE, np.array([rb, hb, rb]), XB, qb)
        assert hit
        np.testing.assert_allclose(want, overlap, rtol=1e-9)  # the closest points are the centres
        np.testing.assert_allclose(depth, want, atol=2e-3 * min(ra, rb))
        assert abs(n @ w) > 0.999
        checked += 1
    assert checked > 100
