"""Seeded synthetic scenes for the collision pipeline.

Arrays follow the layout of the reference API's inputs
(AABBGenerator.cuh:21-30, Narrowphase.cuh:19-34): per shape a type tag T,
local position A (real3), dimensions B (real3), unused C (real3), local
rotation R (real4, scalar first) and body id ID; per body a position, a
rotation and an active flag.  Every workload has one shape per body, local
position 0 and local rotation identity.

Dimension convention (B): sphere (r, 0, 0); box and ellipsoid semi-axes;
cylinder (radius, half-height, radius) about local y; capsule (radius,
half-length of the core segment, radius) about local y; rectangle (half
extents in x and y, height z of its plane); cone (radius, full height, -)
about local y, apex up.  A triangle's vertices are A, B and C.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SPHERE, ELLIPSOID, BOX, CYLINDER, RECT, CONE, TRIANGLEMESH, CAPSULE = range(8)
TYPE_NAMES = {SPHERE: "sphere", ELLIPSOID: "ellipsoid", BOX: "box", CYLINDER: "cylinder", RECT: "rect", CONE: "cone",
              TRIANGLEMESH: "triangle", CAPSULE: "capsule"}

# 1M mixed-body workload: 40 % sphere, 20 % box, 15 % ellipsoid, 15 % capsule, 10 % cylinder.
MIX_TYPES = np.array([SPHERE, BOX, ELLIPSOID, CAPSULE, CYLINDER], dtype=np.int32)
MIX_PROBS = np.array([0.40, 0.20, 0.15, 0.15, 0.10])


@dataclass
class Scene:
    T: np.ndarray        # (n,) int32 shape type
    A: np.ndarray        # (n,3) local position
    B: np.ndarray        # (n,3) dimensions
    C: np.ndarray        # (n,3) unused by the in-scope shapes
    R: np.ndarray        # (n,4) local rotation (w,x,y,z)
    ID: np.ndarray       # (n,) uint32 body id
    pos: np.ndarray      # (nb,3) body position
    rot: np.ndarray      # (nb,4) body rotation (w,x,y,z)
    active: np.ndarray   # (nb,) uint8
    grid: tuple = (16, 16, 16)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def dtype(self):
        return self.A.dtype

    @property
    def n(self):
        return int(self.T.shape[0])

    @property
    def nb(self):
        return int(self.pos.shape[0])

    def astype(self, dtype) -> "Scene":
        c = lambda a: np.ascontiguousarray(a, dtype=dtype)
        return Scene(self.T, c(self.A), c(self.B), c(self.C), c(self.R), self.ID, c(self.pos), c(self.rot),
                     self.active, self.grid, self.name, dict(self.meta))

    def subset(self, idx) -> "Scene":
        """Bodies idx (one shape per body) as a scene of their own."""
        idx = np.asarray(idx)
        n = len(idx)
        return Scene(self.T[idx].copy(), self.A[idx].copy(), self.B[idx].copy(), self.C[idx].copy(),
                     self.R[idx].copy(), np.arange(n, dtype=np.uint32), self.pos[idx].copy(), self.rot[idx].copy(),
                     self.active[idx].copy(), self.grid, self.name + "-subset", dict(self.meta))


def random_quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def _one_shape_per_body(T, B, pos, rot, dtype, grid, name, meta):
    n = len(T)
    z3 = np.zeros((n, 3))
    ident = np.zeros((n, 4))
    ident[:, 0] = 1.0
    c = lambda a: np.ascontiguousarray(a, dtype=dtype)
    return Scene(np.ascontiguousarray(T, dtype=np.int32), c(z3), c(B), c(z3), c(ident),
                 np.arange(n, dtype=np.uint32), c(pos), c(rot), np.ones(n, dtype=np.uint8), tuple(grid), name, meta)


def half_extents(T, B, rot):
    """World AABB half-extents of each shape (|R| times the bounding semi-axes)."""
    T = np.asarray(T)
    B = np.asarray(B, dtype=np.float64)
    sem = B.copy()
    sph = T == SPHERE
    sem[sph] = B[sph, :1]
    cap = T == CAPSULE
    sem[cap, 1] = B[cap, 1] + B[cap, 0]
    q = np.asarray(rot, dtype=np.float64)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    Rm = np.empty((len(q), 3, 3))
    Rm[:, 0, 0] = 1 - 2 * (y * y + z * z); Rm[:, 0, 1] = 2 * (x * y - w * z); Rm[:, 0, 2] = 2 * (x * z + w * y)
    Rm[:, 1, 0] = 2 * (x * y + w * z); Rm[:, 1, 1] = 1 - 2 * (x * x + z * z); Rm[:, 1, 2] = 2 * (y * z - w * x)
    Rm[:, 2, 0] = 2 * (x * z - w * y); Rm[:, 2, 1] = 2 * (y * z + w * x); Rm[:, 2, 2] = 1 - 2 * (x * x + y * y)
    h = np.einsum("nij,nj->ni", np.abs(Rm), sem)
    h[sph] = B[sph, :1]
    return h


def mixed_shapes(rng, n, mu=np.log(0.01), sigma=0.3):
    """Types, dimensions and rotations of the mixed convex workload."""
    T = rng.choice(MIX_TYPES, size=n, p=MIX_PROBS).astype(np.int32)
    s = np.exp(rng.normal(mu, sigma, size=n))
    B = np.zeros((n, 3))
    m = T == SPHERE
    B[m, 0] = s[m]
    for t in (BOX, ELLIPSOID):
        m = T == t
        B[m] = s[m, None] * rng.uniform(0.5, 1.5, size=(m.sum(), 3))
    m = T == CAPSULE
    B[m, 0] = 0.5 * s[m]
    B[m, 1] = s[m]
    B[m, 2] = 0.5 * s[m]
    m = T == CYLINDER
    B[m, 0] = s[m]
    B[m, 1] = s[m] * rng.uniform(0.5, 1.5, size=m.sum())
    B[m, 2] = s[m]
    rot = random_quats(rng, n)
    return T, B, rot


def box_side_for_neighbours(h, n, k, rng, samples=400_000):
    """Cube side L so a body has about k AABB-overlap partners: L^3 = n E[prod(2hA+2hB)] / k."""
    i = rng.integers(0, len(h), samples)
    j = rng.integers(0, len(h), samples)
    e = np.prod(2 * h[i] + 2 * h[j], axis=1).mean()
    return float((n * e / k) ** (1.0 / 3.0))


def grid_for(L_extent, cell):
    g = max(1, int(np.ceil(L_extent / cell)))
    return (g, g, g)


def small_spheres(n=8192, seed=0x5EED0000, dtype=np.float64, grid=None):
    """8,192 FP64 spheres, r ~ U[0.01, 0.02], centres ~ U[0,1]^3, identity rotations."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.01, 0.02, n)
    pos = rng.uniform(0.0, 1.0, (n, 3))
    rot = np.zeros((n, 4)); rot[:, 0] = 1
    B = np.zeros((n, 3)); B[:, 0] = r
    g = grid if grid is not None else grid_for(1.04, 0.04)
    return _one_shape_per_body(np.zeros(n, np.int32), B, pos, rot, dtype, g, "spheres-8k",
                               {"n": n, "seed": seed})


def mixed_bodies(n=1 << 20, seed=0x5EED0001, dtype=np.float64, neighbours=4.0, grid=None):
    """Mixed convex bodies (log-normal size, uniform rotations) in a cube sized for ~4 AABB partners each."""
    rng = np.random.default_rng(seed)
    T, B, rot = mixed_shapes(rng, n)
    h = half_extents(T, B, rot)
    L = box_side_for_neighbours(h, n, neighbours, rng)
    pos = rng.uniform(0.0, L, (n, 3))
    if grid is None:
        cell = 2.0 * float(np.mean(2 * h.max(axis=1)))
        grid = grid_for(L, cell)
    return _one_shape_per_body(T, B, pos, rot, dtype, grid, "mixed",
                               {"n": n, "seed": seed, "L": L, "neighbours": neighbours})


def lattice_spheres(side=256, seed=0x5EED0002, dtype=np.float32, spacing=0.01, grid=None):
    """Dense granular pack: side^3 spheres on a jittered cubic lattice, r in {0.005, 0.0075}."""
    rng = np.random.default_rng(seed)
    n = side ** 3
    idx = np.arange(n, dtype=np.int64)
    ijk = np.stack([idx % side, (idx // side) % side, idx // (side * side)], axis=1).astype(np.float64)
    r = np.where(rng.random(n) < 0.5, 0.005, 0.0075)
    pos = ijk * spacing + rng.uniform(-0.1, 0.1, (n, 3)) * r[:, None]
    del ijk
    rot = np.zeros((n, 4)); rot[:, 0] = 1
    B = np.zeros((n, 3)); B[:, 0] = r
    if grid is None:
        grid = grid_for(side * spacing + 0.02, 0.015)
    return _one_shape_per_body(np.zeros(n, np.int32), B, pos, rot, dtype, grid, "lattice",
                               {"n": n, "seed": seed, "side": side})


def mixed_scene_of_size(n, neighbours, seed, dtype=np.float64):
    """One scene of the batched / served workloads: the mixed shape set at its own density."""
    return mixed_bodies(n=n, seed=seed, dtype=dtype, neighbours=neighbours)


_PAIR_VOLUME = None


def mix_pair_volume(samples=2_000_000, seed=0x5EED00FF):
    """E[prod(2hA + 2hB)] over independent bodies of the mixed shape set (cached)."""
    global _PAIR_VOLUME
    if _PAIR_VOLUME is None:
        rng = np.random.default_rng(seed)
        T, B, rot = mixed_shapes(rng, samples)
        h = half_extents(T, B, rot)
        i = rng.permutation(samples)
        _PAIR_VOLUME = float(np.prod(2 * h + 2 * h[i], axis=1).mean())
    return _PAIR_VOLUME


def fleet_scene(n, neighbours, seed, dtype=np.float64):
    """One scene of the batched / served workloads: n mixed bodies in a cube of side (n E / k)^(1/3).

    E is the mix's expected AABB-overlap volume, so a body has ~k AABB
    partners; the grid is left to the autotuner (bins (0, 0, 0)).
    """
    rng = np.random.default_rng(seed)
    T, B, rot = mixed_shapes(rng, n)
    L = float((n * mix_pair_volume() / neighbours) ** (1.0 / 3.0))
    pos = rng.uniform(0.0, L, (n, 3))
    return _one_shape_per_body(T, B, pos, rot, dtype, (0, 0, 0), "fleet",
                               {"n": n, "seed": seed, "L": L, "neighbours": neighbours})


def batched_sizes(count=512, lo=1024, hi=65536, seed=0x5EED0003):
    """Shape counts log-uniform in [lo, hi] and densities U[2, 8] neighbours per body."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
    k = rng.uniform(2.0, 8.0, count)
    return n, k


def batched_scenes(count=512, lo=1024, hi=65536, seed=0x5EED0003, dtype=np.float64):
    """512 independent mixed scenes, sizes log-uniform in [1024, 65536], U[2, 8] neighbours each."""
    n, k = batched_sizes(count, lo, hi, seed)
    return [fleet_scene(int(n[i]), float(k[i]), seed + 1 + i, dtype) for i in range(count)]


def served_sizes(count=4096, lo=256, hi=262144, seed=0x5EED0004):
    """Shape counts of the served workload: log-uniform in [256, 262144]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


def two_shapes_per_body(s: Scene) -> Scene:
    """Scene s with bodies merged in pairs: shape i rides body i // 2 at a local offset and rotation."""
    n = s.n
    nb = (n + 1) // 2
    rng = np.random.default_rng(n)
    A = s.pos - s.pos[np.arange(n) // 2 * 2]
    R = random_quats(rng, n).astype(s.dtype)
    R[::2] = 0
    R[::2, 0] = 1  # body rotation only for the even shape
    c = lambda a: np.ascontiguousarray(a, dtype=s.dtype)
    ident = np.zeros((nb, 4))
    ident[:, 0] = 1
    return Scene(s.T.copy(), c(A), s.B.copy(), s.C.copy(), c(R), (np.arange(n) // 2).astype(np.uint32),
                 c(s.pos[::2]), c(ident), np.ones(nb, np.uint8), s.grid, s.name + "-2perbody", dict(s.meta))


def quat_mult(a, b):
    """Hamilton product of (n,4) scalar-first quaternions."""
    aw, ax, ay, az = a[:, 0], a[:, 1], a[:, 2], a[:, 3]
    bw, bx, by, bz = b[:, 0], b[:, 1], b[:, 2], b[:, 3]
    return np.stack([aw * bw - ax * bx - ay * by - az * bz,
                     aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw], axis=1)


def pose_step(s: Scene, rng, sigma=0.05, angle=0.05):
    """New absolute poses of every body of s after one served step.

    Displacement ~ N(0, (sigma * size)^2) per axis, size = the body's largest
    dimension; rotation by an angle ~ N(0, angle) rad about a uniform random
    axis, applied on the left of the current rotation.
    """
    nb = s.nb
    size = np.zeros(nb)
    np.maximum.at(size, s.ID, np.abs(s.B).max(axis=1).astype(np.float64))
    pos = s.pos.astype(np.float64) + rng.standard_normal((nb, 3)) * (sigma * size)[:, None]
    axis = rng.standard_normal((nb, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    th = rng.standard_normal(nb) * angle
    dq = np.concatenate([np.cos(th / 2)[:, None], np.sin(th / 2)[:, None] * axis], axis=1)
    rot = quat_mult(dq, s.rot.astype(np.float64))
    rot /= np.linalg.norm(rot, axis=1, keepdims=True)
    return pos.astype(s.dtype), rot.astype(s.dtype)


def all_shape_types(n=4000, seed=0x5EED0010, dtype=np.float64, neighbours=6.0, grid=None):
    """Every shape type of the reference enum plus the capsule, in one cube.

    Placement follows how function_MPR_Store (Narrowphase.cu:356-377) reads
    each type: rectangles and cones sit at their local position A (the
    narrowphase applies no body pose to them) with a random local rotation;
    triangles have vertices A, B, C around the origin and bodies with
    identity rotation at the triangle's place (the reference forms the world
    vertex as R (v + body position)); the other types as in mixed_bodies.
    """
    rng = np.random.default_rng(seed)
    T = rng.choice(np.arange(8, dtype=np.int32), size=n)
    s = np.exp(rng.normal(np.log(0.01), 0.3, size=n))
    B = np.zeros((n, 3))
    A = np.zeros((n, 3))
    Cv = np.zeros((n, 3))
    R = np.zeros((n, 4))
    R[:, 0] = 1.0
    rot = random_quats(rng, n)
    for t in (BOX, ELLIPSOID):
        m = T == t
        B[m] = s[m, None] * rng.uniform(0.5, 1.5, size=(m.sum(), 3))
    m = T == SPHERE
    B[m, 0] = s[m]
    m = T == CAPSULE
    B[m] = np.stack([0.5 * s[m], s[m], 0.5 * s[m]], axis=1)
    m = T == CYLINDER
    B[m] = np.stack([s[m], s[m] * rng.uniform(0.5, 1.5, m.sum()), s[m]], axis=1)
    m = T == RECT
    B[m] = np.stack([s[m] * rng.uniform(0.7, 1.5, m.sum()), s[m] * rng.uniform(0.7, 1.5, m.sum()), np.zeros(m.sum())],
                    axis=1)
    m = T == CONE
    B[m] = np.stack([s[m], 2.0 * s[m] * rng.uniform(0.7, 1.5, m.sum()), np.zeros(m.sum())], axis=1)
    h = np.full((n, 3), 1.2) * s[:, None] * 1.5
    L = box_side_for_neighbours(h, n, neighbours, rng)
    pos = rng.uniform(0.0, L, (n, 3))
    rt = (T == RECT) | (T == CONE)
    A[rt] = pos[rt]
    R[rt] = random_quats(rng, int(rt.sum()))
    tri = T == TRIANGLEMESH
    k = int(tri.sum())
    verts = rng.normal(0.0, 1.0, (k, 3, 3)) * s[tri, None, None] * 1.2
    verts -= verts.mean(axis=1, keepdims=True)
    A[tri], B[tri], Cv[tri] = verts[:, 0], verts[:, 1], verts[:, 2]
    rot[tri] = [1.0, 0.0, 0.0, 0.0]
    if grid is None:
        grid = grid_for(L, 2.0 * float(np.mean(2 * h.max(axis=1))))
    c = lambda a: np.ascontiguousarray(a, dtype=dtype)
    return Scene(T, c(A), c(B), c(Cv), c(R), np.arange(n, dtype=np.uint32), c(pos), c(rot), np.ones(n, np.uint8),
                 tuple(grid), "all-types", {"n": n, "seed": seed, "L": L})
