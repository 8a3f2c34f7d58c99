// Narrowphase kernels (function_MPR_Store, Narrowphase.cu:316-504, and the
// compaction of Narrowphase.cu:580-593).  Included by pipeline.hip only.
#pragma once
#include "cdx_device.hpp"

namespace cdx {
namespace {


// A shape placed in the world for the reference-order MPR: type, centre,
// dimensions, rotation.  The rotation matrix is rebuilt from the quaternion
// at every support call: the same rounded values each time, with fewer live
// registers than a cached 3x3 per shape (caching both matrices per pair
// measured no faster: the kernel stays at 2 waves per SIMD either way).
// A triangle holds its three world vertices in X, B and q.w/x/y instead.
template <class R> struct Placed {
  int type;
  v3<R> X, B;
  q4<R> q;
};

// World support along a unit direction (TransformSupportVert,
// Narrowphase.cu:19-52).
// EXT: rectangles, cones and triangles possible (see support_local).
template <bool EXT, class R> __device__ __forceinline__ v3<R> support_placed(const Placed<R>& s, v3<R> n) {
  if constexpr (EXT)
    if (s.type == TRIANGLEMESH) return support_triangle(s.X, s.B, mk3(s.q.w, s.q.x, s.q.y), n);
  return support_world_unit<R, EXT>(s.type, s.X, s.B, rotmat(s.q), n);
}

// GetCenter (Narrowphase.cu:11-17): the triangle's centroid
// (SupportFunctions.cuh:119-121), else 0 + the centre.
template <bool EXT, class R> __device__ __forceinline__ v3<R> center_of(const Placed<R>& s) {
  if (EXT && s.type == TRIANGLEMESH)
    return mk3((s.X.x + s.B.x + s.q.w) / R(3.0f), (s.X.y + s.B.y + s.q.x) / R(3.0f), (s.X.z + s.B.z + s.q.y) / R(3.0f));
  return mk3<R>(0, 0, 0) + s.X;
}

// sa = sup(A, -n), sb = sup(B, n) with one normalisation of n (normalize_pm).
template <bool EXT, class R>
__device__ __forceinline__ void sup_pair(const Placed<R>& A, const Placed<R>& B, v3<R> n, v3<R>& sa, v3<R>& sb) {
  v3<R> na, nb;
  normalize_pm(n, na, nb);
  sa = support_placed<EXT>(A, na);
  sb = support_placed<EXT>(B, nb);
}

template <class R> __device__ __forceinline__ Placed<R> placed(const ShapeRec<R>& r) {
  Placed<R> s;
  s.type = (int)r.type;
  s.X = mk3(r.x, r.y, r.z);
  s.B = mk3(r.bx, r.by, r.bz);
  s.q.w = r.qw; s.q.x = r.qx; s.q.y = r.qy; s.q.z = r.qz;
  return s;
}

// A shape for the FP32 filter: rotation matrix built once per pair (cheap
// in FP32 registers), approximate normalisation in the support maps.
struct PlacedF {
  int type;
  v3<float> X, B;
  mat3<float> m;
};

__device__ __forceinline__ v3<float> sup_fast(const PlacedF& s, v3<float> b) {
  return support_world_fast(s.type, s.X, s.B, s.m, b);
}

// sa = sup_fast(A, -n), sb = sup_fast(B, n) with one normalisation of n.
__device__ __forceinline__ void sup_fast_pair(const PlacedF& A, const PlacedF& B, v3<float> n, v3<float>& sa,
                                              v3<float>& sb) {
  v3<float> na, nb;
  normalize_fast_pm(n, na, nb);
  sa = support_world_fast_unit(A.type, A.X, A.B, A.m, na);
  sb = support_world_fast_unit(B.type, B.X, B.B, B.m, nb);
}

template <class R> __device__ __forceinline__ void swap3(v3<R>& a, v3<R>& b) { v3<R> t = a; a = b; b = t; }

// CollideAndFindPoint (Narrowphase.cu:122-314), same evaluation order.
template <bool EXT, class R>
__device__ bool mpr_reference(const Placed<R>& A, const Placed<R>& B, v3<R>& rnormal, v3<R>& point, R& depth) {
  v3<R> v01, v02, v0, n, v11, v12, v1, v21, v22, v2;
  v01 = center_of<EXT>(A);
  v02 = center_of<EXT>(B);
  v0 = v02 - v01;
  if (is_zero(v0)) v0 = mk3<R>(1, 0, 0);
  n = -v0;
  sup_pair<EXT>(A, B, n, v11, v12);
  v1 = v12 - v11;
  if (dot(v1, n) <= R(0)) return false;
  n = cross(v1, v0);
  if (is_zero(n)) {
    n = v1 - v0;
    n = normalize(n);
    rnormal = n;
    point = (v11 + v12) * R(.5);
    depth = dot((v12 - v11), n);
    return true;
  }
  sup_pair<EXT>(A, B, n, v21, v22);
  v2 = v22 - v21;
  if (dot(v2, n) <= R(0)) return false;
  n = cross((v1 - v0), (v2 - v0));
  if (dot(n, v0) > R(0)) {
    swap3(v1, v2);
    swap3(v11, v21);
    swap3(v12, v22);
    n = -n;
  }
  v3<R> v31, v32, v3_;
  bool hit = false;
  int phase1 = 0, phase2 = 0;
  const int max_iterations = 34;
  while (true) {
    if (phase1 > max_iterations) return false;
    phase1++;
    sup_pair<EXT>(A, B, n, v31, v32);
    v3_ = v32 - v31;
    if (dot(v3_, n) <= R(0)) return false;
    if (dot(cross(v1, v3_), v0) < R(0)) {
      v2 = v3_; v21 = v31; v22 = v32;
      n = cross((v1 - v0), (v3_ - v0));
      continue;
    }
    if (dot(cross(v3_, v2), v0) < R(0)) {
      v1 = v3_; v11 = v31; v12 = v32;
      n = cross((v3_ - v0), (v2 - v0));
      continue;
    }
    while (true) {
      phase2++;
      n = cross((v2 - v1), (v3_ - v1));
      n = normalize(n);
      if (dot(n, v1) >= R(0) && !hit) hit = true;
      v3<R> v41, v42;
      sup_pair<EXT>(A, B, n, v41, v42);
      v3<R> v4 = v42 - v41;
      R delta = dot((v4 - v3_), n);
      depth = dot(v4, n);
      if ((double)delta <= 1e-20 || depth <= R(0) || phase2 > max_iterations) {
        if (hit) {
          R b0 = dot(cross(v1, v2), v3_);
          R b1 = dot(cross(v3_, v2), v0);
          R b2 = dot(cross(v0, v1), v3_);
          R b3 = dot(cross(v2, v1), v0);
          R sum = b0 + b1 + b2 + b3;
          if (sum <= R(0)) {
            b0 = 0;
            b1 = dot(cross(v2, v3_), n);
            b2 = dot(cross(v3_, v1), n);
            b3 = dot(cross(v1, v2), n);
            sum = b1 + b2 + b3;
          }
          R inv = R(1) / sum;
          point = (b0 * v01 + b1 * v11 + b2 * v21 + b3 * v31) + (b0 * v02 + b1 * v12 + b2 * v22 + b3 * v32);
          point = point * inv * R(.5);
          rnormal = normalize(n);
        }
        return hit;
      }
      if (dot(cross(v4, v1), v0) < R(0)) {
        if (dot(cross(v4, v2), v0) < R(0)) { v1 = v4; v11 = v41; v12 = v42; }
        else { v3_ = v4; v31 = v41; v32 = v42; }
      } else {
        if (dot(cross(v4, v3_), v0) < R(0)) { v2 = v4; v21 = v41; v22 = v42; }
        else { v1 = v4; v11 = v41; v12 = v42; }
      }
    }
  }
}

template <class R> struct Contact {
  v3<R> N, p1, p2;
  R depth;
};

// function_MPR_Store (Narrowphase.cu:316-423) for two shape records in the
// scene's own precision.  Returns true on contact.
// Sphere-sphere closed form (Narrowphase.cu:382-396) plus the envelope and
// depth of :412-415.
template <class R>
__device__ __forceinline__ bool contact_spheres(v3<R> AX, R rad_a, v3<R> BX, R rad_b, R envelope, Contact<R>& out) {
  v3<R> relpos = BX - AX;
  R d2 = dot(relpos, relpos);
  R cd = rad_a + rad_b;
  if (!(d2 <= cd * cd)) return false;
  v3<R> N = normalize(relpos);
  v3<R> p1 = AX + N * rad_a;
  v3<R> p2 = BX - N * rad_b;
  p1 = p1 - N * envelope;
  p2 = p2 + N * envelope;
  out.N = N;
  out.p1 = p1;
  out.p2 = p2;
  out.depth = dot(N, p2 - p1);
  return true;
}
template <class R>
__device__ __forceinline__ bool contact_spheres(const ShapeRec<R>& ra, const ShapeRec<R>& rb, R envelope,
                                                Contact<R>& out) {
  return contact_spheres(mk3(ra.x, ra.y, ra.z), ra.bx, mk3(rb.x, rb.y, rb.z), rb.bx, envelope, out);
}

template <bool EXT, class R>
__device__ bool contact_native(const ShapeRec<R>& ra, const ShapeRec<R>& rb, R envelope, Contact<R>& out) {
  if (!ra.active && !rb.active) return false;
  if (ra.body == rb.body) return false;
  if (EXT && (ra.type == TRIANGLEMESH || rb.type == TRIANGLEMESH)) envelope = 0;  // Narrowphase.cu:364, :373
  v3<R> N = mk3<R>(1, 0, 0), p1 = mk3<R>(0, 0, 0), p2 = mk3<R>(0, 0, 0);
  if (ra.type == SPHERE && rb.type == SPHERE) {
    return contact_spheres(ra, rb, envelope, out);
  } else {
    Placed<R> A = placed(ra), B = placed(rb);
    v3<R> p0 = mk3<R>(0, 0, 0);
    R depth = 0;
    if (!mpr_reference<EXT>(A, B, N, p0, depth)) return false;
    v3<R> sa, sb;
    sup_pair<EXT>(A, B, N, sa, sb);
    p1 = dot((sa - p0), N) * N + p0;
    p2 = dot((sb - p0), N) * N + p0;
    N = -N;
  }
  p1 = p1 - N * envelope;
  p2 = p2 + N * envelope;
  out.N = N;
  out.p1 = p1;
  out.p2 = p2;
  out.depth = dot(N, p2 - p1);
  return true;
}

// Per-pair results before compaction: SoA of 10 reals, body ids, hit flag.
// One contact per pair index, written only for hits: a record per pair, so
// the compaction reads each hit's own bytes (component arrays strided by the
// pair count made it read nearly every line of all ten of them).
template <class R> struct alignas(16) PairRec {
  R v[10];         // N(3) p1(3) p2(3) depth
  int32_t ids[2];  // body ids
};

template <class R> struct PairOut {
  PairRec<R>* rec;  // P records
  uint32_t* hit;    // P hit flags
  uint64_t P;       // capacity
  uint8_t* kind;    // per hit: 0 its contact is in rec, 1 sphere-sphere (recomputed at compaction)
};

__device__ __forceinline__ uint64_t dev_count(uint64_t cap, const uint64_t* d_n) {
  uint64_t v = *d_n;
  return v < cap ? v : cap;
}

template <class R>
__device__ __forceinline__ void store_pair(const PairOut<R>& o, uint64_t P, uint64_t i, const Contact<R>& c,
                                           uint32_t ida, uint32_t idb) {
  (void)P;
  PairRec<R> r;
  r.v[0] = c.N.x; r.v[1] = c.N.y; r.v[2] = c.N.z;
  r.v[3] = c.p1.x; r.v[4] = c.p1.y; r.v[5] = c.p1.z;
  r.v[6] = c.p2.x; r.v[7] = c.p2.y; r.v[8] = c.p2.z;
  r.v[9] = c.depth;
  r.ids[0] = (int32_t)ida;
  r.ids[1] = (int32_t)idb;
  o.rec[i] = r;
  o.hit[i] = 1u;
  o.kind[i] = 0u;
}

// A sphere-sphere hit: no record; the compaction recomputes its contact from
// the shape records with contact_spheres (the same inputs and operations,
// so the same bits), which saves a record write and read per hit.
template <class R> __device__ __forceinline__ void mark_sphere_hit(const PairOut<R>& o, uint64_t i) {
  o.hit[i] = 1u;
  o.kind[i] = 1u;
}

// The hit test of contact_spheres on its own (same operations).
template <class R> __device__ __forceinline__ bool spheres_touch(v3<R> AX, R rad_a, v3<R> BX, R rad_b) {
  const v3<R> relpos = BX - AX;
  const R d2 = dot(relpos, relpos);
  const R cd = rad_a + rad_b;
  return d2 <= cd * cd;
}

// ---- mixed precision: FP32 MPR in the frame of shape A, FP64 re-solve ----

enum : int { MPR_MISS = 0, MPR_HIT = 1, MPR_UNSURE = 2 };

__device__ __forceinline__ float bound_radius(int t, v3<float> b) {
  switch (t) {
    case SPHERE: return b.x;
    case CAPSULE: return b.x + b.y;
    case CYLINDER: return sqrtf(b.x * b.x + b.y * b.y);
    default: return sqrtf(b.x * b.x + b.y * b.y + b.z * b.z);
  }
}

// MPR of the reference's algorithm in FP32, with shape A at the origin.  It
// returns MPR_UNSURE instead of a verdict whenever FP32 rounding could decide
// it differently from FP64: a separation or penetration within tau of zero,
// (nearly) coincident centres, a (nearly) degenerate portal, or an iteration
// cap reached.  With verdict_only (the pair is re-solved in FP64 unless it
// is a clear miss) the cap is VERDICT_ITERATIONS per phase: a pair that has
// not separated by then goes to FP64, so one slow lane does not hold its
// whole wavefront for the full 34 iterations.
constexpr int VERDICT_ITERATIONS = 6;

__device__ int mpr_guarded(const PlacedF& A, const PlacedF& B, float tau, bool verdict_only,
                           v3<float>& rnormal, v3<float>& point) {
  typedef float R;
  v3<R> v01, v02, v0, n, v11, v12, v1, v21, v22, v2;
  v01 = A.X;
  v02 = B.X;
  v0 = v02 - v01;
  if (length_fast(v0) < 1e-4f * tau) return MPR_UNSURE;
  n = -v0;
  sup_fast_pair(A, B, n, v11, v12);
  v1 = v12 - v11;
  R dn = dot(v1, n);
  if (dn <= R(0)) return (-dn < tau * length_fast(n)) ? MPR_UNSURE : MPR_MISS;
  n = cross(v1, v0);
  {
    R nn = dot(n, n), s = 1e-6f * dot(v1, v1) * dot(v0, v0);
    if (nn <= s) return MPR_UNSURE;
  }
  sup_fast_pair(A, B, n, v21, v22);
  v2 = v22 - v21;
  dn = dot(v2, n);
  if (dn <= R(0)) return (-dn < tau * length_fast(n)) ? MPR_UNSURE : MPR_MISS;
  n = cross((v1 - v0), (v2 - v0));
  if (dot(n, v0) > R(0)) {
    swap3(v1, v2);
    swap3(v11, v21);
    swap3(v12, v22);
    n = -n;
  }
  v3<R> v31, v32, v3_;
  bool hit = false;
  int phase1 = 0, phase2 = 0;
  const int max_iterations = verdict_only ? VERDICT_ITERATIONS : 34;
  while (true) {
    if (phase1 > max_iterations) return MPR_UNSURE;
    phase1++;
    sup_fast_pair(A, B, n, v31, v32);
    v3_ = v32 - v31;
    dn = dot(v3_, n);
    if (dn <= R(0)) return (-dn < tau * length_fast(n)) ? MPR_UNSURE : MPR_MISS;
    if (dot(cross(v1, v3_), v0) < R(0)) {
      v2 = v3_; v21 = v31; v22 = v32;
      n = cross((v1 - v0), (v3_ - v0));
      continue;
    }
    if (dot(cross(v3_, v2), v0) < R(0)) {
      v1 = v3_; v11 = v31; v12 = v32;
      n = cross((v3_ - v0), (v2 - v0));
      continue;
    }
    while (true) {
      phase2++;
      n = cross((v2 - v1), (v3_ - v1));
      n = normalize_fast(n);
      R side = dot(n, v1);
      if (side >= R(0) && !hit) hit = true;
      // The origin is inside the portal by more than tau: a certain hit.
      if (verdict_only && side > tau) return MPR_HIT;
      v3<float> v41, v42;
      sup_fast_pair(A, B, n, v41, v42);
      v3<R> v4 = v42 - v41;
      R delta = dot((v4 - v3_), n);
      R depth = dot(v4, n);
      if ((double)delta <= 1e-20 || depth <= R(0) || phase2 > max_iterations) {
        if (phase2 > max_iterations) return MPR_UNSURE;
        if (!hit) {
          R margin = depth <= R(0) ? -depth : -side;
          return margin < tau ? MPR_UNSURE : MPR_MISS;
        }
        R b0 = dot(cross(v1, v2), v3_);
        R b1 = dot(cross(v3_, v2), v0);
        R b2 = dot(cross(v0, v1), v3_);
        R b3 = dot(cross(v2, v1), v0);
        R sum = b0 + b1 + b2 + b3;
        if (sum <= R(0)) {
          b0 = 0;
          b1 = dot(cross(v2, v3_), n);
          b2 = dot(cross(v3_, v1), n);
          b3 = dot(cross(v1, v2), n);
          sum = b1 + b2 + b3;
        }
        R inv = R(1) / sum;
        point = (b0 * v01 + b1 * v11 + b2 * v21 + b3 * v31) + (b0 * v02 + b1 * v12 + b2 * v22 + b3 * v32);
        point = point * inv * R(.5);
        rnormal = normalize_fast(n);
        return MPR_HIT;
      }
      if (dot(cross(v4, v1), v0) < R(0)) {
        if (dot(cross(v4, v2), v0) < R(0)) { v1 = v4; v11 = v41; v12 = v42; }
        else { v3_ = v4; v31 = v41; v32 = v42; }
      } else {
        if (dot(cross(v4, v3_), v0) < R(0)) { v2 = v4; v21 = v41; v22 = v42; }
        else { v1 = v4; v11 = v41; v12 = v42; }
      }
    }
  }
}

__device__ __forceinline__ PlacedF placed_local(const ShapeRecF& r, const ShapeRecF& origin) {
  PlacedF s;
  s.type = (int)r.type;
  s.X = mk3((float)(r.x - origin.x), (float)(r.y - origin.y), (float)(r.z - origin.z));
  s.B = mk3(r.bx, r.by, r.bz);
  q4<float> q;
  q.w = r.qw; q.x = r.qx; q.y = r.qy; q.z = r.qz;
  s.m = rotmat(q);
  return s;
}

// Queue class slot of a shape type: the FP32-filtered types first
// (sphere, ellipsoid, box, cylinder, capsule), then rectangle, cone and
// triangle, which go to the reference-order MPR directly.
__device__ __forceinline__ uint32_t type_slot(uint32_t t) { return t == CAPSULE ? 4u : (t >= RECT && t <= TRIANGLEMESH ? t + 1u : t); }
__device__ __forceinline__ bool fp32_filtered(uint32_t t) { return t <= CYLINDER || t == CAPSULE; }
// Queue class of a pair: type_slot(a) * 8 + type_slot(b), plus EXT_CLASS
// when a rectangle, cone or triangle takes part, so the sorted queue keeps
// those pairs at its end for the resolve kernel that supports them.
constexpr uint32_t EXT_CLASS = 64;
constexpr int NQC = 128;            // queue classes (pair_class < NQC)
constexpr uint32_t NO_CLASS = 255;  // not queued
constexpr int QUEUE_BS = 256;       // class histograms and placement: a tile of
constexpr int QUEUE_IPT = 8;        // QUEUE_BS * QUEUE_IPT pairs per workgroup
constexpr uint64_t QUEUE_TILE = (uint64_t)QUEUE_BS * QUEUE_IPT;
__device__ __forceinline__ uint32_t pair_class(uint32_t ta, uint32_t tb) {
  const uint32_t c = type_slot(ta) * 8u + type_slot(tb);
  return fp32_filtered(ta) && fp32_filtered(tb) ? c : EXT_CLASS + c;
}

// The first portal test of mpr_guarded on its own (same operations, so the
// same verdict): true when the pair is a clear miss along the centre line.
// The front pass drops such pairs before the queue is built and sorted.
__device__ __forceinline__ bool mpr_first_miss(const PlacedF& A, const PlacedF& B, float tau) {
  const v3<float> v0 = B.X - A.X;
  if (length_fast(v0) < 1e-4f * tau) return false;
  const v3<float> n = -v0;
  v3<float> sa, sb;
  sup_fast_pair(A, B, n, sa, sb);
  const v3<float> v1 = sb - sa;
  const float dn = dot(v1, n);
  return dn <= 0.f && !(-dn < tau * length_fast(n));
}

template <class R> __device__ __forceinline__ ShapeRec<double> widen(const ShapeRec<R>& r) {
  ShapeRec<double> d;
  d.x = r.x; d.y = r.y; d.z = r.z; d.qw = r.qw; d.qx = r.qx; d.qy = r.qy; d.qz = r.qz;
  d.bx = r.bx; d.by = r.by; d.bz = r.bz; d.type = r.type; d.body = r.body; d.active = r.active; d.pad = 0;
  return d;
}
template <> __device__ __forceinline__ ShapeRec<double> widen<double>(const ShapeRec<double>& r) { return r; }

template <class CT, class R> __device__ __forceinline__ ShapeRec<CT> as_compute(const ShapeRec<R>& r);
template <> __device__ __forceinline__ ShapeRec<double> as_compute<double, double>(const ShapeRec<double>& r) { return r; }
template <> __device__ __forceinline__ ShapeRec<double> as_compute<double, float>(const ShapeRec<float>& r) { return widen(r); }
template <> __device__ __forceinline__ ShapeRec<float> as_compute<float, float>(const ShapeRec<float>& r) { return r; }

// Front pass over all candidate pairs (function_MPR_Store's early exits and
// sphere-sphere closed form, Narrowphase.cu:338-396).  Pairs that need the
// MPR get their shape-pair class (qcls, NO_CLASS for the others).  The queue
// of (class << 32 | pair index) is then grouped by class with a one-digit
// counting sort over tiles of QUEUE_TILE pairs: k_queue_hist stores each
// tile's class histogram class-major (hist[c * nt + t], nt tiles), one scan
// gives every (class, tile) its first queue slot, and k_queue_scatter places
// the tile's pairs, so every wavefront of the MPR kernels runs one pair of
// support maps.
// MIXED / FAST then run an FP32 MPR in the frame of shape A (k_np_filter):
// clear misses end there, FAST keeps clear FP32 hits, everything else goes
// to the reference-order MPR.  In those modes the front pass already applies
// the FP32 MPR's first portal test, so pairs separated along the centre
// line (most candidate pairs) never enter the queue.
// One candidate pair of the front pass: its hit flag cleared, the early
// exits and the sphere-sphere contact applied, and (return value) whether it
// needs the MPR.  ta / tb receive the shape types.
template <class R>
__device__ __forceinline__ bool np_front_one(uint64_t i, uint32_t ia, uint32_t ib,
                                             const ShapeRec<R>* __restrict__ rec, const ShapeHdr* __restrict__ hdr,
                                             const ShapeRecF* __restrict__ frec, float tau_rel, int prefilter,
                                             R envelope, const PairOut<R>& o, uint32_t& ta, uint32_t& tb) {
  o.hit[i] = 0u;
  bool enq = false;
  auto spheres = [&]() {
    const ShapeRec<R> ra = rec[ia], rb = rec[ib];
    if (spheres_touch(mk3(ra.x, ra.y, ra.z), ra.bx, mk3(rb.x, rb.y, rb.z), rb.bx)) mark_sphere_hit(o, i);
  };
  if (prefilter) {
    // the FP32 records carry type, body and activity: two gathers per pair
    const ShapeRecF fa = frec[ia], fb = frec[ib];
    ta = fa.type;
    tb = fb.type;
    if ((fa.active | fb.active) && fa.body != fb.body) {
      if (ta == SPHERE && tb == SPHERE) {
        // The FP32 records hold the scene-precision centres exactly (widened
        // to FP64) and, in an FP32 scene, the radii too: only an FP64 scene
        // reads its radii from the full records.
        R rad_a, rad_b;
        if constexpr (sizeof(R) == 4) {
          rad_a = fa.bx;
          rad_b = fb.bx;
        } else {
          rad_a = rec[ia].bx;
          rad_b = rec[ib].bx;
        }
        // the FP32 record's centre is the scene-precision centre of the
        // full record, exactly: the compaction's recomputation from the full
        // records sees the same values
        if (spheres_touch(mk3((R)fa.x, (R)fa.y, (R)fa.z), rad_a, mk3((R)fb.x, (R)fb.y, (R)fb.z), rad_b))
          mark_sphere_hit(o, i);
      } else if (!fp32_filtered(ta) || !fp32_filtered(tb)) {
        enq = true;  // no FP32 model of rectangles, cones and triangles: FP64 decides
      } else {
        const PlacedF A = placed_local(fa, fa), B = placed_local(fb, fa);
        const float tau = tau_rel * (bound_radius(A.type, A.B) + bound_radius(B.type, B.B));
        enq = !mpr_first_miss(A, B, tau);
      }
    }
  } else {
    const ShapeHdr ha = hdr[ia], hb = hdr[ib];
    ta = ha.ta & 255u;
    tb = hb.ta & 255u;
    if (((ha.ta | hb.ta) >> 8) && ha.body != hb.body) {
      if (ta == SPHERE && tb == SPHERE) spheres();
      else enq = true;
    }
  }
  return enq;
}

template <class R>
__global__ __launch_bounds__(128) void k_np_front(uint64_t cap, const uint64_t* __restrict__ d_np,
                                                  const uint64_t* __restrict__ pairs,
                                                  const ShapeRec<R>* __restrict__ rec, const ShapeHdr* __restrict__ hdr,
                                                  const ShapeRecF* __restrict__ frec, float tau_rel, int prefilter,
                                                  R envelope, PairOut<R> o, uint8_t* __restrict__ qcls) {
  const uint64_t i = (uint64_t)blockIdx.x * 128 + threadIdx.x;
  if (i >= dev_count(cap, d_np)) return;
  const uint64_t pr = pairs[i];
  const uint32_t ia = (uint32_t)(pr >> 32), ib = (uint32_t)(pr & 0xffffffffu);
  uint32_t ta, tb;
  const bool enq = np_front_one(i, ia, ib, rec, hdr, frec, tau_rel, prefilter, envelope, o, ta, tb);
  qcls[i] = (uint8_t)(enq ? pair_class(ta, tb) : NO_CLASS);
}

// Class histogram of one tile of QUEUE_TILE pairs (pair k * QUEUE_BS + t of
// the tile on thread t), stored class-major: hist[c * nt + tile].  Tile 0
// writes the number of histogram entries (the scan's length).
__global__ __launch_bounds__(QUEUE_BS) void k_queue_hist(uint64_t cap, const uint64_t* __restrict__ d_np,
                                                         const uint8_t* __restrict__ qcls,
                                                         uint32_t* __restrict__ hist, uint64_t* __restrict__ hist_n) {
  __shared__ uint32_t h[NQC];
  const uint64_t n = dev_count(cap, d_np);
  const uint64_t nt = (n + QUEUE_TILE - 1) / QUEUE_TILE;
  if (blockIdx.x == 0 && threadIdx.x == 0) *hist_n = (uint64_t)NQC * nt;
  if (blockIdx.x >= nt) return;  // workgroup-uniform
  if (threadIdx.x < NQC) h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * QUEUE_TILE + threadIdx.x;
#pragma unroll
  for (int k = 0; k < QUEUE_IPT; ++k) {
    const uint64_t i = t0 + (uint64_t)k * QUEUE_BS;
    const uint32_t c = i < n ? qcls[i] : NO_CLASS;
    if (c != NO_CLASS) atomicAdd(&h[c], 1u);
  }
  __syncthreads();
  if (threadIdx.x < NQC) hist[(uint64_t)threadIdx.x * nt + blockIdx.x] = h[threadIdx.x];
}

// Queue of the pairs the front pass kept, grouped by class: a pair of class
// c in tile t takes a slot of the run that starts at the scanned
// hist[c * nt + t].  Within a tile's run of a class, slots follow an LDS
// atomic; order within a class carries no meaning (each pair's result is
// written at the pair's own index).
__global__ __launch_bounds__(QUEUE_BS) void k_queue_scatter(uint64_t cap, const uint64_t* __restrict__ d_np,
                                                            const uint8_t* __restrict__ qcls,
                                                            const uint32_t* __restrict__ base,
                                                            uint64_t* __restrict__ queue) {
  __shared__ uint32_t lcnt[NQC], lbase[NQC];
  const uint64_t n = dev_count(cap, d_np);
  const uint64_t nt = (n + QUEUE_TILE - 1) / QUEUE_TILE;
  if (blockIdx.x >= nt) return;  // workgroup-uniform
  if (threadIdx.x < NQC) {
    lcnt[threadIdx.x] = 0;
    lbase[threadIdx.x] = base[(uint64_t)threadIdx.x * nt + blockIdx.x];
  }
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * QUEUE_TILE + threadIdx.x;
#pragma unroll
  for (int k = 0; k < QUEUE_IPT; ++k) {
    const uint64_t i = t0 + (uint64_t)k * QUEUE_BS;
    const uint32_t c = i < n ? qcls[i] : NO_CLASS;
    if (c != NO_CLASS) queue[lbase[c] + atomicAdd(&lcnt[c], 1u)] = ((uint64_t)c << 32) | i;
  }
}

// FP32 MPR over the class-sorted queue (MIXED / FAST), in the frame of
// shape A: clear misses end here; FAST keeps clear FP32 hits; everything
// else is selected for the reference-order MPR.
// FP32 MPR of one queued pair of the five filtered types: returns whether
// it goes on to the reference-order MPR (KEEP: FAST, which stores clear FP32
// hits itself).
template <class R, bool KEEP>
__device__ __forceinline__ bool np_filter_one(uint64_t i, const ShapeRecF& fa, const ShapeRecF& fb, R envelope,
                                              float tau_rel, const PairOut<R>& o) {
  const uint64_t P = o.P;
  PlacedF A = placed_local(fa, fa), B = placed_local(fb, fa);
  float tau = tau_rel * (bound_radius(A.type, A.B) + bound_radius(B.type, B.B));
  v3<float> N, p0;
  int r = mpr_guarded(A, B, tau, !KEEP, N, p0);
  if (KEEP && r == MPR_HIT) {
    v3<float> sa, sb;
    sup_fast_pair(A, B, N, sa, sb);
    v3<float> p1 = dot((sa - p0), N) * N + p0;
    v3<float> p2 = dot((sb - p0), N) * N + p0;
    N = -N;
    float dl = dot(N, p2 - p1);
    if (fabsf(dl) < tau) {
      r = MPR_UNSURE;
    } else {
      Contact<R> c;
      v3<double> Nd = mk3((double)N.x, (double)N.y, (double)N.z);
      v3<double> o3 = mk3(fa.x, fa.y, fa.z);
      v3<double> q1 = mk3((double)p1.x, (double)p1.y, (double)p1.z) + o3;
      v3<double> q2 = mk3((double)p2.x, (double)p2.y, (double)p2.z) + o3;
      q1 = q1 - Nd * (double)envelope;
      q2 = q2 + Nd * (double)envelope;
      c.N = mk3((R)Nd.x, (R)Nd.y, (R)Nd.z);
      c.p1 = mk3((R)q1.x, (R)q1.y, (R)q1.z);
      c.p2 = mk3((R)q2.x, (R)q2.y, (R)q2.z);
      c.depth = (R)dot(Nd, q2 - q1);
      store_pair(o, P, i, c, fa.body, fb.body);
    }
  }
  // MIXED: FP32 settles clear misses only; hits take the reference path,
  // whose capped iteration sequence FP32 cannot reproduce.
  return r == MPR_UNSURE || (r == MPR_HIT && !KEEP);
}

template <class R, bool KEEP>
__global__ __launch_bounds__(128) void k_np_filter(uint64_t cap, const uint64_t* __restrict__ d_q,
                                                   const uint64_t* __restrict__ queue,
                                                   const uint64_t* __restrict__ pairs,
                                                   const ShapeHdr* __restrict__ hdr,
                                                   const ShapeRecF* __restrict__ frec, R envelope, float tau_rel,
                                                   PairOut<R> o, uint32_t* __restrict__ sel,
                                                   uint64_t* __restrict__ tile_part) {
  const uint64_t j = (uint64_t)blockIdx.x * 128 + threadIdx.x;
  uint32_t keep = 0;
  if (j < dev_count(cap, d_q)) {
    const uint32_t i = (uint32_t)queue[j];
    const uint64_t pr = pairs[i];
    const uint32_t ia = (uint32_t)(pr >> 32), ib = (uint32_t)(pr & 0xffffffffu);
    const ShapeRecF fa = frec[ia], fb = frec[ib];
    // rectangles, cones and triangles: the reference-order MPR decides
    keep = !fp32_filtered(fa.type) || !fp32_filtered(fb.type) || np_filter_one<R, KEEP>(i, fa, fb, envelope, tau_rel, o);
    sel[j] = keep;
  }
  __shared__ uint32_t red[128 / WAVE];
  add_tile_sum<128>(tile_part, j, keep, red);  // the selection's tile sums, for scan_excl_u32_tiled
}

// Stable compaction of the selected queue entries (keeps the class order).
__global__ void k_queue_select(uint64_t cap, const uint64_t* __restrict__ d_q, const uint64_t* __restrict__ queue,
                               const uint32_t* __restrict__ sel, const uint32_t* __restrict__ pos,
                               uint64_t* __restrict__ out) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= dev_count(cap, d_q) || !sel[j]) return;
  out[pos[j]] = queue[j];
}


// MPR on the reference path for one pair, in compute type CT; a hit is
// stored in the scene's real type R.
template <class R, class CT, bool EXT>
__device__ __forceinline__ void np_resolve_one(uint64_t i, uint32_t ia, uint32_t ib,
                                               const ShapeRec<R>* __restrict__ rec, R envelope, const PairOut<R>& o) {
  ShapeRec<CT> ra = as_compute<CT, R>(rec[ia]);
  ShapeRec<CT> rb = as_compute<CT, R>(rec[ib]);
  Contact<CT> c;
  if (contact_native<EXT>(ra, rb, (CT)envelope, c)) {
    Contact<R> cr;
    cr.N = mk3((R)c.N.x, (R)c.N.y, (R)c.N.z);
    cr.p1 = mk3((R)c.p1.x, (R)c.p1.y, (R)c.p1.z);
    cr.p2 = mk3((R)c.p2.x, (R)c.p2.y, (R)c.p2.z);
    cr.depth = (R)c.depth;
    store_pair(o, o.P, i, cr, ra.body, rb.body);
  }
}

// MPR on the reference path for the queued pairs, in compute type CT
// (the scene's real type for NATIVE, FP64 for MIXED / FAST).
// EXT: the launch for the pairs of class EXT_CLASS and above (a rectangle,
// cone or triangle takes part); the other launch takes the rest, one entry
// per thread, skipping the extended ones, which the class sort keeps at the
// end of the queue.
template <class R, class CT, bool EXT>
__global__ __launch_bounds__(128) void k_np_resolve(uint64_t cap, const uint64_t* __restrict__ d_q,
                                                    const uint64_t* __restrict__ pairs,
                                                    const ShapeRec<R>* __restrict__ rec, R envelope, PairOut<R> o,
                                                    const uint64_t* __restrict__ queue) {
  const uint64_t q = dev_count(cap, d_q);
  auto one = [&](uint64_t e) {
    const uint32_t i = (uint32_t)e;
    const uint64_t pr = pairs[i];
    np_resolve_one<R, CT, EXT>(i, (uint32_t)(pr >> 32), (uint32_t)(pr & 0xffffffffu), rec, envelope, o);
  };
  const uint64_t t = (uint64_t)blockIdx.x * 128 + threadIdx.x;
  if constexpr (!EXT) {
    if (t >= q) return;
    const uint64_t e = queue[t];
    if ((uint32_t)(e >> 32) < EXT_CLASS) one(e);
  } else {
    // the extended classes sort last: walk the queue from its end and stop
    // at the first entry of a filtered class
    const uint64_t stride = (uint64_t)gridDim.x * 128;
    for (uint64_t k = t; k < q; k += stride) {
      const uint64_t e = queue[q - 1 - k];
      if ((uint32_t)(e >> 32) < EXT_CLASS) return;
      one(e);
    }
  }
}

// Grid of the extended-class resolve launch (it strides over its entries).
constexpr unsigned EXT_RESOLVE_GRID = 1024;

// Stable partition of hits to the front (Narrowphase.cu:580-593): contacts
// in ascending pair order, and the pair list permuted hits first.  A
// workgroup's hits are consecutive contacts: they are gathered in LDS and the
// interleaved contact arrays leave in coalesced stores.
constexpr int COMPACT_BS = 256;

template <class R>
__global__ __launch_bounds__(COMPACT_BS) void k_np_compact(uint64_t cap, const uint64_t* __restrict__ d_np,
                                                           PairOut<R> o, const uint32_t* __restrict__ pos,
                                                           const uint64_t* __restrict__ pairs,
                                                           const ShapeRec<R>* __restrict__ rec, R envelope,
                                                           const DevStatus* __restrict__ st, R* __restrict__ out,
                                                           uint64_t* __restrict__ perm) {
  __shared__ R lv[10][COMPACT_BS];
  __shared__ int32_t lid[2 * COMPACT_BS];
  __shared__ uint32_t s_k0, s_nh;
  const uint64_t n = dev_count(cap, d_np);
  const uint64_t i0 = (uint64_t)blockIdx.x * COMPACT_BS;
  if (i0 >= n) return;
  const uint64_t i = i0 + threadIdx.x;
  const uint64_t last = min(i0 + COMPACT_BS, n) - 1;
  const uint64_t C = st->C;
  // the contact arrays, one after the other in one buffer, each C rows long
  R* __restrict__ norm = out;
  R* __restrict__ cpta = out + 3 * C;
  R* __restrict__ cptb = out + 6 * C;
  R* __restrict__ dpth = out + 9 * C;
  int32_t* __restrict__ bids = reinterpret_cast<int32_t*>(out + 10 * C);
  if (threadIdx.x == 0) s_k0 = pos[i0];
  if (i == last) s_nh = pos[i] + (o.hit[i] ? 1u : 0u);
  __syncthreads();
  const uint32_t k0 = s_k0, nh = s_nh - k0;
  if (i < n) {
    const uint32_t k = pos[i];
    if (o.hit[i]) {
      const uint32_t sl = k - k0;
      const uint64_t pr = pairs[i];
      if (o.kind[i]) {
        const ShapeRec<R> ra = rec[(uint32_t)(pr >> 32)], rb = rec[(uint32_t)(pr & 0xffffffffu)];
        Contact<R> c;
        contact_spheres(ra, rb, envelope, c);  // a hit by the front pass's test
        lv[0][sl] = c.N.x; lv[1][sl] = c.N.y; lv[2][sl] = c.N.z;
        lv[3][sl] = c.p1.x; lv[4][sl] = c.p1.y; lv[5][sl] = c.p1.z;
        lv[6][sl] = c.p2.x; lv[7][sl] = c.p2.y; lv[8][sl] = c.p2.z;
        lv[9][sl] = c.depth;
        lid[2 * sl] = (int32_t)ra.body;
        lid[2 * sl + 1] = (int32_t)rb.body;
      } else {
        const PairRec<R> r = o.rec[i];
#pragma unroll
        for (int c = 0; c < 10; ++c) lv[c][sl] = r.v[c];
        lid[2 * sl] = r.ids[0];
        lid[2 * sl + 1] = r.ids[1];
      }
      perm[k] = pr;
    } else {
      perm[C + (i - k)] = pairs[i];
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < 3 * nh; t += COMPACT_BS) {
    const uint32_t sl = t / 3, c = t - 3 * sl;
    norm[3 * (uint64_t)k0 + t] = lv[c][sl];
    cpta[3 * (uint64_t)k0 + t] = lv[3 + c][sl];
    cptb[3 * (uint64_t)k0 + t] = lv[6 + c][sl];
  }
  for (uint32_t t = threadIdx.x; t < nh; t += COMPACT_BS) dpth[k0 + t] = lv[9][t];
  for (uint32_t t = threadIdx.x; t < 2 * nh; t += COMPACT_BS) bids[2 * (uint64_t)k0 + t] = lid[t];
}


}  // namespace
}  // namespace cdx
