// ORACLE — test infrastructure only, not part of the product.
//
// A CPU restatement (C++17 + OpenMP) of the reference's shipped default path,
// the OpenMP build of Chrono::Collide (src/collision_detection.h:60 leaves
// SIM_ENABLE_GPU_MODE off).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or as the timed CPU
// baseline; the HIP library never links or calls it.
//
// Each function cites the reference file:line it follows.  The vector and
// quaternion arithmetic the reference takes from its missing gpu_math
// submodule is fixed in ref_math.hpp.  Additions that have no reference
// counterpart are marked EXTENSION: the capsule shape (type 7).
//
// Parity pins (tests/test_oracle_pins.py, tests/test_oracle_kat.py):
//   aabb        — sphere closed form c±r; box |R|h vs 8-corner enumeration;
//                 rotation vs Hamilton product q v q*.
//   broadphase  — reference KATs test/test1.cu (N(N-1)/2 pairs),
//                 test/binTest_Broadphase.cu (n^3 pairs (2i,2i+1)),
//                 test/nSquared_Broadphase.cu (superset of true contacts,
//                 glibc rand() seed 1); pair list vs O(N^2) brute force on the
//                 shifted AABBs; bin counts vs closed form; sorted keys and
//                 occupancy vs an independent histogram.
//   narrowphase — sphere-sphere closed form; MPR sphere pairs vs closed form;
//                 box face contacts vs separating-axis depth; equal-axis
//                 ellipsoid == sphere; capsule / cylinder axial and radial
//                 closed forms; |N|=1; the barycentric contact point vs a
//                 numpy solve of the barycentric system on the final portal
//                 and inside both shapes (planted slips ORACLE_PLANT 1-3
//                 must fail it); p1 / p2 on the shapes' supporting planes;
//                 swap symmetry (a,b) -> (b,a).
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>
#include <omp.h>
#include <parallel/algorithm>

#include "ref_math.hpp"

using namespace orc;

namespace orc {

// Operation-counting real for the FLOP census of the narrowphase
// (oracle_mpr_flops_f64): a double whose +, -, *, / and sqrt each add one
// to a per-thread counter and a fused multiply-add two (negation, fabs and
// comparisons are free, as on the GPU where they fold into operand
// modifiers and compares).
struct Fl {
  double v;
  Fl() = default;
  Fl(double x) : v(x) {}
  explicit operator double() const { return v; }
};
inline thread_local uint64_t t_flops = 0;
inline Fl operator+(Fl a, Fl b) { ++t_flops; return Fl(a.v + b.v); }
inline Fl operator-(Fl a, Fl b) { ++t_flops; return Fl(a.v - b.v); }
inline Fl operator*(Fl a, Fl b) { ++t_flops; return Fl(a.v * b.v); }
inline Fl operator/(Fl a, Fl b) { ++t_flops; return Fl(a.v / b.v); }
inline Fl operator-(Fl a) { return Fl(-a.v); }
inline bool operator<(Fl a, Fl b) { return a.v < b.v; }
inline bool operator>(Fl a, Fl b) { return a.v > b.v; }
inline bool operator<=(Fl a, Fl b) { return a.v <= b.v; }
inline bool operator>=(Fl a, Fl b) { return a.v >= b.v; }
inline bool operator==(Fl a, Fl b) { return a.v == b.v; }
inline bool operator!=(Fl a, Fl b) { return a.v != b.v; }
inline Fl sqrt(Fl a) { ++t_flops; return Fl(std::sqrt(a.v)); }
inline Fl fma(Fl a, Fl b, Fl c) { t_flops += 2; return Fl(std::fma(a.v, b.v, c.v)); }
inline Fl fabs(Fl a) { return Fl(std::fabs(a.v)); }
template <> inline Fl eps_of<Fl>() { return Fl(DBL_EPSILON); }

// Counter value when a pair's world placement is done (function_MPR_Store
// past Narrowphase.cu:377), for the census; a no-op for float / double.
inline thread_local uint64_t t_flops_placed = 0;
template <class R> inline void mark_placed() {}
template <> inline void mark_placed<Fl>() { t_flops_placed = t_flops; }

}  // namespace orc

namespace {

enum : int { SPHERE = 0, ELLIPSOID = 1, BOX = 2, CYLINDER = 3, RECT = 4, CONE = 5, TRIANGLEMESH = 6, CAPSULE = 7 };

template <class R> inline r3<R> ld3(const R* p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
template <class R> inline r4<R> ld4(const R* p, size_t i) { return {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }
template <class R> inline void st3(R* p, size_t i, r3<R> v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

// ---------------------------------------------------------------- AABB ----
// AABBGenerator.cu:11-14
template <class R> void ComputeAABBSphere(R radius, r3<R> position, r3<R>& mn, r3<R>& mx) {
  mn = position - R3<R>(radius, radius, radius);
  mx = position + R3<R>(radius, radius, radius);
}
// AABBGenerator.cu:25-35 (used for ELLIPSOID, BOX, CYLINDER at :108-109)
template <class R> void ComputeAABBBox(r3<R> dim, r3<R> lpos, r3<R> pos, r4<R> lrot, r4<R> rot, r3<R>& mn, r3<R>& mx) {
  r4<R> q1 = mult(rot, lrot);
  m33<R> rotmat = AbsMat(AMat(q1));
  r3<R> temp = MatMult(rotmat, dim);
  r3<R> p = quatRotate(lpos, rot) + pos;
  mn = p - temp;
  mx = p + temp;
}
// AABBGenerator.cu:16-23
template <class R> void ComputeAABBTriangle(r3<R> A, r3<R> B, r3<R> C, r3<R>& mn, r3<R>& mx) {
  mn = R3<R>(std::min(A.x, std::min(B.x, C.x)), std::min(A.y, std::min(B.y, C.y)), std::min(A.z, std::min(B.z, C.z)));
  mx = R3<R>(std::max(A.x, std::max(B.x, C.x)), std::max(A.y, std::max(B.y, C.y)), std::max(A.z, std::max(B.z, C.z)));
}

// AABBGenerator.cu:84-116, with three reference quirks fixed (FIX):
//  * TRIANGLEMESH (6) gets the triangle bounds of :103-107, which the
//    reference computes under type 5 instead;
//  * CONE (5) gets the loose box (r, h/2, r) of its support set instead of
//    the triangle path;
//  * RECT (4), for which the reference writes nothing, gets the box of its
//    rectangle (half extents B.x, B.y at height B.z).
// RECT and CONE are bounded where function_MPR_Store places them: at their
// local position, without the body pose (Narrowphase.cu:359-377 transforms
// only SPHERE, ELLIPSOID, BOX, CYLINDER and TRIANGLEMESH), rotated by
// body (x) local, so the box holds exactly the shape the narrowphase tests.
template <class R>
void function_ComputeAABB(size_t i, const int32_t* T, const R* A, const R* B, const R* Cv, const R* Rq,
                          const uint32_t* ID, const R* bpos, const R* brot, size_t n, R* aabb) {
  int type = T[i];
  uint32_t id = ID[i];
  r3<R> a = ld3(A, i), b = ld3(B, i);
  r3<R> position = ld3(bpos, id);
  r4<R> rotation = ld4(brot, id);
  r3<R> mn, mx;
  if (type == SPHERE) {
    a = quatRotate(a, rotation);
    ComputeAABBSphere(b.x, a + position, mn, mx);
  } else if (type == ELLIPSOID || type == BOX || type == CYLINDER) {
    ComputeAABBBox(b, a, position, ld4(Rq, i), rotation, mn, mx);
  } else if (type == CAPSULE) {
    // EXTENSION: capsule (radius B.x, half-length B.y along local y) bounded
    // by the box of semi-extents (B.x, B.y + B.x, B.z), the same loose-box
    // treatment the reference gives cylinders.
    r3<R> d = R3<R>(b.x, b.y + b.x, b.z);
    ComputeAABBBox(d, a, position, ld4(Rq, i), rotation, mn, mx);
  } else if (type == TRIANGLEMESH) {  // FIX: the reference's type-5 branch
    r4<R> q1 = mult(rotation, ld4(Rq, i));
    r3<R> c = Cv ? ld3(Cv, i) : R3<R>(0, 0, 0);
    ComputeAABBTriangle(quatRotate(a + position, q1), quatRotate(b + position, q1), quatRotate(c + position, q1), mn,
                        mx);
  } else if (type == CONE) {  // FIX: the support set's box, at the narrowphase's placement
    r4<R> q1 = mult(rotation, ld4(Rq, i));
    r3<R> h = MatMult(AbsMat(AMat(q1)), R3<R>(b.x, b.y * R(.5), b.x));
    mn = a - h;
    mx = a + h;
  } else if (type == RECT) {  // FIX: the rectangle's box, at the narrowphase's placement
    r4<R> q1 = mult(rotation, ld4(Rq, i));
    r3<R> ctr = a + quatRotateMat(R3<R>(0, 0, b.z), q1);
    r3<R> h = MatMult(AbsMat(AMat(q1)), R3<R>(b.x, b.y, 0));
    mn = ctr - h;
    mx = ctr + h;
  } else {
    return;
  }
  st3(aabb, i, mn);
  st3(aabb, i + n, mx);
}

// --------------------------------------------------------- broadphase ----
// Broadphase.cu:125-133
template <class R> inline i3 HashMin(r3<R> a, r3<R> inv) {
  i3 t;
  t.x = (int)(a.x * inv.x);
  t.y = (int)(a.y * inv.y);
  t.z = (int)(a.z * inv.z);
  return t;
}
// Broadphase.cu:140-147; called with U3(i,j,k) (:189), so the arithmetic is
// unsigned and wraps.
inline uint32_t Hash_Index(uint32_t i, uint32_t j, uint32_t k, i3 g) {
  return ((k * (uint32_t)g.y) * (uint32_t)g.x) + (j * (uint32_t)g.x) + i;
}

template <class T> void inclusive_scan_seq(std::vector<T>& v) {
  for (size_t i = 1; i < v.size(); ++i) v[i] += v[i - 1];
}
template <class T> void inclusive_scan_par(std::vector<T>& v) {
  size_t n = v.size();
  if (n < 2) return;
  int nt = omp_get_max_threads();
  std::vector<T> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num(), T_ = omp_get_num_threads();
    size_t lo = n * t / T_, hi = n * (t + 1) / T_;
    T s = 0;
    for (size_t i = lo; i < hi; ++i) { s += v[i]; v[i] = s; }
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int k = 1; k <= T_; ++k) part[k] += part[k - 1];
    T off = part[t];
    for (size_t i = lo; i < hi; ++i) v[i] += off;
  }
}
template <class T> void inclusive_scan(std::vector<T>& v, bool par) {
  if (par) inclusive_scan_par(v); else inclusive_scan_seq(v);
}

struct BPFlags {
  bool transform_reduce = true, transform = false, scan1 = true, sort_by_key = false, reduce_by_key = false,
       max_element = false, scan2 = false, scan3 = false, sort = true, unique = false;
};

template <class R> struct BPOut {
  std::vector<uint32_t> bin_counts, bins_intersected, sorted_keys, sorted_bodies, active_keys, active_counts, bin_end,
      num_contactd;
  std::vector<int64_t> pairs;
  size_t pdup = 0;
  float bmin[3], bmax[3];
  r3<R> origin, inv;
};

// Broadphase.cu:150-159
template <class R> inline uint32_t count_bins(size_t i, const R* aabb, r3<R> inv, size_t n) {
  i3 gmin = HashMin(ld3(aabb, i), inv);
  i3 gmax = HashMin(ld3(aabb, i + n), inv);
  return (uint32_t)((gmax.x - gmin.x + 1) * (gmax.y - gmin.y + 1) * (gmax.z - gmin.z + 1));
}
// Broadphase.cu:172-200
template <class R>
inline void store_bins(size_t index, const R* aabb, const uint32_t* bins_intersected, r3<R> inv, i3 grid, size_t n,
                       uint32_t* bin_number, uint32_t* body_number) {
  uint32_t count = 0;
  i3 gmin = HashMin(ld3(aabb, index), inv);
  i3 gmax = HashMin(ld3(aabb, index + n), inv);
  uint32_t mInd = (index == 0) ? 0 : bins_intersected[index - 1];
  for (uint32_t i = gmin.x; i <= (uint32_t)gmax.x; i++)
    for (uint32_t j = gmin.y; j <= (uint32_t)gmax.y; j++)
      for (uint32_t k = gmin.z; k <= (uint32_t)gmax.z; k++) {
        bin_number[mInd + count] = Hash_Index(i, j, k, grid);
        body_number[mInd + count] = (uint32_t)index;
        count++;
      }
}
template <class R> inline bool overlap(const R* aabb, size_t n, uint32_t a, uint32_t b) {
  r3<R> Amin = ld3(aabb, a), Amax = ld3(aabb, a + n), Bmin = ld3(aabb, b), Bmax = ld3(aabb, b + n);
  return (Amin.x <= Bmax.x && Bmin.x <= Amax.x) && (Amin.y <= Bmax.y && Bmin.y <= Amax.y) &&
         (Amin.z <= Bmax.z && Bmin.z <= Amax.z);
}
// Broadphase.cu:215-251
template <class R>
inline uint32_t count_pairs(size_t index, const R* aabb, size_t n, const uint32_t* body_number, const uint32_t* bin_end) {
  uint32_t end = bin_end[index], count = 0, i = index ? bin_end[index - 1] : 0;
  for (; i < end; i++) {
    uint32_t ta = body_number[i];
    for (uint32_t k = i + 1; k < end; k++) {
      uint32_t tb = body_number[k];
      if (ta != tb && overlap(aabb, n, ta, tb)) count++;
    }
  }
  return count;
}
// Broadphase.cu:265-326
template <class R>
inline void store_pairs(size_t index, const R* aabb, size_t n, const uint32_t* body_number, const uint32_t* bin_end,
                        const uint32_t* num_contactd, int64_t* out) {
  uint32_t end = bin_end[index], count = 0, i = index ? bin_end[index - 1] : 0;
  uint32_t offset = index ? num_contactd[index - 1] : 0;
  if (end - i == 1) return;
  for (; i < end; i++) {
    uint32_t ta = body_number[i];
    for (uint32_t k = i + 1; k < end; k++) {
      uint32_t tb = body_number[k];
      if (ta != tb && overlap(aabb, n, ta, tb)) {
        int a = (int)ta, b = (int)tb;
        if (b < a) std::swap(a, b);
        out[offset + count] = ((int64_t)a << 32 | (int64_t)b);
        count++;
      }
    }
  }
}

// Broadphase.cu:344-545 (detectPossibleCollisions).  aabb (2n real3, mins then
// maxes) is shifted in place, as the reference does at :382-386.
template <class R> void broadphase(R* aabb, size_t n, i3 grid, const BPFlags& f, BPOut<R>& o) {
  o = BPOut<R>();
  if (n == 0) return;  // the reference reads aabb_data[0] here (UB on empty input)
  // STEP 2 bounds, Broadphase.cu:368-376: fminf / fmaxf narrow to float.
  float mn[3] = {(float)aabb[0], (float)aabb[1], (float)aabb[2]};
  float mx[3] = {mn[0], mn[1], mn[2]};
  size_t npts = 2 * n;
  if (f.transform_reduce) {
    float a0 = mn[0], a1 = mn[1], a2 = mn[2], b0 = mx[0], b1 = mx[1], b2 = mx[2];
#pragma omp parallel for reduction(min : a0, a1, a2) reduction(max : b0, b1, b2)
    for (size_t i = 0; i < npts; ++i) {
      a0 = std::fmin(a0, (float)aabb[3 * i]); a1 = std::fmin(a1, (float)aabb[3 * i + 1]); a2 = std::fmin(a2, (float)aabb[3 * i + 2]);
      b0 = std::fmax(b0, (float)aabb[3 * i]); b1 = std::fmax(b1, (float)aabb[3 * i + 1]); b2 = std::fmax(b2, (float)aabb[3 * i + 2]);
    }
    mn[0] = a0; mn[1] = a1; mn[2] = a2; mx[0] = b0; mx[1] = b1; mx[2] = b2;
  } else {
    for (size_t i = 0; i < npts; ++i)
      for (int c = 0; c < 3; ++c) {
        mn[c] = fminf(mn[c], (float)aabb[3 * i + c]);
        mx[c] = fmaxf(mx[c], (float)aabb[3 * i + c]);
      }
  }
  for (int c = 0; c < 3; ++c) { o.bmin[c] = mn[c]; o.bmax[c] = mx[c]; }
  r3<R> minb = R3<R>(mn[0], mn[1], mn[2]), maxb = R3<R>(mx[0], mx[1], mx[2]);
  // Broadphase.cu:377-381
  r3<R> origin = fabs3(minb);
  r3<R> ext = fabs3(maxb + fabs3(minb));
  r3<R> inv = R3<R>((R)grid.x, (R)grid.y, (R)grid.z) / ext;
  o.origin = origin;
  o.inv = inv;
  // Broadphase.cu:382-386 (in-place shift)
#pragma omp parallel for if (f.transform)
  for (size_t i = 0; i < npts; ++i) st3(aabb, i, ld3(aabb, i) + origin);

  // STEP 3, Broadphase.cu:393-413
  std::vector<uint32_t> bins(n);
#pragma omp parallel for schedule(guided)
  for (size_t i = 0; i < n; ++i) bins[i] = count_bins(i, aabb, inv, n);
  o.bin_counts = bins;
  inclusive_scan(bins, f.scan1);
  size_t M = bins.back();
  o.bins_intersected = bins;
  std::vector<uint32_t> bin_number(M), body_number(M);
  // STEP 4, Broadphase.cu:415-432
#pragma omp parallel for schedule(guided)
  for (size_t i = 0; i < n; ++i) store_bins(i, aabb, bins.data(), inv, grid, n, bin_number.data(), body_number.data());
  {
    std::vector<std::pair<uint32_t, uint32_t>> kv(M);
    for (size_t i = 0; i < M; ++i) kv[i] = {bin_number[i], body_number[i]};
    auto byKey = [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
      return a.first < b.first;
    };
    if (f.sort_by_key) __gnu_parallel::stable_sort(kv.begin(), kv.end(), byKey);
    else std::stable_sort(kv.begin(), kv.end(), byKey);
    for (size_t i = 0; i < M; ++i) { bin_number[i] = kv[i].first; body_number[i] = kv[i].second; }
  }
  o.sorted_keys = bin_number;
  o.sorted_bodies = body_number;
  // Broadphase.cu:439-443 reduce_by_key (unique keys written over bin_number)
  std::vector<uint32_t> keys, counts;
  for (size_t i = 0; i < M;) {
    size_t j = i;
    while (j < M && bin_number[j] == bin_number[i]) ++j;
    keys.push_back(bin_number[i]);
    counts.push_back((uint32_t)(j - i));
    i = j;
  }
  size_t A = keys.size();
  o.active_keys = keys;
  o.active_counts = counts;
  if (A == 0) return;  // Broadphase.cu:457
  // Broadphase.cu:470-480
  std::vector<uint32_t> bin_end = counts;
  inclusive_scan(bin_end, f.scan2);
  o.bin_end = bin_end;
  // STEP 5, Broadphase.cu:484-500
  std::vector<uint32_t> ncd(A);
#pragma omp parallel for schedule(guided)
  for (size_t b = 0; b < A; ++b) ncd[b] = count_pairs(b, aabb, n, body_number.data(), bin_end.data());
  o.num_contactd = ncd;
  inclusive_scan(ncd, f.scan3);
  size_t P = ncd.back();
  std::vector<int64_t> pairs(P);
  // STEP 6, Broadphase.cu:505-535
#pragma omp parallel for
  for (size_t b = 0; b < A; ++b) store_pairs(b, aabb, n, body_number.data(), bin_end.data(), ncd.data(), pairs.data());
  o.pdup = P;
  if (f.sort) __gnu_parallel::sort(pairs.begin(), pairs.end());
  else std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  o.pairs = std::move(pairs);
}

// -------------------------------------------------------- narrowphase ----
// SupportFunctions.cuh:11-13
template <class R> inline r3<R> sup_sphere(r3<R> B, r3<R> n) { return B.x * n; }
// SupportFunctions.cuh:30-36
template <class R> inline r3<R> sup_box(r3<R> B, r3<R> n) {
  return R3<R>(sign(n.x) * B.x, sign(n.y) * B.y, sign(n.z) * B.z);
}
// SupportFunctions.cuh:37-50
template <class R> inline r3<R> sup_ellipsoid(r3<R> B, r3<R> n) {
  r3<R> nn = normalize(n);
  return B * B * nn * (R(1) / length(B * nn));
}
// SupportFunctions.cuh:52-63
template <class R> inline r3<R> sup_cylinder(r3<R> B, r3<R> n) {
  r3<R> u = R3<R>(0, 1, 0);
  r3<R> w = n - (dot(u, n)) * u;
  if (length(w) != R(0)) return sign(dot(u, n)) * B.y * u + B.x * normalize(w);
  return sign(dot(u, n)) * B.y * u;
}
// EXTENSION: capsule of radius B.x around the segment y in [-B.y, B.y]:
// support = segment support (+) sphere support.
template <class R> inline r3<R> sup_capsule(r3<R> B, r3<R> n) {
  r3<R> u = R3<R>(0, 1, 0);
  return sign(n.y) * B.y * u + B.x * normalize(n);
}

// SupportFunctions.cuh:65-73 (RECT: the rectangle |x| <= B.x, |y| <= B.y at z = B.z)
template <class R> inline r3<R> sup_plane(r3<R> B, r3<R> n) {
  r3<R> r = B;
  if (n.x < R(0)) r.x = -r.x;
  if (n.y < R(0)) r.y = -r.y;
  return r;
}
// SupportFunctions.cuh:74-99 (CONE: radius B.x, height B.y, apex at +B.y/2)
template <class R> inline r3<R> sup_cone(r3<R> B, r3<R> n) {
  R radius = B.x;
  R height = B.y;
  R sin_angle = radius / sqrt(radius * radius + height * height);
  if (n.y > length(n) * sin_angle) return R3<R>(0, height / R(2.0), 0);
  R s = sqrt(n.x * n.x + n.z * n.z);
  if ((double)s > 1e-9) {  // a double literal in the reference: compared in double
    R d = radius / s;
    return R3<R>(n.x * d, -height / R(2.0), n.z * d);
  }
  return R3<R>(0, -height / R(2.0), 0);
}
// SupportFunctions.cuh:14-29 (world-placed vertices; the first maximum wins)
template <class R> inline r3<R> sup_triangle(r3<R> A, r3<R> B, r3<R> C, r3<R> n) {
  R dist = dot(A, n);
  r3<R> point = A;
  if (dot(B, n) > dist) {
    dist = dot(B, n);
    point = B;
  }
  if (dot(C, n) > dist) {
    dist = dot(C, n);
    point = C;
  }
  return point;
}
// Narrowphase.cu:11-17
template <class R> inline r3<R> GetCenter(int type, r3<R> A, r3<R> B, r3<R> C) {
  if (type == TRIANGLEMESH)  // SupportFunctions.cuh:119-121
    return R3<R>((A.x + B.x + C.x) / R(3.0f), (A.y + B.y + C.y) / R(3.0f), (A.z + B.z + C.z) / R(3.0f));
  return R3<R>(0, 0, 0) + A;
}

// Support-map evaluations made by this thread (for the MPR step histogram).
thread_local uint64_t t_support_calls = 0;

// Narrowphase.cu:19-52
template <class R> inline r3<R> TransformSupportVert(int type, r3<R> A, r3<R> B, r3<R> C, r4<R> Rq, r3<R> b) {
  ++t_support_calls;
  r3<R> n = normalize(b);
  if (type == TRIANGLEMESH) return sup_triangle(A, B, C, n);  // :46-48, world vertices
  r3<R> rn = quatRotateMatT(n, Rq);
  r3<R> ls;
  switch (type) {
    case SPHERE: ls = sup_sphere(B, rn); break;
    case ELLIPSOID: ls = sup_ellipsoid(B, rn); break;
    case BOX: ls = sup_box(B, rn); break;
    case CYLINDER: ls = sup_cylinder(B, rn); break;
    case RECT: ls = sup_plane(B, rn); break;
    case CONE: ls = sup_cone(B, rn); break;
    case CAPSULE: ls = sup_capsule(B, rn); break;
    default: ls = R3<R>(0, 0, 0); break;
  }
  return quatRotateMat(ls, Rq) + A;
}

// Final portal of a hit that reaches the barycentric contact point
// (Narrowphase.cu:265-289): the tetrahedron v0..v3 of Minkowski-difference
// points, the support points of A (v*1) and B (v*2) that make each vertex,
// the weights as computed, and whether the sum <= 0 fallback was taken.
// Filled only when a caller asks, for the pin tests; it changes nothing.
template <class R> struct MprPortal {
  r3<R> v[4], a[4], b[4], n;
  R w[4];
  int fallback = 0, filled = 0;
};

// Planted slips (ORACLE_PLANT, compiled only into the test variants
// liboracle_plant<k>.so): 1 drops the halving of the contact point,
// 2 swaps the weights of v2 and v3, 3 pairs v1's weight with v2's A point.
// The pin tests build each variant and check that it fails them.
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

// Narrowphase.cu:122-314 (XenoCollide-derived MPR).  GetCenter (:11-17) is
// the shape position for every in-scope type.
template <class R>
bool CollideAndFindPoint(int tA, r3<R> AX, r3<R> AY, r3<R> AZ, r4<R> AR, int tB, r3<R> BX, r3<R> BY, r3<R> BZ,
                         r4<R> BR, r3<R>& rnormal, r3<R>& point, R& depth, MprPortal<R>* portal = nullptr) {
  r3<R> v01, v02, v0, n, v11, v12, v1, v21, v22, v2;
  v01 = GetCenter(tA, AX, AY, AZ);
  v02 = GetCenter(tB, BX, BY, BZ);
  v0 = v02 - v01;
  if (IsZero(v0)) v0 = R3<R>(1, 0, 0);
  n = -v0;
  v11 = TransformSupportVert(tA, AX, AY, AZ, AR, -n);
  v12 = TransformSupportVert(tB, BX, BY, BZ, BR, n);
  v1 = v12 - v11;
  if (dot(v1, n) <= 0.0) return false;
  n = cross(v1, v0);
  if (IsZero(n)) {
    n = v1 - v0;
    n = normalize(n);
    rnormal = n;
    point = (v11 + v12) * R(.5);
    depth = dot((v12 - v11), n);
    return true;
  }
  v21 = TransformSupportVert(tA, AX, AY, AZ, AR, -n);
  v22 = TransformSupportVert(tB, BX, BY, BZ, BR, n);
  v2 = v22 - v21;
  if (dot(v2, n) <= 0.0) return false;
  n = cross((v1 - v0), (v2 - v0));
  if (dot(n, v0) > 0.0) {
    std::swap(v1, v2);
    std::swap(v11, v21);
    std::swap(v12, v22);
    n = -n;
  }
  r3<R> v31, v32, v3;
  bool hit = false;
  int phase1 = 0, phase2 = 0;
  const int max_iterations = 34;
  while (true) {
    if (phase1 > max_iterations) return false;
    phase1++;
    v31 = TransformSupportVert(tA, AX, AY, AZ, AR, -n);
    v32 = TransformSupportVert(tB, BX, BY, BZ, BR, n);
    v3 = v32 - v31;
    if (dot(v3, n) <= 0.0) return false;
    if (dot(cross(v1, v3), v0) < 0.0) {
      v2 = v3; v21 = v31; v22 = v32;
      n = cross((v1 - v0), (v3 - v0));
      continue;
    }
    if (dot(cross(v3, v2), v0) < 0.0) {
      v1 = v3; v11 = v31; v12 = v32;
      n = cross((v3 - v0), (v2 - v0));
      continue;
    }
    while (true) {
      phase2++;
      n = cross((v2 - v1), (v3 - v1));
      n = normalize(n);
      if (dot(n, v1) >= 0.0 && !hit) hit = true;
      r3<R> v41 = TransformSupportVert(tA, AX, AY, AZ, AR, -n);
      r3<R> v42 = TransformSupportVert(tB, BX, BY, BZ, BR, n);
      r3<R> v4 = v42 - v41;
      R delta = dot((v4 - v3), n);
      depth = dot(v4, n);
      if (delta <= 1e-20 || depth <= 0.0 || phase2 > max_iterations) {
        if (hit) {
          R b0 = dot(cross(v1, v2), v3);
          R b1 = dot(cross(v3, v2), v0);
          R b2 = dot(cross(v0, v1), v3);
          R b3 = dot(cross(v2, v1), v0);
          R sum = b0 + b1 + b2 + b3;
          int fallback = 0;
          if (sum <= 0.) {
            fallback = 1;
            b0 = 0;
            b1 = dot(cross(v2, v3), n);
            b2 = dot(cross(v3, v1), n);
            b3 = dot(cross(v1, v2), n);
            sum = b1 + b2 + b3;
          }
          R inv = R(1.0f) / sum;
#if ORACLE_PLANT == 2
          std::swap(b2, b3);
#endif
#if ORACLE_PLANT == 3
          point = (b0 * v01 + b1 * v21 + b2 * v21 + b3 * v31) + (b0 * v02 + b1 * v12 + b2 * v22 + b3 * v32);
#else
          point = (b0 * v01 + b1 * v11 + b2 * v21 + b3 * v31) + (b0 * v02 + b1 * v12 + b2 * v22 + b3 * v32);
#endif
#if ORACLE_PLANT == 1
          point = point * inv;
#else
          point = point * inv * R(.5);
#endif
          rnormal = normalize(n);
          (void)fallback;
          if (portal) {
            portal->v[0] = v0; portal->v[1] = v1; portal->v[2] = v2; portal->v[3] = v3;
            portal->a[0] = v01; portal->a[1] = v11; portal->a[2] = v21; portal->a[3] = v31;
            portal->b[0] = v02; portal->b[1] = v12; portal->b[2] = v22; portal->b[3] = v32;
            portal->w[0] = b0; portal->w[1] = b1; portal->w[2] = b2; portal->w[3] = b3;
            portal->n = n;
            portal->fallback = fallback;
            portal->filled = 1;
          }
        }
        return hit;
      }
      if (dot(cross(v4, v1), v0) < 0) {
        if (dot(cross(v4, v2), v0) < 0) { v1 = v4; v11 = v41; v12 = v42; }
        else { v3 = v4; v31 = v41; v32 = v42; }
      } else {
        if (dot(cross(v4, v3), v0) < 0) { v2 = v4; v21 = v41; v22 = v42; }
        else { v1 = v4; v11 = v41; v12 = v42; }
      }
    }
  }
}

template <class R> struct NPOut {
  std::vector<R> norm, cpta, cptb, dpth;
  std::vector<int32_t> bids;
  std::vector<int64_t> pairs_permuted;
  std::vector<uint8_t> hit;
  size_t c = 0;
};

// Shapes function_MPR_Store places by the body pose (Narrowphase.cu:359-367);
// RECT and CONE keep their local position as their centre there.
inline bool body_attached(int t) { return t == SPHERE || t == ELLIPSOID || t == BOX || t == CYLINDER || t == CAPSULE; }

// Narrowphase.cu:316-423 (function_MPR_Store); returns true on contact.
template <class R>
bool mpr_store(int64_t p, const int32_t* T, const R* A, const R* B, const R* Cv, const R* Rq, const uint32_t* ID,
               const uint8_t* active, const R* bpos, const R* brot, R envelope, r3<R>& N, r3<R>& p1, r3<R>& p2,
               R& depth, int32_t ids[2]) {
  int ia = (int)(p >> 32), ib = (int)(p & 0xffffffff);
  int tA = T[ia], tB = T[ib];
  uint32_t idA = ID[ia], idB = ID[ib];
  if (!active[idA] && !active[idB]) return false;
  if (idA == idB) return false;
  r3<R> posA = ld3(bpos, idA), posB = ld3(bpos, idB);
  r4<R> rotA = ld4(brot, idA), rotB = ld4(brot, idB);
  r3<R> AX = ld3(A, ia), BX = ld3(A, ib);
  r3<R> AY = ld3(B, ia), BY = ld3(B, ib);
  r3<R> AZ = Cv ? ld3(Cv, ia) : R3<R>(0, 0, 0), BZ = Cv ? ld3(Cv, ib) : R3<R>(0, 0, 0);
  r4<R> AR = mult(rotA, ld4(Rq, ia));
  r4<R> BR = mult(rotB, ld4(Rq, ib));
  if (body_attached(tA)) {
    AX = quatRotate(AX, rotA) + posA;
  } else if (tA == TRIANGLEMESH) {  // Narrowphase.cu:363-367
    envelope = 0;
    AX = quatRotate(AX + posA, AR);
    AY = quatRotate(AY + posA, AR);
    AZ = quatRotate(AZ + posA, AR);
  }
  if (body_attached(tB)) {
    BX = quatRotate(BX, rotB) + posB;
  } else if (tB == TRIANGLEMESH) {  // Narrowphase.cu:372-376
    envelope = 0;
    BX = quatRotate(BX + posB, BR);
    BY = quatRotate(BY + posB, BR);
    BZ = quatRotate(BZ + posB, BR);
  }
  mark_placed<R>();
  N = R3<R>(1, 0, 0);
  p1 = R3<R>(0, 0, 0);
  p2 = R3<R>(0, 0, 0);
  r3<R> p0 = R3<R>(0, 0, 0);
  depth = 0;
  if (tA == SPHERE && tB == SPHERE) {
    r3<R> relpos = BX - AX;
    R d2 = dot(relpos, relpos);
    R cd = AY.x + BY.x;
    if (d2 <= cd * cd) {
      N = normalize(relpos);
      p1 = AX + N * AY.x;
      p2 = BX - N * BY.x;
    } else {
      return false;
    }
  } else {
    if (!CollideAndFindPoint(tA, AX, AY, AZ, AR, tB, BX, BY, BZ, BR, N, p0, depth)) return false;
    p1 = dot((TransformSupportVert(tA, AX, AY, AZ, AR, -N) - p0), N) * N + p0;
    p2 = dot((TransformSupportVert(tB, BX, BY, BZ, BR, N) - p0), N) * N + p0;
    N = -N;
  }
  p1 = p1 - N * envelope;
  p2 = p2 + N * envelope;
  depth = dot(N, p2 - p1);
  ids[0] = (int32_t)idA;
  ids[1] = (int32_t)idB;
  return true;
}

// Narrowphase.cu:511-595 (DoNarrowphase): per-pair loop (:481-503) then a
// stable partition of hits to the front (:580-593).
template <class R>
void narrowphase(size_t P, const int64_t* pairs, const int32_t* T, const R* A, const R* B, const R* Cv, const R* Rq,
                 const uint32_t* ID, const uint8_t* active, const R* bpos, const R* brot, R envelope, NPOut<R>& o) {
  o = NPOut<R>();
  std::vector<R> nrm(3 * P), pa(3 * P), pb(3 * P), dp(P);
  std::vector<int32_t> ids(2 * P);
  std::vector<uint8_t> hit(P, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (size_t i = 0; i < P; ++i) {
    r3<R> N, p1, p2;
    R d;
    int32_t id2[2];
    if (mpr_store(pairs[i], T, A, B, Cv, Rq, ID, active, bpos, brot, envelope, N, p1, p2, d, id2)) {
      hit[i] = 1;
      st3(nrm.data(), i, N);
      st3(pa.data(), i, p1);
      st3(pb.data(), i, p2);
      dp[i] = d;
      ids[2 * i] = id2[0];
      ids[2 * i + 1] = id2[1];
    }
  }
  size_t C = 0;
  for (size_t i = 0; i < P; ++i) C += hit[i];
  o.c = C;
  o.norm.resize(3 * C); o.cpta.resize(3 * C); o.cptb.resize(3 * C); o.dpth.resize(C); o.bids.resize(2 * C);
  o.pairs_permuted.resize(P);
  size_t h = 0, m = C;
  for (size_t i = 0; i < P; ++i) {
    if (hit[i]) {
      for (int c = 0; c < 3; ++c) { o.norm[3 * h + c] = nrm[3 * i + c]; o.cpta[3 * h + c] = pa[3 * i + c]; o.cptb[3 * h + c] = pb[3 * i + c]; }
      o.dpth[h] = dp[i];
      o.bids[2 * h] = ids[2 * i]; o.bids[2 * h + 1] = ids[2 * i + 1];
      o.pairs_permuted[h++] = pairs[i];
    } else {
      o.pairs_permuted[m++] = pairs[i];
    }
  }
  o.hit = std::move(hit);
}

template <class T> T* dup_vec(const std::vector<T>& v) {
  T* p = (T*)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

}  // namespace

extern "C" {

// Result of one broadphase call, arrays malloc'd, freed by oracle_bp_free.
struct oracle_bp_result {
  uint64_t n, m, a, pdup, p;
  uint32_t* bin_counts;        // n, per-AABB bins spanned (Broadphase.cu:150-159)
  uint32_t* bins_intersected;  // n, inclusive scan of bin_counts (:403-407)
  uint32_t* sorted_keys;       // m, bin_number after sort_by_key (:428-432)
  uint32_t* sorted_bodies;     // m, body_number after sort_by_key
  uint32_t* active_keys;       // a, unique keys (reduce_by_key, :439-443)
  uint32_t* active_counts;     // a, occupancy of each active bin
  uint32_t* bin_end;           // a, inclusive scan of active_counts (:476-480)
  uint32_t* num_contactd;      // a, per-bin pre-dedup pair counts (:486-491)
  int64_t* pairs;              // p, sorted unique (a<<32|b) (:522-535)
  float bmin[3], bmax[3];
  double origin[3], inv[3];
};

struct oracle_np_result {
  uint64_t p, c;
  void* norm;   // 3c real
  void* cpta;   // 3c real
  void* cptb;   // 3c real
  void* dpth;   // c real
  int32_t* bids;            // 2c
  int64_t* pairs_permuted;  // p, hits first (stable), then the rest
  uint8_t* hit;             // p, per input pair
};

static BPFlags flags_from(const int32_t* f) {
  BPFlags b;
  if (f) {
    b.transform_reduce = f[0]; b.transform = f[1]; b.scan1 = f[2]; b.sort_by_key = f[3]; b.reduce_by_key = f[4];
    b.max_element = f[5]; b.scan2 = f[6]; b.scan3 = f[7]; b.sort = f[8]; b.unique = f[9];
  }
  return b;
}

#define ORACLE_AABB(NAME, R)                                                                                   \
  int NAME(uint64_t n, const int32_t* T, const R* A, const R* B, const R* C, const R* Rq, const uint32_t* ID,  \
           const R* bpos, const R* brot, R* aabb) {                                                          \
    _Pragma("omp parallel for") for (int64_t i = 0; i < (int64_t)n; ++i)                                      \
        function_ComputeAABB<R>((size_t)i, T, A, B, C, Rq, ID, bpos, brot, (size_t)n, aabb);                  \
    return 0;                                                                                                  \
  }
ORACLE_AABB(oracle_aabb_f64, double)
ORACLE_AABB(oracle_aabb_f32, float)

#define ORACLE_BP(NAME, R)                                                                                     \
  oracle_bp_result* NAME(R* aabb, uint64_t n, int gx, int gy, int gz, const int32_t* flags) {                 \
    BPOut<R> o;                                                                                                \
    broadphase<R>(aabb, (size_t)n, i3{gx, gy, gz}, flags_from(flags), o);                                      \
    auto* r = (oracle_bp_result*)std::calloc(1, sizeof(oracle_bp_result));                                    \
    r->n = n; r->m = o.sorted_keys.size(); r->a = o.active_keys.size(); r->pdup = o.pdup; r->p = o.pairs.size(); \
    r->bin_counts = dup_vec(o.bin_counts); r->bins_intersected = dup_vec(o.bins_intersected);                  \
    r->sorted_keys = dup_vec(o.sorted_keys); r->sorted_bodies = dup_vec(o.sorted_bodies);                      \
    r->active_keys = dup_vec(o.active_keys); r->active_counts = dup_vec(o.active_counts);                      \
    r->bin_end = dup_vec(o.bin_end); r->num_contactd = dup_vec(o.num_contactd); r->pairs = dup_vec(o.pairs);   \
    for (int c = 0; c < 3; ++c) { r->bmin[c] = o.bmin[c]; r->bmax[c] = o.bmax[c]; }                           \
    if (n) {                                                                                                   \
      r->origin[0] = o.origin.x; r->origin[1] = o.origin.y; r->origin[2] = o.origin.z;                         \
      r->inv[0] = o.inv.x; r->inv[1] = o.inv.y; r->inv[2] = o.inv.z;                                           \
    }                                                                                                          \
    return r;                                                                                                  \
  }
ORACLE_BP(oracle_broadphase_f64, double)
ORACLE_BP(oracle_broadphase_f32, float)

void oracle_bp_free(oracle_bp_result* r) {
  if (!r) return;
  std::free(r->bin_counts); std::free(r->bins_intersected); std::free(r->sorted_keys); std::free(r->sorted_bodies);
  std::free(r->active_keys); std::free(r->active_counts); std::free(r->bin_end); std::free(r->num_contactd);
  std::free(r->pairs); std::free(r);
}

#define ORACLE_NP(NAME, R)                                                                                      \
  oracle_np_result* NAME(uint64_t P, const int64_t* pairs, const int32_t* T, const R* A, const R* B, const R* C, \
                         const R* Rq, const uint32_t* ID, const uint8_t* active, const R* bpos, const R* brot,  \
                         R envelope) {                                                                        \
    NPOut<R> o;                                                                                                 \
    narrowphase<R>((size_t)P, pairs, T, A, B, C, Rq, ID, active, bpos, brot, envelope, o);                        \
    auto* r = (oracle_np_result*)std::calloc(1, sizeof(oracle_np_result));                                     \
    r->p = P; r->c = o.c;                                                                                       \
    r->norm = dup_vec(o.norm); r->cpta = dup_vec(o.cpta); r->cptb = dup_vec(o.cptb); r->dpth = dup_vec(o.dpth); \
    r->bids = dup_vec(o.bids); r->pairs_permuted = dup_vec(o.pairs_permuted); r->hit = dup_vec(o.hit);          \
    return r;                                                                                                   \
  }
ORACLE_NP(oracle_narrowphase_f64, double)
ORACLE_NP(oracle_narrowphase_f32, float)

void oracle_np_free(oracle_np_result* r) {
  if (!r) return;
  std::free(r->norm); std::free(r->cpta); std::free(r->cptb); std::free(r->dpth); std::free(r->bids);
  std::free(r->pairs_permuted); std::free(r->hit); std::free(r);
}

// Support steps (pairs of support-map evaluations, the MPR's unit of work,
// including the two that place the contact points) per candidate pair of
// function_MPR_Store (Narrowphase.cu:316-423): 0 for pairs that end before
// the MPR (inactive, same body, sphere-sphere).
void oracle_mpr_steps_f64(uint64_t P, const int64_t* pairs, const int32_t* T, const double* A, const double* B,
                          const double* C, const double* Rq, const uint32_t* ID, const uint8_t* active,
                          const double* bpos, const double* brot, int32_t* steps) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)P; ++i) {
    r3<double> N, p1, p2;
    double d;
    int32_t id2[2];
    t_support_calls = 0;
    mpr_store<double>(pairs[i], T, A, B, C, Rq, ID, active, bpos, brot, 0.0, N, p1, p2, d, id2);
    steps[i] = (int32_t)(t_support_calls / 2);
  }
}

// FLOP census of function_MPR_Store (Narrowphase.cu:316-423) per candidate
// pair in FP64: the +, -, *, / and sqrt operations the reference algorithm
// performs for the pair once both shapes are placed in the world (their
// world centres and rotations are per-shape work), through the MPR and the
// contact points; 0 for pairs that end before the placement (inactive,
// same body).  The pair runs on a two-shape copy of its inputs.
void oracle_mpr_flops_f64(uint64_t P, const int64_t* pairs, const int32_t* T, const double* A, const double* B,
                          const double* C, const double* Rq, const uint32_t* ID, const uint8_t* active,
                          const double* bpos, const double* brot, double envelope, uint64_t* flops) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)P; ++i) {
    const uint64_t ia = (uint64_t)pairs[i] >> 32, ib = (uint64_t)pairs[i] & 0xffffffffu;
    const uint64_t sh[2] = {ia, ib};
    int32_t t2[2];
    uint32_t id2[2] = {0, ID[ia] == ID[ib] ? 0u : 1u};
    uint8_t act2[2];
    Fl a2[6], b2[6], c2[6], r2[8], p2[6], q2[8];
    for (int k = 0; k < 2; ++k) {
      const uint64_t s = sh[k], b = ID[s];
      t2[k] = T[s];
      act2[id2[k]] = active[b];
      for (int c = 0; c < 3; ++c) {
        a2[3 * k + c] = A[3 * s + c];
        b2[3 * k + c] = B[3 * s + c];
        c2[3 * k + c] = C ? C[3 * s + c] : 0.0;
      }
      for (int c = 0; c < 4; ++c) r2[4 * k + c] = Rq[4 * s + c];
      for (int c = 0; c < 3; ++c) p2[3 * id2[k] + c] = bpos[3 * b + c];
      for (int c = 0; c < 4; ++c) q2[4 * id2[k] + c] = brot[4 * b + c];
    }
    r3<Fl> N, pa, pb;
    Fl d;
    int32_t ids[2];
    t_flops = 0;
    t_flops_placed = 0;
    mpr_store<Fl>((int64_t)1, t2, a2, b2, c2, r2, id2, act2, p2, q2, Fl(envelope), N, pa, pb, d, ids);
    flops[i] = t_flops_placed ? t_flops - t_flops_placed : 0;
  }
}

// One MPR query on two world-placed shapes (CollideAndFindPoint,
// Narrowphase.cu:122-314), for closed-form pin tests.
int oracle_mpr_f64(int tA, const double* AX, const double* AY, const double* AR, int tB, const double* BX,
                   const double* BY, const double* BR, double* normal, double* point, double* depth) {
  r3<double> n{0, 0, 0}, p{0, 0, 0};
  double d = 0;
  const r3<double> z{0, 0, 0};
  bool hit = CollideAndFindPoint<double>(tA, ld3(AX, 0), ld3(AY, 0), z, ld4(AR, 0), tB, ld3(BX, 0), ld3(BY, 0), z,
                                         ld4(BR, 0), n, p, d);
  st3(normal, 0, n);
  st3(point, 0, p);
  *depth = d;
  return hit ? 1 : 0;
}

// The same query, also returning the final portal when the hit reaches the
// barycentric contact point (Narrowphase.cu:265-289): portal = 12 v, 12 A
// points, 12 B points (4 vertices x 3), 4 weights, the portal normal (3),
// then the fallback flag and whether the portal was filled (as doubles).
int oracle_mpr_portal_f64(int tA, const double* AX, const double* AY, const double* AR, int tB, const double* BX,
                          const double* BY, const double* BR, double* normal, double* point, double* depth,
                          double* portal) {
  r3<double> n{0, 0, 0}, p{0, 0, 0};
  double d = 0;
  MprPortal<double> pt;
  const r3<double> z{0, 0, 0};
  bool hit = CollideAndFindPoint<double>(tA, ld3(AX, 0), ld3(AY, 0), z, ld4(AR, 0), tB, ld3(BX, 0), ld3(BY, 0), z,
                                         ld4(BR, 0), n, p, d, &pt);
  st3(normal, 0, n);
  st3(point, 0, p);
  *depth = d;
  for (int i = 0; i < 4; ++i) {
    st3(portal, i, pt.v[i]);
    st3(portal + 12, i, pt.a[i]);
    st3(portal + 24, i, pt.b[i]);
    portal[36 + i] = pt.w[i];
  }
  st3(portal + 40, 0, pt.n);
  portal[43] = pt.fallback;
  portal[44] = pt.filled;
  return hit ? 1 : 0;
}

// Support point of one world-placed shape (TransformSupportVert,
// Narrowphase.cu:19-52), for closed-form pin tests.
void oracle_support_f64(int t, const double* X, const double* Y, const double* Rq, const double* dir, double* out) {
  st3(out, 0, TransformSupportVert<double>(t, ld3(X, 0), ld3(Y, 0), r3<double>{0, 0, 0}, ld4(Rq, 0), ld3(dir, 0)));
}

void oracle_quat_rotate_f64(const double* v, const double* q, double* out) {
  st3(out, 0, quatRotate<double>(ld3(v, 0), ld4(q, 0)));
}
void oracle_quat_mult_f64(const double* a, const double* b, double* out) {
  r4<double> q = mult<double>(ld4(a, 0), ld4(b, 0));
  out[0] = q.w; out[1] = q.x; out[2] = q.y; out[3] = q.z;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
