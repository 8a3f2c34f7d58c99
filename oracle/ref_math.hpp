// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this code, and only as the checker.
//
// Vector / quaternion arithmetic the reference takes from its (un-vendored)
// gpu_math submodule (hmazhar/gpu_math, .gitmodules:4-6; pinned version
// unknown: the mount has no gitlink SHA).  The header is absent from the
// reference tree, so every formula below is a stated choice, pinned by the
// mathematics in tests/test_oracle_pins.py (orthonormality of AMat, agreement
// with the Hamilton product q v q*, 8-corner box bounds, closed-form contacts):
//   real3 = three packed reals; real4 = quaternion stored scalar-first (w,x,y,z)
//     (AABBGenerator.cuh:53-57 "[w x y z] - Rotation").
//   mult(a,b)        = Hamilton product a (x) b, applied as body (x) local
//                      (AABBGenerator.cu:27, :92; Narrowphase.cu:356-357).
//   AMat(q)          = rotation matrix of unit q, entries in the
//                      "(e0e0+e1e1)*2-1" form.
//   quatRotate(v,q)  = quatRotateMat(v,q) = AMat(q)*v ; quatRotateMatT = AMat(q)^T*v.
//   dot, cross, AMat and the matrix-vector products fuse each multiply-add
//     (one rounding, fixed order; as nvcc contracts the reference's own
//     device build by default).
//   normalize(v)     = v * (1 / sqrt(dot(v,v))); zero input returns (1,0,0).
//   sign(x)          = -1 / 0 / +1.
//   IsZero(x)        = |x| < FLT_EPSILON (float) or DBL_EPSILON (double);
//                      IsZero(real3) = all three components.
// Every expression keeps a fixed evaluation order and the build uses
// -ffp-contract=off (only the explicit fma calls fuse), so results are
// reproducible to the bit.
#pragma once
#include <cfloat>
#include <cmath>

namespace orc {

// Unqualified below, so that the operation-counting real of the flop
// census (oracle.cpp) finds its own overloads; for float / double these are
// the std functions.
using std::fabs;
using std::fma;
using std::sqrt;

template <class R> struct r3 { R x, y, z; };
template <class R> struct r4 { R w, x, y, z; };
struct i3 { int x, y, z; };

template <class R> inline r3<R> R3(R x, R y, R z) { return r3<R>{x, y, z}; }
template <class R> inline r3<R> operator+(r3<R> a, r3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline r3<R> operator-(r3<R> a, r3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline r3<R> operator-(r3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline r3<R> operator*(r3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline r3<R> operator*(R s, r3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline r3<R> operator*(r3<R> a, r3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline r3<R> operator/(r3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline r3<R> operator/(r3<R> a, r3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }

// Each multiply-add of dot, cross, AMat and the matrix-vector products is
// one fused operation (one rounding) in the order written.
template <class R> inline R dot(r3<R> a, r3<R> b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
template <class R> inline r3<R> cross(r3<R> a, r3<R> b) {
  return {fma(a.y, b.z, -(a.z * b.y)), fma(a.z, b.x, -(a.x * b.z)), fma(a.x, b.y, -(a.y * b.x))};
}
template <class R> inline R length(r3<R> a) { return sqrt(dot(a, a)); }
template <class R> inline r3<R> normalize(r3<R> a) {
  R l = length(a);
  if (l == R(0)) return {R(1), R(0), R(0)};
  return a * (R(1) / l);
}
template <class R> inline r3<R> fabs3(r3<R> a) { return {fabs(a.x), fabs(a.y), fabs(a.z)}; }
template <class R> inline R sign(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0)); }
template <class R> inline R eps_of();
template <> inline float eps_of<float>() { return FLT_EPSILON; }
template <> inline double eps_of<double>() { return DBL_EPSILON; }
template <class R> inline bool IsZero(R x) { return fabs(x) < eps_of<R>(); }
template <class R> inline bool IsZero(r3<R> a) { return IsZero(a.x) && IsZero(a.y) && IsZero(a.z); }

template <class R> inline r4<R> mult(r4<R> a, r4<R> b) {
  r4<R> q;
  q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
  q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
  return q;
}

template <class R> struct m33 { R m[3][3]; };

// (e0e0 + e1e1) * 2 - 1 with e0e0 + e1e1 fused; the doubling is exact, so
// the "- 1" is the single rounding of fma(s, 2, -1).
template <class R> inline m33<R> AMat(r4<R> q) {
  m33<R> r;
  r.m[0][0] = fma(fma(q.w, q.w, q.x * q.x), R(2), R(-1));
  r.m[0][1] = fma(q.x, q.y, -(q.w * q.z)) * R(2);
  r.m[0][2] = fma(q.x, q.z, q.w * q.y) * R(2);
  r.m[1][0] = fma(q.x, q.y, q.w * q.z) * R(2);
  r.m[1][1] = fma(fma(q.w, q.w, q.y * q.y), R(2), R(-1));
  r.m[1][2] = fma(q.y, q.z, -(q.w * q.x)) * R(2);
  r.m[2][0] = fma(q.x, q.z, -(q.w * q.y)) * R(2);
  r.m[2][1] = fma(q.y, q.z, q.w * q.x) * R(2);
  r.m[2][2] = fma(fma(q.w, q.w, q.z * q.z), R(2), R(-1));
  return r;
}
template <class R> inline m33<R> AbsMat(m33<R> a) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a.m[i][j] = fabs(a.m[i][j]);
  return a;
}
template <class R> inline r3<R> MatMult(const m33<R>& a, r3<R> v) {
  return {fma(a.m[0][2], v.z, fma(a.m[0][1], v.y, a.m[0][0] * v.x)),
          fma(a.m[1][2], v.z, fma(a.m[1][1], v.y, a.m[1][0] * v.x)),
          fma(a.m[2][2], v.z, fma(a.m[2][1], v.y, a.m[2][0] * v.x))};
}
template <class R> inline r3<R> MatTMult(const m33<R>& a, r3<R> v) {
  return {fma(a.m[2][0], v.z, fma(a.m[1][0], v.y, a.m[0][0] * v.x)),
          fma(a.m[2][1], v.z, fma(a.m[1][1], v.y, a.m[0][1] * v.x)),
          fma(a.m[2][2], v.z, fma(a.m[1][2], v.y, a.m[0][2] * v.x))};
}
template <class R> inline r3<R> quatRotate(r3<R> v, r4<R> q) { return MatMult(AMat(q), v); }
template <class R> inline r3<R> quatRotateMat(r3<R> v, r4<R> q) { return MatMult(AMat(q), v); }
template <class R> inline r3<R> quatRotateMatT(r3<R> v, r4<R> q) { return MatTMult(AMat(q), v); }

}  // namespace orc
