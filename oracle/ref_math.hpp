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
//   normalize(v)     = v / sqrt(dot(v,v)); zero input returns (1,0,0).
//   sign(x)          = -1 / 0 / +1.
//   IsZero(x)        = |x| < FLT_EPSILON (float) or DBL_EPSILON (double);
//                      IsZero(real3) = all three components.
// Every expression keeps a fixed evaluation order and the build uses
// -ffp-contract=off, so results are reproducible to the bit.
#pragma once
#include <cfloat>
#include <cmath>

namespace orc {

// Unqualified below, so that the operation-counting real of the flop
// census (oracle.cpp) finds its own overloads; for float / double these are
// the std functions.
using std::fabs;
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

template <class R> inline R dot(r3<R> a, r3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline r3<R> cross(r3<R> a, r3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(r3<R> a) { return sqrt(dot(a, a)); }
template <class R> inline r3<R> normalize(r3<R> a) {
  R l = length(a);
  if (l == R(0)) return {R(1), R(0), R(0)};
  return a / l;
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

template <class R> inline m33<R> AMat(r4<R> q) {
  R e0e0 = q.w * q.w, e1e1 = q.x * q.x, e2e2 = q.y * q.y, e3e3 = q.z * q.z;
  R e0e1 = q.w * q.x, e0e2 = q.w * q.y, e0e3 = q.w * q.z;
  R e1e2 = q.x * q.y, e1e3 = q.x * q.z, e2e3 = q.y * q.z;
  m33<R> r;
  r.m[0][0] = (e0e0 + e1e1) * R(2) - R(1);
  r.m[0][1] = (e1e2 - e0e3) * R(2);
  r.m[0][2] = (e1e3 + e0e2) * R(2);
  r.m[1][0] = (e1e2 + e0e3) * R(2);
  r.m[1][1] = (e0e0 + e2e2) * R(2) - R(1);
  r.m[1][2] = (e2e3 - e0e1) * R(2);
  r.m[2][0] = (e1e3 - e0e2) * R(2);
  r.m[2][1] = (e2e3 + e0e1) * R(2);
  r.m[2][2] = (e0e0 + e3e3) * R(2) - R(1);
  return r;
}
template <class R> inline m33<R> AbsMat(m33<R> a) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a.m[i][j] = fabs(a.m[i][j]);
  return a;
}
template <class R> inline r3<R> MatMult(const m33<R>& a, r3<R> v) {
  return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
          a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
          a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
template <class R> inline r3<R> MatTMult(const m33<R>& a, r3<R> v) {
  return {a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z,
          a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z,
          a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z};
}
template <class R> inline r3<R> quatRotate(r3<R> v, r4<R> q) { return MatMult(AMat(q), v); }
template <class R> inline r3<R> quatRotateMat(r3<R> v, r4<R> q) { return MatMult(AMat(q), v); }
template <class R> inline r3<R> quatRotateMatT(r3<R> v, r4<R> q) { return MatTMult(AMat(q), v); }

}  // namespace orc
