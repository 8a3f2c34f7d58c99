// Vector, quaternion and support-map arithmetic for the gfx950 kernels.
//
// The reference takes this arithmetic from its gpu_math submodule, which is
// not vendored.  The conventions used here: real4 quaternions are stored
// scalar first (w,x,y,z) (AABBGenerator.cuh:53-57); mult(a,b) is the Hamilton
// product a (x) b; rotations use the matrix of q in the "(e0e0+e1e1)*2-1"
// form; normalize(0) returns (1,0,0); sign(0) = 0; IsZero compares against
// FLT_EPSILON for float and DBL_EPSILON for double.
//
// Every expression keeps one fixed evaluation order and the library is built
// with -ffp-contract=off and correctly rounded division and square root, so
// the kernels reproduce an IEEE host evaluation of the same expressions to
// the bit.  That is what makes AABBs, bin keys and the pair list bit-exact.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace cdx {

enum : int { SPHERE = 0, ELLIPSOID = 1, BOX = 2, CYLINDER = 3, RECT = 4, CONE = 5, TRIANGLEMESH = 6, CAPSULE = 7 };

#define CDX_HD __host__ __device__ __forceinline__

template <class R> struct v3 { R x, y, z; };
template <class R> struct q4 { R w, x, y, z; };

template <class R> CDX_HD v3<R> mk3(R x, R y, R z) { v3<R> r; r.x = x; r.y = y; r.z = z; return r; }
template <class R> CDX_HD v3<R> operator+(v3<R> a, v3<R> b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> CDX_HD v3<R> operator-(v3<R> a, v3<R> b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> CDX_HD v3<R> operator-(v3<R> a) { return mk3(-a.x, -a.y, -a.z); }
template <class R> CDX_HD v3<R> operator*(v3<R> a, R s) { return mk3(a.x * s, a.y * s, a.z * s); }
template <class R> CDX_HD v3<R> operator*(R s, v3<R> a) { return mk3(s * a.x, s * a.y, s * a.z); }
template <class R> CDX_HD v3<R> operator*(v3<R> a, v3<R> b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> CDX_HD v3<R> operator/(v3<R> a, R s) { return mk3(a.x / s, a.y / s, a.z / s); }

// Fused multiply-add, one rounding (v_fma_f32 / v_fma_f64).  The products
// and sums of dot, cross, the rotation matrix and matrix-vector products are
// written as explicit fused operations in a fixed order (the oracle states
// the same ones), so the build's -ffp-contract=off changes none of them.
CDX_HD float cfma(float a, float b, float c) { return fmaf(a, b, c); }
CDX_HD double cfma(double a, double b, double c) { return fma(a, b, c); }

template <class R> CDX_HD R dot(v3<R> a, v3<R> b) { return cfma(a.z, b.z, cfma(a.y, b.y, a.x * b.x)); }
template <class R> CDX_HD v3<R> cross(v3<R> a, v3<R> b) {
  return mk3(cfma(a.y, b.z, -(a.z * b.y)), cfma(a.z, b.x, -(a.x * b.z)), cfma(a.x, b.y, -(a.y * b.x)));
}
CDX_HD float csqrt(float x) { return sqrtf(x); }
CDX_HD double csqrt(double x) { return sqrt(x); }
CDX_HD float cabs(float x) { return fabsf(x); }
CDX_HD double cabs(double x) { return fabs(x); }
template <class R> CDX_HD R length(v3<R> a) { return csqrt(dot(a, a)); }
// a / |a| as a times one reciprocal of the length.
template <class R> CDX_HD v3<R> normalize(v3<R> a) {
  R l = length(a);
  if (l == R(0)) return mk3<R>(1, 0, 0);
  return a * (R(1) / l);
}
template <class R> CDX_HD R sgn(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0)); }
template <class R> CDX_HD R eps_of();
template <> CDX_HD float eps_of<float>() { return FLT_EPSILON; }
template <> CDX_HD double eps_of<double>() { return DBL_EPSILON; }
template <class R> CDX_HD bool is_zero(R x) { return cabs(x) < eps_of<R>(); }
template <class R> CDX_HD bool is_zero(v3<R> a) { return is_zero(a.x) && is_zero(a.y) && is_zero(a.z); }

// Hamilton product a (x) b.
template <class R> CDX_HD q4<R> qmul(q4<R> a, q4<R> b) {
  q4<R> q;
  q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  q.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
  q.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
  return q;
}

// Rotation matrix of a unit quaternion, row major.
template <class R> struct mat3 { R a00, a01, a02, a10, a11, a12, a20, a21, a22; };
// Entries in the "(e0e0+e1e1)*2-1" form, each product pair fused: the
// doubling is exact, so (s * 2) - 1 is the single rounding of fma(s, 2, -1).
template <class R> CDX_HD mat3<R> rotmat(q4<R> q) {
  mat3<R> m;
  m.a00 = cfma(cfma(q.w, q.w, q.x * q.x), R(2), R(-1));
  m.a01 = cfma(q.x, q.y, -(q.w * q.z)) * R(2);
  m.a02 = cfma(q.x, q.z, q.w * q.y) * R(2);
  m.a10 = cfma(q.x, q.y, q.w * q.z) * R(2);
  m.a11 = cfma(cfma(q.w, q.w, q.y * q.y), R(2), R(-1));
  m.a12 = cfma(q.y, q.z, -(q.w * q.x)) * R(2);
  m.a20 = cfma(q.x, q.z, -(q.w * q.y)) * R(2);
  m.a21 = cfma(q.y, q.z, q.w * q.x) * R(2);
  m.a22 = cfma(cfma(q.w, q.w, q.z * q.z), R(2), R(-1));
  return m;
}
template <class R> CDX_HD v3<R> mul(const mat3<R>& m, v3<R> v) {
  return mk3(cfma(m.a02, v.z, cfma(m.a01, v.y, m.a00 * v.x)), cfma(m.a12, v.z, cfma(m.a11, v.y, m.a10 * v.x)),
             cfma(m.a22, v.z, cfma(m.a21, v.y, m.a20 * v.x)));
}
template <class R> CDX_HD v3<R> mulT(const mat3<R>& m, v3<R> v) {
  return mk3(cfma(m.a20, v.z, cfma(m.a10, v.y, m.a00 * v.x)), cfma(m.a21, v.z, cfma(m.a11, v.y, m.a01 * v.x)),
             cfma(m.a22, v.z, cfma(m.a12, v.y, m.a02 * v.x)));
}
template <class R> CDX_HD mat3<R> absm(mat3<R> m) {
  m.a00 = cabs(m.a00); m.a01 = cabs(m.a01); m.a02 = cabs(m.a02);
  m.a10 = cabs(m.a10); m.a11 = cabs(m.a11); m.a12 = cabs(m.a12);
  m.a20 = cabs(m.a20); m.a21 = cabs(m.a21); m.a22 = cabs(m.a22);
  return m;
}
template <class R> CDX_HD v3<R> rotate(v3<R> v, q4<R> q) { return mul(rotmat(q), v); }

// Local-frame support maps (SupportFunctions.cuh:11-99), n of unit length.
// EXT compiles in the rectangle and the cone; the kernels that never meet
// them leave them out (a kernel's register allocation covers every case it
// holds).
template <class R, bool EXT = true> CDX_HD v3<R> support_local(int type, v3<R> B, v3<R> n) {
  switch (type) {
    case SPHERE: return B.x * n;
    case BOX: return mk3(sgn(n.x) * B.x, sgn(n.y) * B.y, sgn(n.z) * B.z);
    case ELLIPSOID: {
      v3<R> nn = normalize(n);
      return B * B * nn * (R(1) / length(B * nn));
    }
    case CYLINDER: {
      // Axis y.  w = n - (u.n)u with u = (0,1,0) evaluates exactly to
      // (n.x, 0, n.z), and every zero term of the reference's vector form
      // adds exactly, so this is the same value to the bit.
      v3<R> w = mk3(n.x, R(0), n.z);
      R s = sgn(n.y) * B.y;
      if (length(w) != R(0)) {
        v3<R> wn = normalize(w);
        return mk3(B.x * wn.x, s, B.x * wn.z);
      }
      return mk3(R(0) * s, s, R(0) * s);
    }
    case CAPSULE: {
      // Segment y in [-B.y, B.y] swept by a sphere of radius B.x.
      v3<R> nn = normalize(n);
      R s = sgn(n.y) * B.y;
      return mk3(R(0) * s + B.x * nn.x, s + B.x * nn.y, R(0) * s + B.x * nn.z);
    }
    case RECT:  // SupportFunctions.cuh:65-73: |x| <= B.x, |y| <= B.y at z = B.z
      if constexpr (!EXT) return mk3<R>(0, 0, 0);
      return mk3(n.x < R(0) ? -B.x : B.x, n.y < R(0) ? -B.y : B.y, B.z);
    case CONE: {  // SupportFunctions.cuh:74-99: radius B.x, height B.y, apex at +B.y/2
      if constexpr (!EXT) return mk3<R>(0, 0, 0);
      const R radius = B.x, height = B.y;
      const R sin_angle = radius / csqrt(radius * radius + height * height);
      if (n.y > length(n) * sin_angle) return mk3(R(0), height / R(2.0), R(0));
      const R s = csqrt(n.x * n.x + n.z * n.z);
      if ((double)s > 1e-9) {  // the reference's double literal: compared in double
        const R d = radius / s;
        return mk3(n.x * d, -height / R(2.0), n.z * d);
      }
      return mk3(R(0), -height / R(2.0), R(0));
    }
    default: return mk3<R>(0, 0, 0);
  }
}

// SupportFunctions.cuh:14-29: the vertex furthest along n, the first of
// equals (world-placed vertices; TRIANGLEMESH is not rotated per query).
template <class R> CDX_HD v3<R> support_triangle(v3<R> A, v3<R> B, v3<R> C, v3<R> n) {
  R dist = dot(A, n);
  v3<R> point = A;
  if (dot(B, n) > dist) {
    dist = dot(B, n);
    point = B;
  }
  if (dot(C, n) > dist) point = C;
  return point;
}

// World support of a placed shape along b (Narrowphase.cu:19-52):
// normalize, rotate into the shape frame, support, rotate back, translate.
template <class R, bool EXT = true>
CDX_HD v3<R> support_world_unit(int type, v3<R> X, v3<R> B, const mat3<R>& m, v3<R> n) {
  v3<R> rn = mulT(m, n);
  return mul(m, support_local<R, EXT>(type, B, rn)) + X;
}
template <class R> CDX_HD v3<R> support_world(int type, v3<R> X, v3<R> B, const mat3<R>& m, v3<R> b) {
  return support_world_unit(type, X, B, m, normalize(b));
}

// normalize(-b) and normalize(b) from one square root and one reciprocal:
// -b has the same squared length, and rounding commutes with negation, so
// normalize(-b) == -normalize(b) bit for bit, except for
// b = 0, which normalises to (1, 0, 0) either way.  The MPR always asks for
// the support of one shape along -b and of the other along b.
template <class R> CDX_HD void normalize_pm(v3<R> b, v3<R>& neg, v3<R>& pos) {
  const R l = length(b);
  if (l == R(0)) {
    neg = pos = mk3<R>(1, 0, 0);
    return;
  }
  pos = b * (R(1) / l);
  neg = -pos;
}

// Approximate FP32 variants (hardware reciprocal / reciprocal square root,
// about 1 ulp) for the FP32 filter, whose verdicts keep a margin of tau far
// above that error.  A zero vector normalises to (1, 0, 0) as above.
__device__ __forceinline__ v3<float> normalize_fast(v3<float> v) {
  const float d = dot(v, v);
  if (d == 0.f) return mk3(1.f, 0.f, 0.f);
  return v * __builtin_amdgcn_rsqf(d);
}
__device__ __forceinline__ float length_fast(v3<float> v) { return __builtin_amdgcn_sqrtf(dot(v, v)); }

__device__ __forceinline__ v3<float> support_local_fast(int type, v3<float> B, v3<float> n) {
  switch (type) {
    case SPHERE: return B.x * n;
    case BOX: return mk3(sgn(n.x) * B.x, sgn(n.y) * B.y, sgn(n.z) * B.z);
    case ELLIPSOID: {
      v3<float> nn = normalize_fast(n);
      return B * B * nn * __builtin_amdgcn_rcpf(length_fast(B * nn));
    }
    case CYLINDER: {
      const float l2 = n.x * n.x + n.z * n.z;
      const float s = sgn(n.y) * B.y;
      if (l2 != 0.f) {
        const float r = B.x * __builtin_amdgcn_rsqf(l2);
        return mk3(n.x * r, s, n.z * r);
      }
      return mk3(0.f, s, 0.f);
    }
    case CAPSULE: {
      v3<float> nn = normalize_fast(n);
      const float s = sgn(n.y) * B.y;
      return mk3(B.x * nn.x, s + B.x * nn.y, B.x * nn.z);
    }
    default: return mk3(0.f, 0.f, 0.f);
  }
}

__device__ __forceinline__ v3<float> support_world_fast_unit(int type, v3<float> X, v3<float> B,
                                                             const mat3<float>& m, v3<float> n) {
  v3<float> rn = mulT(m, n);
  return mul(m, support_local_fast(type, B, rn)) + X;
}
__device__ __forceinline__ v3<float> support_world_fast(int type, v3<float> X, v3<float> B, const mat3<float>& m,
                                                        v3<float> b) {
  return support_world_fast_unit(type, X, B, m, normalize_fast(b));
}

// normalize_fast of -b and of b from one reciprocal square root ((-b) * r
// is -(b * r) exactly), as normalize_pm.
__device__ __forceinline__ void normalize_fast_pm(v3<float> b, v3<float>& neg, v3<float>& pos) {
  const float d = dot(b, b);
  if (d == 0.f) {
    neg = pos = mk3(1.f, 0.f, 0.f);
    return;
  }
  pos = b * __builtin_amdgcn_rsqf(d);
  neg = -pos;
}

// Order-preserving map float -> uint32 used for atomic min / max of bounds.
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
CDX_HD float ord2f(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  union { uint32_t u; float f; } c;
  c.u = b;
  return c.f;
}

}  // namespace cdx
