// gfx950 collision pipeline: AABB generation, uniform-grid broadphase and
// MPR narrowphase, with the host orchestration of one scene.
//
// Reference path (uwsbel/collision-detection):
//   AABB         AABBGenerator.cu:84-130   -> k_aabb
//   bounds       Broadphase.cu:368-381     -> k_aabb (fused) / k_bounds, k_params
//   count        Broadphase.cu:150-166     -> k_count
//   scan #1      Broadphase.cu:403-407     -> scan_excl_u32
//   key store    Broadphase.cu:172-209     -> k_emit_keys
//   sort_by_key  Broadphase.cu:428-432     -> radix_sort_u64 on (key<<32 | body)
//   reduce_by_key + scan #2  :439-480      -> k_count_runs (run boundaries inside k_pairs)
//   pair count / store / sort / unique :484-535
//                                          -> k_pairs (owner-bin dedup), radix_sort_u64
//   MPR store    Narrowphase.cu:316-504    -> k_np_front + class sort + k_np_resolve
//   compaction   Narrowphase.cu:580-593    -> scan + k_np_compact
#include <algorithm>
#include <cstring>

#include "cdx_math.hpp"
#include "cdx_pipeline.hpp"
#include "cdx_prims.hpp"
#include "cdx_util.hpp"

namespace cdx {

namespace {

struct DevStatus {
  uint32_t bmin_enc[3], bmax_enc[3];
  uint32_t bad, key_max;
  uint64_t M, A, P, C, Q, Q2;
};

template <class R> struct Params { R origin[3]; R inv[3]; };

// Compact per-shape header for pair classification: type | active << 8, body.
struct ShapeHdr {
  uint32_t ta;
  uint32_t body;
};

// Per-shape record of the FP32 filter: FP64 centre (the local frame is
// formed by an FP64 subtraction), FP32 rotation and dimensions.
struct alignas(16) ShapeRecF {
  double x, y, z;
  float qw, qx, qy, qz;
  float bx, by, bz;
  uint32_t type;
};

template <class R> struct alignas(16) ShapeRec {
  R x, y, z;          // world centre: quatRotate(A, body_rot) + body_pos
  R qw, qx, qy, qz;   // world rotation: body_rot (x) local rot
  R bx, by, bz;       // dimensions
  uint32_t type, body, active, pad;
};

inline unsigned blocks_for(uint64_t n, unsigned bs) {
  uint64_t b = (n + bs - 1) / bs;
  return (unsigned)(b ? b : 1);
}

__device__ __forceinline__ uint32_t hash_index(uint32_t i, uint32_t j, uint32_t k, int gx, int gy) {
  // Broadphase.cu:140-147, evaluated in unsigned arithmetic (U3 argument).
  return ((k * (uint32_t)gy) * (uint32_t)gx) + (j * (uint32_t)gx) + i;
}

template <class R> __device__ __forceinline__ int cell_of(R x, R inv) { return (int)(x * inv); }

// ------------------------------------------------------------------ AABB --

__global__ void k_init_status(DevStatus* st) {
  if (threadIdx.x == 0) {
    for (int c = 0; c < 3; ++c) { st->bmin_enc[c] = 0xffffffffu; st->bmax_enc[c] = 0u; }
    st->bad = 0; st->key_max = 0;
    st->M = st->A = st->P = st->C = st->Q = st->Q2 = 0;
  }
}

// Per-workgroup bounds partials (reduced by k_params): same-address
// atomics from every workgroup serialise badly across the XCDs.
template <int BS>
__device__ __forceinline__ void block_bounds_commit(uint32_t mn[3], uint32_t mx[3], uint32_t* part) {
  __shared__ uint32_t smn[3][BS / WAVE], smx[3][BS / WAVE];
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uint32_t a = mn[c], b = mx[c];
    for (int o = 32; o > 0; o >>= 1) {
      a = min(a, (uint32_t)__shfl_xor((int)a, o));
      b = max(b, (uint32_t)__shfl_xor((int)b, o));
    }
    if (l == 0) { smn[c][w] = a; smx[c][w] = b; }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int c = threadIdx.x;
    uint32_t a = 0xffffffffu, b = 0u;
    for (int i = 0; i < BS / WAVE; ++i) { a = min(a, smn[c][i]); b = max(b, smx[c][i]); }
    part[6 * (uint64_t)blockIdx.x + c] = a;
    part[6 * (uint64_t)blockIdx.x + 3 + c] = b;
  }
}

// One pass over shapes (SoA) and body poses: world AABB (AABBGenerator.cu:
// 84-116), the narrowphase shape record, and the float scene bounds over all
// 2n corners (Broadphase.cu:368-376, fminf / fmaxf).
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_aabb(uint64_t n, const uint8_t* __restrict__ type,
                                              const uint32_t* __restrict__ body, const R* __restrict__ lpos,
                                              const R* __restrict__ dims, const R* __restrict__ lrot,
                                              const R* __restrict__ bpos, const R* __restrict__ brot,
                                              const uint8_t* __restrict__ active, uint64_t nb, R* __restrict__ aabb,
                                              ShapeRec<R>* __restrict__ rec, ShapeHdr* __restrict__ hdr,
                                              ShapeRecF* __restrict__ frec, uint32_t* __restrict__ bpart,
                                              DevStatus* st) {
  uint64_t i = (uint64_t)blockIdx.x * BS + threadIdx.x;
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  if (i < n) {
    int t = type[i];
    uint32_t id = body[i];
    if (id >= nb) {
      atomicOr(&st->bad, 2u);
    } else if (!(t == SPHERE || t == ELLIPSOID || t == BOX || t == CYLINDER || t == CAPSULE)) {
      atomicOr(&st->bad, 1u);
    } else {
      v3<R> a = mk3(lpos[i], lpos[n + i], lpos[2 * n + i]);
      v3<R> b = mk3(dims[i], dims[n + i], dims[2 * n + i]);
      q4<R> lr; lr.w = lrot[i]; lr.x = lrot[n + i]; lr.y = lrot[2 * n + i]; lr.z = lrot[3 * n + i];
      v3<R> p = mk3(bpos[id], bpos[nb + id], bpos[2 * nb + id]);
      q4<R> rq; rq.w = brot[id]; rq.x = brot[nb + id]; rq.y = brot[2 * nb + id]; rq.z = brot[3 * nb + id];
      v3<R> c = rotate(a, rq) + p;
      q4<R> q1 = qmul(rq, lr);
      v3<R> lo, hi;
      if (t == SPHERE) {
        v3<R> r3 = mk3(b.x, b.x, b.x);
        lo = c - r3;
        hi = c + r3;
      } else {
        // Loose box of the semi-extents for ellipsoid and cylinder, as the
        // reference does; the capsule's is (r, h + r, r).
        v3<R> d = (t == CAPSULE) ? mk3(b.x, b.y + b.x, b.z) : b;
        v3<R> h = mul(absm(rotmat(q1)), d);
        lo = c - h;
        hi = c + h;
      }
      aabb[i] = lo.x; aabb[2 * n + i] = lo.y; aabb[4 * n + i] = lo.z;
      aabb[n + i] = hi.x; aabb[3 * n + i] = hi.y; aabb[5 * n + i] = hi.z;
      ShapeRec<R> r;
      r.x = c.x; r.y = c.y; r.z = c.z;
      r.qw = q1.w; r.qx = q1.x; r.qy = q1.y; r.qz = q1.z;
      r.bx = b.x; r.by = b.y; r.bz = b.z;
      r.type = (uint32_t)t; r.body = id; r.active = active[id]; r.pad = 0;
      rec[i] = r;
      ShapeHdr h;
      h.ta = (uint32_t)t | ((uint32_t)r.active << 8);
      h.body = id;
      hdr[i] = h;
      ShapeRecF fr;
      fr.x = (double)c.x; fr.y = (double)c.y; fr.z = (double)c.z;
      fr.qw = (float)q1.w; fr.qx = (float)q1.x; fr.qy = (float)q1.y; fr.qz = (float)q1.z;
      fr.bx = (float)b.x; fr.by = (float)b.y; fr.bz = (float)b.z;
      fr.type = (uint32_t)t;
      frec[i] = fr;
      float f;
      f = (float)lo.x; mn[0] = f2ord(f); mx[0] = f2ord(f);
      f = (float)lo.y; mn[1] = f2ord(f); mx[1] = f2ord(f);
      f = (float)lo.z; mn[2] = f2ord(f); mx[2] = f2ord(f);
      uint32_t u;
      u = f2ord((float)hi.x); mn[0] = min(mn[0], u); mx[0] = max(mx[0], u);
      u = f2ord((float)hi.y); mn[1] = min(mn[1], u); mx[1] = max(mx[1], u);
      u = f2ord((float)hi.z); mn[2] = min(mn[2], u); mx[2] = max(mx[2], u);
    }
  }
  block_bounds_commit<BS>(mn, mx, bpart);
}

// Bounds of AABBs supplied by the caller (SoA: x[2n], y[2n], z[2n]).
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_bounds(uint64_t n2, const R* __restrict__ aabb, uint32_t* __restrict__ bpart) {
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  for (uint64_t i = (uint64_t)blockIdx.x * BS + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * BS) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint32_t u = f2ord((float)aabb[c * n2 + i]);
      mn[c] = min(mn[c], u);
      mx[c] = max(mx[c], u);
    }
  }
  block_bounds_commit<BS>(mn, mx, bpart);
}

// Scene bounds from the workgroup partials, then origin and inverse bin
// size (Broadphase.cu:377-381).  One workgroup of 256 threads.
template <class R>
__global__ __launch_bounds__(256) void k_params(const uint32_t* __restrict__ bpart, uint64_t npart, DevStatus* st,
                                                int gx, int gy, int gz, Params<R>* prm) {
  __shared__ uint32_t red[6][256];
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (uint64_t b = threadIdx.x; b < npart; b += 256) {
    for (int c = 0; c < 3; ++c) {
      v[c] = min(v[c], bpart[6 * b + c]);
      v[3 + c] = max(v[3 + c], bpart[6 * b + 3 + c]);
    }
  }
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = v[c];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int t = 1; t < 256; ++t)
    for (int c = 0; c < 3; ++c) {
      v[c] = min(v[c], red[c][t]);
      v[3 + c] = max(v[3 + c], red[3 + c][t]);
    }
  int g[3] = {gx, gy, gz};
  for (int c = 0; c < 3; ++c) {
    st->bmin_enc[c] = v[c];
    st->bmax_enc[c] = v[3 + c];
    R mn = (R)ord2f(v[c]);
    R mx = (R)ord2f(v[3 + c]);
    R origin = cabs(mn);
    R ext = cabs(mx + cabs(mn));
    prm->origin[c] = origin;
    prm->inv[c] = (R)g[c] / ext;
  }
}

// ------------------------------------------------------------ broadphase --

template <class R> struct Cells { int x0, y0, z0, x1, y1, z1; };

template <class R>
__device__ __forceinline__ Cells<R> cells_of(const R* __restrict__ aabb, uint64_t n, uint64_t i, const Params<R>& p) {
  // Shifted coordinates (Broadphase.cu:382-386) then HashMin (:125-133).
  Cells<R> c;
  c.x0 = cell_of(aabb[i] + p.origin[0], p.inv[0]);
  c.y0 = cell_of(aabb[2 * n + i] + p.origin[1], p.inv[1]);
  c.z0 = cell_of(aabb[4 * n + i] + p.origin[2], p.inv[2]);
  c.x1 = cell_of(aabb[n + i] + p.origin[0], p.inv[0]);
  c.y1 = cell_of(aabb[3 * n + i] + p.origin[1], p.inv[1]);
  c.z1 = cell_of(aabb[5 * n + i] + p.origin[2], p.inv[2]);
  return c;
}

// Finite AABBs always map into cells 0..grid on every axis (the shift makes
// coordinates non-negative up to rounding, and truncation sends tiny
// negatives to 0).  Anything else comes from non-finite input or a
// degenerate scene extent; such AABBs are reported instead of binned.
template <class R> __device__ __forceinline__ bool cells_ok(const Cells<R>& c, int gx, int gy, int gz) {
  return c.x0 >= 0 && c.y0 >= 0 && c.z0 >= 0 && c.x0 <= c.x1 && c.y0 <= c.y1 && c.z0 <= c.z1 && c.x1 <= gx &&
         c.y1 <= gy && c.z1 <= gz;
}

// Per-body broadphase record: shifted AABB and its minimum cell, gathered as
// one aligned block by the pair search.
template <class R> struct alignas(16) BpRec {
  R lo[3], hi[3];
  int g[3];
  int pad;
};

// Bins spanned per AABB (Broadphase.cu:150-159), the per-body record, and
// the largest key (the max corner's, when the key space does not wrap) to
// bound the number of radix passes.
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_count(uint64_t n, const R* __restrict__ aabb, const Params<R>* __restrict__ prm,
                                               int gx, int gy, int gz, uint32_t* __restrict__ counts,
                                               BpRec<R>* __restrict__ brec, uint32_t* __restrict__ kpart,
                                               DevStatus* st) {
  __shared__ uint32_t skm[BS / WAVE];
  uint64_t i = (uint64_t)blockIdx.x * BS + threadIdx.x;
  uint32_t km = 0;
  if (i < n) {
    Params<R> p = *prm;
    Cells<R> c = cells_of(aabb, n, i, p);
    BpRec<R> r;
    r.lo[0] = aabb[i] + p.origin[0];
    r.lo[1] = aabb[2 * n + i] + p.origin[1];
    r.lo[2] = aabb[4 * n + i] + p.origin[2];
    r.hi[0] = aabb[n + i] + p.origin[0];
    r.hi[1] = aabb[3 * n + i] + p.origin[1];
    r.hi[2] = aabb[5 * n + i] + p.origin[2];
    r.g[0] = c.x0; r.g[1] = c.y0; r.g[2] = c.z0; r.pad = 0;
    brec[i] = r;
    if (cells_ok(c, gx, gy, gz)) {
      counts[i] = (uint32_t)((c.x1 - c.x0 + 1) * (c.y1 - c.y0 + 1) * (c.z1 - c.z0 + 1));
      km = hash_index((uint32_t)c.x1, (uint32_t)c.y1, (uint32_t)c.z1, gx, gy);
    } else {
      counts[i] = 0;
      atomicOr(&st->bad, 4u);
    }
  }
  for (int o = 32; o > 0; o >>= 1) km = max(km, (uint32_t)__shfl_xor((int)km, o));
  if (threadIdx.x % WAVE == 0) skm[threadIdx.x / WAVE] = km;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int w = 0; w < BS / WAVE; ++w) m = max(m, skm[w]);
    kpart[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(256) void k_reduce_keymax(const uint32_t* __restrict__ kpart, uint64_t npart,
                                                       DevStatus* st) {
  __shared__ uint32_t red[256];
  uint32_t m = 0;
  for (uint64_t b = threadIdx.x; b < npart; b += 256) m = max(m, kpart[b]);
  red[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < 256; ++t) m = max(m, red[t]);
    st->key_max = m;
  }
}

// (bin, body) entries in the reference's store order (Broadphase.cu:
// 172-200), packed as key << 32 | body so one 64-bit sort orders by bin and
// keeps bodies ascending inside a bin.
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_emit_keys(uint64_t n, const R* __restrict__ aabb,
                                                   const Params<R>* __restrict__ prm, int gx, int gy, int gz,
                                                   const uint32_t* __restrict__ offs, uint64_t* __restrict__ keys) {
  uint64_t i = (uint64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  Params<R> p = *prm;
  Cells<R> c = cells_of(aabb, n, i, p);
  if (!cells_ok(c, gx, gy, gz)) return;
  uint64_t o = offs[i];
  for (uint32_t x = (uint32_t)c.x0; x <= (uint32_t)c.x1; ++x)
    for (uint32_t y = (uint32_t)c.y0; y <= (uint32_t)c.y1; ++y)
      for (uint32_t z = (uint32_t)c.z0; z <= (uint32_t)c.z1; ++z)
        keys[o++] = ((uint64_t)hash_index(x, y, z, gx, gy) << 32) | (uint64_t)i;
}

// Active bins (reduce_by_key, Broadphase.cu:439-443): the number of runs of
// equal keys, counted one atomic per wavefront.  Their keys and occupancy
// are derived from the sorted entries on readback.
__global__ __launch_bounds__(256) void k_count_runs(uint64_t m, const uint64_t* __restrict__ keys, DevStatus* st) {
  __shared__ uint32_t red[256 / WAVE];
  uint32_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256)
    c += (i == 0 || (keys[i] >> 32) != (keys[i - 1] >> 32)) ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 256 / WAVE; ++w) t += red[w];
    atomicAdd((unsigned long long*)&st->A, (unsigned long long)t);
  }
}

// Pair search with de-duplication by owning bin.  A pair (a, b) of
// overlapping AABBs (closed test on shifted coordinates, Broadphase.cu:237 /
// :296) is reported only from the bin whose key is the hash of the cell
// max(gmin_a, gmin_b): that cell lies in both AABBs' cell ranges, so both
// bodies appear in its run of sorted entries, and every other run skips the
// pair.  Repeated entries of one body in a run (hash aliasing at index ==
// grid size, Broadphase.cu:140-147) are skipped, so each pair is emitted once
// and the reference's sort + unique (:522-535) reduces to a sort.
//
// One thread per sorted entry i: it pairs body(i) with the later entries of
// the same run (bodies ascend inside a run, so body(i) < body(k)).  Work
// per thread is the distance to the end of its run, which keeps waves far
// better balanced than a thread per bin.
constexpr int PAIR_BS = 256;   // sorted entries per workgroup
constexpr int PAIR_AHEAD = 128; // entries staged past the workgroup's end

template <class R, bool EMIT>
__global__ __launch_bounds__(PAIR_BS) void k_pairs(uint64_t m, const uint64_t* __restrict__ keys,
                                                   const BpRec<R>* __restrict__ brec, int gx, int gy,
                                                   uint32_t* __restrict__ pcount, const uint32_t* __restrict__ poffs,
                                                   uint64_t* __restrict__ pairs) {
  // The workgroup's entries and the first PAIR_AHEAD after them are staged
  // in LDS with their body records, so the candidate loop reads LDS; only
  // bins reaching past the staged window fall back to global loads.
  constexpr int W = PAIR_BS + PAIR_AHEAD;
  __shared__ BpRec<R> srec[W];
  __shared__ uint64_t skey[W];
  const uint64_t b0 = (uint64_t)blockIdx.x * PAIR_BS;
  for (int t = threadIdx.x; t < W; t += PAIR_BS) {
    uint64_t e = b0 + t;
    if (e < m) {
      uint64_t kk = keys[e];
      skey[t] = kk;
      srec[t] = brec[(uint32_t)kk];
    }
  }
  __syncthreads();
  const uint64_t i = b0 + threadIdx.x;
  if (i >= m) return;
  const uint64_t ki = skey[threadIdx.x];
  uint32_t cnt = 0;
  const uint64_t kprev = i == 0 ? ~ki : (threadIdx.x > 0 ? skey[threadIdx.x - 1] : keys[i - 1]);
  if (kprev != ki) {
    const uint32_t K = (uint32_t)(ki >> 32), a = (uint32_t)ki;
    const BpRec<R> ra = srec[threadIdx.x];
    uint64_t o = EMIT ? poffs[i] : 0;
    uint32_t prev = a;
    for (uint64_t k = i + 1; k < m; ++k) {
      const uint64_t t = k - b0;
      const bool staged = t < (uint64_t)W;
      const uint64_t kk = staged ? skey[t] : keys[k];
      if ((uint32_t)(kk >> 32) != K) break;
      uint32_t b = (uint32_t)kk;
      if (b == prev) continue;
      prev = b;
      const BpRec<R> rb = staged ? srec[t] : brec[b];
      bool hit = (ra.lo[0] <= rb.hi[0] && rb.lo[0] <= ra.hi[0]) && (ra.lo[1] <= rb.hi[1] && rb.lo[1] <= ra.hi[1]) &&
                 (ra.lo[2] <= rb.hi[2] && rb.lo[2] <= ra.hi[2]);
      if (!hit) continue;
      uint32_t ox = (uint32_t)max(ra.g[0], rb.g[0]);
      uint32_t oy = (uint32_t)max(ra.g[1], rb.g[1]);
      uint32_t oz = (uint32_t)max(ra.g[2], rb.g[2]);
      if (hash_index(ox, oy, oz, gx, gy) != K) continue;
      if (EMIT) pairs[o++] = ((uint64_t)a << 32) | (uint64_t)b;
      else ++cnt;
    }
  }
  if (!EMIT) pcount[i] = cnt;
}

// ----------------------------------------------------------- narrowphase --

template <class R> struct Placed {
  int type;
  v3<R> X, B;
  q4<R> q;
};

// The rotation matrix is rebuilt from the quaternion at every support call:
// the same rounded values each time, and far fewer live registers than a
// cached 3x3 per shape (the FP64 MPR is occupancy-bound).
template <class R> __device__ __forceinline__ v3<R> sup(const Placed<R>& s, v3<R> b) {
  return support_world(s.type, s.X, s.B, rotmat(s.q), b);
}

template <class R> __device__ __forceinline__ Placed<R> placed(const ShapeRec<R>& r) {
  Placed<R> s;
  s.type = (int)r.type;
  s.X = mk3(r.x, r.y, r.z);
  s.B = mk3(r.bx, r.by, r.bz);
  s.q.w = r.qw; s.q.x = r.qx; s.q.y = r.qy; s.q.z = r.qz;
  return s;
}

template <class R> __device__ __forceinline__ void swap3(v3<R>& a, v3<R>& b) { v3<R> t = a; a = b; b = t; }

// CollideAndFindPoint (Narrowphase.cu:122-314), same evaluation order.
template <class R>
__device__ bool mpr_reference(const Placed<R>& A, const Placed<R>& B, v3<R>& rnormal, v3<R>& point, R& depth) {
  v3<R> v01, v02, v0, n, v11, v12, v1, v21, v22, v2;
  v01 = mk3<R>(0, 0, 0) + A.X;
  v02 = mk3<R>(0, 0, 0) + B.X;
  v0 = v02 - v01;
  if (is_zero(v0)) v0 = mk3<R>(1, 0, 0);
  n = -v0;
  v11 = sup(A, -n);
  v12 = sup(B, n);
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
  v21 = sup(A, -n);
  v22 = sup(B, n);
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
    v31 = sup(A, -n);
    v32 = sup(B, n);
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
      v3<R> v41 = sup(A, -n);
      v3<R> v42 = sup(B, n);
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
__device__ __forceinline__ bool contact_spheres(const ShapeRec<R>& ra, const ShapeRec<R>& rb, R envelope,
                                                Contact<R>& out) {
  v3<R> AX = mk3(ra.x, ra.y, ra.z), BX = mk3(rb.x, rb.y, rb.z);
  v3<R> relpos = BX - AX;
  R d2 = dot(relpos, relpos);
  R cd = ra.bx + rb.bx;
  if (!(d2 <= cd * cd)) return false;
  v3<R> N = normalize(relpos);
  v3<R> p1 = AX + N * ra.bx;
  v3<R> p2 = BX - N * rb.bx;
  p1 = p1 - N * envelope;
  p2 = p2 + N * envelope;
  out.N = N;
  out.p1 = p1;
  out.p2 = p2;
  out.depth = dot(N, p2 - p1);
  return true;
}

template <class R>
__device__ bool contact_native(const ShapeRec<R>& ra, const ShapeRec<R>& rb, R envelope, Contact<R>& out) {
  if (!ra.active && !rb.active) return false;
  if (ra.body == rb.body) return false;
  v3<R> N = mk3<R>(1, 0, 0), p1 = mk3<R>(0, 0, 0), p2 = mk3<R>(0, 0, 0);
  if (ra.type == SPHERE && rb.type == SPHERE) {
    return contact_spheres(ra, rb, envelope, out);
  } else {
    Placed<R> A = placed(ra), B = placed(rb);
    v3<R> p0 = mk3<R>(0, 0, 0);
    R depth = 0;
    if (!mpr_reference(A, B, N, p0, depth)) return false;
    p1 = dot((sup(A, -N) - p0), N) * N + p0;
    p2 = dot((sup(B, N) - p0), N) * N + p0;
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
template <class R> struct PairOut {
  R* v;          // 10 * P: N(3) p1(3) p2(3) depth
  int32_t* ids;  // 2 * P
  uint32_t* hit; // P
  uint64_t P;
};

template <class R>
__device__ __forceinline__ void store_pair(const PairOut<R>& o, uint64_t P, uint64_t i, const Contact<R>& c,
                                           uint32_t ida, uint32_t idb) {
  R* v = o.v;
  v[i] = c.N.x; v[P + i] = c.N.y; v[2 * P + i] = c.N.z;
  v[3 * P + i] = c.p1.x; v[4 * P + i] = c.p1.y; v[5 * P + i] = c.p1.z;
  v[6 * P + i] = c.p2.x; v[7 * P + i] = c.p2.y; v[8 * P + i] = c.p2.z;
  v[9 * P + i] = c.depth;
  o.ids[2 * i] = (int32_t)ida;
  o.ids[2 * i + 1] = (int32_t)idb;
  o.hit[i] = 1u;
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
// cap reached.
__device__ int mpr_guarded(const Placed<float>& A, const Placed<float>& B, float tau, bool verdict_only,
                           v3<float>& rnormal, v3<float>& point) {
  typedef float R;
  v3<R> v01, v02, v0, n, v11, v12, v1, v21, v22, v2;
  v01 = A.X;
  v02 = B.X;
  v0 = v02 - v01;
  if (length(v0) < 1e-4f * tau) return MPR_UNSURE;
  n = -v0;
  v11 = sup(A, -n);
  v12 = sup(B, n);
  v1 = v12 - v11;
  R dn = dot(v1, n);
  if (dn <= R(0)) return (-dn < tau * length(n)) ? MPR_UNSURE : MPR_MISS;
  n = cross(v1, v0);
  {
    R nn = dot(n, n), s = 1e-6f * dot(v1, v1) * dot(v0, v0);
    if (nn <= s) return MPR_UNSURE;
  }
  v21 = sup(A, -n);
  v22 = sup(B, n);
  v2 = v22 - v21;
  dn = dot(v2, n);
  if (dn <= R(0)) return (-dn < tau * length(n)) ? MPR_UNSURE : MPR_MISS;
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
    if (phase1 > max_iterations) return MPR_UNSURE;
    phase1++;
    v31 = sup(A, -n);
    v32 = sup(B, n);
    v3_ = v32 - v31;
    dn = dot(v3_, n);
    if (dn <= R(0)) return (-dn < tau * length(n)) ? MPR_UNSURE : MPR_MISS;
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
      R side = dot(n, v1);
      if (side >= R(0) && !hit) hit = true;
      // The origin is inside the portal by more than tau: a certain hit.
      if (verdict_only && side > tau) return MPR_HIT;
      v3<R> v41 = sup(A, -n);
      v3<R> v42 = sup(B, n);
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
        rnormal = normalize(n);
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

__device__ __forceinline__ Placed<float> placed_local(const ShapeRecF& r, const ShapeRecF& origin) {
  Placed<float> s;
  s.type = (int)r.type;
  s.X = mk3((float)(r.x - origin.x), (float)(r.y - origin.y), (float)(r.z - origin.z));
  s.B = mk3(r.bx, r.by, r.bz);
  s.q.w = r.qw; s.q.x = r.qx; s.q.y = r.qy; s.q.z = r.qz;
  return s;
}

__device__ __forceinline__ uint32_t type_slot(uint32_t t) { return t == CAPSULE ? 4u : t; }

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
// MPR are appended to a queue as (shape-pair class << 32 | pair index), one
// atomic per wavefront; a one-digit radix pass then groups the queue by
// class so every wavefront of the MPR kernel runs one pair of support maps.
// MIXED / FAST first run an FP32 MPR in the frame of shape A: clear misses
// end here, FAST keeps clear FP32 hits, everything else is queued for the
// reference-order MPR.
template <class R>
__global__ __launch_bounds__(128) void k_np_front(uint64_t P, const uint64_t* __restrict__ pairs,
                                                  const ShapeRec<R>* __restrict__ rec, const ShapeHdr* __restrict__ hdr,
                                                  R envelope, PairOut<R> o, uint32_t* __restrict__ qsel,
                                                  uint8_t* __restrict__ qcls) {
  uint64_t i = (uint64_t)blockIdx.x * 128 + threadIdx.x;
  if (i >= P) return;
  uint64_t pr = pairs[i];
  const uint32_t ia = (uint32_t)(pr >> 32), ib = (uint32_t)(pr & 0xffffffffu);
  const ShapeHdr ha = hdr[ia], hb = hdr[ib];
  const uint32_t ta = ha.ta & 255u, tb = hb.ta & 255u;
  o.hit[i] = 0u;
  bool enq = false;
  if (((ha.ta | hb.ta) >> 8) && ha.body != hb.body) {
    if (ta == SPHERE && tb == SPHERE) {
      Contact<R> c;
      ShapeRec<R> ra = rec[ia], rb = rec[ib];
      if (contact_spheres(ra, rb, envelope, c)) store_pair(o, P, i, c, ra.body, rb.body);
    } else {
      enq = true;
    }
  }
  qsel[i] = enq ? 1u : 0u;
  qcls[i] = (uint8_t)(type_slot(ta) * 5u + type_slot(tb));
}

// FP32 MPR over the class-sorted queue (MIXED / FAST), in the frame of
// shape A: clear misses end here; FAST keeps clear FP32 hits; everything
// else is selected for the reference-order MPR.
template <class R, bool KEEP>
__global__ __launch_bounds__(128) void k_np_filter(uint64_t q, const uint64_t* __restrict__ queue,
                                                   const uint64_t* __restrict__ pairs,
                                                   const ShapeHdr* __restrict__ hdr,
                                                   const ShapeRecF* __restrict__ frec, R envelope, float tau_rel,
                                                   PairOut<R> o, uint32_t* __restrict__ sel) {
  uint64_t j = (uint64_t)blockIdx.x * 128 + threadIdx.x;
  if (j >= q) return;
  const uint64_t P = o.P;
  const uint32_t i = (uint32_t)queue[j];
  const uint64_t pr = pairs[i];
  const uint32_t ia = (uint32_t)(pr >> 32), ib = (uint32_t)(pr & 0xffffffffu);
  const ShapeRecF fa = frec[ia], fb = frec[ib];
  Placed<float> A = placed_local(fa, fa), B = placed_local(fb, fa);
  float tau = tau_rel * (bound_radius(A.type, A.B) + bound_radius(B.type, B.B));
  v3<float> N, p0;
  int r = mpr_guarded(A, B, tau, !KEEP, N, p0);
  if (KEEP && r == MPR_HIT) {
    v3<float> p1 = dot((sup(A, -N) - p0), N) * N + p0;
    v3<float> p2 = dot((sup(B, N) - p0), N) * N + p0;
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
      store_pair(o, P, i, c, hdr[ia].body, hdr[ib].body);
    }
  }
  // MIXED: FP32 settles clear misses only; hits take the reference path,
  // whose capped iteration sequence FP32 cannot reproduce.
  sel[j] = (r == MPR_UNSURE || (r == MPR_HIT && !KEEP)) ? 1u : 0u;
}

// Stable compaction of the selected queue entries (keeps the class order).
__global__ void k_queue_select(uint64_t q, const uint64_t* __restrict__ queue, const uint32_t* __restrict__ sel,
                               const uint32_t* __restrict__ pos, uint64_t* __restrict__ out) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= q || !sel[j]) return;
  out[pos[j]] = queue[j];
}

// Queue of the pairs that need the MPR, in pair order: slot = exclusive scan
// of the selection flags (no same-address atomics).
__global__ void k_queue_compact(uint64_t P, const uint32_t* __restrict__ qsel, const uint32_t* __restrict__ qpos,
                                const uint8_t* __restrict__ qcls, uint64_t* __restrict__ queue) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P || !qsel[i]) return;
  queue[qpos[i]] = ((uint64_t)qcls[i] << 32) | i;
}

// MPR on the reference path for the queued pairs, in compute type CT
// (the scene's real type for NATIVE, FP64 for MIXED / FAST).
template <class R, class CT>
__global__ __launch_bounds__(128) void k_np_resolve(uint64_t q, const uint64_t* __restrict__ pairs,
                                                    const ShapeRec<R>* __restrict__ rec, R envelope, PairOut<R> o,
                                                    const uint64_t* __restrict__ queue) {
  uint64_t j = (uint64_t)blockIdx.x * 128 + threadIdx.x;
  if (j >= q) return;
  uint64_t P = o.P;
  uint32_t i = (uint32_t)queue[j];
  uint64_t pr = pairs[i];
  ShapeRec<CT> ra = as_compute<CT, R>(rec[(uint32_t)(pr >> 32)]);
  ShapeRec<CT> rb = as_compute<CT, R>(rec[(uint32_t)(pr & 0xffffffffu)]);
  Contact<CT> c;
  if (contact_native(ra, rb, (CT)envelope, c)) {
    Contact<R> cr;
    cr.N = mk3((R)c.N.x, (R)c.N.y, (R)c.N.z);
    cr.p1 = mk3((R)c.p1.x, (R)c.p1.y, (R)c.p1.z);
    cr.p2 = mk3((R)c.p2.x, (R)c.p2.y, (R)c.p2.z);
    cr.depth = (R)c.depth;
    store_pair(o, P, i, cr, ra.body, rb.body);
  }
}

// Stable partition of hits to the front (Narrowphase.cu:580-593): contacts
// in ascending pair order, and the pair list permuted hits first.
template <class R>
__global__ void k_np_compact(uint64_t P, PairOut<R> o, const uint32_t* __restrict__ pos,
                             const uint64_t* __restrict__ pairs, const DevStatus* __restrict__ st, R* __restrict__ norm,
                             R* __restrict__ cpta, R* __restrict__ cptb, R* __restrict__ dpth, int32_t* __restrict__ bids,
                             uint64_t* __restrict__ perm) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  uint64_t C = st->C;
  uint32_t k = pos[i];
  if (o.hit[i]) {
    const R* v = o.v;
    norm[3 * k] = v[i]; norm[3 * k + 1] = v[P + i]; norm[3 * k + 2] = v[2 * P + i];
    cpta[3 * k] = v[3 * P + i]; cpta[3 * k + 1] = v[4 * P + i]; cpta[3 * k + 2] = v[5 * P + i];
    cptb[3 * k] = v[6 * P + i]; cptb[3 * k + 1] = v[7 * P + i]; cptb[3 * k + 2] = v[8 * P + i];
    dpth[k] = v[9 * P + i];
    bids[2 * k] = o.ids[2 * i];
    bids[2 * k + 1] = o.ids[2 * i + 1];
    perm[k] = pairs[i];
  } else {
    perm[C + (i - k)] = pairs[i];
  }
}

// ------------------------------------------------------------- utilities --

__global__ void k_zero_np(DevStatus* st) {
  if (threadIdx.x == 0) { st->Q = 0; st->C = 0; }
}

template <class R>
__global__ void k_deinterleave(uint64_t n, int k, const R* __restrict__ aos, R* __restrict__ soa) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < k; ++c) soa[c * n + i] = aos[i * k + c];
}

template <class R>
__global__ void k_scatter_bodies(uint64_t k, uint64_t nb, const uint32_t* __restrict__ idx, const R* __restrict__ pos,
                                 const R* __restrict__ rot, R* __restrict__ bpos, R* __restrict__ brot) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  uint32_t b = idx[j];
  if (b >= nb) return;
  for (int c = 0; c < 3; ++c) bpos[c * nb + b] = pos[3 * j + c];
  for (int c = 0; c < 4; ++c) brot[c * nb + b] = rot[4 * j + c];
}

template <class R>
__global__ void k_aabb_aos(uint64_t n, const R* __restrict__ aabb, const Params<R>* __restrict__ prm, int shifted,
                           R* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  R ox = 0, oy = 0, oz = 0;
  R x = aabb[i], y = aabb[2 * n + i], z = aabb[4 * n + i];
  if (shifted) { x = x + prm->origin[0]; y = y + prm->origin[1]; z = z + prm->origin[2]; }
  (void)ox; (void)oy; (void)oz;
  out[3 * i] = x; out[3 * i + 1] = y; out[3 * i + 2] = z;
}

// ------------------------------------------------------------ host side --

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void reserve(size_t b) {
    if (b <= bytes) return;
    if (p) {
      // Growing: kernels queued on any stream may still read the old block.
      CDX_HIP_CHECK(hipDeviceSynchronize());
      CDX_HIP_CHECK(hipFree(p));
      p = nullptr;
      bytes = 0;
    }
    size_t nb = b < 256 ? 256 : b;
    CDX_HIP_CHECK(hipMalloc(&p, nb));
    bytes = nb;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

template <class R> class Scene final : public SceneBase {
 public:
  explicit Scene(int device) : dev_(device) {
    CDX_HIP_CHECK(hipSetDevice(dev_));
    CDX_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    status_.reserve(sizeof(DevStatus));
    params_.reserve(sizeof(Params<R>));
    CDX_HIP_CHECK(hipHostMalloc(&hstat_, sizeof(DevStatus), hipHostMallocDefault));
    for (int i = 0; i < 2 * ST_COUNT; ++i) CDX_HIP_CHECK(hipEventCreate(&ev_[i]));
    std::memset(&counts_, 0, sizeof(counts_));
  }
  ~Scene() override {
    (void)hipSetDevice(dev_);
    (void)hipStreamSynchronize(s_);
    for (int i = 0; i < 2 * ST_COUNT; ++i) (void)hipEventDestroy(ev_[i]);
    (void)hipHostFree(hstat_);
    (void)hipStreamDestroy(s_);
  }

  int real_kind() const override { return sizeof(R) == 8 ? REAL_F64 : REAL_F32; }
  hipStream_t stream() const override { return s_; }

  void set_shapes(uint64_t n, const int32_t* T, const void* A, const void* B, const void* Rq,
                  const uint32_t* ID) override {
    CDX_REQUIRE(n < (1ull << 31), ERR_INVALID, "too many shapes (limit 2^31 - 1)");
    for (uint64_t i = 0; i < n; ++i) {
      int t = T[i];
      CDX_REQUIRE(t == SPHERE || t == ELLIPSOID || t == BOX || t == CYLINDER || t == CAPSULE, ERR_UNSUPPORTED_SHAPE,
                  "shape " + std::to_string(i) + " has type " + std::to_string(t) +
                      ": only SPHERE, ELLIPSOID, BOX, CYLINDER and CAPSULE are supported");
    }
    sync_internal();
    n_ = n;
    std::vector<uint8_t> t8(n);
    for (uint64_t i = 0; i < n; ++i) t8[i] = (uint8_t)T[i];
    type_.reserve(n);
    body_.reserve(n * 4);
    lpos_.reserve(n * 3 * sizeof(R));
    dims_.reserve(n * 3 * sizeof(R));
    lrot_.reserve(n * 4 * sizeof(R));
    if (n) {
      CDX_HIP_CHECK(hipMemcpy(type_.p, t8.data(), n, hipMemcpyHostToDevice));
      CDX_HIP_CHECK(hipMemcpy(body_.p, ID, n * 4, hipMemcpyHostToDevice));
      upload_soa(A, n, 3, lpos_.as<R>());
      upload_soa(B, n, 3, dims_.as<R>());
      upload_soa(Rq, n, 4, lrot_.as<R>());
    }
    have_aabb_ = false; shapes_set_ = true; records_valid_ = false;
  }

  void set_bodies(uint64_t nb, const void* pos, const void* rot, const uint8_t* active) override {
    sync_internal();
    nb_ = nb;
    bpos_.reserve(nb * 3 * sizeof(R));
    brot_.reserve(nb * 4 * sizeof(R));
    active_.reserve(nb);
    if (nb) {
      upload_soa(pos, nb, 3, bpos_.as<R>());
      upload_soa(rot, nb, 4, brot_.as<R>());
      if (active) CDX_HIP_CHECK(hipMemcpy(active_.p, active, nb, hipMemcpyHostToDevice));
      else CDX_HIP_CHECK(hipMemset(active_.p, 1, nb));
    }
    have_aabb_ = false; bodies_set_ = true; records_valid_ = false;
  }

  void update_bodies(uint64_t k, const uint32_t* idx, const void* pos, const void* rot) override {
    if (k == 0) return;
    const uint64_t idx_bytes = (k * 4 + 15) & ~15ull;
    upd_.reserve(idx_bytes + k * 7 * sizeof(R));
    uint32_t* didx = upd_.as<uint32_t>();
    R* dpos = reinterpret_cast<R*>(upd_.as<char>() + idx_bytes);
    R* drot = dpos + 3 * k;
    CDX_HIP_CHECK(hipMemcpyAsync(didx, idx, k * 4, hipMemcpyHostToDevice, s_));
    CDX_HIP_CHECK(hipMemcpyAsync(dpos, pos, k * 3 * sizeof(R), hipMemcpyHostToDevice, s_));
    CDX_HIP_CHECK(hipMemcpyAsync(drot, rot, k * 4 * sizeof(R), hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL(k_scatter_bodies<R>, dim3(blocks_for(k, 256)), dim3(256), 0, s_, k, nb_, (const uint32_t*)didx,
                       (const R*)dpos, (const R*)drot, bpos_.as<R>(), brot_.as<R>());
    have_aabb_ = false; records_valid_ = false;
  }

  void set_aabbs(uint64_t n, const void* aabb) override {
    sync_internal();
    CDX_REQUIRE(n < (1ull << 31), ERR_INVALID, "too many AABBs (limit 2^31 - 1)");
    n_ = n;
    aabb_.reserve(n * 6 * sizeof(R));
    if (n) upload_soa(aabb, 2 * n, 3, aabb_.as<R>());
    have_aabb_ = true;
    bounds_fused_ = false;
  }

  void set_pairs(uint64_t p, const int64_t* pairs) override {
    sync_internal();
    P_ = p;
    pairs_.reserve(p * 8);
    if (p) CDX_HIP_CHECK(hipMemcpy(pairs_.p, pairs, p * 8, hipMemcpyHostToDevice));
    counts_.p = p;
    // Shape indices must address uploaded shapes; validate on the host.
    for (uint64_t i = 0; i < p; ++i) {
      uint64_t a = (uint64_t)pairs[i] >> 32, b = (uint64_t)pairs[i] & 0xffffffffu;
      CDX_REQUIRE(a < n_ && b < n_, ERR_INVALID, "pair " + std::to_string(i) + " addresses a shape out of range");
    }
  }

  void set_grid(int gx, int gy, int gz) override {
    CDX_REQUIRE(gx > 0 && gy > 0 && gz > 0, ERR_INVALID, "bins per axis must be positive");
    g_[0] = gx; g_[1] = gy; g_[2] = gz;
  }
  void set_envelope(double e) override { envelope_ = (R)e; }
  void set_np_mode(int mode, double tau) override {
    CDX_REQUIRE(mode == NP_NATIVE || mode == NP_MIXED || mode == NP_FAST, ERR_INVALID, "unknown narrowphase mode");
    mode_ = mode;
    tau_ = (float)tau;
  }
  void set_profiling(bool on) override { prof_ = on; }

  void run_aabb() override {
    CDX_REQUIRE(nb_ > 0 || n_ == 0, ERR_INVALID, "no bodies uploaded");
    reset_stage_marks();
    hipLaunchKernelGGL(k_init_status, dim3(1), dim3(64), 0, s_, status_.as<DevStatus>());
    aabb_.reserve(n_ * 6 * sizeof(R));
    rec_.reserve(n_ * sizeof(ShapeRec<R>));
    hdr_.reserve(n_ * sizeof(ShapeHdr));
    frec_.reserve(n_ * sizeof(ShapeRecF));
    bpart_.reserve((uint64_t)blocks_for(n_, 256) * 6 * 4);
    npart_ = blocks_for(n_, 256);
    mark(ST_AABB, true);
    if (n_)
      hipLaunchKernelGGL((k_aabb<R, 256>), dim3(blocks_for(n_, 256)), dim3(256), 0, s_, n_, type_.as<uint8_t>(),
                         body_.as<uint32_t>(), lpos_.as<R>(), dims_.as<R>(), lrot_.as<R>(), bpos_.as<R>(),
                         brot_.as<R>(), active_.as<uint8_t>(), nb_, aabb_.as<R>(), rec_.as<ShapeRec<R>>(),
                         hdr_.as<ShapeHdr>(), frec_.as<ShapeRecF>(), bpart_.as<uint32_t>(), status_.as<DevStatus>());
    mark(ST_AABB, false);
    have_aabb_ = true;
    bounds_fused_ = true;
    records_valid_ = true;
  }

  void run_broadphase() override {
    CDX_REQUIRE(have_aabb_, ERR_NOT_READY, "broadphase needs AABBs: run the AABB stage or upload AABBs first");
    if (!bounds_fused_) reset_stage_marks();
    const uint64_t n = n_;
    DevStatus* st = status_.as<DevStatus>();
    if (n == 0) {
      M_ = A_ = P_ = 0;
      counts_.n = 0; counts_.m = counts_.a = counts_.p = 0;
      return;
    }
    mark(ST_BP_BOUNDS, true);
    if (!bounds_fused_) {
      hipLaunchKernelGGL(k_init_status, dim3(1), dim3(64), 0, s_, st);
      unsigned nbk = std::min<unsigned>(blocks_for(2 * n, 256), 2048u);
      bpart_.reserve((uint64_t)nbk * 6 * 4);
      npart_ = nbk;
      hipLaunchKernelGGL((k_bounds<R, 256>), dim3(nbk), dim3(256), 0, s_, 2 * n, (const R*)aabb_.as<R>(),
                         bpart_.as<uint32_t>());
    }
    hipLaunchKernelGGL(k_params<R>, dim3(1), dim3(256), 0, s_, (const uint32_t*)bpart_.as<uint32_t>(), npart_, st,
                       g_[0], g_[1], g_[2],
                       params_.as<Params<R>>());
    mark(ST_BP_BOUNDS, false);
    bounds_fused_ = false;  // a second broadphase on the same AABBs recomputes bounds

    // count + scan (host reads M back, as the reference reads .back())
    counts_buf_.reserve(n * 4);
    offs_.reserve(n * 4);
    brec_.reserve(n * sizeof(BpRec<R>));
    scan_tmp_.reserve(std::max<size_t>(scan_tmp_words(n), 16) * 4);
    mark(ST_BP_COUNT, true);
    kpart_.reserve((uint64_t)blocks_for(n, 256) * 4);
    hipLaunchKernelGGL((k_count<R, 256>), dim3(blocks_for(n, 256)), dim3(256), 0, s_, n, (const R*)aabb_.as<R>(),
                       (const Params<R>*)params_.as<Params<R>>(), g_[0], g_[1], g_[2], counts_buf_.as<uint32_t>(),
                       brec_.as<BpRec<R>>(), kpart_.as<uint32_t>(), st);
    hipLaunchKernelGGL(k_reduce_keymax, dim3(1), dim3(256), 0, s_, (const uint32_t*)kpart_.as<uint32_t>(),
                       (uint64_t)blocks_for(n, 256), st);
    mark(ST_BP_COUNT, false);
    mark(ST_BP_SCAN, true);
    scan_excl_u32(counts_buf_.as<uint32_t>(), offs_.as<uint32_t>(), n, &st->M, scan_tmp_.as<uint32_t>(), s_);
    mark(ST_BP_SCAN, false);
    read_status();
    M_ = hstat_->M;
    CDX_REQUIRE(M_ < (1ull << 32), ERR_CAPACITY, "more than 2^32 bin entries: use fewer bins per axis");

    // key emit + sort by bin
    keys_.reserve(M_ * 8);
    keys_tmp_.reserve(M_ * 8);
    mark(ST_BP_EMIT, true);
    hipLaunchKernelGGL((k_emit_keys<R, 256>), dim3(blocks_for(n, 256)), dim3(256), 0, s_, n, (const R*)aabb_.as<R>(),
                       (const Params<R>*)params_.as<Params<R>>(), g_[0], g_[1], g_[2], (const uint32_t*)offs_.as<uint32_t>(),
                       keys_.as<uint64_t>());
    mark(ST_BP_EMIT, false);
    uint64_t key_space = (uint64_t)g_[0] * ((uint64_t)g_[1] * g_[2] + g_[1] + 1) + 1;
    int key_bits = key_space > 0xffffffffull ? 32 : bits_for(hstat_->key_max);
    std::vector<int> sh;
    for (int b = 0; b < key_bits; b += 8) sh.push_back(32 + b);
    radix_tmp_.reserve(std::max<size_t>(radix_tmp_words(M_), scan_tmp_words(M_) + 16) * 4);
    mark(ST_BP_SORT, true);
    radix_sort_u64(keys_.as<uint64_t>(), keys_tmp_.as<uint64_t>(), M_, sh.data(), (int)sh.size(),
                   radix_tmp_.as<uint32_t>(), s_);
    mark(ST_BP_SORT, false);

    // active bins
    mark(ST_BP_RUNS, true);
    hipLaunchKernelGGL(k_count_runs, dim3(std::min<unsigned>(blocks_for(M_, 256), 1024u)), dim3(256), 0, s_, M_,
                       (const uint64_t*)keys_.as<uint64_t>(), st);
    mark(ST_BP_RUNS, false);

    // pair count (per sorted entry), scan, emit
    pcount_.reserve((M_ + 1) * 4);
    poffs_.reserve((M_ + 1) * 4);
    mark(ST_BP_PAIRS, true);
    hipLaunchKernelGGL((k_pairs<R, false>), dim3(blocks_for(M_, PAIR_BS)), dim3(PAIR_BS), 0, s_, M_,
                       (const uint64_t*)keys_.as<uint64_t>(), (const BpRec<R>*)brec_.as<BpRec<R>>(), g_[0], g_[1],
                       pcount_.as<uint32_t>(), (const uint32_t*)nullptr, (uint64_t*)nullptr);
    scan_excl_u32(pcount_.as<uint32_t>(), poffs_.as<uint32_t>(), M_, &st->P, radix_tmp_.as<uint32_t>(), s_);
    read_status();
    P_ = hstat_->P;
    A_ = hstat_->A;
    CDX_REQUIRE(P_ < (1ull << 32), ERR_CAPACITY, "more than 2^32 candidate pairs");
    pairs_.reserve(P_ * 8);
    pairs_tmp_.reserve(P_ * 8);
    hipLaunchKernelGGL((k_pairs<R, true>), dim3(blocks_for(M_, PAIR_BS)), dim3(PAIR_BS), 0, s_, M_,
                       (const uint64_t*)keys_.as<uint64_t>(), (const BpRec<R>*)brec_.as<BpRec<R>>(), g_[0], g_[1],
                       (uint32_t*)nullptr, (const uint32_t*)poffs_.as<uint32_t>(), pairs_.as<uint64_t>());
    mark(ST_BP_PAIRS, false);
    int nbits = bits_for((uint32_t)(n - 1));
    std::vector<int> ps;
    for (int b = 0; b < nbits; b += 8) ps.push_back(b);
    for (int b = 0; b < nbits; b += 8) ps.push_back(32 + b);
    radix_tmp_.reserve(std::max<size_t>(radix_tmp_words(P_), 16) * 4);
    mark(ST_BP_PAIRSORT, true);
    radix_sort_u64(pairs_.as<uint64_t>(), pairs_tmp_.as<uint64_t>(), P_, ps.data(), (int)ps.size(),
                   radix_tmp_.as<uint32_t>(), s_);
    mark(ST_BP_PAIRSORT, false);
    counts_.n = n; counts_.m = M_; counts_.a = A_; counts_.p = P_;
    counts_.key_max = hstat_->key_max;
    counts_.bad = hstat_->bad;
    for (int c = 0; c < 3; ++c) { counts_.bmin[c] = ord2f(hstat_->bmin_enc[c]); counts_.bmax[c] = ord2f(hstat_->bmax_enc[c]); }
  }

  void run_narrowphase() override {
    CDX_REQUIRE(shapes_set_ && bodies_set_, ERR_NOT_READY, "narrowphase needs shapes and bodies");
    const uint64_t P = P_;
    DevStatus* st = status_.as<DevStatus>();
    if (!records_valid_) {
      // Shape records come from the AABB pass; build them when the caller
      // runs the narrowphase on its own (Narrowphase::DoNarrowphase).
      hipLaunchKernelGGL(k_init_status, dim3(1), dim3(64), 0, s_, st);
      run_records_only();
      records_valid_ = true;
    }
    np_v_.reserve(std::max<uint64_t>(P, 1) * 10 * sizeof(R));
    np_ids_.reserve(std::max<uint64_t>(P, 1) * 8);
    np_hit_.reserve(std::max<uint64_t>(P, 1) * 4);
    np_pos_.reserve(std::max<uint64_t>(P, 1) * 4);
    queue_.reserve(std::max<uint64_t>(P, 1) * 8);
    queue_tmp_.reserve(std::max<uint64_t>(P, 1) * 8);
    qsel_.reserve(std::max<uint64_t>(P, 1) * 4);
    qcls_.reserve(std::max<uint64_t>(P, 1));
    PairOut<R> o{np_v_.as<R>(), np_ids_.as<int32_t>(), np_hit_.as<uint32_t>(), P};
    hipLaunchKernelGGL(k_zero_np, dim3(1), dim3(64), 0, s_, st);
    mark(ST_NP_MAIN, true);
    if (P) {
      hipLaunchKernelGGL(k_np_front<R>, dim3(blocks_for(P, 128)), dim3(128), 0, s_, P,
                         (const uint64_t*)pairs_.as<uint64_t>(), (const ShapeRec<R>*)rec_.as<ShapeRec<R>>(),
                         (const ShapeHdr*)hdr_.as<ShapeHdr>(), envelope_, o, qsel_.as<uint32_t>(), qcls_.as<uint8_t>());
      scan_excl_u32(qsel_.as<uint32_t>(), np_pos_.as<uint32_t>(), P, &st->Q, scan_tmp_for(P), s_);
      hipLaunchKernelGGL(k_queue_compact, dim3(blocks_for(P, 256)), dim3(256), 0, s_, P,
                         (const uint32_t*)qsel_.as<uint32_t>(), (const uint32_t*)np_pos_.as<uint32_t>(),
                         (const uint8_t*)qcls_.as<uint8_t>(), queue_.as<uint64_t>());
    }
    read_status();
    uint64_t q = hstat_->Q;
    if (q) {
      // group the MPR queue by shape-pair class (one 8-bit digit, stable)
      const int cls_shift[1] = {32};
      radix_tmp_.reserve(std::max<size_t>(radix_tmp_words(q), 16) * 4);
      radix_sort_u64(queue_.as<uint64_t>(), queue_tmp_.as<uint64_t>(), q, cls_shift, 1, radix_tmp_.as<uint32_t>(), s_);
    }
    mark(ST_NP_MAIN, false);
    mark(ST_NP_RESOLVE, true);
    uint64_t* rq = queue_.as<uint64_t>();
    if (q && mode_ != NP_NATIVE) {
      auto filt = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(blocks_for(q, 128)), dim3(128), 0, s_, q, (const uint64_t*)queue_.as<uint64_t>(),
                           (const uint64_t*)pairs_.as<uint64_t>(), (const ShapeHdr*)hdr_.as<ShapeHdr>(),
                           (const ShapeRecF*)frec_.as<ShapeRecF>(), envelope_, tau_, o, qsel_.as<uint32_t>());
      };
      if (mode_ == NP_FAST) filt(k_np_filter<R, true>);
      else filt(k_np_filter<R, false>);
      scan_excl_u32(qsel_.as<uint32_t>(), np_pos_.as<uint32_t>(), q, &st->Q2, scan_tmp_for(q), s_);
      hipLaunchKernelGGL(k_queue_select, dim3(blocks_for(q, 256)), dim3(256), 0, s_, q,
                         (const uint64_t*)queue_.as<uint64_t>(), (const uint32_t*)qsel_.as<uint32_t>(),
                         (const uint32_t*)np_pos_.as<uint32_t>(), queue_tmp_.as<uint64_t>());
      read_status();
      q = hstat_->Q2;
      rq = queue_tmp_.as<uint64_t>();
    }
    resolved_ = mode_ == NP_NATIVE ? 0 : q;
    if (q) {
      if (mode_ == NP_NATIVE)
        hipLaunchKernelGGL((k_np_resolve<R, R>), dim3(blocks_for(q, 128)), dim3(128), 0, s_, q,
                           (const uint64_t*)pairs_.as<uint64_t>(), (const ShapeRec<R>*)rec_.as<ShapeRec<R>>(), envelope_,
                           o, (const uint64_t*)rq);
      else
        hipLaunchKernelGGL((k_np_resolve<R, double>), dim3(blocks_for(q, 128)), dim3(128), 0, s_, q,
                           (const uint64_t*)pairs_.as<uint64_t>(), (const ShapeRec<R>*)rec_.as<ShapeRec<R>>(), envelope_,
                           o, (const uint64_t*)rq);
    }
    mark(ST_NP_RESOLVE, false);
    mark(ST_NP_COMPACT, true);
    scan_excl_u32(np_hit_.as<uint32_t>(), np_pos_.as<uint32_t>(), P, &st->C,
                  scan_tmp_for(P), s_);
    read_status();
    C_ = hstat_->C;
    out_norm_.reserve(std::max<uint64_t>(C_, 1) * 3 * sizeof(R));
    out_pa_.reserve(std::max<uint64_t>(C_, 1) * 3 * sizeof(R));
    out_pb_.reserve(std::max<uint64_t>(C_, 1) * 3 * sizeof(R));
    out_d_.reserve(std::max<uint64_t>(C_, 1) * sizeof(R));
    out_ids_.reserve(std::max<uint64_t>(C_, 1) * 8);
    perm_.reserve(std::max<uint64_t>(P, 1) * 8);
    if (P)
      hipLaunchKernelGGL(k_np_compact<R>, dim3(blocks_for(P, 256)), dim3(256), 0, s_, P, o,
                         (const uint32_t*)np_pos_.as<uint32_t>(), (const uint64_t*)pairs_.as<uint64_t>(),
                         (const DevStatus*)st, out_norm_.as<R>(), out_pa_.as<R>(), out_pb_.as<R>(), out_d_.as<R>(),
                         out_ids_.as<int32_t>(), perm_.as<uint64_t>());
    mark(ST_NP_COMPACT, false);
    counts_.c = C_;
    counts_.resolved = resolved_;
    counts_.p = P;
  }

  void sync() override {
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
    CDX_HIP_CHECK(hipGetLastError());
  }

  const StepCounts& counts() const override { return counts_; }

  void read_aabb(void* out, bool shifted) override {
    if (n_ == 0) return;
    tmp_out_.reserve(n_ * 6 * sizeof(R));
    hipLaunchKernelGGL(k_aabb_aos<R>, dim3(blocks_for(2 * n_, 256)), dim3(256), 0, s_, n_, (const R*)aabb_.as<R>(),
                       (const Params<R>*)params_.as<Params<R>>(), shifted ? 1 : 0, tmp_out_.as<R>());
    CDX_HIP_CHECK(hipMemcpyAsync(out, tmp_out_.p, n_ * 6 * sizeof(R), hipMemcpyDeviceToHost, s_));
    sync();
  }

  void read_bins(uint32_t* bin_counts, uint32_t* sorted_keys, uint32_t* sorted_bodies, uint32_t* active_keys,
                 uint32_t* active_counts) override {
    sync();
    if (bin_counts && n_) CDX_HIP_CHECK(hipMemcpy(bin_counts, counts_buf_.p, n_ * 4, hipMemcpyDeviceToHost));
    if ((sorted_keys || sorted_bodies || active_keys || active_counts) && M_) {
      std::vector<uint64_t> k(M_);
      CDX_HIP_CHECK(hipMemcpy(k.data(), keys_.p, M_ * 8, hipMemcpyDeviceToHost));
      uint64_t r = 0;
      for (uint64_t i = 0; i < M_; ++i) {
        uint32_t key = (uint32_t)(k[i] >> 32);
        if (sorted_keys) sorted_keys[i] = key;
        if (sorted_bodies) sorted_bodies[i] = (uint32_t)k[i];
        if (i == 0 || key != (uint32_t)(k[i - 1] >> 32)) {
          if (active_keys) active_keys[r] = key;
          if (active_counts) active_counts[r] = 0;
          ++r;
        }
        if (active_counts) active_counts[r - 1] += 1;
      }
    }
  }

  void read_pairs(int64_t* out) override {
    sync();
    if (P_) CDX_HIP_CHECK(hipMemcpy(out, pairs_.p, P_ * 8, hipMemcpyDeviceToHost));
  }
  void read_pairs_permuted(int64_t* out) override {
    sync();
    if (P_) CDX_HIP_CHECK(hipMemcpy(out, perm_.p, P_ * 8, hipMemcpyDeviceToHost));
  }
  void read_contacts(void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids) override {
    sync();
    if (!C_) return;
    if (norm) CDX_HIP_CHECK(hipMemcpy(norm, out_norm_.p, C_ * 3 * sizeof(R), hipMemcpyDeviceToHost));
    if (cpta) CDX_HIP_CHECK(hipMemcpy(cpta, out_pa_.p, C_ * 3 * sizeof(R), hipMemcpyDeviceToHost));
    if (cptb) CDX_HIP_CHECK(hipMemcpy(cptb, out_pb_.p, C_ * 3 * sizeof(R), hipMemcpyDeviceToHost));
    if (dpth) CDX_HIP_CHECK(hipMemcpy(dpth, out_d_.p, C_ * sizeof(R), hipMemcpyDeviceToHost));
    if (bids) CDX_HIP_CHECK(hipMemcpy(bids, out_ids_.p, C_ * 8, hipMemcpyDeviceToHost));
  }

  void stage_ms(double* out) override {
    sync();
    for (int s = 0; s < ST_COUNT; ++s) {
      out[s] = 0.0;
      if (!used_[s]) continue;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev_[2 * s], ev_[2 * s + 1]) == hipSuccess) out[s] = ms;
    }
  }

  uint64_t device_bytes() const override {
    const DevBuf* all[] = {&type_, &body_, &lpos_, &dims_, &lrot_, &bpos_, &brot_, &active_, &aabb_, &rec_, &brec_, &hdr_, &frec_,
                           &counts_buf_, &offs_, &keys_, &keys_tmp_,
                           &pcount_, &poffs_, &pairs_, &pairs_tmp_, &np_v_, &np_ids_, &np_hit_, &np_pos_, &queue_, &queue_tmp_, &qsel_, &qcls_,
                           &out_norm_, &out_pa_, &out_pb_, &out_d_, &out_ids_, &perm_, &scan_tmp_, &radix_tmp_};
    uint64_t t = 0;
    for (auto* b : all) t += b->bytes;
    return t;
  }

 private:
  static int bits_for(uint32_t v) {
    int b = 0;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
  }
  void run_records_only() {
    rec_.reserve(n_ * sizeof(ShapeRec<R>));
    hdr_.reserve(n_ * sizeof(ShapeHdr));
    frec_.reserve(n_ * sizeof(ShapeRecF));
    bpart_.reserve((uint64_t)blocks_for(n_, 256) * 6 * 4);
    npart_ = blocks_for(n_, 256);
    aabb_.reserve(n_ * 6 * sizeof(R));
    if (n_)
      hipLaunchKernelGGL((k_aabb<R, 256>), dim3(blocks_for(n_, 256)), dim3(256), 0, s_, n_, type_.as<uint8_t>(),
                         body_.as<uint32_t>(), lpos_.as<R>(), dims_.as<R>(), lrot_.as<R>(), bpos_.as<R>(),
                         brot_.as<R>(), active_.as<uint8_t>(), nb_, aabb_.as<R>(), rec_.as<ShapeRec<R>>(),
                         hdr_.as<ShapeHdr>(), frec_.as<ShapeRecF>(), bpart_.as<uint32_t>(), status_.as<DevStatus>());
  }
  uint32_t* scan_tmp_for(uint64_t n) {
    scan_tmp_.reserve(std::max<size_t>(scan_tmp_words(n), 16) * 4);
    return scan_tmp_.as<uint32_t>();
  }
  void upload_soa(const void* host, uint64_t n, int k, R* dst) {
    stage_.reserve(n * k * sizeof(R));
    CDX_HIP_CHECK(hipMemcpy(stage_.p, host, n * k * sizeof(R), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_deinterleave<R>, dim3(blocks_for(n, 256)), dim3(256), 0, s_, n, k, (const R*)stage_.as<R>(), dst);
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
  }
  void sync_internal() { CDX_HIP_CHECK(hipStreamSynchronize(s_)); }
  void read_status() {
    CDX_HIP_CHECK(hipMemcpyAsync(hstat_, status_.p, sizeof(DevStatus), hipMemcpyDeviceToHost, s_));
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
    uint32_t bad = hstat_->bad;
    CDX_REQUIRE(!(bad & 1u), ERR_UNSUPPORTED_SHAPE, "unsupported shape type");
    CDX_REQUIRE(!(bad & 2u), ERR_INVALID, "shape body id out of range");
    CDX_REQUIRE(!(bad & 4u), ERR_INVALID,
                "an AABB maps outside the grid: non-finite coordinates or a zero scene extent");
  }
  void reset_stage_marks() {
    for (int s = 0; s < ST_COUNT; ++s) used_[s] = false;
  }
  void mark(int stage, bool start) {
    if (!prof_) return;
    CDX_HIP_CHECK(hipEventRecord(ev_[2 * stage + (start ? 0 : 1)], s_));
    if (start) used_[stage] = true;
  }

  int dev_;
  hipStream_t s_ = nullptr;
  DevStatus* hstat_ = nullptr;
  hipEvent_t ev_[2 * ST_COUNT];
  bool used_[ST_COUNT] = {};
  bool prof_ = false;
  uint64_t n_ = 0, nb_ = 0, M_ = 0, A_ = 0, P_ = 0, C_ = 0;
  int g_[3] = {16, 16, 16};
  R envelope_ = 0;
  int mode_ = NP_NATIVE;
  float tau_ = 1e-2f;
  bool have_aabb_ = false, bounds_fused_ = false;
  bool shapes_set_ = false, bodies_set_ = false, records_valid_ = false;
  StepCounts counts_;
  DevBuf status_, params_, stage_, upd_, tmp_out_;
  DevBuf type_, body_, lpos_, dims_, lrot_, bpos_, brot_, active_;
  DevBuf aabb_, rec_, brec_, hdr_, frec_, bpart_, kpart_;
  uint64_t npart_ = 0, resolved_ = 0;
  DevBuf counts_buf_, offs_, keys_, keys_tmp_, pcount_, poffs_, pairs_, pairs_tmp_;
  DevBuf np_v_, np_ids_, np_hit_, np_pos_, queue_, queue_tmp_, qsel_, qcls_;
  DevBuf out_norm_, out_pa_, out_pb_, out_d_, out_ids_, perm_;
  DevBuf scan_tmp_, radix_tmp_;
};

}  // namespace

std::unique_ptr<SceneBase> make_scene(int real_kind, int device) {
  if (real_kind == REAL_F64) return std::unique_ptr<SceneBase>(new Scene<double>(device));
  if (real_kind == REAL_F32) return std::unique_ptr<SceneBase>(new Scene<float>(device));
  throw Error(ERR_INVALID, "real kind must be CDX_REAL_F64 or CDX_REAL_F32");
}

}  // namespace cdx
