// AABB generation and uniform-grid broadphase kernels.  Included by
// pipeline.hip only.
//
// Every per-shape kernel runs one workgroup per Chunk (CHUNK consecutive
// shapes of one scene), so a step covers one scene or any set of resident
// scenes with the same kernels.  Per-shape arrays are SoA with the stride of
// the resident capacity NS; per-step arrays (bin counts, offsets) live in
// "slot" space: slot = chunk * CHUNK + lane.
#pragma once
#include <type_traits>

#include "cdx_device.hpp"

namespace cdx {
namespace {

__device__ __forceinline__ uint64_t cap_count(uint64_t cap, const uint64_t* d_n) {
  uint64_t v = *d_n;
  return v < cap ? v : cap;
}

// Per-workgroup bounds partials (reduced per scene by k_params), BPART_WORDS
// per workgroup: the float min / max over the workgroup's AABB corners, as
// order-preserving uints, then the widest AABB extent of the workgroup (float
// bits: non-negative floats order as unsigned integers), which sizes the cell
// search's grid, with bit 31 (WIDTH_INVERTED) set when an AABB's max corner
// is not at or above its min corner on some axis (such a scene takes the bin
// search, whose count reports the AABBs the reference grid cannot bin).
constexpr int BPART_WORDS = 7;
constexpr uint32_t WIDTH_INVERTED = 0x80000000u;

__device__ __forceinline__ void block_bounds_commit(uint32_t mn[3], uint32_t mx[3], uint32_t wm, uint32_t* part) {
  __shared__ uint32_t smn[3][CHUNK / WAVE], smx[3][CHUNK / WAVE], swm[CHUNK / WAVE];
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
  for (int o = 32; o > 0; o >>= 1) wm = max(wm, (uint32_t)__shfl_xor((int)wm, o));
  if (l == 0) swm[w] = wm;
  __syncthreads();
  uint32_t* out = part + (uint64_t)BPART_WORDS * blockIdx.x;
  if (threadIdx.x < 3) {
    int c = threadIdx.x;
    uint32_t a = 0xffffffffu, b = 0u;
    for (int i = 0; i < CHUNK / WAVE; ++i) { a = min(a, smn[c][i]); b = max(b, smx[c][i]); }
    out[c] = a;
    out[3 + c] = b;
  } else if (threadIdx.x == 3) {
    uint32_t m = 0u;
    for (int i = 0; i < CHUNK / WAVE; ++i) m = max(m, swm[i]);
    out[6] = m;
  }
}

// Folds one AABB into the corner bounds and the widest extent (NaN extents
// count as 0: such AABBs make the scene bounds non-finite and are reported),
// flagging an inverted or NaN AABB (WIDTH_INVERTED).
__device__ __forceinline__ void bounds_fold(uint32_t mn[3], uint32_t mx[3], uint32_t& wm, float lx, float ly,
                                            float lz, float hx, float hy, float hz) {
  uint32_t u;
  u = f2ord(lx); mn[0] = min(mn[0], u); mx[0] = max(mx[0], u);
  u = f2ord(ly); mn[1] = min(mn[1], u); mx[1] = max(mx[1], u);
  u = f2ord(lz); mn[2] = min(mn[2], u); mx[2] = max(mx[2], u);
  u = f2ord(hx); mn[0] = min(mn[0], u); mx[0] = max(mx[0], u);
  u = f2ord(hy); mn[1] = min(mn[1], u); mx[1] = max(mx[1], u);
  u = f2ord(hz); mn[2] = min(mn[2], u); mx[2] = max(mx[2], u);
  const float w = fmaxf(fmaxf(hx - lx, hy - ly), hz - lz);
  if (w > 0.f) wm = max(wm, __float_as_uint(w));
  if (!(hx >= lx) || !(hy >= ly) || !(hz >= lz)) wm |= WIDTH_INVERTED;
}

// One pass over shapes (SoA) and body poses: world AABB (AABBGenerator.cu:
// 84-116), the narrowphase records, and the float scene bounds over all 2n
// corners (Broadphase.cu:368-376, fminf / fmaxf narrowing).
template <class R>
__global__ __launch_bounds__(CHUNK) void k_aabb(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                uint64_t NS, const uint8_t* __restrict__ type,
                                                const uint32_t* __restrict__ body, const R* __restrict__ lpos,
                                                const R* __restrict__ dims, const R* __restrict__ cpos,
                                                const R* __restrict__ lrot, uint64_t NB,
                                                const R* __restrict__ bpos, const R* __restrict__ brot,
                                                const uint8_t* __restrict__ active, R* __restrict__ aabb,
                                                ShapeRec<R>* __restrict__ rec, ShapeHdr* __restrict__ hdr,
                                                ShapeRecF* __restrict__ frec, uint32_t* __restrict__ bpart,
                                                const SceneDesc<R>* __restrict__ desc, DevStatus* st) {
  // the chunk's narrowphase records are built in LDS and leave in coalesced
  // 16-byte stores (a record per thread would store at a 64-112 B stride)
  __shared__ ShapeRec<R> lrec[CHUNK];
  __shared__ ShapeRecF lfrec[CHUNK];
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  const Chunk ch = chunks[blockIdx.x];
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u}, wm = 0u;
  if (threadIdx.x < ch.len) {
    int t = type[i];
    uint32_t id = body[i];
    if (id < desc[ch.scene].body_begin || id >= desc[ch.scene].body_end) {
      atomicOr(&st->bad, 2u);
    } else if (t < SPHERE || t > CAPSULE) {
      atomicOr(&st->bad, 1u);
    } else {
      v3<R> a = mk3(lpos[i], lpos[NS + i], lpos[2 * NS + i]);
      v3<R> b = mk3(dims[i], dims[NS + i], dims[2 * NS + i]);
      q4<R> lr; lr.w = lrot[i]; lr.x = lrot[NS + i]; lr.y = lrot[2 * NS + i]; lr.z = lrot[3 * NS + i];
      v3<R> p = mk3(bpos[id], bpos[NB + id], bpos[2 * NB + id]);
      q4<R> rq; rq.w = brot[id]; rq.x = brot[NB + id]; rq.y = brot[2 * NB + id]; rq.z = brot[3 * NB + id];
      q4<R> q1 = qmul(rq, lr);
      v3<R> lo, hi, c;
      // the narrowphase record: centre, rotation, dimensions; a triangle
      // keeps its three world vertices there instead (x,y,z | bx,by,bz |
      // qw,qx,qy), the narrowphase's support of it
      v3<R> rb = b;
      q4<R> rq1 = q1;
      if (t == SPHERE) {
        c = rotate(a, rq) + p;
        v3<R> r3 = mk3(b.x, b.x, b.x);
        lo = c - r3;
        hi = c + r3;
      } else if (t == TRIANGLEMESH) {
        // function_MPR_Store's placement (Narrowphase.cu:363-367) and the
        // reference's triangle bounds (AABBGenerator.cu:16-23, :103-107)
        const v3<R> cc = mk3(cpos[i], cpos[NS + i], cpos[2 * NS + i]);
        const v3<R> va = rotate(a + p, q1), vb = rotate(b + p, q1), vc = rotate(cc + p, q1);
        lo = mk3(min(va.x, min(vb.x, vc.x)), min(va.y, min(vb.y, vc.y)), min(va.z, min(vb.z, vc.z)));
        hi = mk3(max(va.x, max(vb.x, vc.x)), max(va.y, max(vb.y, vc.y)), max(va.z, max(vb.z, vc.z)));
        c = va;
        rb = vb;
        rq1.w = vc.x; rq1.x = vc.y; rq1.y = vc.z; rq1.z = R(0);
      } else if (t == RECT || t == CONE) {
        // placed at the local position, rotated by body (x) local, as
        // function_MPR_Store does for these types (it applies no body pose);
        // bounded by the rectangle at height B.z / the cone's (r, h/2, r) box
        c = a;
        const mat3<R> m = rotmat(q1);
        const v3<R> ctr = t == RECT ? a + mul(m, mk3(R(0), R(0), b.z)) : a;
        const v3<R> h = mul(absm(m), t == RECT ? mk3(b.x, b.y, R(0)) : mk3(b.x, b.y * R(.5), b.x));
        lo = ctr - h;
        hi = ctr + h;
      } else {
        c = rotate(a, rq) + p;
        // Loose box of the semi-extents for ellipsoid and cylinder, as the
        // reference does; the capsule's is (r, h + r, r).
        v3<R> d = (t == CAPSULE) ? mk3(b.x, b.y + b.x, b.z) : b;
        v3<R> h = mul(absm(rotmat(q1)), d);
        lo = c - h;
        hi = c + h;
      }
      aabb[i] = lo.x; aabb[2 * NS + i] = lo.y; aabb[4 * NS + i] = lo.z;
      aabb[NS + i] = hi.x; aabb[3 * NS + i] = hi.y; aabb[5 * NS + i] = hi.z;
      ShapeRec<R> r;
      r.x = c.x; r.y = c.y; r.z = c.z;
      r.qw = rq1.w; r.qx = rq1.x; r.qy = rq1.y; r.qz = rq1.z;
      r.bx = rb.x; r.by = rb.y; r.bz = rb.z;
      r.type = (uint32_t)t; r.body = id; r.active = active[id]; r.pad = 0;
      lrec[threadIdx.x] = r;
      ShapeHdr h;
      h.ta = (uint32_t)t | ((uint32_t)r.active << 8);
      h.body = id;
      hdr[i] = h;
      ShapeRecF fr;
      fr.x = (double)c.x; fr.y = (double)c.y; fr.z = (double)c.z;
      fr.qw = (float)q1.w; fr.qx = (float)q1.x; fr.qy = (float)q1.y; fr.qz = (float)q1.z;
      fr.bx = (float)b.x; fr.by = (float)b.y; fr.bz = (float)b.z;
      fr.type = (uint32_t)t;
      fr.body = id;
      fr.active = r.active;
      lfrec[threadIdx.x] = fr;
      bounds_fold(mn, mx, wm, (float)lo.x, (float)lo.y, (float)lo.z, (float)hi.x, (float)hi.y, (float)hi.z);
    }
  }
  block_bounds_commit(mn, mx, wm, bpart);  // synchronises: the LDS records are complete
  {
    const uint32_t nr = ch.len * (uint32_t)(sizeof(ShapeRec<R>) / 16);
    const uint4* src = reinterpret_cast<const uint4*>(lrec);
    uint4* dst = reinterpret_cast<uint4*>(rec + ch.begin);
    for (uint32_t t = threadIdx.x; t < nr; t += CHUNK) dst[t] = src[t];
    const uint32_t nf = ch.len * (uint32_t)(sizeof(ShapeRecF) / 16);
    const uint4* fsrc = reinterpret_cast<const uint4*>(lfrec);
    uint4* fdst = reinterpret_cast<uint4*>(frec + ch.begin);
    for (uint32_t t = threadIdx.x; t < nf; t += CHUNK) fdst[t] = fsrc[t];
  }
}

// Bounds of AABBs supplied by the caller (Broadphase::detectPossibleCollisions
// on its own): the same partials over both corners of every AABB.
template <class R>
__global__ __launch_bounds__(CHUNK) void k_bounds(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st,
                                                  uint64_t NS, const R* __restrict__ aabb,
                                                  uint32_t* __restrict__ bpart) {
  if ((uint64_t)blockIdx.x * CHUNK >= st->nslots) return;
  const Chunk ch = chunks[blockIdx.x];
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u}, wm = 0u;
  if (threadIdx.x < ch.len)
    bounds_fold(mn, mx, wm, (float)aabb[i], (float)aabb[2 * NS + i], (float)aabb[4 * NS + i], (float)aabb[NS + i],
                (float)aabb[3 * NS + i], (float)aabb[5 * NS + i]);
  block_bounds_commit(mn, mx, wm, bpart);
}

// Per stepped scene: bounds from its chunks' partials, then origin and
// inverse bin size (Broadphase.cu:377-381).  One workgroup per scene; the
// workgroup of the first stepped scene also stores that scene's widest AABB
// extent in st->wmax.
constexpr int PARAMS_BS = 1024;

template <class R>
__global__ __launch_bounds__(PARAMS_BS) void k_params(const uint32_t* __restrict__ sel, DevStatus* __restrict__ st,
                                                      const uint32_t* __restrict__ bpart,
                                                      SceneDesc<R>* __restrict__ desc) {
  if (blockIdx.x >= st->nsel) return;
  __shared__ uint32_t red[BPART_WORDS][PARAMS_BS / WAVE];
  SceneDesc<R>& d = desc[sel[blockIdx.x]];
  uint32_t v[BPART_WORDS] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u, 0u};
#pragma unroll 4
  for (uint64_t b = d.chunk_begin + threadIdx.x; b < d.chunk_end; b += PARAMS_BS) {
    const uint32_t* p = bpart + BPART_WORDS * b;
    for (int c = 0; c < 3; ++c) {
      v[c] = min(v[c], p[c]);
      v[3 + c] = max(v[3 + c], p[3 + c]);
    }
    v[6] = max(v[6], p[6]);
  }
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  for (int o = 32; o > 0; o >>= 1) {
    for (int c = 0; c < 3; ++c) {
      v[c] = min(v[c], (uint32_t)__shfl_xor((int)v[c], o));
      v[3 + c] = max(v[3 + c], (uint32_t)__shfl_xor((int)v[3 + c], o));
    }
    v[6] = max(v[6], (uint32_t)__shfl_xor((int)v[6], o));
  }
  if (l == 0)
    for (int c = 0; c < BPART_WORDS; ++c) red[c][w] = v[c];
  __syncthreads();
  if (threadIdx.x >= WAVE) return;
  // the first wavefront reduces the per-wavefront values by shuffles
  const bool has = threadIdx.x < PARAMS_BS / WAVE;
  for (int c = 0; c < 3; ++c) {
    v[c] = has ? red[c][threadIdx.x] : 0xffffffffu;
    v[3 + c] = has ? red[3 + c][threadIdx.x] : 0u;
  }
  v[6] = has ? red[6][threadIdx.x] : 0u;
  for (int o = 32; o > 0; o >>= 1) {
    for (int c = 0; c < 3; ++c) {
      v[c] = min(v[c], (uint32_t)__shfl_xor((int)v[c], o));
      v[3 + c] = max(v[3 + c], (uint32_t)__shfl_xor((int)v[3 + c], o));
    }
    v[6] = max(v[6], (uint32_t)__shfl_xor((int)v[6], o));
  }
  if (threadIdx.x != 0) return;
  if (blockIdx.x == 0) st->wmax = v[6];
  int g[3] = {d.gx, d.gy, d.gz};
  for (int c = 0; c < 3; ++c) {
    d.bmin_enc[c] = v[c];
    d.bmax_enc[c] = v[3 + c];
    R mn = (R)ord2f(v[c]);
    R mx = (R)ord2f(v[3 + c]);
    R origin = cabs(mn);
    R ext = cabs(mx + cabs(mn));
    d.origin[c] = origin;
    d.inv[c] = (R)g[c] / ext;
  }
}

struct Cells { int x0, y0, z0, x1, y1, z1; };

// Finite AABBs always map into cells 0..grid on every axis (the shift makes
// coordinates non-negative up to rounding, and truncation sends tiny
// negatives to 0).  Anything else comes from non-finite input or a
// degenerate scene extent; such AABBs are reported instead of binned.
__device__ __forceinline__ bool cells_ok(const Cells& c, int gx, int gy, int gz) {
  return c.x0 >= 0 && c.y0 >= 0 && c.z0 >= 0 && c.x0 <= c.x1 && c.y0 <= c.y1 && c.z0 <= c.z1 && c.x1 <= gx &&
         c.y1 <= gy && c.z1 <= gz;
}

// Shifted AABB of shape i (coordinates plus the scene origin,
// Broadphase.cu:382-386).  k_count and k_emit_keys both build it with the
// same single additions, so their cells agree bit for bit.
template <class R>
__device__ __forceinline__ BpRec<R> shifted_rec(const R* __restrict__ aabb, uint64_t NS, uint64_t i,
                                                const SceneDesc<R>& d, uint32_t scene) {
  BpRec<R> r;
  r.lo[0] = aabb[i] + d.origin[0];
  r.lo[1] = aabb[2 * NS + i] + d.origin[1];
  r.lo[2] = aabb[4 * NS + i] + d.origin[2];
  r.hi[0] = aabb[NS + i] + d.origin[0];
  r.hi[1] = aabb[3 * NS + i] + d.origin[1];
  r.hi[2] = aabb[5 * NS + i] + d.origin[2];
  r.scene = scene;
  r.offs = 0;
  return r;
}

template <class R> __device__ __forceinline__ Cells cells_of(const BpRec<R>& r, const SceneDesc<R>& d) {
  Cells c;
  c.x0 = cell_of(r.lo[0], d.inv[0]); c.y0 = cell_of(r.lo[1], d.inv[1]); c.z0 = cell_of(r.lo[2], d.inv[2]);
  c.x1 = cell_of(r.hi[0], d.inv[0]); c.y1 = cell_of(r.hi[1], d.inv[1]); c.z1 = cell_of(r.hi[2], d.inv[2]);
  return c;
}

// Launches of at most this many workgroups fold their per-workgroup totals
// into the step status by atomics: that many same-address atomics cost less
// than the extra reduction kernel (a few microseconds per launch); larger
// grids leave per-workgroup partials to a reduction kernel, where the
// atomics would serialise.
constexpr uint64_t FOLD_WGS = 128;

// Bins spanned per AABB (Broadphase.cu:150-159), on coordinates shifted by
// the scene origin (Broadphase.cu:382-386) and hashed by truncation
// (HashMin, :125-133); plus the largest key per chunk (the max corner's, keys
// do not wrap in the dense key range), into kpart for k_reduce_keymax or,
// with kpart null, into st->key_max by atomicMax.  With tile_part, also the
// counts' tile sums for scan_excl_u32_tiled.
template <class R>
__global__ __launch_bounds__(CHUNK) void k_count(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                 uint64_t NS, const R* __restrict__ aabb,
                                                 const SceneDesc<R>* __restrict__ desc, uint32_t* __restrict__ counts,
                                                 uint32_t* __restrict__ kpart, uint64_t* __restrict__ tile_part,
                                                 DevStatus* st) {
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  __shared__ uint32_t skm[CHUNK / WAVE];
  const Chunk ch = chunks[blockIdx.x];
  const uint64_t slot = (uint64_t)blockIdx.x * CHUNK + threadIdx.x;
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  uint32_t km = 0, cv = 0;
  if (threadIdx.x < ch.len) {
    const SceneDesc<R>& d = desc[ch.scene];
    const BpRec<R> r = shifted_rec(aabb, NS, i, d, ch.scene);
    const Cells c = cells_of(r, d);
    if (cells_ok(c, d.gx, d.gy, d.gz)) {
      cv = (uint32_t)((c.x1 - c.x0 + 1) * (c.y1 - c.y0 + 1) * (c.z1 - c.z0 + 1));
      km = d.key_base + hash_index((uint32_t)c.x1, (uint32_t)c.y1, (uint32_t)c.z1, d.gx, d.gy);
    } else {
      atomicOr(&st->bad, 4u);
    }
  }
  counts[slot] = cv;
  if (tile_part) {
    __shared__ uint32_t red[CHUNK / WAVE];
    add_tile_sum<CHUNK>(tile_part, slot, cv, red);
  }
  for (int o = 32; o > 0; o >>= 1) km = max(km, (uint32_t)__shfl_xor((int)km, o));
  if (threadIdx.x % WAVE == 0) skm[threadIdx.x / WAVE] = km;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int w = 0; w < CHUNK / WAVE; ++w) m = max(m, skm[w]);
    if (kpart) kpart[blockIdx.x] = m;
    else if (m) atomicMax(&st->key_max, m);
  }
}

constexpr int KEYMAX_BS = 1024;

// The largest key over the chunk maxima, by one workgroup of BS threads.
template <int BS>
__device__ __forceinline__ void block_keymax(const uint32_t* __restrict__ kpart, uint64_t npart, DevStatus* st) {
  __shared__ uint32_t red[BS / WAVE];
  uint32_t m = 0;
  for (uint64_t b = threadIdx.x; b < npart; b += BS) m = max(m, kpart[b]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = m;
  __syncthreads();
  if (threadIdx.x < WAVE) {
    m = threadIdx.x < BS / WAVE ? red[threadIdx.x] : 0u;
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if (threadIdx.x == 0) st->key_max = m;
  }
}

__global__ __launch_bounds__(KEYMAX_BS) void k_reduce_keymax(const uint32_t* __restrict__ kpart, DevStatus* st) {
  block_keymax<KEYMAX_BS>(kpart, st->nslots / CHUNK, st);
}

// (bin, body) entries in the reference's store order (Broadphase.cu:
// 172-200: x, then y, then z), packed as key << 32 | shape so one 64-bit
// stable sort orders by bin and keeps shapes ascending inside a bin.  Keys
// of a scene are offset by its key base, so the bins of all stepped scenes
// sort into one array, scene after scene.  The pass also writes each
// shape's broadphase record (shifted AABB, scene, first entry position),
// built in LDS and stored in coalesced 16-byte runs, so the pair search
// finds a shape's emission range in the record it gathers anyway.
constexpr int EMIT_LDS = 4096;  // entries of one chunk staged for coalesced stores

template <class R>
__global__ __launch_bounds__(CHUNK) void k_emit_keys(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                     uint64_t NS, const R* __restrict__ aabb,
                                                     const SceneDesc<R>* __restrict__ desc,
                                                     const uint32_t* __restrict__ offs,
                                                     const uint32_t* __restrict__ counts, uint64_t mcap,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ pcount,
                                                     BpRec<R>* __restrict__ brec, const uint32_t* __restrict__ kpart,
                                                     DevStatus* st) {
  static_assert(sizeof(BpRec<R>) * CHUNK <= sizeof(uint64_t) * EMIT_LDS, "records fit the key staging area");
  __shared__ alignas(64) uint64_t sk[EMIT_LDS];  // the chunk's entries, then its records
  __shared__ uint64_t s_total;
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  // with kpart, workgroup 0 also reduces the count pass's chunk maxima
  // (steps whose key width is fixed beforehand need the maximum only as a
  // reported total, so no kernel of its own runs for it)
  if (kpart && blockIdx.x == 0) block_keymax<CHUNK>(kpart, st_in->nslots / CHUNK, st);
  const Chunk ch = chunks[blockIdx.x];
  if (ch.len == 0) return;
  const uint64_t slot0 = (uint64_t)blockIdx.x * CHUNK;
  const uint64_t e0 = offs[slot0];
  const bool mine = threadIdx.x < ch.len;
  const uint64_t o0 = mine ? offs[slot0 + threadIdx.x] : 0;
  if (threadIdx.x == ch.len - 1) s_total = o0 + counts[slot0 + threadIdx.x] - e0;
  __syncthreads();
  const uint64_t total = s_total;
  const bool staged = total <= EMIT_LDS;
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  BpRec<R> r;
  if (mine) {
    const SceneDesc<R>& d = desc[ch.scene];
    r = shifted_rec(aabb, NS, i, d, ch.scene);
    r.offs = (uint32_t)o0;
    const Cells c = cells_of(r, d);
    if (cells_ok(c, d.gx, d.gy, d.gz)) {
      uint64_t o = o0;
      for (uint32_t x = (uint32_t)c.x0; x <= (uint32_t)c.x1; ++x)
        for (uint32_t y = (uint32_t)c.y0; y <= (uint32_t)c.y1; ++y)
          for (uint32_t z = (uint32_t)c.z0; z <= (uint32_t)c.z1; ++z) {
            const uint64_t v = ((uint64_t)(d.key_base + hash_index(x, y, z, d.gx, d.gy)) << 32) | i;
            if (staged) {
              sk[o - e0] = v;
            } else if (o < mcap) {
              keys[o] = v;
              pcount[o] = 0;  // pair count of this entry, set by the pair search
            } else {
              atomicOr(&st->overflow, 1u);
            }
            ++o;
          }
    }
  }
  if (staged) {
    __syncthreads();
    // coalesced stores of the chunk's entries (and zeroed pair counts)
    for (uint64_t t = threadIdx.x; t < total; t += CHUNK) {
      const uint64_t o = e0 + t;
      if (o < mcap) {
        keys[o] = sk[t];
        pcount[o] = 0;
      } else {
        atomicOr(&st->overflow, 1u);
      }
    }
  }
  __syncthreads();  // the staging area now takes the records
  BpRec<R>* lrec = reinterpret_cast<BpRec<R>*>(sk);
  if (mine) lrec[threadIdx.x] = r;
  __syncthreads();
  const uint32_t nr = ch.len * (uint32_t)(sizeof(BpRec<R>) / 16);
  const uint4* src = reinterpret_cast<const uint4*>(lrec);
  uint4* dst = reinterpret_cast<uint4*>(brec + ch.begin);
  for (uint32_t t = threadIdx.x; t < nr; t += CHUNK) dst[t] = src[t];
}

// Position, in the emission order of k_emit_keys (x outer, z inner), of the
// cell with scene-relative key kk among the cells c of one shape.  A max
// corner at index == grid size makes up to four cells share a key
// (Broadphase.cu:140-147); the first one inside the shape's range is taken.
// Returns ~0u if no candidate lies in the range (never for emitted keys).
__device__ __forceinline__ uint32_t local_cell_index(uint32_t kk, const Cells& c, int gx, int gy, uint64_t rgx,
                                                     uint64_t rgy) {
  const uint32_t q = div_by_recip(kk, (uint32_t)gx, rgx), i0 = kk - q * (uint32_t)gx;
  for (int ai = 0; ai < 2; ++ai) {
    uint32_t i, qq;
    if (ai == 0) { i = i0; qq = q; }
    else { if (i0 != 0 || q == 0) break; i = (uint32_t)gx; qq = q - 1; }
    const uint32_t l0 = div_by_recip(qq, (uint32_t)gy, rgy), j0 = qq - l0 * (uint32_t)gy;
    for (int aj = 0; aj < 2; ++aj) {
      uint32_t j, l;
      if (aj == 0) { j = j0; l = l0; }
      else { if (j0 != 0 || l0 == 0) break; j = (uint32_t)gy; l = l0 - 1; }
      if ((int)i >= c.x0 && (int)i <= c.x1 && (int)j >= c.y0 && (int)j <= c.y1 && (int)l >= c.z0 && (int)l <= c.z1)
        return ((i - (uint32_t)c.x0) * (uint32_t)(c.y1 - c.y0 + 1) + (j - (uint32_t)c.y0)) *
                   (uint32_t)(c.z1 - c.z0 + 1) + (l - (uint32_t)c.z0);
    }
  }
  return ~0u;
}

// Pair search with de-duplication by owning bin.  A pair (a, b) of
// overlapping AABBs (closed test on shifted coordinates, Broadphase.cu:237 /
// :296) is reported only from the bin whose key is the hash of the cell
// max(gmin_a, gmin_b): that cell lies in both AABBs' cell ranges, so both
// shapes appear in its run of sorted entries, and every other run skips the
// pair.  Repeated entries of one shape in a run (hash aliasing at index ==
// grid size, Broadphase.cu:140-147) are skipped, so each pair is emitted
// once and the reference's sort + unique (:522-535) reduces to ordering.
//
// Count pass (k_pairs): each sorted entry i pairs shape(i) with the later
// entries of its run (shapes ascend inside a run, so shape(i) < shape(k)).
// A workgroup stages its 256 entries and the next PAIR_AHEAD keys in LDS,
// with the records of its own entries and of the look-ahead entries that
// continue its last run, and splits the (entry, later entry) candidates of
// its own entries evenly over its lanes, so short and long runs cost the
// same per lane.  Per entry it stores the pair count and a hit mask (bit d:
// the entry d+1 places later is a partner; bit 63: partners beyond 62
// places) at the entry's emission position j (the first entry of its shape,
// carried in the shape's record, + its cell's index among the shape's
// cells), so pairs can be gathered per first shape.  The run starts among
// its own entries (the active bins) go to runs_part, summed by sum_runs.
// k_shape_pairs sums each shape's counts, a scan gives each shape's pair
// range, and k_pair_emit writes each shape's partners from the masks and
// sorts them in registers: the pair list comes out ascending per scene
// without a global sort.
constexpr int PAIR_BS = 256;
constexpr int PAIR_AHEAD = 64;
constexpr int PAIR_W = PAIR_BS + PAIR_AHEAD;
constexpr uint64_t MASK_TAIL = 1ull << 63;

// Per entry with pairs, 16 bytes: either the partners themselves (at most
// META_IDS of them and none beyond the window: the emit pass then reads no
// sorted entries; in any order, groups are sorted after), or the hit mask
// and the entry's sorted position.  w0 = pair count, bit 31 set for the
// mask form; ids form: w1..w3 = partner ids; mask form: w1 = sorted
// position, w2 / w3 = low / high half of the hit mask.
constexpr int META_IDS = 3;
constexpr uint32_t META_MASK_FORM = 0x80000000u;
struct alignas(16) PairMeta {
  uint32_t w0, w1, w2, w3;
};

// Shifted AABB alone, as the pair search stages it in LDS (the scene and
// first entry are kept only for the workgroup's own entries).  FP32 boxes
// are padded to 32 bytes so an entry's LDS address is a shift, not a
// quarter-rate multiply.
template <class R> struct alignas(sizeof(R) == 4 ? 32 : 16) BpBox {
  R lo[3], hi[3];
};

template <class R> __device__ __forceinline__ BpBox<R> box_of(const BpRec<R>& r) {
  BpBox<R> b;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = r.lo[k];
    b.hi[k] = r.hi[k];
  }
  return b;
}

template <class A, class B> __device__ __forceinline__ bool overlaps(const A& a, const B& b) {
  return (a.lo[0] <= b.hi[0]) & (b.lo[0] <= a.hi[0]) & (a.lo[1] <= b.hi[1]) & (b.lo[1] <= a.hi[1]) &
         (a.lo[2] <= b.hi[2]) & (b.lo[2] <= a.hi[2]);
}

// Whether the overlapping pair belongs to the run of key K: the hash of the
// cell max(gmin_a, gmin_b) (both AABBs cover it) must be K.
// SMALL: the scene's grid satisfies grid_fits_u24, so the hash can use
// full-rate 24-bit multiplies (same low 32 bits as the 32-bit products).
template <class R, class B, bool SMALL = false>
__device__ __forceinline__ bool owner_bin(const int ga[3], const B& rb, const R inv[3], uint32_t K, uint32_t kb,
                                          int gx, int gy) {
  const uint32_t ox = (uint32_t)max(ga[0], cell_of(rb.lo[0], inv[0]));
  const uint32_t oy = (uint32_t)max(ga[1], cell_of(rb.lo[1], inv[1]));
  const uint32_t oz = (uint32_t)max(ga[2], cell_of(rb.lo[2], inv[2]));
  if constexpr (SMALL)
    return kb + mad_u24(mad_u24(oz, (uint32_t)gy, oy), (uint32_t)gx, ox) == K;
  else
    return kb + hash_index(ox, oy, oz, gx, gy) == K;
}

// Cells run 0..grid size per axis (the max corner may sit on index == grid
// size), so with these bounds every factor of the hash, and z * gy + y,
// stays below 2^24.
__device__ __forceinline__ bool grid_fits_u24(int gx, int gy, int gz) {
  return (uint64_t)gx < (1u << 24) && (uint64_t)(gz + 1) * (uint64_t)(gy + 1) < (1u << 24);
}

template <class R, class A, class B>
__device__ __forceinline__ bool pair_test(const A& ra, const int ga[3], const B& rb, const R inv[3], uint32_t K,
                                          uint32_t kb, int gx, int gy) {
  return overlaps(ra, rb) && owner_bin(ga, rb, inv, K, kb, gx, gy);
}

// Scene parameters and cells of one entry's shape.
template <class R> struct EntryCtx {
  BpBox<R> r;
  R inv[3];
  int g0[3], g1[3];
  int gx, gy;
  uint64_t rgx, rgy;
  uint32_t kb, K, a;
};

template <class R, class B>
__device__ __forceinline__ void entry_ctx_d(EntryCtx<R>& c, uint32_t K, uint32_t a, const B& r,
                                            const SceneDesc<R>& d) {
  c.K = K;
  c.a = a;
  c.gx = d.gx;
  c.gy = d.gy;
  c.rgx = d.rgx;
  c.rgy = d.rgy;
  c.kb = d.key_base;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c.r.lo[k] = r.lo[k];
    c.r.hi[k] = r.hi[k];
    c.inv[k] = d.inv[k];
    c.g0[k] = cell_of(r.lo[k], c.inv[k]);
    c.g1[k] = cell_of(r.hi[k], c.inv[k]);
  }
}

template <class R>
__device__ __forceinline__ void entry_ctx(EntryCtx<R>& c, uint32_t K, uint32_t a, const BpRec<R>& r,
                                          const SceneDesc<R>* __restrict__ desc) {
  entry_ctx_d(c, K, a, r, desc[r.scene]);
}

// entry_ctx with the scene descriptor from LDS when the whole staged window
// belongs to one scene (block-uniform branch), else from global memory.
template <class R>
__device__ __forceinline__ void entry_ctx_w(EntryCtx<R>& c, uint32_t K, uint32_t a, const BpBox<R>& r,
                                            uint32_t scene, bool one, const SceneDesc<R>& sd,
                                            const SceneDesc<R>* __restrict__ desc) {
  if (one) entry_ctx_d(c, K, a, r, sd);
  else entry_ctx_d(c, K, a, r, desc[scene]);
}

template <class R>
__global__ __launch_bounds__(PAIR_BS) void k_pairs(uint64_t mcap, const uint64_t* __restrict__ keys,
                                                   const BpRec<R>* __restrict__ brec,
                                                   const SceneDesc<R>* __restrict__ desc, uint32_t* __restrict__ pcount,
                                                   PairMeta* __restrict__ meta, uint32_t* __restrict__ runs_part,
                                                   DevStatus* st) {
  __shared__ BpBox<R> sbox[PAIR_W];
  __shared__ uint32_t skey[PAIR_W], sbody[PAIR_W];
  __shared__ uint32_t sscene[PAIR_BS];  // scenes of the own entries
  __shared__ uint64_t send[PAIR_W / WAVE];
  __shared__ uint32_t sstart[PAIR_BS + 1];
  __shared__ uint64_t smask[PAIR_BS];  // hit mask per own entry, set by LDS atomics
  __shared__ uint32_t scan_scratch[PAIR_BS / WAVE + 1];
  __shared__ SceneDesc<R> sdesc;  // the window's scene, when it has only one
  __shared__ int sone;
  __shared__ uint32_t sruns[PAIR_BS / WAVE];
  const uint64_t m = cap_count(mcap, &st->M);
  const uint64_t b0 = (uint64_t)blockIdx.x * PAIR_BS;
  if (b0 >= m) return;
  const int tid = threadIdx.x, lane = tid % WAVE;
  const int nloc = (int)min((uint64_t)PAIR_W, m - b0);
  const int nown = min(nloc, PAIR_BS);
  uint32_t my_offs = 0;  // first entry of the own entry's shape
  if (tid < nloc) {
    const uint64_t kk = keys[b0 + tid];
    skey[tid] = (uint32_t)(kk >> 32);
    sbody[tid] = (uint32_t)kk;
    const BpRec<R> r = brec[(uint32_t)kk];
    sbox[tid] = box_of(r);
    sscene[tid] = r.scene;
    my_offs = r.offs;
  }
  static_assert(PAIR_AHEAD <= PAIR_BS, "one look-ahead entry per thread");
  if (tid < PAIR_AHEAD && PAIR_BS + tid < nloc) {
    // A look-ahead entry is a candidate of an own entry only if it continues
    // the run of the last own entry (keys ascend, so those entries come
    // first); only their records are gathered.
    const int t = PAIR_BS + tid;
    const uint64_t kk = keys[b0 + t];
    const uint32_t klast = (uint32_t)(keys[b0 + PAIR_BS - 1] >> 32);
    skey[t] = (uint32_t)(kk >> 32);
    sbody[t] = (uint32_t)kk;
    if ((uint32_t)(kk >> 32) == klast) sbox[t] = box_of(brec[(uint32_t)kk]);
  }
  __syncthreads();
  if (tid == 0) {
    // keys ascend and scenes own disjoint key ranges: one scene at both ends
    // of the own entries means one scene for every staged record
    const uint32_t s0 = sscene[0];
    sone = s0 == sscene[nown - 1] ? 1 : 0;
    if (sone) sdesc = desc[s0];
  }
  // run ends inside the window (the last staged entry always ends one)
  for (int t = tid; t < PAIR_W; t += PAIR_BS) {
    const bool f = t < nloc && (t + 1 == nloc || skey[t + 1] != skey[t]);
    const uint64_t bal = __ballot(f);
    if (lane == 0) send[t / WAVE] = bal;
  }
  __syncthreads();
  const bool one = sone != 0;
  // own entry: candidates inside the window
  const bool valid = tid < nloc;
  const uint64_t i = b0 + tid;
  bool dup = false, beyond = false;
  uint32_t n_i = 0;
  int u = tid;  // last entry of the own entry's run inside the window
  bool run_start = false;  // the entry opens a run of equal keys (an active bin)
  if (valid) {
    if (tid > 0) {
      dup = skey[tid - 1] == skey[tid] && sbody[tid - 1] == sbody[tid];
      run_start = skey[tid - 1] != skey[tid];
    } else if (i > 0) {
      const uint64_t prev = keys[i - 1];
      dup = prev == (((uint64_t)skey[0] << 32) | sbody[0]);
      run_start = (uint32_t)(prev >> 32) != skey[0];
    } else {
      run_start = true;
    }
    int w = tid / WAVE;
    uint64_t fm = send[w] & (~0ull << (tid % WAVE));
    while (!fm && ++w < PAIR_W / WAVE) fm = send[w];
    u = fm ? w * WAVE + __ffsll((unsigned long long)fm) - 1 : nloc - 1;
    if (!dup) {
      n_i = (uint32_t)(u - tid);
      beyond = u == nloc - 1 && b0 + nloc < m && (uint32_t)(keys[b0 + nloc] >> 32) == skey[tid];
    }
  }
  {
    const uint64_t rb = __ballot(run_start);
    if (lane == 0) sruns[tid / WAVE] = (uint32_t)__popcll(rb);
  }
  uint32_t total;
  const uint32_t my_start = block_excl_scan<PAIR_BS>(n_i, scan_scratch, &total);  // synchronises: sruns complete
  sstart[tid] = my_start;
  if (tid == PAIR_BS - 1) sstart[PAIR_BS] = total;
  if (tid == 0) {
    uint32_t r = 0;
#pragma unroll
    for (int w = 0; w < PAIR_BS / WAVE; ++w) r += sruns[w];
    runs_part[blockIdx.x] = r;
  }
  smask[tid] = 0;
  __syncthreads();
  // balanced pass over the candidates of the workgroup's entries
  auto balanced_pass = [&](auto small_grid) {
    constexpr bool SMALL = decltype(small_grid)::value;
    uint32_t f = (uint32_t)(((uint64_t)total * tid) / PAIR_BS);
    const uint32_t fend = (uint32_t)(((uint64_t)total * (tid + 1)) / PAIR_BS);
    if (f >= fend) return;
    int lo = 0, hi = PAIR_BS - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sstart[mid] <= f) lo = mid; else hi = mid - 1;
    }
    int e = lo;
    uint32_t ebeg = sstart[e], eend = sstart[e + 1];
    EntryCtx<R> c;
    entry_ctx_w(c, skey[e], sbody[e], sbox[e], sscene[e], one, sdesc, desc);
    uint32_t prev = f == ebeg ? c.a : sbody[e + (int)(f - ebeg)];
    for (; f < fend; ++f) {
      if (f >= eend) {
        do { ++e; ebeg = sstart[e]; eend = sstart[e + 1]; } while (eend == ebeg);
        e &= PAIR_BS - 1;  // e < PAIR_BS: as for k below
        entry_ctx_w(c, skey[e], sbody[e], sbox[e], sscene[e], one, sdesc, desc);
        prev = c.a;
      }
      // k < PAIR_W <= 1024: the mask changes nothing but tells the compiler
      // the index is small, so the record address uses a full-rate 24-bit
      // multiply
      const int k = (e + 1 + (int)(f - ebeg)) & 1023;
      const uint32_t b = sbody[k];
      const BpBox<R> rk = sbox[k];
      // repeat entries of one shape (b == prev) are skipped; the record is
      // read and tested with the body id rather than after it
      const bool fresh = b != prev;
      prev = b;
      if ((fresh & overlaps(c.r, rk)) &&
          owner_bin<R, BpBox<R>, SMALL>(c.g0, rk, c.inv, c.K, c.kb, c.gx, c.gy)) {
        // a hit goes straight to the entry's mask: lanes sharing an entry
        // need no merge step, and the partners are read back from the
        // mask bits and the staged bodies (partners 63 or more places on
        // are only flagged, and counted again below)
        // bit dd for partners up to 62 places on, MASK_TAIL (bit 63) beyond
        const int dd = min(k - e - 1, 63);
        atomicOr((unsigned long long*)&smask[e], 1ull << dd);
      }
    }
  };
  if (one && grid_fits_u24(sdesc.gx, sdesc.gy, sdesc.gz)) balanced_pass(std::true_type{});
  else balanced_pass(std::false_type{});
  __syncthreads();
  if (!valid || dup) return;
  uint64_t mk = smask[tid];
  uint32_t cnt = (uint32_t)__popcll(mk & ~MASK_TAIL);
  EntryCtx<R> c;
  entry_ctx_w(c, skey[tid], sbody[tid], sbox[tid], sscene[tid], one, sdesc, desc);
  if (mk & MASK_TAIL) {  // partners 63 or more places on inside the window (long runs only)
    for (int k = tid + 64; k <= u; ++k) {
      const uint32_t b = sbody[k];
      if (b != sbody[k - 1] && pair_test(c.r, c.g0, sbox[k], c.inv, c.K, c.kb, c.gx, c.gy)) ++cnt;
    }
  }
  if (beyond) {  // run continues past the window: the rest from global memory (rare)
    uint32_t prev = sbody[nloc - 1];
    for (uint64_t k = b0 + nloc; k < m; ++k) {
      const uint64_t kk = keys[k];
      if ((uint32_t)(kk >> 32) != c.K) break;
      const uint32_t b = (uint32_t)kk;
      if (b == prev) continue;
      prev = b;
      if (pair_test(c.r, c.g0, brec[b], c.inv, c.K, c.kb, c.gx, c.gy)) {
        ++cnt;
        mk |= MASK_TAIL;
      }
    }
  }
  if (!cnt) return;
  Cells cl;
  cl.x0 = c.g0[0]; cl.y0 = c.g0[1]; cl.z0 = c.g0[2];
  cl.x1 = c.g1[0]; cl.y1 = c.g1[1]; cl.z1 = c.g1[2];
  const uint32_t li = local_cell_index(c.K - c.kb, cl, c.gx, c.gy, c.rgx, c.rgy);
  const uint64_t j = (uint64_t)my_offs + li;
  if (li == ~0u || j >= m) {
    atomicOr(&st->bad, 8u);
    return;
  }
  pcount[j] = cnt;
  PairMeta pm;
  if (cnt <= (uint32_t)META_IDS && !(mk & MASK_TAIL)) {
    // partners in ascending position: the lowest set bits of the mask
    uint32_t id[META_IDS];
    uint64_t bits = mk;
#pragma unroll
    for (int q = 0; q < META_IDS; ++q) {
      id[q] = bits ? sbody[tid + __ffsll((unsigned long long)bits)] : 0u;
      bits &= bits - 1;
    }
    pm.w0 = cnt;
    pm.w1 = id[0];
    pm.w2 = id[1];
    pm.w3 = id[2];
  } else {
    pm.w0 = cnt | META_MASK_FORM;
    pm.w1 = (uint32_t)i;
    pm.w2 = (uint32_t)mk;
    pm.w3 = (uint32_t)(mk >> 32);
  }
  meta[j] = pm;
}

// Active bins (reduce_by_key, Broadphase.cu:439-443): the pair count pass
// counts the run starts among each workgroup's own entries; the first
// workgroups of the pair offset pass (k_shape_pairs) sum those counts,
// striding over them, so at most RUNS_GRID atomics meet at the one counter
// (same-address atomics are serialised device-wide) and no kernel of its
// own runs for it.
constexpr unsigned RUNS_GRID = 64;

__device__ __forceinline__ void sum_runs(uint64_t nparts, const uint32_t* __restrict__ runs_part, uint64_t nblocks,
                                         DevStatus* st) {
  const uint64_t g = min((uint64_t)RUNS_GRID, nblocks);
  if (blockIdx.x >= g) return;
  __shared__ uint64_t red[CHUNK / WAVE];
  uint64_t c = 0;
  for (uint64_t b = (uint64_t)blockIdx.x * CHUNK + threadIdx.x; b < nparts; b += g * CHUNK) c += runs_part[b];
  for (int o = 32; o > 0; o >>= 1) c += (uint64_t)__shfl_xor((long long)c, o);
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < CHUNK / WAVE; ++w) t += red[w];
    if (t) atomicAdd((unsigned long long*)&st->A, (unsigned long long)t);
  }
}

// reduce_by_key materialised for a caller that reads the bins
// (Broadphase.cu:428-443, :476-480): the sorted (bin, body) entries split
// into their two arrays, one head flag per run of equal keys; after an
// exclusive scan of the heads, each head writes its key and start to its
// run's slot, and each run's occupancy is the distance to the next start.
// The step itself never runs these (the pair search finds runs on its own).
__global__ __launch_bounds__(256) void k_bins_split(uint64_t m, const uint64_t* __restrict__ entries,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ bodies,
                                                    uint32_t* __restrict__ heads) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t e = entries[i];
  const uint32_t k = (uint32_t)(e >> 32);
  keys[i] = k;
  bodies[i] = (uint32_t)e;
  heads[i] = (i == 0 || (uint32_t)(entries[i - 1] >> 32) != k) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_bins_runs(uint64_t m, const uint32_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ heads,
                                                   const uint32_t* __restrict__ run_of,
                                                   uint32_t* __restrict__ active_keys,
                                                   uint32_t* __restrict__ run_start) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !heads[i]) return;
  active_keys[run_of[i]] = keys[i];
  run_start[run_of[i]] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_bins_occupancy(uint64_t a, uint64_t m, const uint32_t* __restrict__ run_start,
                                                        uint32_t* __restrict__ counts) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a) return;
  const uint64_t end = r + 1 < a ? run_start[r + 1] : m;
  counts[r] = (uint32_t)(end - run_start[r]);
}

// Pairs per shape: the sum of its entries' counts.  A chunk's entries are
// contiguous, so the workgroup first reads their counts in one coalesced
// sweep into LDS (one memory latency), then each shape sums its own.
constexpr int SHAPE_PAIRS_LDS = 4096;

// With runs_part, the first workgroups also sum the active bins (sum_runs).
__global__ __launch_bounds__(CHUNK) void k_shape_pairs(const Chunk* __restrict__ chunks, DevStatus* __restrict__ st,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ offs,
                                                       const uint32_t* __restrict__ pcount, uint64_t mcap,
                                                       uint32_t* __restrict__ scount, uint64_t* __restrict__ tile_part,
                                                       const uint32_t* __restrict__ runs_part) {
  __shared__ uint32_t sp[SHAPE_PAIRS_LDS];
  __shared__ uint64_t s_e0, s_e1;
  const uint64_t slot = (uint64_t)blockIdx.x * CHUNK + threadIdx.x;
  if ((uint64_t)blockIdx.x * CHUNK >= st->nslots) return;
  const uint64_t m = cap_count(mcap, &st->M);
  if (runs_part) sum_runs((m + PAIR_BS - 1) / PAIR_BS, runs_part, st->nslots / CHUNK, st);
  const Chunk ch = chunks[blockIdx.x];
  if (ch.len == 0) {
    scount[slot] = 0;
    return;
  }
  const bool mine = threadIdx.x < ch.len;
  const uint64_t e0 = mine ? offs[slot] : 0, e1 = mine ? min(e0 + counts[slot], m) : 0;
  if (threadIdx.x == 0) s_e0 = min(e0, m);
  if (threadIdx.x == ch.len - 1) s_e1 = e1;
  __syncthreads();
  const uint64_t c0 = s_e0, n = s_e1 > c0 ? s_e1 - c0 : 0;
  uint32_t sum = 0;
  if (n <= (uint64_t)SHAPE_PAIRS_LDS) {
    for (uint64_t t = threadIdx.x; t < n; t += CHUNK) sp[t] = pcount[c0 + t];
    __syncthreads();
    if (mine)
      for (uint64_t e = e0; e < e1; ++e) sum += sp[e - c0];
  } else if (mine) {
    for (uint64_t e = e0; e < e1; ++e) sum += pcount[e];
  }
  scount[slot] = sum;
  if (tile_part) {  // a chunk with no shapes adds nothing
    __shared__ uint32_t red[CHUNK / WAVE];
    add_tile_sum<CHUNK>(tile_part, slot, sum, red);
  }
}

// In-register bitonic sort of N values (compile-time indices only).
template <int N> __device__ __forceinline__ void sort_net(uint32_t (&v)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const uint32_t x = v[i], y = v[l];
          const bool up = (i & k) == 0;
          v[i] = up ? min(x, y) : max(x, y);
          v[l] = up ? max(x, y) : min(x, y);
        }
      }
}

template <int N>
__device__ __forceinline__ void sort_pairs_in_place(uint64_t* __restrict__ pairs, uint64_t beg, uint32_t len,
                                                    uint32_t a) {
  uint32_t v[N];
#pragma unroll
  for (int t = 0; t < N; ++t) v[t] = (uint32_t)t < len ? (uint32_t)pairs[beg + t] : ~0u;
  sort_net<N>(v);
#pragma unroll
  for (int t = 0; t < N; ++t)
    if ((uint32_t)t < len) pairs[beg + t] = ((uint64_t)a << 32) | v[t];
}

constexpr int GROUP_LONG = 32;
constexpr int GROUP_LDS = 4096;

template <int N>
__device__ __forceinline__ void sort_lds_group(uint32_t* __restrict__ v, uint32_t len) {
  uint32_t r[N];
#pragma unroll
  for (int t = 0; t < N; ++t) r[t] = (uint32_t)t < len ? v[t] : ~0u;
  sort_net<N>(r);
#pragma unroll
  for (int t = 0; t < N; ++t)
    if ((uint32_t)t < len) v[t] = r[t];
}

// Partners of one shape from the hit masks of its entries (the rest of runs
// longer than the mask by repeating the tests), appended through put(b).
template <class R, class PC, class Put>
__device__ __forceinline__ void shape_partners(uint64_t e0, uint64_t e1, uint32_t a, PC pcount,
                                               const PairMeta* __restrict__ meta, const uint64_t* __restrict__ keys,
                                               const BpRec<R>* __restrict__ brec, const SceneDesc<R>* __restrict__ desc, uint64_t m, Put put) {
  // The next entry's metadata is loaded before this one is unpacked, and
  // mask-form partners are read four at a time, so the dependent loads of
  // one shape overlap.
  auto next = [&](uint64_t e) {
    while (e < e1 && !pcount(e)) ++e;
    return e;
  };
  uint64_t e = next(e0);
  PairMeta pm = {0u, 0u, 0u, 0u};
  if (e < e1) pm = meta[e];
  while (e < e1) {
    const uint64_t en = next(e + 1);
    PairMeta pn = {0u, 0u, 0u, 0u};
    if (en < e1) pn = meta[en];
    if (!(pm.w0 & META_MASK_FORM)) {
      if (pm.w0 > 0) put(pm.w1);
      if (pm.w0 > 1) put(pm.w2);
      if (pm.w0 > 2) put(pm.w3);
    } else {
      const uint64_t ip = pm.w1;
      const uint64_t mask = ((uint64_t)pm.w3 << 32) | pm.w2;
      uint64_t bits = mask & ~MASK_TAIL;
      while (bits) {
        uint32_t id[4];
        int got = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (bits) {
            const int dd = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            id[q] = (uint32_t)keys[ip + 1 + dd];
            got = q + 1;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < got) put(id[q]);
      }
      if (mask & MASK_TAIL) {
        const uint64_t ki = keys[ip];
        EntryCtx<R> c;
        entry_ctx(c, (uint32_t)(ki >> 32), a, brec[a], desc);
        uint32_t prev = (uint32_t)keys[ip + 63];
        for (uint64_t k = ip + 64; k < m; ++k) {
          const uint64_t kk = keys[k];
          if ((uint32_t)(kk >> 32) != c.K) break;
          const uint32_t b = (uint32_t)kk;
          if (b == prev) continue;
          prev = b;
          if (pair_test(c.r, c.g0, brec[b], c.inv, c.K, c.kb, c.gx, c.gy)) put(b);
        }
      }
    }
    e = en;
    pm = pn;
  }
}

constexpr int EMIT_PAIRS_LDS = 3584;
constexpr int EMIT_COUNTS_LDS = 2048;

// Writes each shape's pairs (a << 32 | b) into its range, sorted by b.  A
// workgroup handles one chunk: partners go to LDS, each shape's group is
// sorted in registers, and the chunk's pairs leave in one coalesced store.
// Chunks with more pairs than fit in LDS write in place; groups longer than
// GROUP_LONG are listed for k_pair_groups_long.  The kernel waits on
// dependent loads (entry metadata, then sorted entries), so it is held to
// 80 VGPRs and 26 KB of LDS: 6 waves per SIMD.
template <class R>
__global__ __launch_bounds__(CHUNK, 6) void k_pair_emit(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                     const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offs,
                                                     const uint32_t* __restrict__ pcount,
                                                     const PairMeta* __restrict__ meta,
                                                     const uint32_t* __restrict__ scount,
                                                     const uint32_t* __restrict__ soff, const uint64_t* __restrict__ keys,
                                                     const BpRec<R>* __restrict__ brec,
                                                     const SceneDesc<R>* __restrict__ desc, uint64_t mcap,
                                                     uint64_t pcap, uint64_t* __restrict__ pairs,
                                                     uint32_t* __restrict__ longs, uint64_t long_cap, DevStatus* st) {
  __shared__ uint32_t sb[EMIT_PAIRS_LDS];
  __shared__ uint8_t sa[EMIT_PAIRS_LDS];
  __shared__ uint32_t spc[EMIT_COUNTS_LDS];  // the chunk's entry counts, read in one coalesced sweep
  __shared__ uint64_t s_span, s_e0, s_e1;
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  const Chunk ch = chunks[blockIdx.x];
  if (ch.len == 0) return;
  const uint64_t slot0 = (uint64_t)blockIdx.x * CHUNK;
  const uint64_t p0 = soff[slot0];
  const bool mine = threadIdx.x < ch.len;
  const uint64_t slot = slot0 + threadIdx.x;
  const uint32_t tot = mine ? scount[slot] : 0;
  const uint64_t beg = mine ? soff[slot] : 0;
  const uint64_t m = cap_count(mcap, &st_in->M);
  const uint64_t e0 = mine ? offs[slot] : 0, e1 = mine ? min(e0 + counts[slot], m) : 0;
  if (threadIdx.x == 0) s_e0 = min(e0, m);
  if (threadIdx.x == ch.len - 1) {
    s_span = beg + tot - p0;
    s_e1 = e1;
  }
  __syncthreads();
  const uint64_t span = s_span;
  if (p0 + span > pcap) {
    if (threadIdx.x == 0) atomicOr(&st->overflow, 2u);
    return;
  }
  if (span == 0) return;
  const bool staged = span <= EMIT_PAIRS_LDS;
  const uint32_t a = ch.begin + threadIdx.x;
  const uint64_t c0 = s_e0, nc = s_e1 > c0 ? s_e1 - c0 : 0;
  const bool counts_staged = staged && nc <= (uint64_t)EMIT_COUNTS_LDS;
  if (counts_staged) {
    for (uint64_t t = threadIdx.x; t < nc; t += CHUNK) spc[t] = pcount[c0 + t];
    __syncthreads();
  }
  if (mine && tot) {
    uint64_t o = beg;
    if (staged) {
      uint32_t* v = sb + (beg - p0);
      uint32_t q = 0;
      auto put = [&](uint32_t b) { v[q++] = b; };
      if (counts_staged)
        shape_partners<R>(e0, e1, a, [&](uint64_t e) { return spc[e - c0]; }, meta, keys, brec, desc, m, put);
      else
        shape_partners<R>(e0, e1, a, [&](uint64_t e) { return pcount[e]; }, meta, keys, brec, desc, m, put);
      for (uint32_t t = 0; t < tot; ++t) sa[beg - p0 + t] = (uint8_t)threadIdx.x;
      if (tot <= 4) sort_lds_group<4>(v, tot);
      else if (tot <= 8) sort_lds_group<8>(v, tot);
      else if (tot <= 16) sort_lds_group<16>(v, tot);
      else if (tot <= GROUP_LONG) sort_lds_group<GROUP_LONG>(v, tot);
      else {
        const uint32_t g = atomicAdd(&st->nbig, 1u);
        if (g < long_cap) longs[g] = (uint32_t)slot;
        else atomicOr(&st->overflow, 4u);
      }
    } else {
      shape_partners<R>(e0, e1, a, [&](uint64_t e) { return pcount[e]; }, meta, keys, brec, desc, m,
                        [&](uint32_t b) { pairs[o++] = ((uint64_t)a << 32) | b; });
      if (tot <= 4) sort_pairs_in_place<4>(pairs, beg, tot, a);
      else if (tot <= 8) sort_pairs_in_place<8>(pairs, beg, tot, a);
      else if (tot <= 16) sort_pairs_in_place<16>(pairs, beg, tot, a);
      else if (tot <= GROUP_LONG) sort_pairs_in_place<GROUP_LONG>(pairs, beg, tot, a);
      else {
        const uint32_t g = atomicAdd(&st->nbig, 1u);
        if (g < long_cap) longs[g] = (uint32_t)slot;
        else atomicOr(&st->overflow, 4u);
      }
    }
  }
  if (!staged) return;
  __syncthreads();
  for (uint64_t q = threadIdx.x; q < span; q += CHUNK)
    pairs[p0 + q] = ((uint64_t)(ch.begin + sa[q]) << 32) | sb[q];
}

// Long groups: one workgroup per group, bitonic sort in LDS (groups beyond
// GROUP_LDS, i.e. one shape overlapping thousands of others, take a serial
// insertion sort).
__global__ __launch_bounds__(256) void k_pair_groups_long(const Chunk* __restrict__ chunks,
                                                          const DevStatus* __restrict__ st,
                                                          const uint32_t* __restrict__ scount,
                                                          const uint32_t* __restrict__ soff,
                                                          uint64_t* __restrict__ pairs,
                                                          const uint32_t* __restrict__ longs, uint64_t long_cap) {
  __shared__ uint32_t v[GROUP_LDS];
  const uint64_t nlong = min((uint64_t)st->nbig, long_cap);
  for (uint64_t g = blockIdx.x; g < nlong; g += gridDim.x) {
    const uint32_t slot = longs[g];
    const uint64_t beg = soff[slot];
    const uint32_t len = scount[slot];
    const uint32_t a = chunks[slot / CHUNK].begin + slot % CHUNK;
    if (len <= GROUP_LDS) {
      uint32_t n2 = 1;
      while (n2 < len) n2 <<= 1;
      for (uint32_t t = threadIdx.x; t < n2; t += blockDim.x) v[t] = t < len ? (uint32_t)pairs[beg + t] : ~0u;
      __syncthreads();
      for (uint32_t k = 2; k <= n2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
          for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
            const uint32_t l = i ^ j;
            if (l > i) {
              const uint32_t x = v[i], y = v[l];
              const bool up = (i & k) == 0;
              if ((x > y) == up) { v[i] = y; v[l] = x; }
            }
          }
          __syncthreads();
        }
      for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) pairs[beg + t] = ((uint64_t)a << 32) | v[t];
      __syncthreads();
    } else if (threadIdx.x == 0) {
      for (uint32_t t = 1; t < len; ++t) {
        const uint64_t x = pairs[beg + t];
        uint32_t u = t;
        while (u > 0 && pairs[beg + u - 1] > x) {
          pairs[beg + u] = pairs[beg + u - 1];
          --u;
        }
        pairs[beg + u] = x;
      }
    }
  }
}

// Per stepped scene: its ranges in the step's sorted entries, pair list and
// contact list (pairs are sorted by their first shape, and shapes of a
// scene are contiguous, so each scene's results are contiguous).
struct SceneOut {
  uint64_t m_begin, m_end, pair_begin, pair_end, contact_begin, contact_end;
};

template <class R>
__global__ void k_scene_ranges(const uint32_t* __restrict__ sel, const SceneDesc<R>* __restrict__ desc,
                               const DevStatus* __restrict__ st, uint64_t mcap, uint64_t pcap,
                               const uint32_t* __restrict__ offs, const uint32_t* __restrict__ soff,
                               const uint32_t* __restrict__ hitpos, SceneOut* __restrict__ out) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= st->nsel) return;
  const SceneDesc<R>& d = desc[sel[j]];
  const uint64_t nslot = st->nslots;
  const uint64_t M = cap_count(mcap, &st->M);
  const uint64_t P = cap_count(pcap, &st->P);
  auto entry_at = [&](uint64_t slot) -> uint64_t { return slot < nslot ? min((uint64_t)offs[slot], M) : M; };
  auto pair_at = [&](uint64_t slot) -> uint64_t { return slot < nslot ? min((uint64_t)soff[slot], P) : P; };
  SceneOut o;
  o.m_begin = entry_at((uint64_t)d.chunk_begin * CHUNK);
  o.m_end = entry_at((uint64_t)d.chunk_end * CHUNK);
  o.pair_begin = pair_at((uint64_t)d.chunk_begin * CHUNK);
  o.pair_end = pair_at((uint64_t)d.chunk_end * CHUNK);
  o.contact_begin = o.pair_begin < P ? hitpos[o.pair_begin] : st->C;
  o.contact_end = o.pair_end < P ? hitpos[o.pair_end] : st->C;
  out[j] = o;
}

}  // namespace
}  // namespace cdx
