// AABB generation and uniform-grid broadphase kernels.  Included by
// pipeline.hip only.
//
// Every per-shape kernel runs one workgroup per Chunk (CHUNK consecutive
// shapes of one scene), so a step covers one scene or any set of resident
// scenes with the same kernels.  Per-shape arrays are SoA with the stride of
// the resident capacity NS; per-step arrays (bin counts, offsets) live in
// "slot" space: slot = chunk * CHUNK + lane.
#pragma once
#include "cdx_device.hpp"

namespace cdx {
namespace {

__device__ __forceinline__ uint64_t cap_count(uint64_t cap, const uint64_t* d_n) {
  uint64_t v = *d_n;
  return v < cap ? v : cap;
}

// Per-workgroup bounds partials (reduced per scene by k_params): the float
// min / max over the workgroup's AABB corners, as order-preserving uints.
__device__ __forceinline__ void block_bounds_commit(uint32_t mn[3], uint32_t mx[3], uint32_t* part) {
  __shared__ uint32_t smn[3][CHUNK / WAVE], smx[3][CHUNK / WAVE];
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
    for (int i = 0; i < CHUNK / WAVE; ++i) { a = min(a, smn[c][i]); b = max(b, smx[c][i]); }
    part[6 * (uint64_t)blockIdx.x + c] = a;
    part[6 * (uint64_t)blockIdx.x + 3 + c] = b;
  }
}

__device__ __forceinline__ void bounds_fold(uint32_t mn[3], uint32_t mx[3], float lx, float ly, float lz, float hx,
                                            float hy, float hz) {
  uint32_t u;
  u = f2ord(lx); mn[0] = min(mn[0], u); mx[0] = max(mx[0], u);
  u = f2ord(ly); mn[1] = min(mn[1], u); mx[1] = max(mx[1], u);
  u = f2ord(lz); mn[2] = min(mn[2], u); mx[2] = max(mx[2], u);
  u = f2ord(hx); mn[0] = min(mn[0], u); mx[0] = max(mx[0], u);
  u = f2ord(hy); mn[1] = min(mn[1], u); mx[1] = max(mx[1], u);
  u = f2ord(hz); mn[2] = min(mn[2], u); mx[2] = max(mx[2], u);
}

// One pass over shapes (SoA) and body poses: world AABB (AABBGenerator.cu:
// 84-116), the narrowphase records, and the float scene bounds over all 2n
// corners (Broadphase.cu:368-376, fminf / fmaxf narrowing).
template <class R>
__global__ __launch_bounds__(CHUNK) void k_aabb(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                uint64_t NS, const uint8_t* __restrict__ type,
                                                const uint32_t* __restrict__ body, const R* __restrict__ lpos,
                                                const R* __restrict__ dims, const R* __restrict__ lrot, uint64_t NB,
                                                const R* __restrict__ bpos, const R* __restrict__ brot,
                                                const uint8_t* __restrict__ active, R* __restrict__ aabb,
                                                ShapeRec<R>* __restrict__ rec, ShapeHdr* __restrict__ hdr,
                                                ShapeRecF* __restrict__ frec, uint32_t* __restrict__ bpart,
                                                const SceneDesc<R>* __restrict__ desc, DevStatus* st) {
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  const Chunk ch = chunks[blockIdx.x];
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  if (threadIdx.x < ch.len) {
    int t = type[i];
    uint32_t id = body[i];
    if (id < desc[ch.scene].body_begin || id >= desc[ch.scene].body_end) {
      atomicOr(&st->bad, 2u);
    } else if (!(t == SPHERE || t == ELLIPSOID || t == BOX || t == CYLINDER || t == CAPSULE)) {
      atomicOr(&st->bad, 1u);
    } else {
      v3<R> a = mk3(lpos[i], lpos[NS + i], lpos[2 * NS + i]);
      v3<R> b = mk3(dims[i], dims[NS + i], dims[2 * NS + i]);
      q4<R> lr; lr.w = lrot[i]; lr.x = lrot[NS + i]; lr.y = lrot[2 * NS + i]; lr.z = lrot[3 * NS + i];
      v3<R> p = mk3(bpos[id], bpos[NB + id], bpos[2 * NB + id]);
      q4<R> rq; rq.w = brot[id]; rq.x = brot[NB + id]; rq.y = brot[2 * NB + id]; rq.z = brot[3 * NB + id];
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
      aabb[i] = lo.x; aabb[2 * NS + i] = lo.y; aabb[4 * NS + i] = lo.z;
      aabb[NS + i] = hi.x; aabb[3 * NS + i] = hi.y; aabb[5 * NS + i] = hi.z;
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
      bounds_fold(mn, mx, (float)lo.x, (float)lo.y, (float)lo.z, (float)hi.x, (float)hi.y, (float)hi.z);
    }
  }
  block_bounds_commit(mn, mx, bpart);
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
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  if (threadIdx.x < ch.len)
    bounds_fold(mn, mx, (float)aabb[i], (float)aabb[2 * NS + i], (float)aabb[4 * NS + i], (float)aabb[NS + i],
                (float)aabb[3 * NS + i], (float)aabb[5 * NS + i]);
  block_bounds_commit(mn, mx, bpart);
}

// Per stepped scene: bounds from its chunks' partials, then origin and
// inverse bin size (Broadphase.cu:377-381).  One workgroup per scene.
template <class R>
__global__ __launch_bounds__(256) void k_params(const uint32_t* __restrict__ sel, const DevStatus* __restrict__ st,
                                                const uint32_t* __restrict__ bpart, SceneDesc<R>* __restrict__ desc) {
  if (blockIdx.x >= st->nsel) return;
  __shared__ uint32_t red[6][256];
  SceneDesc<R>& d = desc[sel[blockIdx.x]];
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (uint64_t b = d.chunk_begin + threadIdx.x; b < d.chunk_end; b += 256) {
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

// Bins spanned per AABB (Broadphase.cu:150-159), on coordinates shifted by
// the scene origin (Broadphase.cu:382-386) and hashed by truncation
// (HashMin, :125-133); plus the per-shape broadphase record and the largest
// key (the max corner's, keys do not wrap in the dense key range).
template <class R>
__global__ __launch_bounds__(CHUNK) void k_count(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                 uint64_t NS, const R* __restrict__ aabb,
                                                 const SceneDesc<R>* __restrict__ desc, uint32_t* __restrict__ counts,
                                                 BpRec<R>* __restrict__ brec, uint32_t* __restrict__ kpart,
                                                 DevStatus* st) {
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  __shared__ uint32_t skm[CHUNK / WAVE];
  const Chunk ch = chunks[blockIdx.x];
  const uint64_t slot = (uint64_t)blockIdx.x * CHUNK + threadIdx.x;
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  uint32_t km = 0;
  if (threadIdx.x < ch.len) {
    const SceneDesc<R>& d = desc[ch.scene];
    BpRec<R> r;
    r.lo[0] = aabb[i] + d.origin[0];
    r.lo[1] = aabb[2 * NS + i] + d.origin[1];
    r.lo[2] = aabb[4 * NS + i] + d.origin[2];
    r.hi[0] = aabb[NS + i] + d.origin[0];
    r.hi[1] = aabb[3 * NS + i] + d.origin[1];
    r.hi[2] = aabb[5 * NS + i] + d.origin[2];
    Cells c;
    c.x0 = cell_of(r.lo[0], d.inv[0]); c.y0 = cell_of(r.lo[1], d.inv[1]); c.z0 = cell_of(r.lo[2], d.inv[2]);
    c.x1 = cell_of(r.hi[0], d.inv[0]); c.y1 = cell_of(r.hi[1], d.inv[1]); c.z1 = cell_of(r.hi[2], d.inv[2]);
    r.g[0] = c.x0; r.g[1] = c.y0; r.g[2] = c.z0;
    r.scene = ch.scene;
    brec[i] = r;
    if (cells_ok(c, d.gx, d.gy, d.gz)) {
      counts[slot] = (uint32_t)((c.x1 - c.x0 + 1) * (c.y1 - c.y0 + 1) * (c.z1 - c.z0 + 1));
      km = d.key_base + hash_index((uint32_t)c.x1, (uint32_t)c.y1, (uint32_t)c.z1, d.gx, d.gy);
    } else {
      counts[slot] = 0;
      atomicOr(&st->bad, 4u);
    }
  } else {
    counts[slot] = 0;
  }
  for (int o = 32; o > 0; o >>= 1) km = max(km, (uint32_t)__shfl_xor((int)km, o));
  if (threadIdx.x % WAVE == 0) skm[threadIdx.x / WAVE] = km;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int w = 0; w < CHUNK / WAVE; ++w) m = max(m, skm[w]);
    kpart[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(256) void k_reduce_keymax(const uint32_t* __restrict__ kpart, DevStatus* st) {
  __shared__ uint32_t red[256];
  const uint64_t npart = st->nslots / CHUNK;
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
// 172-200: x, then y, then z), packed as key << 32 | shape so one 64-bit
// stable sort orders by bin and keeps shapes ascending inside a bin.  Keys
// of a scene are offset by its key base, so the bins of all stepped scenes
// sort into one array, scene after scene.
template <class R>
__global__ __launch_bounds__(CHUNK) void k_emit_keys(const Chunk* __restrict__ chunks, const DevStatus* __restrict__ st_in,
                                                     const SceneDesc<R>* __restrict__ desc,
                                                     const BpRec<R>* __restrict__ brec, const uint32_t* __restrict__ offs,
                                                     uint64_t mcap, uint64_t* __restrict__ keys, DevStatus* st) {
  if ((uint64_t)blockIdx.x * CHUNK >= st_in->nslots) return;
  const Chunk ch = chunks[blockIdx.x];
  if (threadIdx.x >= ch.len) return;
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  const SceneDesc<R>& d = desc[ch.scene];
  const BpRec<R> r = brec[i];
  Cells c;
  c.x0 = r.g[0]; c.y0 = r.g[1]; c.z0 = r.g[2];
  c.x1 = cell_of(r.hi[0], d.inv[0]); c.y1 = cell_of(r.hi[1], d.inv[1]); c.z1 = cell_of(r.hi[2], d.inv[2]);
  if (!cells_ok(c, d.gx, d.gy, d.gz)) return;
  uint64_t o = offs[(uint64_t)blockIdx.x * CHUNK + threadIdx.x];
  for (uint32_t x = (uint32_t)c.x0; x <= (uint32_t)c.x1; ++x)
    for (uint32_t y = (uint32_t)c.y0; y <= (uint32_t)c.y1; ++y)
      for (uint32_t z = (uint32_t)c.z0; z <= (uint32_t)c.z1; ++z) {
        if (o < mcap) keys[o] = ((uint64_t)(d.key_base + hash_index(x, y, z, d.gx, d.gy)) << 32) | i;
        else atomicOr(&st->overflow, 1u);
        ++o;
      }
}

// Active bins (reduce_by_key, Broadphase.cu:439-443): the number of runs of
// equal keys, one atomic per workgroup of a capped grid.
__global__ __launch_bounds__(256) void k_count_runs(uint64_t mcap, const uint64_t* __restrict__ keys, DevStatus* st) {
  __shared__ uint32_t red[256 / WAVE];
  const uint64_t m = cap_count(mcap, &st->M);
  uint32_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256)
    c += (i == 0 || (keys[i] >> 32) != (keys[i - 1] >> 32)) ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 256 / WAVE; ++w) t += red[w];
    if (t) atomicAdd((unsigned long long*)&st->A, (unsigned long long)t);
  }
}

// Pair search with de-duplication by owning bin.  A pair (a, b) of
// overlapping AABBs (closed test on shifted coordinates, Broadphase.cu:237 /
// :296) is reported only from the bin whose key is the hash of the cell
// max(gmin_a, gmin_b): that cell lies in both AABBs' cell ranges, so both
// shapes appear in its run of sorted entries, and every other run skips the
// pair.  Repeated entries of one shape in a run (hash aliasing at index ==
// grid size, Broadphase.cu:140-147) are skipped, so each pair is emitted
// once and the reference's sort + unique (:522-535) reduces to a sort.
//
// One thread per sorted entry i pairs shape(i) with the later entries of
// the same run (shapes ascend inside a run, so shape(i) < shape(k)).  The
// workgroup's entries and the next PAIR_AHEAD are staged in LDS with their
// records; only runs reaching past the window read global memory.
constexpr int PAIR_BS = 256;
constexpr int PAIR_AHEAD = 128;

template <class R, bool EMIT>
__global__ __launch_bounds__(PAIR_BS) void k_pairs(uint64_t mcap, const uint64_t* __restrict__ keys,
                                                   const BpRec<R>* __restrict__ brec,
                                                   const SceneDesc<R>* __restrict__ desc,
                                                   uint32_t* __restrict__ pcount, const uint32_t* __restrict__ poffs,
                                                   uint64_t pcap, uint64_t* __restrict__ pairs, DevStatus* st) {
  constexpr int W = PAIR_BS + PAIR_AHEAD;
  __shared__ BpRec<R> srec[W];
  __shared__ uint64_t skey[W];
  const uint64_t m = cap_count(mcap, &st->M);
  const uint64_t b0 = (uint64_t)blockIdx.x * PAIR_BS;
  if (b0 >= m) return;
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
    const SceneDesc<R>& d = desc[ra.scene];
    const int gx = d.gx, gy = d.gy;
    const uint32_t kb = d.key_base;
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
      if (kb + hash_index(ox, oy, oz, gx, gy) != K) continue;
      if (EMIT) {
        if (o < pcap) pairs[o] = ((uint64_t)a << 32) | (uint64_t)b;
        else atomicOr(&st->overflow, 2u);
        ++o;
      } else {
        ++cnt;
      }
    }
  }
  if (!EMIT) pcount[i] = cnt;
}

// Per stepped scene: its ranges in the step's sorted entries, pair list and
// contact list (pairs are sorted by their first shape, and shapes of a
// scene are contiguous, so each scene's results are contiguous).
struct SceneOut {
  uint64_t m_begin, m_end, pair_begin, pair_end, contact_begin, contact_end;
};

template <class R>
__global__ void k_scene_ranges(const uint32_t* __restrict__ sel, const SceneDesc<R>* __restrict__ desc,
                               const DevStatus* __restrict__ st, uint64_t pcap, const uint32_t* __restrict__ offs,
                               const uint64_t* __restrict__ pairs, const uint32_t* __restrict__ hitpos,
                               SceneOut* __restrict__ out) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= st->nsel) return;
  const SceneDesc<R>& d = desc[sel[j]];
  const uint64_t nslot = st->nslots;
  const uint64_t P = cap_count(pcap, &st->P);
  auto entry_at = [&](uint64_t slot) -> uint64_t { return slot < nslot ? offs[slot] : st->M; };
  auto lower = [&](uint64_t key) -> uint64_t {
    uint64_t lo = 0, hi = P;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (pairs[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  SceneOut o;
  o.m_begin = entry_at((uint64_t)d.chunk_begin * CHUNK);
  o.m_end = entry_at((uint64_t)d.chunk_end * CHUNK);
  o.pair_begin = lower((uint64_t)d.shape_begin << 32);
  o.pair_end = lower((uint64_t)d.shape_end << 32);
  o.contact_begin = o.pair_begin < P ? hitpos[o.pair_begin] : st->C;
  o.contact_end = o.pair_end < P ? hitpos[o.pair_end] : st->C;
  out[j] = o;
}

}  // namespace
}  // namespace cdx
