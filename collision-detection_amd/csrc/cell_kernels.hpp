// Cell-list pair search for scenes of near-uniform shape size.  Included by
// pipeline.hip only.
//
// The reference finds candidate pairs through the bins every AABB spans
// (Broadphase.cu:393-535): M entries, a sort of all of them, per-bin tests,
// then sort + unique of the pairs.  The list that comes out does not depend
// on the grid: it is the set of shape pairs whose shifted AABBs overlap
// (closed test, Broadphase.cu:237, :296), because two overlapping boxes
// share a point and hence the bin of that point (the bin of a coordinate is
// a rounded multiply and a truncation, both monotone).  So any search that
// finds every overlapping pair reproduces it.
//
// Here every shape has one entry, the cell of its shifted minimum corner in
// an internal grid whose cells are at least as wide as the widest AABB along
// y and z, and 1/kx of that along x.  Two overlapping AABBs then have
// minimum corners at most one cell apart along y and z and kx cells apart
// along x, so a search of 9 rows of 2 kx + 1 consecutive cells around a
// shape finds all its partners.  The cell table is a histogram of the dense
// cell keys and its scan; the entries are put in cell order by a counting
// sort (each shape's rank in its cell comes from the histogram's atomic,
// so no sort pass runs).  Each shape keeps its partners of higher index, so
// each pair is found once, by its first shape, and the pair list comes out
// grouped by first shape; each group is sorted, which leaves the list
// ascending as sort + unique leaves it.
//
// Rounding: cells are floor(x * inv) in double, inv = 1 / edge along y and
// z and kx / edge along x, edge = (1 + 1e-3) * (W + 2^-20 * C), W the widest
// AABB extent as the bounds pass computes it (float differences of the
// unshifted corners) and C the largest coordinate magnitude of the scene,
// shifted or not.  The widest shifted extent, in exact arithmetic on the
// stored shifted corners, is below W * (1 + 2^-23) + 2^-22 * C (the float
// rounding of the corners, of their difference and of the shift), so
// below edge / (1 + 1e-3) * (1 + 1e-6).  For overlapping boxes lo_b - lo_a
// <= hi_a - lo_a, so the two products differ by less than (1 + 1e-6) /
// (1 + 1e-3) < 1 (kx times that along x) plus the double rounding of the
// products (below 1e-6 for any grid here): their floors differ by at most 1
// (kx along x).  Clamping to the grid keeps that.
#pragma once
#include "bp_kernels.hpp"

namespace cdx {
namespace {

// Search entry of a shape, in two arrays: its shifted minimum corner with
// its index (CellLo), then its shifted maximum corner (CellHi); the values
// are the additions of shifted_rec, so the overlap test is bit for bit the
// bin path's.
template <class R> struct alignas(16) CellLo {
  R lo[3];
  uint32_t id;
};
template <class R> struct alignas(16) CellHi {
  R hi[3];
};

// Cells per axis and inverse cell edges of the internal grid.  Cells are
// kx times finer along x than along y and z, so the x window of a search
// is 2 kx + 1 fine cells (about (2 + 1/kx) widest extents) instead of 3.
struct CellGrid {
  uint32_t g[3];
  uint32_t kx;
  double inv[3];
};

__device__ __forceinline__ uint32_t cell_axis(double x, double inv, uint32_t g) {
  const double f = floor(x * inv);
  if (!(f > 0.0)) return 0u;  // also NaN (such AABBs are reported before this search runs)
  if (f >= (double)(g - 1u)) return g - 1u;
  return (uint32_t)f;
}

template <class R> __device__ __forceinline__ void cell_of_box(const R lo[3], const CellGrid& cg, uint32_t c[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = cell_axis((double)lo[k], cg.inv[k], cg.g[k]);
}

__device__ __forceinline__ uint32_t cell_key(const uint32_t c[3], const CellGrid& cg) {
  return (c[2] * cg.g[1] + c[1]) * cg.g[0] + c[0];
}

// Per shape: one count in its cell's histogram slot (ccount zero
// beforehand), the shape's rank among the shapes of its cell (the count
// before its own), and the largest shape index of the cell (last, zero
// beforehand).  After a scan of the histogram, cell start + rank is the
// shape's place in cell order: a counting sort with no sort pass.  The
// order within a cell is the order of the atomics; nothing depends on it
// (every shape's partner group is sorted, and the walk reads whole cells).
template <class R>
__global__ __launch_bounds__(CHUNK) void k_cell_keys(const Chunk* __restrict__ chunks,
                                                     const DevStatus* __restrict__ st, uint64_t NS,
                                                     const R* __restrict__ aabb,
                                                     const SceneDesc<R>* __restrict__ desc, CellGrid cg,
                                                     uint32_t* __restrict__ ccount, uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ last) {
  if ((uint64_t)blockIdx.x * CHUNK >= st->nslots) return;
  const Chunk ch = chunks[blockIdx.x];
  if (threadIdx.x >= ch.len) return;
  const uint64_t i = (uint64_t)ch.begin + threadIdx.x;
  const BpRec<R> r = shifted_rec(aabb, NS, i, desc[ch.scene], ch.scene);
  uint32_t c[3];
  cell_of_box(r.lo, cg, c);
  const uint32_t key = cell_key(c, cg);
  rank[i] = atomicAdd(&ccount[key], 1u);
  atomicMax(&last[key], (uint32_t)i);
}

// Search entries in cell order (scene 0, whose shapes start at 0): shape
// i goes to cell start + rank.
template <class R>
__global__ __launch_bounds__(256) void k_cell_scatter(uint64_t n, const uint32_t* __restrict__ rank,
                                                      const uint32_t* __restrict__ cstart, uint64_t NS,
                                                      const R* __restrict__ aabb,
                                                      const SceneDesc<R>* __restrict__ desc, CellGrid cg,
                                                      CellLo<R>* __restrict__ slo, CellHi<R>* __restrict__ shi) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const BpRec<R> r = shifted_rec(aabb, NS, i, desc[0], 0u);
  uint32_t c[3];
  cell_of_box(r.lo, cg, c);
  const uint64_t p = (uint64_t)cstart[cell_key(c, cg)] + rank[i];
  CellLo<R> l{};
  CellHi<R> h{};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    l.lo[k] = r.lo[k];
    h.hi[k] = r.hi[k];
  }
  l.id = (uint32_t)i;
  slo[p] = l;
  shi[p] = h;
}

// Largest shape index over the window of 2 kx + 1 fine cells starting at c
// (cut at the end of c's grid row): a search row whose window maximum is
// not above the searching shape's index holds no partner of it.
__global__ __launch_bounds__(256) void k_cell_window(uint64_t cells, CellGrid cg, const uint32_t* __restrict__ last,
                                                     uint32_t* __restrict__ wmax) {
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= cells) return;
  const uint32_t x = (uint32_t)(c % cg.g[0]);
  const uint32_t w = min(2 * cg.kx, cg.g[0] - 1 - x);
  uint32_t m = 0;
  for (uint32_t k = 0; k <= w; ++k) m = max(m, last[c + k]);
  wmax[c] = m;
}

// Mean shapes per (coarse) cell up to which the automatic choice takes the
// cell search (a shape then tests about (2 + 1/kx) * 9 times as many).
constexpr double CELL_AUTO_DENSITY = 6.0;
constexpr int CELL_ROWS = 9;    // (dz, dy) rows of neighbour cells
constexpr int CELL_BS = 256;
constexpr int CELL_SLAB = 16;   // partners a shape keeps in the slab
constexpr int CELL_KX = 2;      // fine cells along x per coarse cell
constexpr int CELL_ROUND = 16;  // candidates loaded per round

// The shape being searched for: its corners and index.
template <class R> struct CellSelf {
  R lo[3], hi[3];
  uint32_t id;
};

// Candidate ranges of a shape in cell c: per (dz, dy) row, the sorted
// positions of cells x-kx .. x+kx (consecutive keys, so one range); rows
// outside the grid are empty.  The 18 table loads are independent.
// Rows whose window maximum (wmax) is not above a are left empty.
__device__ __forceinline__ void cell_ranges(const uint32_t c[3], const CellGrid& cg,
                                            const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ wmax,
                                            uint32_t a, uint32_t s[CELL_ROWS], uint32_t len[CELL_ROWS]) {
  const uint32_t gx = cg.g[0], gy = cg.g[1], gz = cg.g[2];
  const uint32_t x0 = c[0] > cg.kx ? c[0] - cg.kx : 0, x1 = min(c[0] + cg.kx, gx - 1);
#pragma unroll
  for (int r = 0; r < CELL_ROWS; ++r) {
    const int zz = (int)c[2] + r / 3 - 1, yy = (int)c[1] + r % 3 - 1;
    const bool in = zz >= 0 && zz < (int)gz && yy >= 0 && yy < (int)gy;
    const uint32_t row = in ? ((uint32_t)zz * gy + (uint32_t)yy) * gx : 0u;
    const uint32_t b = cstart[row + x0], e = cstart[row + x1 + 1], m = wmax[row + x0];
    s[r] = b;
    len[r] = in && m > a ? e - b : 0u;
  }
}

// Candidate range of one search row r (as in cell_ranges).
__device__ __forceinline__ void cell_row_range(const uint32_t c[3], const CellGrid& cg,
                                               const uint32_t* __restrict__ cstart,
                                               const uint32_t* __restrict__ wmax, uint32_t a, int r, uint32_t& s,
                                               uint32_t& len) {
  const uint32_t gx = cg.g[0], gy = cg.g[1], gz = cg.g[2];
  const uint32_t x0 = c[0] > cg.kx ? c[0] - cg.kx : 0, x1 = min(c[0] + cg.kx, gx - 1);
  const int zz = (int)c[2] + r / 3 - 1, yy = (int)c[1] + r % 3 - 1;
  const bool in = zz >= 0 && zz < (int)gz && yy >= 0 && yy < (int)gy;
  const uint32_t row = in ? ((uint32_t)zz * gy + (uint32_t)yy) * gx : 0u;
  const uint32_t b = cstart[row + x0], e = cstart[row + x1 + 1], m = wmax[row + x0];
  s = b;
  len = in && m > a ? e - b : 0u;
}

// Partners b > a of shape A, in walk order: hit(b) for each (closed
// overlap test, Broadphase.cu:237).  Row by row, CELL_ROUND candidates per
// round with all their loads issued first.  Going row by row keeps the
// lanes of a wavefront on neighbouring entries of the same row, so a load
// instruction touches few cache lines: the number of lines a wavefront's
// loads touch, not the bytes tested, is what this search's time follows
// (as measured: a flat walk across rows, and tile staging in LDS, both ran
// slower).
template <class R, class Hit>
__device__ __forceinline__ void cell_find(const CellSelf<R>& A, const uint32_t s[CELL_ROWS],
                                          const uint32_t len[CELL_ROWS], const CellLo<R>* __restrict__ slo,
                                          const CellHi<R>* __restrict__ shi, Hit hit) {
  constexpr int U = CELL_ROUND;
#pragma unroll
  for (int r = 0; r < CELL_ROWS; ++r) {
    const uint32_t b0 = s[r], nr = len[r];
    for (uint32_t j = 0; j < nr; j += U) {
      CellLo<R> L[U];
      CellHi<R> H[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < nr) {
          L[u] = slo[b0 + j + u];
          H[u] = shi[b0 + j + u];
        }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if ((j + u < nr) && (L[u].id > A.id) & (L[u].lo[0] <= A.hi[0]) & (L[u].lo[1] <= A.hi[1]) &
                                (L[u].lo[2] <= A.hi[2]) & (A.lo[0] <= H[u].hi[0]) & (A.lo[1] <= H[u].hi[1]) &
                                (A.lo[2] <= H[u].hi[2]))
          hit(L[u].id);
    }
  }
}

template <class R>
__device__ __forceinline__ CellSelf<R> cell_self(const CellLo<R>* __restrict__ slo, const CellHi<R>* __restrict__ shi,
                                                 uint64_t p) {
  const CellLo<R> l = slo[p];
  const CellHi<R> h = shi[p];
  CellSelf<R> A;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A.lo[k] = l.lo[k];
    A.hi[k] = h.hi[k];
  }
  A.id = l.id;
  return A;
}

// Tile of a workgroup: the blocks of one XCD (dispatched round-robin over
// the 8) take one contiguous run of tiles, so tiles whose candidates
// overlap share an L2.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t nb) {
  const uint64_t x = b % 8, k = b / 8, q = nb / 8, rem = nb % 8;
  return x * q + min(x, rem) + k;
}

// Wavefront-wide minimum and maximum (every lane active).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}

// LDS written by one lane and read by another of the same wavefront: the
// wavefront's LDS operations run in order; this keeps the compiler from
// moving them across each other.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The closed overlap test of candidate (L, H) against shape A with L's
// index above A's (Broadphase.cu:237), on the active lanes: the seven
// comparisons are combined as wavefront lane masks (scalar ANDs), which
// the compiler otherwise packs into per-lane bit vectors at twice the
// vector instructions.
template <class R>
__device__ __forceinline__ bool cell_hit(const CellSelf<R>& A, const CellLo<R>& L, const CellHi<R>& H) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(L.id > A.id) & __builtin_amdgcn_ballot_w64(L.lo[0] <= A.hi[0]) &
                     __builtin_amdgcn_ballot_w64(L.lo[1] <= A.hi[1]) & __builtin_amdgcn_ballot_w64(L.lo[2] <= A.hi[2]) &
                     __builtin_amdgcn_ballot_w64(A.lo[0] <= H.hi[0]) & __builtin_amdgcn_ballot_w64(A.lo[1] <= H.hi[1]) &
                     __builtin_amdgcn_ballot_w64(A.lo[2] <= H.hi[2]);
  return __builtin_amdgcn_inverse_ballot_w64(m);
}

// Entries a wavefront stages in LDS per window.
constexpr int CELL_WIN = 2 * WAVE;

// Search pass: per sorted position, the shape's partner count (scount)
// and its first CELL_SLAB partners in walk order (slab row of the shape);
// shapes with more are listed (ovf) to be searched again at placement.
//
// The 64 lanes of a wavefront hold neighbouring shapes in cell order, so
// their candidate ranges in a search row overlap.  Per row the wavefront
// loads the union of its lanes' ranges once, coalesced, in windows of
// CELL_WIN entries starting at the lowest position some lane still needs
// (lanes whose ranges lie further on wait for a later window: a wavefront
// that spans two grid rows takes two windows), and each lane tests its
// own range from LDS.  Each candidate record is then read from L2 once per
// wavefront instead of once per lane.
template <class R>
__global__ __launch_bounds__(CELL_BS) void k_cell_find(uint64_t n, const CellLo<R>* __restrict__ slo,
                                                       const CellHi<R>* __restrict__ shi,
                                                       const uint32_t* __restrict__ cstart,
                                                       const uint32_t* __restrict__ wmax, CellGrid cg,
                                                       uint32_t* __restrict__ scount, uint32_t* __restrict__ slab,
                                                       uint32_t* __restrict__ ovf, DevStatus* st) {
  // records move as 16-byte words (both record types are multiples of 16 bytes)
  constexpr int QL = sizeof(CellLo<R>) / sizeof(uint4), QH = sizeof(CellHi<R>) / sizeof(uint4);
  static_assert(sizeof(CellLo<R>) % sizeof(uint4) == 0 && sizeof(CellHi<R>) % sizeof(uint4) == 0, "record size");
  __shared__ uint4 l_lo[CELL_BS / WAVE][CELL_WIN * QL];
  __shared__ uint4 l_hi[CELL_BS / WAVE][CELL_WIN * QH];
  const uint64_t p = xcd_tile(blockIdx.x, gridDim.x) * CELL_BS + threadIdx.x;
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  uint4* wlo = l_lo[wv];
  uint4* whi = l_hi[wv];
  const uint4* glo = reinterpret_cast<const uint4*>(slo);
  const uint4* ghi = reinterpret_cast<const uint4*>(shi);
  const bool valid = p < n;  // no early exit: every lane takes part in the wavefront reductions
  CellSelf<R> A{};
  uint32_t c[3] = {0, 0, 0};
  uint32_t sn = 0, ln = 0;  // range of the next row, loaded one row ahead
  if (valid) {
    A = cell_self(slo, shi, p);
    cell_of_box(A.lo, cg, c);
    cell_row_range(c, cg, cstart, wmax, A.id, 0, sn, ln);
  }
  // partners go to the slab row four at a time: hb is a shift register whose
  // last (cnt & 3) words hold the partners not yet stored, oldest first
  uint32_t cnt = 0;
  uint4* row4 = reinterpret_cast<uint4*>(slab + (uint64_t)A.id * CELL_SLAB);
  uint4 hb = make_uint4(0u, 0u, 0u, 0u);
  for (int r = 0; r < CELL_ROWS; ++r) {
    uint32_t cur = sn;
    const uint32_t end = sn + ln;
    const uint64_t busy = __builtin_amdgcn_ballot_w64(ln != 0u);
    if (valid && r + 1 < CELL_ROWS) cell_row_range(c, cg, cstart, wmax, A.id, r + 1, sn, ln);
    if (!busy) continue;  // wave-uniform
    // Lanes hold consecutive positions in cell order and a row's start and
    // end are monotone in the cell key, so the busy lanes' ranges do not
    // decrease with the lane: the union ends at the last busy lane's end and
    // the lowest position still needed is the first unfinished lane's.  A
    // wavefront reduction backs both up (never taken when the order holds).
    uint32_t last = (uint32_t)__builtin_amdgcn_readlane((int)end, 63 - __builtin_clzll(busy));
    if (__builtin_amdgcn_ballot_w64(ln != 0u && end > last)) last = wave_max_u32(ln ? end : 0u);
    for (;;) {
      const uint64_t act = __builtin_amdgcn_ballot_w64(cur < end);
      if (!act) break;  // wave-uniform
      uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)cur, __builtin_ctzll(act));
      if (__builtin_amdgcn_ballot_w64(cur < end && cur < w0)) w0 = wave_min_u32(cur < end ? cur : ~0u);
      const uint32_t w1 = min(w0 + (uint32_t)CELL_WIN, last);
      const uint32_t k0 = w0 + lane, k1 = k0 + WAVE;
      if constexpr (QL == 1 && QH == 1) {
        // one-word records: loaded before the wavefront sync (measured ~1 %
        // faster than loading after it)
        uint4 L0 = {}, L1 = {}, H0 = {}, H1 = {};
        if (k0 < w1) {
          L0 = glo[k0];
          H0 = ghi[k0];
        }
        if (k1 < w1) {
          L1 = glo[k1];
          H1 = ghi[k1];
        }
        wave_lds_sync();  // the previous window's reads are done
        if (k0 < w1) {
          wlo[lane] = L0;
          whi[lane] = H0;
        }
        if (k1 < w1) {
          wlo[lane + WAVE] = L1;
          whi[lane + WAVE] = H1;
        }
      } else {
        // wider records go straight from memory to LDS after the sync (held
        // in registers across it, they were kept in per-lane scratch)
        wave_lds_sync();  // the previous window's reads are done
        if (k0 < w1) {
#pragma unroll
          for (int q = 0; q < QL; ++q) wlo[lane * QL + q] = glo[(uint64_t)k0 * QL + q];
#pragma unroll
          for (int q = 0; q < QH; ++q) whi[lane * QH + q] = ghi[(uint64_t)k0 * QH + q];
        }
        if (k1 < w1) {
#pragma unroll
          for (int q = 0; q < QL; ++q) wlo[(lane + WAVE) * QL + q] = glo[(uint64_t)k1 * QL + q];
#pragma unroll
          for (int q = 0; q < QH; ++q) whi[(lane + WAVE) * QH + q] = ghi[(uint64_t)k1 * QH + q];
        }
      }
      wave_lds_sync();
      if (cur < end && cur < w1) {
        const uint32_t e = min(end, w1);
        const CellLo<R>* tlo = reinterpret_cast<const CellLo<R>*>(wlo) - w0;
        const CellHi<R>* thi = reinterpret_cast<const CellHi<R>*>(whi) - w0;
        for (uint32_t j = cur; j < e; ++j) {
          const CellLo<R>& L = tlo[j];
          const CellHi<R>& H = thi[j];
          if (cell_hit(A, L, H)) {
            hb = make_uint4(hb.y, hb.z, hb.w, L.id);
            if ((cnt & 3u) == 3u && cnt < (uint32_t)CELL_SLAB) row4[cnt >> 2] = hb;
            ++cnt;
          }
        }
        cur = e;
      }
    }
  }
  // the last partial group: its k = cnt & 3 partners are hb's last k words,
  // moved to the front (slots past cnt are never read)
  const uint32_t k = cnt & 3u;
  if (k && cnt < (uint32_t)CELL_SLAB)
    row4[cnt >> 2] = make_uint4(k == 1u ? hb.w : (k == 2u ? hb.z : hb.y), k == 2u ? hb.w : hb.z, hb.w, 0u);
  if (!valid) return;
  scount[A.id] = cnt;
  if (cnt > (uint32_t)CELL_SLAB) ovf[atomicAdd(&st->novf, 1u)] = (uint32_t)p;
}

// Placement, one thread per shape: each group from the slab, sorted in
// registers, goes to LDS at its place in the workgroup's run of the pair
// list (the groups of consecutive shapes are consecutive there), and the
// run is written with coalesced stores.  Groups of listed shapes are left
// to k_cell_overflow; a workgroup whose run holds one, or does not fit
// LDS, writes its groups directly.
constexpr int CELL_PLACE_CAP = CELL_BS * CELL_SLAB;

__global__ __launch_bounds__(CELL_BS) void k_cell_place(uint64_t n, const uint32_t* __restrict__ scount,
                                                        const uint32_t* __restrict__ soff,
                                                        const uint32_t* __restrict__ slab, uint64_t pcap,
                                                        uint64_t* __restrict__ pairs, DevStatus* st) {
  __shared__ uint64_t l_run[CELL_PLACE_CAP];
  __shared__ uint32_t s_base, s_end, s_ovf;
  const uint64_t a0 = (uint64_t)blockIdx.x * CELL_BS;
  const uint64_t a = a0 + threadIdx.x;
  const uint64_t alast = min(a0 + CELL_BS, n) - 1;
  const bool valid = a < n;
  const uint32_t cnt = valid ? scount[a] : 0u;
  const uint64_t o = valid ? soff[a] : 0u;
  if (threadIdx.x == 0) {
    s_base = (uint32_t)o;
    s_ovf = 0;
  }
  if (a == alast) s_end = (uint32_t)(o + cnt);
  __syncthreads();
  if (cnt > (uint32_t)CELL_SLAB) s_ovf = 1;
  const bool fits = o + cnt <= pcap;
  if (valid && !fits) atomicOr(&st->overflow, 2u);
  __syncthreads();
  const uint64_t base = s_base, run = (uint64_t)s_end - base;
  const bool staged = !s_ovf && run <= (uint64_t)CELL_PLACE_CAP && base + run <= pcap;
  if (valid && fits && cnt > 0 && cnt <= (uint32_t)CELL_SLAB) {
    // every lane runs the same network: no divergence
    const uint32_t* row = slab + a * CELL_SLAB;
    uint32_t v[CELL_SLAB];
#pragma unroll
    for (int q = 0; q < CELL_SLAB; ++q) v[q] = (uint32_t)q < cnt ? row[q] : ~0u;
    sort_net<CELL_SLAB>(v);
    const uint64_t hi = a << 32;
#pragma unroll
    for (int q = 0; q < CELL_SLAB; ++q)
      if ((uint32_t)q < cnt) {
        if (staged) l_run[o - base + q] = hi | v[q];
        else pairs[o + q] = hi | v[q];
      }
  }
  if (!staged) return;
  __syncthreads();
  for (uint64_t q = threadIdx.x; q < run; q += CELL_BS) pairs[base + q] = l_run[q];
}

// Groups of the shapes with more than CELL_SLAB partners: searched again,
// written at soff[shape] and sorted there (groups longer than GROUP_LONG
// are listed for k_pair_groups_long).  Grid-stride over the list.
template <class R>
__global__ __launch_bounds__(CELL_BS) void k_cell_overflow(const uint32_t* __restrict__ ovf,
                                                           const CellLo<R>* __restrict__ slo,
                                                           const CellHi<R>* __restrict__ shi,
                                                           const uint32_t* __restrict__ cstart,
                                                           const uint32_t* __restrict__ wmax, CellGrid cg,
                                                           const uint32_t* __restrict__ scount,
                                                           const uint32_t* __restrict__ soff, uint64_t pcap,
                                                           uint64_t* __restrict__ pairs, uint32_t* __restrict__ longs,
                                                           uint64_t long_cap, DevStatus* st) {
  const uint32_t nov = st->novf;
  for (uint64_t k = (uint64_t)blockIdx.x * CELL_BS + threadIdx.x; k < nov; k += (uint64_t)gridDim.x * CELL_BS) {
    const uint64_t p = ovf[k];
    const CellSelf<R> A = cell_self(slo, shi, p);
    const uint32_t a = A.id, cnt = scount[a];
    const uint64_t o = soff[a];
    if (o + cnt > pcap) continue;  // reported by k_cell_place
    uint32_t c[3];
    cell_of_box(A.lo, cg, c);
    uint32_t s[CELL_ROWS], len[CELL_ROWS];
    cell_ranges(c, cg, cstart, wmax, a, s, len);
    const uint64_t hi = (uint64_t)a << 32;
    uint32_t w = 0;
    cell_find<R>(A, s, len, slo, shi, [&](uint32_t b) {
      if (w < cnt) pairs[o + w] = hi | b;
      ++w;
    });
    if (cnt <= (uint32_t)GROUP_LONG) {
      sort_pairs_in_place<GROUP_LONG>(pairs, o, cnt, a);
    } else {
      const uint32_t g = atomicAdd(&st->nbig, 1u);
      if (g < long_cap) longs[g] = a;  // scene 0 starts at shape 0: the shape is its slot
      else atomicOr(&st->overflow, 4u);
    }
  }
}

}  // namespace
}  // namespace cdx
