// Scan and sort primitives (declarations in cdx_prims.hpp).
#include "cdx_prims.hpp"
#include "cdx_util.hpp"

#include <algorithm>

namespace cdx {

namespace {

constexpr int SCAN_BS = 256;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IPT;
static_assert((uint64_t)SCAN_TILE == SCAN_TILE_ELEMS, "producers sum the scan's own tiles");

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ uint64_t eff_n(uint64_t cap, const uint64_t* d_n) {
  if (!d_n) return cap;
  uint64_t v = *d_n;
  return v < cap ? v : cap;
}

// SCAN_IPT consecutive uint32 of one thread: four 16-byte accesses when the
// run is whole and aligned, else one element at a time (the tile's tail).
__device__ __forceinline__ void load_run(const uint32_t* __restrict__ in, uint64_t base, uint64_t n,
                                         uint32_t (&v)[SCAN_IPT]) {
  if (base + SCAN_IPT <= n && (((uintptr_t)(in + base)) & 15) == 0) {
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; ++q) {
      const uint4 x = p[q];
      v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) v[i] = base + i < n ? in[base + i] : 0u;
  }
}

__device__ __forceinline__ void store_run(uint32_t* __restrict__ out, uint64_t base, uint64_t n,
                                          const uint32_t (&r)[SCAN_IPT]) {
  if (base + SCAN_IPT <= n && (((uintptr_t)(out + base)) & 15) == 0) {
    uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; ++q) p[q] = make_uint4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i)
      if (base + i < n) out[base + i] = r[i];
  }
}

// Per-tile sums (64-bit, so a tile of large counts cannot wrap).
__global__ __launch_bounds__(SCAN_BS) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t cap,
                                                         const uint64_t* __restrict__ d_n, uint64_t* __restrict__ part) {
  __shared__ uint64_t ws[SCAN_BS / WAVE];
  const uint64_t n = eff_n(cap, d_n);
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
  if (base + SCAN_TILE <= n && (((uintptr_t)(in + base)) & 15) == 0) {
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; ++q) {
      const uint4 x = p[q * SCAN_BS + threadIdx.x];
      s += (uint64_t)x.x + x.y + x.z + x.w;
    }
  } else if (base < n) {
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
      uint64_t e = base + (uint64_t)i * SCAN_BS + threadIdx.x;
      if (e < n) s += in[e];
    }
  }
  s = wave_sum_u64(s);
  if (threadIdx.x % WAVE == 0) ws[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < SCAN_BS / WAVE; ++w) t += ws[w];
    part[blockIdx.x] = t;
  }
}

// Exclusive scan of the tile sums by one block; writes the grand total.
// Each of the 1024 threads owns a contiguous slice of the sums: it adds its
// slice, one block scan gives the slice offsets, and it rewrites its slice.
// Slices are read eight sums at a time, so a long slice costs a few memory
// round trips rather than one per sum.
constexpr int PART_UNROLL = 8;

__global__ __launch_bounds__(1024) void k_scan_partials(uint64_t* __restrict__ part, uint64_t nb,
                                                        uint64_t* __restrict__ d_total) {
  __shared__ uint64_t ws[1024 / WAVE + 1];
  const uint64_t per = (nb + 1023) / 1024;
  const uint64_t b0 = min((uint64_t)threadIdx.x * per, nb), b1 = min(b0 + per, nb);
  uint64_t sum = 0;
  uint64_t i = b0;
  for (; i + PART_UNROLL <= b1; i += PART_UNROLL) {
    uint64_t x[PART_UNROLL];
#pragma unroll
    for (int k = 0; k < PART_UNROLL; ++k) x[k] = part[i + k];
#pragma unroll
    for (int k = 0; k < PART_UNROLL; ++k) sum += x[k];
  }
  for (; i < b1; ++i) sum += part[i];
  // block exclusive scan of the slice sums (64-bit)
  uint64_t inc = sum;
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint64_t t = __shfl_up(inc, o);
    if ((int)(threadIdx.x % WAVE) >= o) inc += t;
  }
  if (threadIdx.x % WAVE == WAVE - 1) ws[threadIdx.x / WAVE] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t run = 0;
    for (int w = 0; w < 1024 / WAVE; ++w) { const uint64_t c = ws[w]; ws[w] = run; run += c; }
    ws[1024 / WAVE] = run;
  }
  __syncthreads();
  uint64_t run = ws[threadIdx.x / WAVE] + inc - sum;
  i = b0;
  for (; i + PART_UNROLL <= b1; i += PART_UNROLL) {
    uint64_t x[PART_UNROLL];
#pragma unroll
    for (int k = 0; k < PART_UNROLL; ++k) x[k] = part[i + k];
#pragma unroll
    for (int k = 0; k < PART_UNROLL; ++k) {
      part[i + k] = run;
      run += x[k];
    }
  }
  for (; i < b1; ++i) {
    const uint64_t c = part[i];
    part[i] = run;
    run += c;
  }
  if (threadIdx.x == 0) *d_total = ws[1024 / WAVE];
}

// Tile scan: each thread owns SCAN_IPT consecutive elements.
__global__ __launch_bounds__(SCAN_BS) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        uint64_t cap, const uint64_t* __restrict__ d_n,
                                                        const uint64_t* __restrict__ part) {
  __shared__ uint32_t scratch[SCAN_BS / WAVE + 1];
  const uint64_t n = eff_n(cap, d_n);
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_IPT;
  uint32_t v[SCAN_IPT];
  load_run(in, base, n, v);
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) s += v[i];
  uint32_t tot;
  uint32_t pre = block_excl_scan<SCAN_BS>(s, scratch, &tot);
  uint32_t run = (uint32_t)part[blockIdx.x] + pre;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    const uint32_t x = v[i];
    v[i] = run;
    run += x;
  }
  store_run(out, base, n, v);
}

// Tile scan that also forms its own tile offset: the sum of the preceding
// tile sums, read by the whole workgroup (a few KB from L2 while the tile
// count is small), so no separate kernel scans the tile sums.  The last
// tile writes the grand total.  With no tile sums (part null) the input is
// one tile, whose 64-bit total the workgroup forms itself.
constexpr uint64_t SCAN_DIRECT_TILES = 2048;

__global__ __launch_bounds__(SCAN_BS) void k_scan_tiles_direct(const uint32_t* __restrict__ in,
                                                               uint32_t* __restrict__ out, uint64_t cap,
                                                               const uint64_t* __restrict__ d_n,
                                                               const uint64_t* __restrict__ part,
                                                               uint64_t* __restrict__ d_total) {
  __shared__ uint32_t scratch[SCAN_BS / WAVE + 1];
  __shared__ uint64_t ws[SCAN_BS / WAVE];
  const uint64_t n = eff_n(cap, d_n);
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (n == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_total = 0;
    return;
  }
  if (blockIdx.x >= nb) return;
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_IPT;
  uint32_t v[SCAN_IPT];
  load_run(in, base, n, v);
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) s += v[i];
  uint64_t pre_tiles = 0;
  if (part) {
    for (uint64_t i = threadIdx.x; i < blockIdx.x; i += SCAN_BS) pre_tiles += part[i];
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) pre_tiles += v[i];
  }
  pre_tiles = wave_sum_u64(pre_tiles);
  if (threadIdx.x % WAVE == 0) ws[threadIdx.x / WAVE] = pre_tiles;
  uint32_t tot;
  uint32_t pre = block_excl_scan<SCAN_BS>(s, scratch, &tot);  // synchronises: ws is complete
  uint64_t offset = 0;
#pragma unroll
  for (int w = 0; w < SCAN_BS / WAVE; ++w) offset += ws[w];
  const uint64_t total = part ? 0 : offset;  // one tile: ws held its total
  if (!part) offset = 0;
  uint32_t run = (uint32_t)offset + pre;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    const uint32_t x = v[i];
    v[i] = run;
    run += x;
  }
  store_run(out, base, n, v);
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) *d_total = part ? offset + part[blockIdx.x] : total;
}

__global__ void k_copy_u64(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, uint64_t cap,
                           const uint64_t* __restrict__ d_n) {
  const uint64_t n = eff_n(cap, d_n);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr int RS_BS = 256;
constexpr int RS_IPT = 16;
constexpr int RS_TILE = RS_BS * RS_IPT;
constexpr int RS_NW = RS_BS / WAVE;
constexpr int RS_WTILE = RS_IPT * WAVE;  // elements per wavefront: a contiguous run of the tile

// Lanes of the wavefront holding the same DB-bit digit d (DB ballots).
template <int DB> __device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < DB; ++b) {
    const uint64_t bb = __ballot((d >> b) & 1u);
    peers &= ((d >> b) & 1u) ? bb : ~bb;
  }
  return peers;
}

// Per-tile digit histogram: one LDS atomic per element.  (Counting equal
// digits of a wavefront once, by ballots over the digit bits, costs about
// six VALU instructions per digit bit per element and made this pass
// VALU-bound; a full 64-way same-address conflict costs less than that.)
template <int DB>
__global__ __launch_bounds__(RS_BS) void k_radix_upsweep(const uint64_t* __restrict__ in, uint64_t cap,
                                                         const uint64_t* __restrict__ d_n, int shift,
                                                         uint32_t* __restrict__ hist, uint32_t nb) {
  constexpr uint32_t ND = 1u << DB;
  __shared__ uint32_t h[ND];
  const uint64_t n = eff_n(cap, d_n);
  for (uint32_t d = threadIdx.x; d < ND; d += RS_BS) h[d] = 0;
  __syncthreads();
  const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE + (uint64_t)w * RS_WTILE + lane;
#pragma unroll 4
  for (int i = 0; i < RS_IPT; ++i) {
    const uint64_t e = base + (uint64_t)i * WAVE;
    if (e < n) atomicAdd(&h[(uint32_t)(in[e] >> shift) & (ND - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < ND; d += RS_BS) hist[(uint64_t)d * nb + blockIdx.x] = h[d];
}

// Stable scatter.  Tile order is (wavefront, item, lane), each wavefront a
// contiguous run of the input.  Pass 1 ranks every element among equal
// digits of its wavefront (ballots + a wavefront-private LDS counter per
// digit, no barriers).  The tile is then ordered by digit in LDS (a scan of
// the tile's digit counts gives each digit's place) and written out in LDS
// order: consecutive threads store consecutive positions of one digit's
// run, so the stores coalesce instead of scattering one element per lane.
template <int DB>
__global__ __launch_bounds__(RS_BS) void k_radix_downsweep(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                           uint64_t cap, const uint64_t* __restrict__ d_n, int shift,
                                                           const uint32_t* __restrict__ offs, uint32_t nb) {
  constexpr uint32_t ND = 1u << DB;
  constexpr int DPT = ND / RS_BS;  // digits per thread in the digit scan
  __shared__ uint32_t wcnt[RS_NW][ND];
  __shared__ uint32_t gbase[ND];  // global position of the digit's run minus its place in the tile
  __shared__ uint64_t sk[RS_TILE];
  __shared__ uint32_t scratch[RS_BS / WAVE + 1];
  const uint64_t n = eff_n(cap, d_n);
  const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  for (uint32_t d = threadIdx.x; d < ND; d += RS_BS)
#pragma unroll
    for (int ww = 0; ww < RS_NW; ++ww) wcnt[ww][d] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * RS_TILE;
  const uint64_t base = t0 + (uint64_t)w * RS_WTILE + lane;
  const uint64_t lt = lanemask_lt();
  uint64_t v[RS_IPT];
  uint32_t rk[RS_IPT];
#pragma unroll
  for (int i = 0; i < RS_IPT; ++i) {
    const uint64_t e = base + (uint64_t)i * WAVE;
    const bool valid = e < n;
    v[i] = valid ? in[e] : 0ull;
  }
#pragma unroll
  for (int i = 0; i < RS_IPT; ++i) {
    const uint64_t e = base + (uint64_t)i * WAVE;
    const bool valid = e < n;
    const uint32_t d = (uint32_t)(v[i] >> shift) & (ND - 1);
    const uint64_t peers = digit_peers<DB>(d, valid);
    const uint32_t before = wcnt[w][d];  // same value in every lane of the digit (no lane wrote it yet)
    rk[i] = before + (uint32_t)__popcll(peers & lt);
    // lanes of one wavefront execute in lockstep: every peer has read the
    // counter before the leader updates it
    __builtin_amdgcn_wave_barrier();
    if (valid && (peers & lt) == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // place of each digit in the tile (exclusive scan of the tile's digit
  // counts, DPT consecutive digits per thread), per-wavefront places inside
  // each digit, and the digit's global offset
  {
    uint32_t tot[DPT];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const uint32_t d = threadIdx.x * DPT + j;
      uint32_t c = 0;
#pragma unroll
      for (int ww = 0; ww < RS_NW; ++ww) c += wcnt[ww][d];
      tot[j] = c;
      sum += c;
    }
    uint32_t all;
    uint32_t place = block_excl_scan<RS_BS>(sum, scratch, &all);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const uint32_t d = threadIdx.x * DPT + j;
      gbase[d] = offs[(uint64_t)d * nb + blockIdx.x] - place;
      uint32_t run = place;
#pragma unroll
      for (int ww = 0; ww < RS_NW; ++ww) {
        const uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = run;
        run += c;
      }
      place += tot[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_IPT; ++i) {
    const uint64_t e = base + (uint64_t)i * WAVE;
    if (e < n) sk[wcnt[w][(uint32_t)(v[i] >> shift) & (ND - 1)] + rk[i]] = v[i];
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)min((uint64_t)RS_TILE, n > t0 ? n - t0 : 0);
  for (uint32_t t = threadIdx.x; t < cnt; t += RS_BS) {
    const uint64_t x = sk[t];
    out[gbase[(uint32_t)(x >> shift) & (ND - 1)] + t] = x;
  }
}

// Digit width per pass: the fewest passes whose digits have at most 11 bits
// and whose per-tile histograms stay within RS_HIST_CAP counts.
constexpr int RS_MIN_DB = 8, RS_MAX_DB = 11;
constexpr uint64_t RS_HIST_CAP = 8ull << 20;

struct RadixPlan {
  int passes, db;
};

RadixPlan radix_plan(uint64_t n, int nbits) {
  if (nbits <= 0) return {0, RS_MIN_DB};
  const uint64_t nb = std::max<uint64_t>((n + RS_TILE - 1) / RS_TILE, 1);
  const int p8 = (nbits + RS_MIN_DB - 1) / RS_MIN_DB;
  for (int p = 1; p < p8; ++p) {
    const int db = (nbits + p - 1) / p;
    if (db <= RS_MAX_DB && (1ull << db) * nb <= RS_HIST_CAP) return {p, db};
  }
  return {p8, RS_MIN_DB};
}

template <int DB>
void radix_pass(const uint64_t* src, uint64_t* dst, uint64_t n, const uint64_t* d_n, int shift, uint32_t* hist,
                uint64_t* total, uint32_t* stmp, uint32_t nb, hipStream_t s) {
  const uint64_t h = (uint64_t)(1u << DB) * nb;
  hipLaunchKernelGGL(k_radix_upsweep<DB>, dim3(nb), dim3(RS_BS), 0, s, src, n, d_n, shift, hist, nb);
  scan_excl_u32(hist, hist, h, total, stmp, s, nullptr);
  hipLaunchKernelGGL(k_radix_downsweep<DB>, dim3(nb), dim3(RS_BS), 0, s, src, dst, n, d_n, shift,
                     (const uint32_t*)hist, nb);
}

}  // namespace

size_t scan_tmp_words(uint64_t n) {
  uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  return 2 * (nb + 1);  // uint64 partials
}

void scan_excl_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, uint32_t* tmp, hipStream_t s,
                   const uint64_t* d_n) {
  uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint64_t* part = reinterpret_cast<uint64_t*>(tmp);
  if (n == 0) {
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, part, (uint64_t)0, d_total);
    return;
  }
  if (nb == 1) {  // one tile: one kernel, no tile sums
    hipLaunchKernelGGL(k_scan_tiles_direct, dim3(1), dim3(SCAN_BS), 0, s, in, out, n, d_n, (const uint64_t*)nullptr,
                       d_total);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(SCAN_BS), 0, s, in, n, d_n, part);
  if (nb <= SCAN_DIRECT_TILES) {
    hipLaunchKernelGGL(k_scan_tiles_direct, dim3((uint32_t)nb), dim3(SCAN_BS), 0, s, in, out, n, d_n,
                       (const uint64_t*)part, d_total);
    return;
  }
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, part, nb, d_total);
  hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)nb), dim3(SCAN_BS), 0, s, in, out, n, d_n, (const uint64_t*)part);
}

void scan_excl_u32_tiled(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, uint64_t* part,
                         hipStream_t s, const uint64_t* d_n) {
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (n == 0) {
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, part, (uint64_t)0, d_total);
    return;
  }
  if (nb <= SCAN_DIRECT_TILES) {
    hipLaunchKernelGGL(k_scan_tiles_direct, dim3((uint32_t)nb), dim3(SCAN_BS), 0, s, in, out, n, d_n,
                       (const uint64_t*)part, d_total);
    return;
  }
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, part, nb, d_total);
  hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)nb), dim3(SCAN_BS), 0, s, in, out, n, d_n, (const uint64_t*)part);
}

size_t radix_tmp_words(uint64_t n, int nbits) {
  const RadixPlan pl = radix_plan(n, nbits);
  const uint64_t nb = std::max<uint64_t>((n + RS_TILE - 1) / RS_TILE, 1);
  const uint64_t h = (uint64_t)(1u << pl.db) * nb;
  return h + scan_tmp_words(h) + 2;
}

bool radix_sort_u64(uint64_t* data, uint64_t* tmp, uint64_t n, int lo_bit, int nbits, uint32_t* hist_tmp,
                    hipStream_t s, const uint64_t* d_n) {
  const RadixPlan pl = radix_plan(n, nbits);
  if (n <= 1 || pl.passes == 0) return false;
  const uint32_t nb = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  const uint64_t h = (uint64_t)(1u << pl.db) * nb;
  uint32_t* hist = hist_tmp;
  uint32_t* stmp = hist_tmp + h;
  uint64_t* total = reinterpret_cast<uint64_t*>(hist_tmp + h + scan_tmp_words(h));
  uint64_t* src = data;
  uint64_t* dst = tmp;
  for (int p = 0; p < pl.passes; ++p) {
    const int shift = lo_bit + p * pl.db;
    switch (pl.db) {
      case 8: radix_pass<8>(src, dst, n, d_n, shift, hist, total, stmp, nb, s); break;
      case 9: radix_pass<9>(src, dst, n, d_n, shift, hist, total, stmp, nb, s); break;
      case 10: radix_pass<10>(src, dst, n, d_n, shift, hist, total, stmp, nb, s); break;
      default: radix_pass<11>(src, dst, n, d_n, shift, hist, total, stmp, nb, s); break;
    }
    uint64_t* t = src; src = dst; dst = t;
  }
  return src != data;
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, v, n);
}

}  // namespace cdx
