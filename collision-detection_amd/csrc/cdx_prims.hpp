// Data-parallel primitives for gfx950 (64-wide wavefronts): wavefront scans
// built from DPP row shifts / row broadcasts, an LDS-staged block scan, a
// reduce-then-scan device scan (two kernels up to 2048 tiles of 4096), and a stable LSD radix sort of 64-bit values.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cdx {

constexpr int WAVE = 64;

// Inclusive add-scan across one 64-lane wavefront.  row_shr:1,2,4,8 scans
// each 16-lane row; row_bcast:15 carries row 0 into row 1 and row 2 into
// row 3; row_bcast:31 carries rows 0-1 into rows 2-3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  uint32_t t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false); v += t;
  return v;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  uint32_t l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Exclusive block scan of one value per thread; returns the prefix and
// stores the block total in *total.  scratch holds BS/64 + 1 uint32.
template <int BS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* scratch, uint32_t* total) {
  constexpr int NW = BS / WAVE;
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  uint32_t inc = wave_incl_scan(v);
  if (l == WAVE - 1) scratch[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < NW; ++i) { uint32_t c = scratch[i]; scratch[i] = run; run += c; }
    scratch[NW] = run;
  }
  __syncthreads();
  uint32_t r = scratch[w] + inc - v;
  *total = scratch[NW];
  __syncthreads();
  return r;
}

// ---- host launchers (prims.hip) ------------------------------------------
// Element counts are capacities known on the host; when d_n is given, the
// kernels process min(*d_n, capacity) elements, so a launch sequence can be
// recorded once (HIP graph) while the data-dependent sizes stay on the GPU.

// Exclusive scan of n uint32 (in may equal out).  The 64-bit total goes to
// *d_total (device memory).  tmp must hold scan_tmp_words(n) uint32.
size_t scan_tmp_words(uint64_t n);
void scan_excl_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, uint32_t* tmp, hipStream_t s,
                   const uint64_t* d_n = nullptr);

// The same scan when the kernel that produced `in` has already summed its
// tiles: part[t] holds the sum of elements [t * SCAN_TILE_ELEMS, (t + 1) *
// SCAN_TILE_ELEMS) (zero for tiles past the input), so the pass that reads
// the input only to sum it is skipped.  part holds scan_tile_count(n)
// uint64 and is scanned in place.
constexpr uint64_t SCAN_TILE_ELEMS = 4096;
inline uint64_t scan_tile_count(uint64_t n) { return (n + SCAN_TILE_ELEMS - 1) / SCAN_TILE_ELEMS + 1; }
void scan_excl_u32_tiled(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, uint64_t* part,
                         hipStream_t s, const uint64_t* d_n = nullptr);

// For producers of such an input: adds the values the workgroup's threads
// hold (element index i on each thread; one workgroup's elements lie in
// one tile) to the tile's partial with one atomic.  Every thread of the
// workgroup calls it (it synchronises the workgroup); red holds BS / WAVE
// words of LDS.
template <int BS>
__device__ __forceinline__ void add_tile_sum(uint64_t* part, uint64_t i, uint32_t v, uint32_t* red) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
#pragma unroll
    for (int w = 0; w < BS / WAVE; ++w) t += red[w];
    if (t) atomicAdd((unsigned long long*)&part[i / SCAN_TILE_ELEMS], (unsigned long long)t);
  }
}

// Stable LSD radix sort of n 64-bit values by bits [lo_bit, lo_bit + nbits),
// in the fewest passes of 8 to 11-bit digits whose per-tile digit counts
// stay small, ping-ponging between data and tmp.  Returns true when the
// result ends in tmp (an odd number of passes): the caller swaps the buffers
// instead of copying.  hist_tmp must hold radix_tmp_words(n, nbits).
size_t radix_tmp_words(uint64_t n, int nbits);
bool radix_sort_u64(uint64_t* data, uint64_t* tmp, uint64_t n, int lo_bit, int nbits, uint32_t* hist_tmp,
                    hipStream_t s, const uint64_t* d_n = nullptr);

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);

}  // namespace cdx
