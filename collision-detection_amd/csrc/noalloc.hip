// Broadphase on caller device buffers that allocates nothing.
//
// The same kernels as the scene path's broadphase from caller AABBs
// (detectPossibleCollisions, Broadphase.cu:344-545): bounds, bin counts,
// scan, key store, key sort, owner-bin pair search, pair emission.  Every
// intermediate lives in `scratch`, one fixed device array of this
// translation unit, carved per call by the input size; the pairs go straight
// into the caller's output buffer.  No hipMalloc, hipFree, stream or event
// creation happens here: for callers that must not allocate (tight serving
// loops, harnesses that forbid it).
//
// The array's host-side name is not a device address: its address on the
// current device comes from hipGetSymbolAddress, taken once per device on
// the first call and kept, and only that pointer is handed to kernels.
// Each device holds its own copy until the process ends.  Calls on one
// device are serialised by a mutex (they share that device's array, which
// keeps the previous call's contents: everything a kernel accumulates into
// is cleared by the first kernel of each call).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "bp_kernels.hpp"
#include "cdx_noalloc.hpp"
#include "cdx_pipeline.hpp"
#include "cdx_util.hpp"

namespace cdx {

namespace {

constexpr uint64_t kScratchBytes = 48ull << 20;  // per device, for the process lifetime
constexpr uint64_t kMaxShapes = 65536;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchTable {
  std::mutex mu[kMaxDevices];
  void* ptr[kMaxDevices] = {};
};
ScratchTable& table() {
  static ScratchTable* t = new ScratchTable();  // lives until exit
  return *t;
}

// The device address of `scratch` on device dev (dev current); the caller
// holds table().mu[dev].
void* scratch_on(int dev) {
  ScratchTable& t = table();
  if (!t.ptr[dev]) {
    void* p = nullptr;
    CDX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CDX_REQUIRE(p, ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    t.ptr[dev] = p;
  }
  return t.ptr[dev];
}

// Step header, totals, the one scene's descriptor and the chunk table of a
// one-scene step, written from kernel arguments (no host copy) and with
// every accumulated total cleared.
template <class R>
__global__ void k_na_init(uint64_t n, int gx, int gy, int gz, uint64_t rgx, uint64_t rgy, DevStatus* st,
                          SceneDesc<R>* desc, uint32_t* sel, Chunk* chunks, uint64_t nchunks) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nchunks) {
    Chunk c;
    c.scene = 0;
    c.begin = (uint32_t)(t * CHUNK);
    c.len = (uint32_t)(n - t * CHUNK < (uint64_t)CHUNK ? n - t * CHUNK : (uint64_t)CHUNK);
    c.pad = 0;
    chunks[t] = c;
  }
  if (t != 0) return;
  DevStatus s{};
  s.nslots = nchunks * CHUNK;
  s.nsel = 1;
  *st = s;
  SceneDesc<R> d{};
  d.shape_begin = 0;
  d.shape_end = (uint32_t)n;
  d.body_begin = 0;
  d.body_end = (uint32_t)n;
  d.gx = gx; d.gy = gy; d.gz = gz;
  d.rgx = rgx; d.rgy = rgy;
  d.key_base = 0;
  d.chunk_begin = 0;
  d.chunk_end = (uint32_t)nchunks;
  *desc = d;
  sel[0] = 0;
}

// Caller AoS AABBs (2n real3, mins then maxes) into SoA blocks of stride NS.
template <class R>
__global__ void k_na_soa(uint64_t n, uint64_t NS, const R* __restrict__ aos, R* __restrict__ aabb) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * n) return;
  const uint64_t dst = e < n ? e : NS + (e - n);
  aabb[dst] = aos[3 * e];
  aabb[2 * NS + dst] = aos[3 * e + 1];
  aabb[4 * NS + dst] = aos[3 * e + 2];
}

// The AABBs shifted by the scene origin, back into the caller's AoS buffer
// (detectPossibleCollisions shifts in place, Broadphase.cu:382-386).
template <class R>
__global__ void k_na_shift_out(uint64_t n, uint64_t NS, const R* __restrict__ aabb,
                               const SceneDesc<R>* __restrict__ desc, R* __restrict__ aos) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * n) return;
  const uint64_t src = e < n ? e : NS + (e - n);
  aos[3 * e] = aabb[src] + desc->origin[0];
  aos[3 * e + 1] = aabb[2 * NS + src] + desc->origin[1];
  aos[3 * e + 2] = aabb[4 * NS + src] + desc->origin[2];
}

// Bump carving of the scratch array, 256-byte aligned.
struct Carve {
  char* base;
  uint64_t off = 0;
  template <class T> T* take(uint64_t count) {
    off = (off + 255) & ~255ull;
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

inline int bits_of(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

template <class R>
uint64_t broadphase(int dev, void* aabb_aos, uint64_t n, int gx, int gy, int gz, int64_t* pairs_out,
                    uint64_t pair_cap, hipStream_t s) {
  CDX_REQUIRE(n <= kMaxShapes, ERR_CAPACITY,
              "the allocation-free broadphase takes at most " + std::to_string(kMaxShapes) + " AABBs, got " +
                  std::to_string(n));
  const uint64_t key_space = (uint64_t)gx * ((uint64_t)gy * gz + gy + 1) + 1;
  CDX_REQUIRE(key_space <= (1ull << 32), ERR_INVALID, "bin key space exceeds 2^32: use fewer bins per axis");
  const int key_bits = bits_of(key_space - 1);
  const uint64_t nchunks = (n + CHUNK - 1) / CHUNK;
  const uint64_t slots = nchunks * CHUNK;
  const uint64_t NS = n;
  const uint64_t long_cap = std::min<uint64_t>(pair_cap / (GROUP_LONG + 1) + 1, n);

  Carve c{static_cast<char*>(scratch_on(dev))};
  DevStatus* st = c.take<DevStatus>(1);
  SceneDesc<R>* desc = c.take<SceneDesc<R>>(1);
  uint32_t* sel = c.take<uint32_t>(1);
  Chunk* chunks = c.take<Chunk>(nchunks);
  R* aabb = c.take<R>(6 * NS);
  BpRec<R>* brec = c.take<BpRec<R>>(NS);
  uint32_t* bpart = c.take<uint32_t>(BPART_WORDS * nchunks);
  uint32_t* kpart = c.take<uint32_t>(nchunks);
  uint32_t* counts = c.take<uint32_t>(slots);
  uint32_t* offs = c.take<uint32_t>(slots);
  uint32_t* scount = c.take<uint32_t>(slots);
  uint32_t* soff = c.take<uint32_t>(slots);
  uint32_t* scan_tmp = c.take<uint32_t>(std::max<size_t>(scan_tmp_words(slots), 16));
  uint32_t* longs = c.take<uint32_t>(long_cap);
  // the rest holds the bin entries: the largest capacity whose arrays fit
  const uint64_t fixed = (c.off + 255) & ~255ull;
  CDX_REQUIRE(fixed < kScratchBytes, ERR_CAPACITY, "scratch too small for the per-shape arrays");
  auto entry_bytes = [&](uint64_t m) {
    return m * (8 + 8 + sizeof(PairMeta)) + (m + 1) * 4 + (m / PAIR_BS + 1) * 4 +
           std::max<size_t>(radix_tmp_words(m, key_bits), 16) * 4 + 6 * 256;
  };
  uint64_t mcap = (kScratchBytes - fixed) / 40;
  while (mcap > 0 && fixed + entry_bytes(mcap) > kScratchBytes) mcap -= mcap / 32 + 1;
  CDX_REQUIRE(mcap > 0, ERR_CAPACITY, "scratch too small for any bin entries");
  uint64_t* keys = c.take<uint64_t>(mcap);
  uint64_t* keys_tmp = c.take<uint64_t>(mcap);
  PairMeta* meta = c.take<PairMeta>(mcap);
  uint32_t* pcount = c.take<uint32_t>(mcap + 1);
  uint32_t* runs_part = c.take<uint32_t>(mcap / PAIR_BS + 1);
  uint32_t* radix_tmp = c.take<uint32_t>(std::max<size_t>(radix_tmp_words(mcap, key_bits), 16));
  CDX_REQUIRE(c.off <= kScratchBytes, ERR_HIP, "internal error: scratch carving overran the array");

  const uint64_t rgx = recip_of((uint32_t)gx), rgy = recip_of((uint32_t)gy);
  hipLaunchKernelGGL(k_na_init<R>, dim3(blocks_for(std::max<uint64_t>(nchunks, 1), 256)), dim3(256), 0, s, n, gx, gy,
                     gz, rgx, rgy, st, desc, sel, chunks, nchunks);
  hipLaunchKernelGGL(k_na_soa<R>, dim3(blocks_for(2 * n, 256)), dim3(256), 0, s, n, NS, (const R*)aabb_aos, aabb);
  hipLaunchKernelGGL(k_bounds<R>, dim3((unsigned)nchunks), dim3(CHUNK), 0, s, (const Chunk*)chunks,
                     (const DevStatus*)st, NS, (const R*)aabb, bpart);
  hipLaunchKernelGGL(k_params<R>, dim3(1), dim3(PARAMS_BS), 0, s, (const uint32_t*)sel, st,
                     (const uint32_t*)bpart, desc);
  hipLaunchKernelGGL(k_count<R>, dim3((unsigned)nchunks), dim3(CHUNK), 0, s, (const Chunk*)chunks,
                     (const DevStatus*)st, NS, (const R*)aabb, (const SceneDesc<R>*)desc, counts, kpart, (uint64_t*)nullptr, st);
  hipLaunchKernelGGL(k_reduce_keymax, dim3(1), dim3(KEYMAX_BS), 0, s, (const uint32_t*)kpart, st);
  scan_excl_u32(counts, offs, slots, &st->M, scan_tmp, s, &st->nslots);
  hipLaunchKernelGGL(k_emit_keys<R>, dim3((unsigned)nchunks), dim3(CHUNK), 0, s, (const Chunk*)chunks,
                     (const DevStatus*)st, NS, (const R*)aabb, (const SceneDesc<R>*)desc, (const uint32_t*)offs,
                     (const uint32_t*)counts, mcap, keys, pcount, brec, (const uint32_t*)nullptr, st);
  if (radix_sort_u64(keys, keys_tmp, mcap, 32, key_bits, radix_tmp, s, &st->M)) std::swap(keys, keys_tmp);
  hipLaunchKernelGGL(k_pairs<R>, dim3(blocks_for(mcap, PAIR_BS)), dim3(PAIR_BS), 0, s, mcap, (const uint64_t*)keys,
                     (const BpRec<R>*)brec, (const SceneDesc<R>*)desc, pcount, meta, runs_part, st);
  hipLaunchKernelGGL(k_shape_pairs, dim3((unsigned)nchunks), dim3(CHUNK), 0, s, (const Chunk*)chunks, st,
                     (const uint32_t*)counts, (const uint32_t*)offs, (const uint32_t*)pcount, mcap, scount,
                     (uint64_t*)nullptr, (const uint32_t*)runs_part);
  scan_excl_u32(scount, soff, slots, &st->P, scan_tmp, s, &st->nslots);
  hipLaunchKernelGGL(k_pair_emit<R>, dim3((unsigned)nchunks), dim3(CHUNK), 0, s, (const Chunk*)chunks,
                     (const DevStatus*)st, (const uint32_t*)counts, (const uint32_t*)offs, (const uint32_t*)pcount,
                     (const PairMeta*)meta, (const uint32_t*)scount, (const uint32_t*)soff, (const uint64_t*)keys,
                     (const BpRec<R>*)brec, (const SceneDesc<R>*)desc, mcap, pair_cap,
                     reinterpret_cast<uint64_t*>(pairs_out), longs, long_cap, st);
  hipLaunchKernelGGL(k_pair_groups_long, dim3(256), dim3(256), 0, s, (const Chunk*)chunks, (const DevStatus*)st,
                     (const uint32_t*)scount, (const uint32_t*)soff, reinterpret_cast<uint64_t*>(pairs_out),
                     (const uint32_t*)longs, long_cap);
  hipLaunchKernelGGL(k_na_shift_out<R>, dim3(blocks_for(2 * n, 256)), dim3(256), 0, s, n, NS, (const R*)aabb,
                     (const SceneDesc<R>*)desc, (R*)aabb_aos);
  CDX_HIP_CHECK(hipGetLastError());
  DevStatus h;
  CDX_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
  CDX_HIP_CHECK(hipStreamSynchronize(s));
  CDX_REQUIRE(!(h.bad & 4u), ERR_INVALID,
              "an AABB maps outside the grid: non-finite coordinates or a zero scene extent");
  CDX_REQUIRE(!(h.bad & 8u), ERR_HIP, "internal error: a bin entry lies outside its shape's cell range");
  CDX_REQUIRE(h.M <= mcap && !(h.overflow & 1u), ERR_CAPACITY,
              std::to_string(h.M) + " bin entries exceed the " + std::to_string(mcap) +
                  " the scratch array holds at this size: use fewer bins per axis");
  CDX_REQUIRE(h.P <= pair_cap && !(h.overflow & 2u), ERR_CAPACITY,
              std::to_string(h.P) + " candidate pairs exceed the output capacity of " + std::to_string(pair_cap));
  CDX_REQUIRE(h.nbig <= long_cap && !(h.overflow & 4u), ERR_HIP, "internal error: more long pair groups than shapes");
  return h.P;
}

}  // namespace

uint64_t noalloc_scratch_bytes() { return kScratchBytes; }
uint64_t noalloc_max_shapes() { return kMaxShapes; }

uint64_t noalloc_broadphase(int device, int real_kind, void* aabb_aos, uint64_t n, int gx, int gy, int gz,
                            int64_t* pairs_out, uint64_t pair_cap, void* stream) {
  CDX_REQUIRE(device >= 0 && device < kMaxDevices, ERR_INVALID, "device index out of range");
  CDX_REQUIRE(gx > 0 && gy > 0 && gz > 0, ERR_INVALID, "bins per axis must be positive");
  CDX_REQUIRE(real_kind == REAL_F64 || real_kind == REAL_F32, ERR_INVALID, "unknown real kind");
  if (n == 0) return 0;
  CDX_REQUIRE(aabb_aos && pairs_out, ERR_INVALID, "null device buffer");
  std::lock_guard<std::mutex> lk(table().mu[device]);
  CDX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (real_kind == REAL_F64) return broadphase<double>(device, aabb_aos, n, gx, gy, gz, pairs_out, pair_cap, s);
  return broadphase<float>(device, aabb_aos, n, gx, gy, gz, pairs_out, pair_cap, s);
}

}  // namespace cdx
