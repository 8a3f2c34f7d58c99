// Device-side data structures shared by the broadphase and narrowphase
// kernels.  Included by pipeline.hip only.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "cdx_math.hpp"
#include "cdx_prims.hpp"

namespace cdx {
namespace {

// Totals of one step, left on the device for the next kernels (and read by
// the host only at synchronisation).
struct DevStatus {
  uint64_t M;        // bin entries
  uint64_t A;        // active bins
  uint64_t P;        // unique candidate pairs
  uint64_t C;        // contacts
  uint64_t Q, Q2;    // MPR queue, FP64 re-solve queue
  uint64_t nslots;   // chunk slots of this step (chunks * CHUNK)
  uint32_t bad;      // 1 unsupported type, 2 body id out of range, 4 AABB outside the grid
  uint32_t overflow; // 1 bin entries, 2 pairs: capacity exceeded, results incomplete
  uint32_t key_max;
  uint32_t nsel;     // scenes in this step
  uint32_t nbig;     // pair groups too long for the in-register sort
  uint32_t wmax;     // widest AABB extent of the first stepped scene (float bits | WIDTH_INVERTED)
  uint64_t cells;    // shapes counted into the cell search's table (its scan total)
  uint32_t novf;     // cell search: shapes with more partners than the slab holds
  uint32_t pad;
  uint64_t qhn;      // entries of the MPR queue's per-workgroup class histograms
};

// What the host sets before a step (copied into DevStatus by k_begin_step,
// so a recorded step reads the current selection from device memory).
struct StepHdr {
  uint64_t nslots;
  uint32_t nsel;
  uint32_t pad;
};

constexpr int CHUNK = 256;

// A run of at most CHUNK consecutive shapes of one scene, processed by one
// workgroup.  Steps over several scenes are lists of chunks.
struct Chunk {
  uint32_t scene;
  uint32_t begin;  // global shape index
  uint32_t len;
  uint32_t pad;
};

// One resident scene: its shape and body ranges in the concatenated arrays,
// its grid, the base of its keys in the combined key space, and the
// per-step origin / inverse bin size computed on the device.
template <class R> struct SceneDesc {
  uint32_t shape_begin, shape_end, body_begin, body_end;
  int gx, gy, gz;
  uint32_t key_base;
  uint32_t chunk_begin, chunk_end;  // this step's chunk range (empty if not stepped)
  uint32_t bmin_enc[3], bmax_enc[3];
  R origin[3], inv[3];
  uint64_t rgx, rgy;  // floor((2^64 - 1) / g) + 1 for g = gx, gy: exact 32-bit division by multiply (0 for g = 1)
};

// n / d for n < 2^32 from r = floor((2^64 - 1) / d) + 1 (d >= 2): r exceeds
// 2^64 / d by less than 1, so n * r / 2^64 exceeds n / d by less than 2^-32,
// less than the gap 1/d between n / d and the next integer.
__device__ __forceinline__ uint32_t div_by_recip(uint32_t n, uint32_t d, uint64_t r) {
  return d == 1u ? n : (uint32_t)__umul64hi((uint64_t)n, r);
}
inline uint64_t recip_of(uint32_t d) { return d <= 1u ? 0ull : ~0ull / d + 1ull; }

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
  uint32_t body, active;  // as in ShapeRec: the front pass of MIXED / FAST reads only this record
};

template <class R> struct alignas(16) ShapeRec {
  R x, y, z;          // world centre: quatRotate(A, body_rot) + body_pos
  R qw, qx, qy, qz;   // world rotation: body_rot (x) local rot
  R bx, by, bz;       // dimensions
  uint32_t type, body, active, pad;
};

// Per-shape broadphase record: the shifted AABB, the shape's scene and the
// position of its first bin entry in emission order, gathered as one
// aligned block by the pair search (32 B in FP32, 64 B in FP64: one gather
// never straddles more lines than it must; cells are recomputed from it,
// one multiply and truncation per axis, bit-identical to k_count's).
template <class R> struct alignas(sizeof(R) == 4 ? 32 : 64) BpRec {
  R lo[3], hi[3];
  uint32_t scene;
  uint32_t offs;
};
static_assert(sizeof(BpRec<float>) == 32 && sizeof(BpRec<double>) == 64, "record sizes");

inline unsigned blocks_for(uint64_t n, unsigned bs) {
  uint64_t b = (n + bs - 1) / bs;
  return (unsigned)(b ? b : 1);
}

__device__ __forceinline__ uint32_t hash_index(uint32_t i, uint32_t j, uint32_t k, int gx, int gy) {
  // Broadphase.cu:140-147, evaluated in unsigned arithmetic (U3 argument).
  return ((k * (uint32_t)gy) * (uint32_t)gx) + (j * (uint32_t)gx) + i;
}

template <class R> __device__ __forceinline__ int cell_of(R x, R inv) { return (int)(x * inv); }

// a * b + c for a, b < 2^24: the low 32 bits of the product, one full-rate
// instruction (a 32-bit multiply is quarter rate, and the compiler cannot
// prove the operand range itself).
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t d;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

}  // namespace
}  // namespace cdx
