// Internal interface between the C ABI (capi.cpp) and the device pipeline
// (pipeline.hip).  A Scene owns one HIP stream and the device-resident state
// of one collision scene: shapes and bodies in SoA layout, AABBs, the sorted
// (bin, body) entries, the pair list and the contacts.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace cdx {

enum RealKind : int { REAL_F64 = 0, REAL_F32 = 1 };
enum NpMode : int { NP_NATIVE = 0, NP_MIXED = 1, NP_FAST = 2 };

// Pipeline stages timed with HIP events on the scene's stream.
enum Stage : int {
  ST_AABB = 0,
  ST_BP_BOUNDS,
  ST_BP_COUNT,
  ST_BP_SCAN,
  ST_BP_EMIT,
  ST_BP_SORT,
  ST_BP_RUNS,   // kept for the stage numbering: active bins are summed inside ST_BP_PAIRSCAN (reads 0)
  ST_BP_PAIRS,  // pair counting pass
  ST_BP_PAIRSORT,
  ST_NP_MAIN,
  ST_NP_RESOLVE,
  ST_NP_COMPACT,
  ST_BP_PAIRSCAN,  // scan of per-entry pair counts (between the two pair passes)
  ST_BP_PAIREMIT,  // pair emission pass
  ST_NP_MPR,       // the reference-order MPR launch alone (inside ST_NP_RESOLVE)
  ST_COUNT
};

// Totals a step leaves on the device and copies back at synchronisation.
struct StepCounts {
  uint64_t n, m, a, p, c, resolved;
  uint32_t key_max;
  uint32_t bad;  // bit 0: unsupported shape type, bit 1: body id out of range
  float bmin[3], bmax[3];
  uint32_t info = 0;  // INFO_* bits
};

// StepCounts::info (reported with the flags of cdx_counts).
constexpr uint32_t INFO_CELL_SEARCH = 0x100;     // pairs found by the cell search
constexpr uint32_t INFO_NO_ACTIVE_BINS = 0x200;  // bin totals and active bins not computed (yet)

// Pair search of a scene's broadphase.
enum : int { PAIRS_AUTO = 0, PAIRS_BINS = 1, PAIRS_CELLS = 2 };

class SceneBase {
 public:
  virtual ~SceneBase() = default;
  virtual int real_kind() const = 0;

  // Inputs (host pointers, reference AoS layouts).
  // C (obj_data_C, the third triangle vertex) may be null: zeros.
  virtual void set_shapes(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                          const uint32_t* ID) = 0;
  virtual void set_bodies(uint64_t nb, const void* pos, const void* rot, const uint8_t* active) = 0;
  virtual void update_bodies(uint64_t k, const uint32_t* idx, const void* pos, const void* rot) = 0;
  virtual void set_aabbs(uint64_t n, const void* aabb) = 0;
  virtual void set_pairs(uint64_t p, const int64_t* pairs) = 0;

  // Settings.
  virtual void set_grid(int gx, int gy, int gz) = 0;
  virtual void set_envelope(double e) = 0;
  virtual void set_np_mode(int mode, double tau) = 0;
  virtual void set_profiling(bool on) = 0;
  // PAIRS_AUTO: the cell search for near-uniform shape sizes, else the
  // bins; PAIRS_BINS / PAIRS_CELLS force one (the cell search only where its
  // grid fits).
  virtual void set_pair_search(int mode) = 0;

  // Stages.  Each enqueues on the scene stream; the broadphase reads its
  // entry and pair totals back to size buffers, as the reference does.
  virtual void run_aabb() = 0;
  virtual void run_broadphase() = 0;
  virtual void run_narrowphase() = 0;
  virtual void sync() = 0;

  // Results (after sync()).
  virtual const StepCounts& counts() const = 0;
  // counts(), with the active bins of a cell-search broadphase computed.
  virtual const StepCounts& counts_full() = 0;
  virtual void read_aabb(void* out, bool shifted) = 0;
  virtual void read_bins(uint32_t* bin_counts, uint32_t* sorted_keys, uint32_t* sorted_bodies, uint32_t* active_keys,
                         uint32_t* active_counts) = 0;
  virtual void read_pairs(int64_t* out) = 0;
  virtual void read_pairs_permuted(int64_t* out) = 0;
  // Per input pair (narrowphase order before compaction): 1 for a contact.
  virtual void read_pair_hits(uint32_t* out) = 0;
  // Pairs (indices into the pair list) the reference-order MPR launch of the
  // last narrowphase ran on, in queue order; returns their count and writes
  // at most cap of them (out may be null).
  virtual uint64_t read_mpr_queue(uint32_t* out, uint64_t cap) = 0;
  virtual void read_contacts(void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids) = 0;
  virtual void stage_ms(double* out) = 0;
  virtual hipStream_t stream() const = 0;
  virtual uint64_t device_bytes() const = 0;
};

// Many resident scenes stepped together: each step covers a selection of
// scenes with one launch sequence (recorded as a HIP graph and replayed),
// buffers sized by capacities, totals read back once at finish().

class BatchBase {
 public:
  virtual ~BatchBase() = default;
  virtual int real_kind() const = 0;
  // Appends a scene (host pointers, reference layouts; body ids local to the
  // scene).  Returns its index.
  virtual uint32_t add_scene(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C,
                             const void* R, const uint32_t* ID, uint64_t nb, const void* pos, const void* rot,
                             const uint8_t* active, int gx, int gy, int gz) = 0;
  virtual uint64_t scene_count() const = 0;
  virtual void update_scene_bodies(uint32_t scene, uint64_t k, const uint32_t* idx, const void* pos,
                                   const void* rot) = 0;
  virtual void set_envelope(double e) = 0;
  virtual void set_np_mode(int mode, double tau) = 0;
  // Scenes of the following steps (null = all resident scenes).
  virtual void select(const uint32_t* scenes, uint64_t k) = 0;
  virtual uint64_t selected_count() const = 0;
  // Bin entries and pairs per selected shape the step buffers are sized for.
  virtual void set_capacity_factors(double m_per_shape, double p_per_shape) = 0;
  // Sizes every step buffer for steps of up to `shapes` shapes in `scenes`
  // scenes now, so that later steps neither allocate nor re-record.
  virtual void reserve_step(uint64_t shapes, uint64_t scenes) = 0;
  virtual void step(bool use_graph) = 0;  // enqueue AABB + broadphase + narrowphase
  virtual void finish() = 0;              // wait, re-run on capacity overflow, read totals
  virtual const StepCounts& totals() const = 0;
  // Per selected scene j (position in the selection), after finish().
  virtual void scene_counts(uint64_t j, StepCounts* out) = 0;
  virtual void read_scene_pairs(uint64_t j, int64_t* out) = 0;
  virtual void read_scene_contacts(uint64_t j, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids) = 0;
  virtual double last_step_ms() = 0;
  // Serving helpers.  Global body index of a scene's first body.
  virtual uint64_t scene_body_begin(uint32_t scene) const = 0;
  // Pose updates of bodies anywhere in the batch (global body indices),
  // from pinned host memory that stays untouched until finish() returns:
  // the copies are asynchronous on the batch stream.
  // The block holds k uint32 ids, then k real3 positions at byte offset
  // pose_block_pos_offset(k), then k real4 rotations at
  // pose_block_rot_offset(k).  first_body >= 0: the updates cover global
  // bodies first_body .. first_body+k-1 in order and the id list is unused.
  virtual void update_bodies_pinned(uint64_t k, const void* block, int64_t first_body) = 0;
  // The same update from a block already copied to device memory (same
  // layout) by another stream: the batch stream waits for `ready`, recorded
  // after that copy.  The block stays untouched until finish() returns.
  virtual void update_bodies_device(uint64_t k, const void* dev_block, int64_t first_body, hipEvent_t ready) = 0;
  // After finish(): contact range of selected scene j, and every contact of
  // the step (global body ids) in one copy of contacts_packed_bytes() bytes:
  // with C = totals().c and R the real type, normals (3C R), points A (3C R),
  // points B (3C R), depths (C R), then body id pairs (2C int32).
  virtual void scene_contact_range(uint64_t j, uint64_t* begin, uint64_t* end) const = 0;
  virtual uint64_t contacts_packed_bytes() const = 0;
  virtual void read_all_contacts_packed(void* dst) = 0;
  // The next step() also copies the first `rows` contact rows of that
  // packed layout (as laid out for the step's own count) to dst, pinned
  // host memory of at least rows rows untouched until finish() returns,
  // right behind the step on its stream.  After finish(),
  // contacts_prefetched() is true when the step's contacts all fit those
  // rows: dst then holds what read_all_contacts_packed would copy.
  virtual void prefetch_contacts(void* dst, uint64_t rows) = 0;
  virtual bool contacts_prefetched() const = 0;
  virtual uint64_t scene_bytes_estimate(uint64_t n_shapes, uint64_t n_bodies) const = 0;
  // Bytes a scene keeps resident, and the step buffers of a step over n shapes.
  virtual uint64_t resident_bytes_estimate(uint64_t n_shapes, uint64_t n_bodies) const = 0;
  virtual uint64_t step_bytes_estimate(uint64_t n_shapes) const = 0;
  virtual uint64_t device_bytes() const = 0;
  virtual hipStream_t stream() const = 0;
};

inline uint64_t pose_block_pos_offset(uint64_t k) { return (k * 4 + 255) & ~255ull; }
inline uint64_t pose_block_rot_offset(uint64_t k, int real_kind) {
  const uint64_t r = real_kind == REAL_F64 ? 8 : 4;
  return pose_block_pos_offset(k) + ((k * 3 * r + 255) & ~255ull);
}
inline uint64_t pose_block_bytes(uint64_t k, int real_kind) {
  return pose_block_rot_offset(k, real_kind) + k * 4 * (real_kind == REAL_F64 ? 8 : 4);
}

std::unique_ptr<SceneBase> make_scene(int real_kind, int device);
std::unique_ptr<BatchBase> make_batch(int real_kind, int device);

}  // namespace cdx
