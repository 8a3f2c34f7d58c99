// gfx950 collision pipeline: host orchestration of steps over resident
// scenes (one scene for the drop-in API, many for batched / served steps).
//
// Reference path (uwsbel/collision-detection):
//   AABB         AABBGenerator.cu:84-130   -> k_aabb
//   bounds       Broadphase.cu:368-381     -> k_aabb / k_bounds partials, k_params
//   count        Broadphase.cu:150-166     -> k_count
//   scan #1      Broadphase.cu:403-407     -> scan_excl_u32 over chunk slots
//   key store    Broadphase.cu:172-209     -> k_emit_keys
//   sort_by_key  Broadphase.cu:428-432     -> radix_sort_u64 on (key<<32 | shape)
//   reduce_by_key + scan #2  :439-480      -> run starts counted in k_pairs, summed in k_shape_pairs
//   pair count / store / sort / unique :484-535
//                                          -> k_pairs (owner-bin dedup), k_shape_pairs + scan,
//                                             k_pair_emit (per-shape sorted groups), k_pair_groups_long
//   MPR store    Narrowphase.cu:316-504    -> k_np_front, class sort, k_np_filter, k_np_resolve
//   compaction   Narrowphase.cu:580-593    -> scan + k_np_compact
//
// Two execution modes share every kernel:
//  - exact mode (the drop-in calls): the host reads the entry, pair and
//    queue totals back between stages and sizes every buffer exactly, as the
//    reference reads .back() after its scans;
//  - capacity mode (batched / served steps): buffers are sized by
//    capacities, every kernel reads its element count from device memory,
//    and the whole step is recorded once as a HIP graph and replayed.  A
//    step whose totals exceed the capacities is re-run with larger ones.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#include "bp_kernels.hpp"
#include "cdx_pipeline.hpp"
#include "cell_kernels.hpp"
#include "cdx_util.hpp"
#include "hostcopy.hpp"
#include "np_kernels.hpp"

namespace cdx {

namespace {

// ------------------------------------------------------------- utilities --

enum : int { ZERO_AABB = 1, ZERO_BP = 2, ZERO_NP = 4 };

// Opens a stage: copies the step header, zeroes the totals the stage
// produces (a kernel, so the sequence can be recorded as a graph).
// Also zeroes the tile sums the stage's producers add to (zbp: the
// broadphase's, znp: the narrowphase's).
__global__ void k_stage_begin(const StepHdr* __restrict__ hdr, DevStatus* st, uint64_t* __restrict__ zbp,
                              uint32_t nbp, uint64_t* __restrict__ znp, uint32_t nnp, int what) {
  if (what & ZERO_BP)
    for (uint32_t k = threadIdx.x; k < nbp; k += blockDim.x) zbp[k] = 0;
  if (what & ZERO_NP)
    for (uint32_t k = threadIdx.x; k < nnp; k += blockDim.x) znp[k] = 0;
  if (threadIdx.x != 0) return;
  st->nslots = hdr->nslots;
  st->nsel = hdr->nsel;
  if (what & ZERO_AABB) st->bad = 0;
  if (what & ZERO_BP) { st->M = st->A = st->P = 0; st->overflow = 0; st->key_max = 0; st->nbig = 0; st->wmax = 0; st->novf = 0; }
  if (what & ZERO_NP) { st->Q = st->Q2 = st->C = 0; }
}

// Chunk ranges and step-local key bases of the selected scenes, written
// into their descriptors (rng: scene, first chunk, end chunk, key base).
template <class R>
__global__ void k_apply_selection(uint64_t nsel, const uint32_t* __restrict__ rng, SceneDesc<R>* __restrict__ desc) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsel) return;
  SceneDesc<R>& d = desc[rng[4 * j]];
  d.chunk_begin = rng[4 * j + 1];
  d.chunk_end = rng[4 * j + 2];
  d.key_base = rng[4 * j + 3];
}

template <class R>
__global__ void k_deinterleave(uint64_t n, int k, const R* __restrict__ aos, R* __restrict__ soa, uint64_t stride,
                               uint64_t offset) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < k; ++c) soa[c * stride + offset + i] = aos[i * k + c];
}

template <class R>
__global__ void k_scatter_bodies(uint64_t k, uint64_t base, uint64_t nb_scene, uint64_t stride,
                                 const uint32_t* __restrict__ idx, const R* __restrict__ pos,
                                 const R* __restrict__ rot, R* __restrict__ bpos, R* __restrict__ brot) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  uint64_t b = idx ? idx[j] : j;  // no index list: bodies 0..k-1 of the scene
  if (b >= nb_scene) return;
  uint64_t g = base + b;
  for (int c = 0; c < 3; ++c) bpos[c * stride + g] = pos[3 * j + c];
  for (int c = 0; c < 4; ++c) brot[c * stride + g] = rot[4 * j + c];
}

// AABBs of scene 0 in the reference's AoS layout (2n real3), optionally
// shifted by the scene origin as detectPossibleCollisions leaves them.
template <class R>
__global__ void k_aabb_aos(uint64_t n, uint64_t NS, const R* __restrict__ aabb, const SceneDesc<R>* __restrict__ desc,
                           int shifted, R* __restrict__ out) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * n) return;
  uint64_t src = e < n ? e : NS + (e - n);
  R x = aabb[src], y = aabb[2 * NS + src], z = aabb[4 * NS + src];
  if (shifted) { x = x + desc[0].origin[0]; y = y + desc[0].origin[1]; z = z + desc[0].origin[2]; }
  out[3 * e] = x; out[3 * e + 1] = y; out[3 * e + 2] = z;
}

// Caller AABBs (2n real3, mins then maxes) into the SoA blocks of stride 2NS.
template <class R>
__global__ void k_aabb_from_aos(uint64_t n, uint64_t NS, const R* __restrict__ aos, R* __restrict__ aabb) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * n) return;
  uint64_t dst = e < n ? e : NS + (e - n);
  aabb[dst] = aos[3 * e];
  aabb[2 * NS + dst] = aos[3 * e + 1];
  aabb[4 * NS + dst] = aos[3 * e + 2];
}

// ------------------------------------------------------------ host side --

// Buffers grow only inside library calls, which hold the HIP gate shared,
// so no other thread is recording a graph meanwhile.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) {
      GateShared g;
      (void)hipFree(p);
    }
  }
  // Grows to at least b bytes; contents are not kept.
  void reserve(size_t b) {
    if (b <= bytes) return;
    GateShared g;
    release();
    size_t nb = b < 256 ? 256 : b;
    CDX_HIP_CHECK(hipMalloc(&p, nb));
    bytes = nb;
  }
  void release() {
    if (p) {
      GateShared g;
      // kernels queued on any stream may still read the old block
      CDX_HIP_CHECK(hipDeviceSynchronize());
      CDX_HIP_CHECK(hipFree(p));
      p = nullptr;
      bytes = 0;
    }
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Exchanges two buffers (a sort that ends in its ping-pong buffer).
inline void swap_bufs(DevBuf& a, DevBuf& b) {
  std::swap(a.p, b.p);
  std::swap(a.bytes, b.bytes);
}

struct HostScene {
  uint64_t shape_begin, n_shapes, body_begin, n_bodies;
  int g[3];
  uint64_t key_base, key_space;
};

inline uint64_t key_space_of(const int g[3]) {
  // Hash_Index values reach gx*(gy*gz + gy + 1) with max-corner aliasing.
  return (uint64_t)g[0] * ((uint64_t)g[1] * g[2] + g[1] + 1) + 1;
}

inline int bits_for(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

template <class R> class Engine final : public SceneBase, public BatchBase {
 public:
  explicit Engine(int device) : dev_(device) {
    CDX_HIP_CHECK(hipSetDevice(dev_));
    CDX_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    status_.reserve(sizeof(DevStatus));
    stephdr_.reserve(sizeof(StepHdr));
    CDX_HIP_CHECK(hipHostMalloc(&hstat_, sizeof(DevStatus), hipHostMallocDefault));
    CDX_HIP_CHECK(hipHostMalloc(&hhdr_, sizeof(StepHdr), hipHostMallocDefault));
    std::memset(hstat_, 0, sizeof(DevStatus));
    CDX_HIP_CHECK(hipMemsetAsync(status_.p, 0, sizeof(DevStatus), s_));
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
    for (int i = 0; i < 2 * ST_COUNT + 2; ++i) CDX_HIP_CHECK(hipEventCreate(&ev_[i]));
    std::memset(&counts_, 0, sizeof(counts_));
  }
  ~Engine() override {
    GateShared gate;
    (void)hipSetDevice(dev_);
    (void)hipStreamSynchronize(s_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    for (int i = 0; i < 2 * ST_COUNT + 2; ++i) (void)hipEventDestroy(ev_[i]);
    (void)hipHostFree(hstat_);
    (void)hipHostFree(hhdr_);
    if (ev_stage_) (void)hipEventDestroy(ev_stage_);
    if (stage_host_) (void)hipHostFree(stage_host_);
    if (souts_pinned_) (void)hipHostFree(souts_pinned_);
    (void)hipStreamDestroy(s_);
  }

  int real_kind() const override { return sizeof(R) == 8 ? REAL_F64 : REAL_F32; }
  hipStream_t stream() const override { return s_; }

  // ---------------------------------------------------- single scene API --

  void set_shapes(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* Rq,
                  const uint32_t* ID) override {
    const bool ext = check_types(n, T);
    sync_internal();
    single_ensure_scene0();
    const bool had = has_ext_;
    has_ext_ = scenes_.size() == 1 ? ext : (has_ext_ || ext);  // sticky over several scenes
    if (has_ext_ != had) graph_valid_ = false;
    ensure_shape_capacity(n);
    scenes_[0].n_shapes = n;
    upload_shapes(0, n, T, A, B, C, Rq, ID, 0);
    shapes_set_ = true;
    aabb_current_ = false;
    records_valid_ = false;
    update_desc(0);
  }

  void set_bodies(uint64_t nb, const void* pos, const void* rot, const uint8_t* active) override {
    sync_internal();
    single_ensure_scene0();
    ensure_body_capacity(nb);
    scenes_[0].n_bodies = nb;
    upload_bodies(0, nb, pos, rot, active);
    bodies_set_ = true;
    aabb_current_ = false;
    records_valid_ = false;
    update_desc(0);
  }

  void update_bodies(uint64_t k, const uint32_t* idx, const void* pos, const void* rot) override {
    update_scene_bodies(0, k, idx, pos, rot);
  }

  void set_aabbs(uint64_t n, const void* aabb) override {
    CDX_REQUIRE(n < (1ull << 31), ERR_INVALID, "too many AABBs (limit 2^31 - 1)");
    sync_internal();
    aabbs_changed();
    single_ensure_scene0();
    ensure_shape_capacity(n);
    scenes_[0].n_shapes = n;
    if (n) {
      stage_.reserve(n * 6 * sizeof(R));
      stager_.h2d(stage_.p, aabb, n * 6 * sizeof(R), s_);
      hipLaunchKernelGGL(k_aabb_from_aos<R>, dim3(blocks_for(2 * n, 256)), dim3(256), 0, s_, n, NSC_,
                         (const R*)stage_.as<R>(), aabb_.as<R>());
      sync_internal();
    }
    aabb_current_ = true;
    aabb_from_caller_ = true;
    records_valid_ = false;
    update_desc(0);
  }

  void set_pairs(uint64_t p, const int64_t* pairs) override {
    sync_internal();
    const uint64_t n = scenes_.empty() ? 0 : scenes_[0].n_shapes;
    for (uint64_t i = 0; i < p; ++i) {
      uint64_t a = (uint64_t)pairs[i] >> 32, b = (uint64_t)pairs[i] & 0xffffffffu;
      CDX_REQUIRE(a < n && b < n, ERR_INVALID, "pair " + std::to_string(i) + " addresses a shape out of range");
    }
    pcap_ = std::max<uint64_t>(p, 1);
    pairs_.reserve(pcap_ * 8);
    if (p) stager_.h2d(pairs_.p, pairs, p * 8, s_);
    given_pairs_ = p;
    pairs_given_ = true;
  }

  void set_grid(int gx, int gy, int gz) override {
    CDX_REQUIRE(gx > 0 && gy > 0 && gz > 0, ERR_INVALID, "bins per axis must be positive");
    single_ensure_scene0();
    scenes_[0].g[0] = gx; scenes_[0].g[1] = gy; scenes_[0].g[2] = gz;
    aabbs_changed();  // deferred bins of a cell-search broadphase would mix the two grids
    update_desc(0);
  }
  void set_envelope(double e) override {
    if ((R)e != envelope_) graph_valid_ = false;
    envelope_ = (R)e;
  }
  void set_np_mode(int mode, double tau) override {
    CDX_REQUIRE(mode == NP_NATIVE || mode == NP_MIXED || mode == NP_FAST, ERR_INVALID, "unknown narrowphase mode");
    if (mode != mode_ || (float)tau != tau_) graph_valid_ = false;
    mode_ = mode;
    tau_ = (float)tau;
  }
  void set_profiling(bool on) override { prof_ = on; }
  void set_pair_search(int mode) override {
    CDX_REQUIRE(mode == PAIRS_AUTO || mode == PAIRS_BINS || mode == PAIRS_CELLS, ERR_INVALID,
                "unknown pair search mode");
    pair_search_ = mode;
  }

  void run_aabb() override {
    CDX_REQUIRE(shapes_set_ && bodies_set_, ERR_NOT_READY, "the AABB stage needs shapes and bodies");
    select_single();
    exact_ = true;
    reset_stage_marks();
    stage_aabb();
    aabb_current_ = true;
    aabb_from_caller_ = false;
    records_valid_ = true;
  }

  void run_broadphase() override {
    CDX_REQUIRE(aabb_current_, ERR_NOT_READY, "broadphase needs AABBs: run the AABB stage or upload AABBs first");
    select_single();
    if (aabb_from_caller_) reset_stage_marks();
    exact_ = true;
    pairs_given_ = false;
    if (scenes_[0].n_shapes == 0) {
      M_ = A_ = P_ = 0;
      counts_.n = 0; counts_.m = counts_.a = counts_.p = 0;
      counts_.info = 0;
      cell_ = bins_pending_ = bins_lost_ = false;
      bins_counted_ = true;
      return;
    }
    stage_broadphase();
  }

  void run_narrowphase() override {
    CDX_REQUIRE(shapes_set_ && bodies_set_, ERR_NOT_READY, "narrowphase needs shapes and bodies");
    select_single();
    exact_ = true;
    if (!records_valid_) {
      // Records come from the AABB pass; build them when the caller runs
      // the narrowphase on its own (Narrowphase::DoNarrowphase).
      stage_aabb();
      records_valid_ = true;
    }
    if (pairs_given_) P_ = given_pairs_;
    if (P_ == 0) {
      C_ = 0;
      counts_.c = 0; counts_.resolved = 0; counts_.p = 0;
      return;
    }
    // the narrowphase reads its pair count from the device
    hstat_->P = P_;
    CDX_HIP_CHECK(hipMemcpyAsync(&st()->P, &hstat_->P, 8, hipMemcpyHostToDevice, s_));
    pcap_ = P_;
    stage_narrowphase();
  }

  void sync() override {
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
    CDX_HIP_CHECK(hipGetLastError());
  }

  const StepCounts& counts() const override { return counts_; }

  void read_aabb(void* out, bool shifted) override {
    const uint64_t n = scenes_.empty() ? 0 : scenes_[0].n_shapes;
    if (n == 0) return;
    tmp_out_.reserve(n * 6 * sizeof(R));
    hipLaunchKernelGGL(k_aabb_aos<R>, dim3(blocks_for(2 * n, 256)), dim3(256), 0, s_, n, NSC_,
                       (const R*)aabb_.as<R>(), (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), shifted ? 1 : 0,
                       tmp_out_.as<R>());
    stager_.d2h(out, tmp_out_.p, n * 6 * sizeof(R), s_);
  }

  void read_bins(uint32_t* bin_counts, uint32_t* sorted_keys, uint32_t* sorted_bodies, uint32_t* active_keys,
                 uint32_t* active_counts) override {
    sync();
    ensure_bin_counts();
    const uint64_t n = scenes_.empty() ? 0 : scenes_[0].n_shapes;
    // scene 0 starts at shape 0, so its slots are its shape indices
    if (bin_counts && n) CDX_HIP_CHECK(hipMemcpy(bin_counts, counts_buf_.p, n * 4, hipMemcpyDeviceToHost));
    if ((sorted_keys || sorted_bodies || active_keys || active_counts) && M_) {
      const uint64_t m = M_;
      const uint64_t runs = reduce_bins();
      const uint32_t* k = bins_.as<uint32_t>();
      const uint32_t* b = k + m;
      const uint32_t* occupancy = b + m;
      const uint32_t* akeys = occupancy + 2 * m;
      if (sorted_keys) CDX_HIP_CHECK(hipMemcpyAsync(sorted_keys, k, m * 4, hipMemcpyDeviceToHost, s_));
      if (sorted_bodies) CDX_HIP_CHECK(hipMemcpyAsync(sorted_bodies, b, m * 4, hipMemcpyDeviceToHost, s_));
      if (active_keys && runs) CDX_HIP_CHECK(hipMemcpyAsync(active_keys, akeys, runs * 4, hipMemcpyDeviceToHost, s_));
      if (active_counts && runs)
        CDX_HIP_CHECK(hipMemcpyAsync(active_counts, occupancy, runs * 4, hipMemcpyDeviceToHost, s_));
      sync();
    }
  }

  // reduce_by_key over the sorted bin entries on the device
  // (Broadphase.cu:439-443), only for callers that read the bins: bins_
  // then holds the sorted keys, the sorted shapes, the occupancy of each
  // active bin, a scratch block and the active keys, m words each.
  // Returns the number of active bins.
  uint64_t reduce_bins() {
    ensure_bins();
    const uint64_t m = M_;
    const size_t tmp_words = std::max<size_t>(scan_tmp_words(m), 16);
    bins_.reserve(m * 4 * 6 + 64 + tmp_words * 4);
    uint32_t* k = bins_.as<uint32_t>();
    uint32_t* b = k + m;
    uint32_t* heads = b + m;
    uint32_t* run_of = heads + m;
    uint32_t* akeys = run_of + m;
    uint32_t* rstart = akeys + m;
    uint64_t* d_runs = reinterpret_cast<uint64_t*>(rstart + m);  // 24m bytes in: 8-byte aligned
    uint32_t* tmp = reinterpret_cast<uint32_t*>(d_runs + 8);
    hipLaunchKernelGGL(k_bins_split, dim3(blocks_for(m, 256)), dim3(256), 0, s_, m,
                       (const uint64_t*)keys_.as<uint64_t>(), k, b, heads);
    scan_excl_u32(heads, run_of, m, d_runs, tmp, s_);
    hipLaunchKernelGGL(k_bins_runs, dim3(blocks_for(m, 256)), dim3(256), 0, s_, m, (const uint32_t*)k,
                       (const uint32_t*)heads, (const uint32_t*)run_of, akeys, rstart);
    uint64_t runs = 0;
    CDX_HIP_CHECK(hipMemcpyAsync(&runs, d_runs, 8, hipMemcpyDeviceToHost, s_));
    sync();
    if (cell_) {
      A_ = runs;
      counts_.a = runs;
      counts_.info &= ~INFO_NO_ACTIVE_BINS;
    }
    CDX_REQUIRE(runs == A_, ERR_HIP, "internal error: active bins of the pair search and reduce_by_key differ");
    if (runs)  // heads are no longer needed: the occupancy goes there
      hipLaunchKernelGGL(k_bins_occupancy, dim3(blocks_for(runs, 256)), dim3(256), 0, s_, runs, m,
                         (const uint32_t*)rstart, heads);
    return runs;
  }

  const StepCounts& counts_full() override {
    if (cell_ && (counts_.info & INFO_NO_ACTIVE_BINS) && !bins_lost_) {
      sync();
      ensure_bin_counts();
      if (M_) {
        reduce_bins();
      } else {
        counts_.a = A_ = 0;
        counts_.info &= ~INFO_NO_ACTIVE_BINS;
      }
      sync();
    }
    return counts_;
  }

  void read_pairs(int64_t* out) override {
    sync();
    if (P_) stager_.d2h(out, pairs_.p, P_ * 8, s_);
  }
  void read_pairs_permuted(int64_t* out) override {
    sync();
    if (P_) stager_.d2h(out, perm_.p, P_ * 8, s_);
  }
  uint64_t read_mpr_queue(uint32_t* out, uint64_t cap) override {
    sync();
    const uint64_t q = mpr_q_;  // set by exact-mode (scene) narrowphases only
    if (!out || !q) return q;
    const uint64_t k = std::min(q, cap);
    std::vector<uint64_t> e(k);
    const DevBuf& b = mpr_in_tmp_ ? queue_tmp_ : queue_;
    CDX_HIP_CHECK(hipMemcpy(e.data(), b.p, k * 8, hipMemcpyDeviceToHost));
    for (uint64_t j = 0; j < k; ++j) out[j] = (uint32_t)e[j];
    return q;
  }
  void read_pair_hits(uint32_t* out) override {
    sync();
    if (P_) CDX_HIP_CHECK(hipMemcpy(out, np_hit_.p, P_ * 4, hipMemcpyDeviceToHost));
  }
  void read_contacts(void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids) override {
    sync();
    copy_contacts(0, C_, norm, cpta, cptb, dpth, bids, 0);
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
    const DevBuf* all[] = {&type_, &body_, &lpos_, &dims_, &cpos_, &lrot_, &bpos_, &brot_, &active_, &aabb_, &rec_,
                           &brec_, &hdr_, &frec_, &bpart_, &kpart_, &counts_buf_, &offs_, &keys_, &keys_tmp_,
                           &pcount_, &pairs_, &np_v_, &np_hit_, &np_kind_, &np_pos_, &queue_,
                           &queue_tmp_, &qsel_, &qcls_, &tile_parts_, &out_, &perm_,
                           &scan_tmp_, &radix_tmp_, &runs_part_, &chunks_, &sel_, &desc_, &souts_, &longs_, &meta_, &scount_, &soff_, &bins_,
                           &cell_tab_, &cell_rank_, &cell_lo_, &cell_hi_, &cell_slab_, &cell_last_, &cell_wmax_, &cell_ovf_};
    uint64_t t = 0;
    for (auto* b : all) t += b->bytes;
    return t;
  }

  // ----------------------------------------------------------- batch API --

  uint32_t add_scene(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* Rq,
                     const uint32_t* ID,
                     uint64_t nb, const void* pos, const void* rot, const uint8_t* active, int gx, int gy,
                     int gz) override {
    CDX_REQUIRE(gx > 0 && gy > 0 && gz > 0, ERR_INVALID, "bins per axis must be positive");
    CDX_REQUIRE(n > 0 && nb > 0, ERR_INVALID, "a scene needs shapes and bodies");
    const bool ext = check_types(n, T);
    for (uint64_t i = 0; i < n; ++i)
      CDX_REQUIRE(ID[i] < nb, ERR_INVALID, "shape " + std::to_string(i) + " has a body id out of range");
    sync_internal();
    HostScene sc;
    sc.shape_begin = scenes_.empty() ? 0 : scenes_.back().shape_begin + scenes_.back().n_shapes;
    sc.body_begin = scenes_.empty() ? 0 : scenes_.back().body_begin + scenes_.back().n_bodies;
    sc.n_shapes = n;
    sc.n_bodies = nb;
    sc.g[0] = gx; sc.g[1] = gy; sc.g[2] = gz;
    sc.key_space = key_space_of(sc.g);
    sc.key_base = scenes_.empty() ? 0 : scenes_.back().key_base + scenes_.back().key_space;
    CDX_REQUIRE(sc.key_base + sc.key_space <= (1ull << 32), ERR_CAPACITY,
                "the combined bin key space of resident scenes exceeds 2^32");
    CDX_REQUIRE(sc.shape_begin + n < (1ull << 31), ERR_CAPACITY, "too many resident shapes (limit 2^31)");
    ensure_shape_capacity(sc.shape_begin + n);
    ensure_body_capacity(sc.body_begin + nb);
    scenes_.push_back(sc);
    const uint32_t s = (uint32_t)(scenes_.size() - 1);
    upload_shapes(sc.shape_begin, n, T, A, B, C, Rq, ID, sc.body_begin);
    upload_bodies(sc.body_begin, nb, pos, rot, active);
    update_desc(s);
    shapes_set_ = bodies_set_ = true;
    has_ext_ |= ext;
    graph_valid_ = false;
    return s;
  }

  uint64_t scene_count() const override { return scenes_.size(); }

  void update_scene_bodies(uint32_t s, uint64_t k, const uint32_t* idx, const void* pos, const void* rot) override {
    CDX_REQUIRE(s < scenes_.size(), ERR_INVALID, "scene index out of range");
    if (k == 0) return;
    const HostScene& sc = scenes_[s];
    for (uint64_t j = 0; j < k; ++j)
      CDX_REQUIRE(idx[j] < sc.n_bodies, ERR_INVALID, "body index out of range in a pose update");
    const uint64_t idx_bytes = (k * 4 + 15) & ~15ull;
    upd_.reserve(idx_bytes + k * 7 * sizeof(R));
    uint32_t* didx = upd_.as<uint32_t>();
    R* dpos = reinterpret_cast<R*>(upd_.as<char>() + idx_bytes);
    R* drot = dpos + 3 * k;
    CDX_HIP_CHECK(hipMemcpyAsync(didx, idx, k * 4, hipMemcpyHostToDevice, s_));
    CDX_HIP_CHECK(hipMemcpyAsync(dpos, pos, k * 3 * sizeof(R), hipMemcpyHostToDevice, s_));
    CDX_HIP_CHECK(hipMemcpyAsync(drot, rot, k * 4 * sizeof(R), hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL(k_scatter_bodies<R>, dim3(blocks_for(k, 256)), dim3(256), 0, s_, k, sc.body_begin,
                       sc.n_bodies, NBC_, (const uint32_t*)didx, (const R*)dpos, (const R*)drot, bpos_.as<R>(),
                       brot_.as<R>());
    // pageable sources: the copies above must finish before the caller reuses them
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
    aabb_current_ = false;
    records_valid_ = false;
  }

  void select(const uint32_t* scenes, uint64_t k) override {
    std::vector<uint32_t> sel;
    if (scenes) sel.assign(scenes, scenes + k);
    else for (uint32_t s = 0; s < scenes_.size(); ++s) sel.push_back(s);
    for (uint32_t s : sel) CDX_REQUIRE(s < scenes_.size(), ERR_INVALID, "scene index out of range");
    // no stream-wide wait: set_selection waits for the last copies out of
    // its staging block, and everything else it writes is ordered on the
    // stream (pose updates queued just before keep running meanwhile)
    set_selection(sel);
    explicit_selection_ = true;
  }

  uint64_t selected_count() const override { return selection_.size(); }

  void set_capacity_factors(double m_per_shape, double p_per_shape) override {
    CDX_REQUIRE(m_per_shape > 0 && p_per_shape > 0, ERR_INVALID, "capacity factors must be positive");
    cap_m_per_shape_ = m_per_shape;
    cap_p_per_shape_ = p_per_shape;
  }

  void step(bool use_graph) override {
    CDX_REQUIRE(!scenes_.empty(), ERR_NOT_READY, "no resident scenes");
    if (selection_dirty_ || !explicit_selection_) {
      if (selection_dirty_ || selection_.size() != scenes_.size()) {
        std::vector<uint32_t> all;
        for (uint32_t s = 0; s < scenes_.size(); ++s) all.push_back(s);
        sync_internal();
        set_selection(all);
      }
    }
    exact_ = false;
    aabb_from_caller_ = false;
    ensure_capacities();
    launch_capacity_step(use_graph);
    pending_step_ = true;
    pending_graph_ = use_graph;
  }

  void finish() override {
    sync();
    if (!pending_step_) return;
    pending_step_ = false;
    check_status();  // the totals were copied behind the step
    // Re-run with larger capacities until every total fits.
    for (int attempt = 0; attempt < 4; ++attempt) {
      const DevStatus& st = *hstat_;
      bool over = st.overflow != 0 || st.M > mcap_ || st.P > pcap_;
      if (!over) break;
      mcap_ = std::max<uint64_t>(mcap_, st.M + st.M / 4 + 1024);
      pcap_ = std::max<uint64_t>(pcap_, st.P + st.P / 4 + 1024);
      graph_valid_ = false;
      launch_capacity_step(pending_graph_);  // queues its copies again
      sync();
      check_status();
    }
    CDX_REQUIRE(hstat_->overflow == 0 && hstat_->M <= mcap_ && hstat_->P <= pcap_, ERR_CAPACITY,
                "step totals exceed the buffer capacities");
    M_ = hstat_->M; A_ = hstat_->A; P_ = hstat_->P; C_ = hstat_->C;
    m_seen_ = std::max(m_seen_, M_);
    p_seen_ = std::max(p_seen_, P_);
    souts_host_.resize(selection_.size());
    if (!selection_.empty())  // copied behind the step
      std::memcpy(souts_host_.data(), souts_pinned_, selection_.size() * sizeof(SceneOut));
    pf_ok_ = pf_dst_ != nullptr && C_ <= pf_rows_;
    pf_dst_ = nullptr;  // one step
    pf_rows_ = 0;
    counts_.n = total_selected_shapes();
    counts_.m = M_; counts_.a = A_; counts_.p = P_; counts_.c = C_;
    counts_.resolved = mode_ == NP_NATIVE ? 0 : hstat_->Q2;
    counts_.key_max = hstat_->key_max;
    counts_.bad = hstat_->bad;
  }

  const StepCounts& totals() const override { return counts_; }

  void scene_counts(uint64_t j, StepCounts* out) override {
    CDX_REQUIRE(j < souts_host_.size(), ERR_INVALID, "selection index out of range");
    const SceneOut& o = souts_host_[j];
    std::memset(out, 0, sizeof(*out));
    out->n = scenes_[selection_[j]].n_shapes;
    out->m = o.m_end - o.m_begin;
    out->p = o.pair_end - o.pair_begin;
    out->c = o.contact_end - o.contact_begin;
  }

  void read_scene_pairs(uint64_t j, int64_t* out) override {
    CDX_REQUIRE(j < souts_host_.size(), ERR_INVALID, "selection index out of range");
    const SceneOut& o = souts_host_[j];
    uint64_t np = o.pair_end - o.pair_begin;
    if (!np) return;
    CDX_HIP_CHECK(hipMemcpy(out, pairs_.as<uint64_t>() + o.pair_begin, np * 8, hipMemcpyDeviceToHost));
    const uint64_t sb = scenes_[selection_[j]].shape_begin;
    for (uint64_t i = 0; i < np; ++i) {
      uint64_t v = (uint64_t)out[i];
      out[i] = (int64_t)((((v >> 32) - sb) << 32) | ((v & 0xffffffffu) - sb));
    }
  }

  void read_scene_contacts(uint64_t j, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids) override {
    CDX_REQUIRE(j < souts_host_.size(), ERR_INVALID, "selection index out of range");
    const SceneOut& o = souts_host_[j];
    copy_contacts(o.contact_begin, o.contact_end - o.contact_begin, norm, cpta, cptb, dpth, bids,
                  scenes_[selection_[j]].body_begin);
  }

  double last_step_ms() override {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev_[2 * ST_COUNT], ev_[2 * ST_COUNT + 1]) != hipSuccess) return -1.0;
    return ms;
  }

  uint64_t scene_body_begin(uint32_t s) const override {
    CDX_REQUIRE(s < scenes_.size(), ERR_INVALID, "scene index out of range");
    return scenes_[s].body_begin;
  }

  void update_bodies_pinned(uint64_t k, const void* block, int64_t first_body) override {
    if (k == 0) return;
    const int kind = sizeof(R) == 8 ? REAL_F64 : REAL_F32;
    const uint64_t bytes = pose_block_bytes(k, kind);
    upd_.reserve(bytes);
    const char* hb = static_cast<const char*>(block);
    const uint64_t from = first_body >= 0 ? pose_block_pos_offset(k) : 0;  // contiguous: no id list
    CDX_HIP_CHECK(hipMemcpyAsync(upd_.as<char>() + from, hb + from, bytes - from, hipMemcpyHostToDevice, s_));
    scatter_bodies(upd_.as<char>(), k, first_body);
  }

  void update_bodies_device(uint64_t k, const void* dev_block, int64_t first_body, hipEvent_t ready) override {
    if (k == 0) return;
    CDX_REQUIRE(dev_block && ready, ERR_INVALID, "device pose block without its copy event");
    CDX_HIP_CHECK(hipStreamWaitEvent(s_, ready, 0));
    scatter_bodies(static_cast<const char*>(dev_block), k, first_body);
  }

  // Applies a pose block in device memory at `base` (layout of
  // update_bodies_pinned) on the batch stream.
  void scatter_bodies(const char* base, uint64_t k, int64_t first_body) {
    const uint64_t nbt = scenes_.empty() ? 0 : scenes_.back().body_begin + scenes_.back().n_bodies;
    const int kind = sizeof(R) == 8 ? REAL_F64 : REAL_F32;
    const uint32_t* didx = first_body >= 0 ? nullptr : (const uint32_t*)base;
    const uint64_t b0 = first_body >= 0 ? (uint64_t)first_body : 0;
    CDX_REQUIRE(first_body < 0 || b0 + k <= nbt, ERR_INVALID, "pose update past the resident bodies");
    // listed indices past the resident bodies are dropped by the kernel
    hipLaunchKernelGGL(k_scatter_bodies<R>, dim3(blocks_for(k, 256)), dim3(256), 0, s_, k, b0,
                       first_body >= 0 ? k : nbt, NBC_, didx, (const R*)(base + pose_block_pos_offset(k)),
                       (const R*)(base + pose_block_rot_offset(k, kind)), bpos_.as<R>(), brot_.as<R>());
    aabb_current_ = false;
    records_valid_ = false;
  }

  void scene_contact_range(uint64_t j, uint64_t* begin, uint64_t* end) const override {
    CDX_REQUIRE(j < souts_host_.size(), ERR_INVALID, "selection index out of range");
    *begin = souts_host_[j].contact_begin;
    *end = souts_host_[j].contact_end;
  }

  uint64_t contacts_packed_bytes() const override { return C_ * (10 * sizeof(R) + 8); }

  void prefetch_contacts(void* dst, uint64_t rows) override {
    pf_dst_ = rows ? dst : nullptr;
    pf_rows_ = dst ? rows : 0;
  }
  bool contacts_prefetched() const override { return pf_ok_; }

  void read_all_contacts_packed(void* dst) override {
    if (!C_) return;
    CDX_HIP_CHECK(hipMemcpyAsync(dst, out_.p, contacts_packed_bytes(), hipMemcpyDeviceToHost, s_));
    sync();
  }

  void reserve_step(uint64_t shapes, uint64_t scenes) override {
    sync_internal();
    const uint64_t need_m = (uint64_t)(cap_m_per_shape_ * (double)shapes) + 4096;
    const uint64_t need_p = (uint64_t)(cap_p_per_shape_ * (double)shapes) + 4096;
    if (need_m > mcap_) { mcap_ = need_m; graph_valid_ = false; }
    if (need_p > pcap_) { pcap_ = need_p; graph_valid_ = false; }
    const uint64_t chunks = shapes / CHUNK + scenes + 1;
    if (chunks > chunk_cap_) grow_chunk_cap(chunks);
    if (scenes > sel_cap_) grow_sel_cap(scenes);
    prepare_capacity_buffers();
  }

  uint64_t resident_bytes_estimate(uint64_t n, uint64_t nb) const override {
    // inputs in SoA plus the per-shape records each step rewrites
    const uint64_t per_shape = 1 + 4 + 13 * sizeof(R) + 6 * sizeof(R) + sizeof(ShapeRec<R>) + sizeof(BpRec<R>) +
                               sizeof(ShapeHdr) + sizeof(ShapeRecF);
    const uint64_t per_body = 7 * sizeof(R) + 1;
    return n * per_shape + nb * per_body;
  }

  uint64_t step_bytes_estimate(uint64_t n) const override {
    // keys + ping-pong and per-entry pair counts/offsets; pairs, narrowphase
    // records, queues and contact outputs per pair; slot counts/offsets
    const double per_m = 2 * 8 + 4 + 16;
    const double per_p = 2 * 8 + sizeof(PairRec<R>) + 4 + 1 + 4 + 2 * 8 + 4 + 1 + 10 * sizeof(R) + 8 + 8;
    return (uint64_t)((double)n * (cap_m_per_shape_ * per_m + cap_p_per_shape_ * per_p + 8.0)) + (1u << 20);
  }

  uint64_t scene_bytes_estimate(uint64_t n, uint64_t nb) const override {
    return resident_bytes_estimate(n, nb) + step_bytes_estimate(n);
  }

 private:
  // ------------------------------------------------------------- uploads --
  // Returns whether a rectangle, cone or triangle is among the shapes.
  static bool check_types(uint64_t n, const int32_t* T) {
    CDX_REQUIRE(n < (1ull << 31), ERR_INVALID, "too many shapes (limit 2^31 - 1)");
    bool ext = false;
    for (uint64_t i = 0; i < n; ++i) {
      int t = T[i];
      CDX_REQUIRE(t >= SPHERE && t <= CAPSULE, ERR_UNSUPPORTED_SHAPE,
                  "shape " + std::to_string(i) + " has type " + std::to_string(t) +
                      ": shape types are SPHERE..TRIANGLEMESH (0-6) and CAPSULE (7)");
      ext |= t >= RECT && t <= TRIANGLEMESH;
    }
    return ext;
  }

  uint64_t total_selected_shapes() const {
    uint64_t n = 0;
    for (uint32_t s : selection_) n += scenes_[s].n_shapes;
    return n;
  }

  void single_ensure_scene0() {
    if (!scenes_.empty()) return;
    HostScene sc{};
    sc.g[0] = sc.g[1] = sc.g[2] = 16;  // Broadphase.cu:37 default bins per axis
    sc.key_space = key_space_of(sc.g);
    scenes_.push_back(sc);
    desc_host_.resize(1);
    selection_dirty_ = true;
  }

  // Grows the SoA arrays of resident shapes / bodies, keeping contents.
  void ensure_shape_capacity(uint64_t need) {
    if (need <= NSC_ && type_.p) return;
    uint64_t nc = std::max<uint64_t>(need, std::max<uint64_t>(NSC_ * 2, 1024));
    uint64_t keep = 0;
    for (auto& sc : scenes_) keep = std::max(keep, sc.shape_begin + sc.n_shapes);
    grow_soa(type_, 1, 1, nc, keep);
    grow_soa(body_, 4, 1, nc, keep);
    grow_soa(lpos_, sizeof(R), 3, nc, keep);
    grow_soa(dims_, sizeof(R), 3, nc, keep);
    grow_soa(cpos_, sizeof(R), 3, nc, keep);
    grow_soa(lrot_, sizeof(R), 4, nc, keep);
    aabb_.release();
    aabb_.reserve(nc * 6 * sizeof(R));
    aabb_current_ = false;
    rec_.reserve(nc * sizeof(ShapeRec<R>));
    brec_.reserve(nc * sizeof(BpRec<R>));
    hdr_.reserve(nc * sizeof(ShapeHdr));
    frec_.reserve(nc * sizeof(ShapeRecF));
    NSC_ = nc;
    graph_valid_ = false;
  }
  void ensure_body_capacity(uint64_t need) {
    if (need <= NBC_ && active_.p) return;
    uint64_t nc = std::max<uint64_t>(need, std::max<uint64_t>(NBC_ * 2, 1024));
    uint64_t keep = 0;
    for (auto& sc : scenes_) keep = std::max(keep, sc.body_begin + sc.n_bodies);
    grow_soa(bpos_, sizeof(R), 3, nc, keep);
    grow_soa(brot_, sizeof(R), 4, nc, keep);
    grow_soa(active_, 1, 1, nc, keep);
    NBC_ = nc;
    graph_valid_ = false;
  }
  // k components of `elem` bytes, stride old capacity -> new capacity.
  void grow_soa(DevBuf& buf, size_t elem, int k, uint64_t ncap, uint64_t keep) {
    DevBuf nb;
    nb.reserve(ncap * k * elem);
    if (buf.p && keep) {
      const uint64_t ocap = buf.bytes / (k * elem);
      for (int c = 0; c < k; ++c)
        CDX_HIP_CHECK(hipMemcpy(nb.as<char>() + c * ncap * elem, buf.as<char>() + c * ocap * elem, keep * elem,
                                hipMemcpyDeviceToDevice));
    }
    buf.release();
    buf.p = nb.p;
    buf.bytes = ncap * k * elem;
    nb.p = nullptr;
    nb.bytes = 0;
  }

  void upload_soa(const void* host, uint64_t n, int k, R* dst, uint64_t stride, uint64_t offset) {
    stage_.reserve(n * k * sizeof(R));
    stager_.h2d(stage_.p, host, n * k * sizeof(R), s_);
    hipLaunchKernelGGL(k_deinterleave<R>, dim3(blocks_for(n, 256)), dim3(256), 0, s_, n, k, (const R*)stage_.as<R>(),
                       dst, stride, offset);
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
  }

  void upload_shapes(uint64_t at, uint64_t n, const int32_t* T, const void* A, const void* B, const void* C,
                     const void* Rq, const uint32_t* ID, uint64_t body_base) {
    if (!n) return;
    std::vector<uint8_t> t8(n);
    std::vector<uint32_t> ids(n);
    for (uint64_t i = 0; i < n; ++i) {
      t8[i] = (uint8_t)T[i];
      ids[i] = (uint32_t)(ID[i] + body_base);
    }
    stager_.h2d(type_.as<uint8_t>() + at, t8.data(), n, s_);
    stager_.h2d(body_.as<uint32_t>() + at, ids.data(), n * 4, s_);
    CDX_HIP_CHECK(hipStreamSynchronize(s_));  // t8 and ids are freed on return
    upload_soa(A, n, 3, lpos_.as<R>(), NSC_, at);
    upload_soa(B, n, 3, dims_.as<R>(), NSC_, at);
    if (C) {
      upload_soa(C, n, 3, cpos_.as<R>(), NSC_, at);
    } else {
      for (int c = 0; c < 3; ++c) CDX_HIP_CHECK(hipMemset(cpos_.as<R>() + c * NSC_ + at, 0, n * sizeof(R)));
    }
    upload_soa(Rq, n, 4, lrot_.as<R>(), NSC_, at);
  }

  void upload_bodies(uint64_t at, uint64_t nb, const void* pos, const void* rot, const uint8_t* active) {
    if (!nb) return;
    upload_soa(pos, nb, 3, bpos_.as<R>(), NBC_, at);
    upload_soa(rot, nb, 4, brot_.as<R>(), NBC_, at);
    if (active) {
      stager_.h2d(active_.as<uint8_t>() + at, active, nb, s_);
      CDX_HIP_CHECK(hipStreamSynchronize(s_));
    }
    else CDX_HIP_CHECK(hipMemset(active_.as<uint8_t>() + at, 1, nb));
  }

  void update_desc(uint32_t s) {
    HostScene& sc = scenes_[s];
    sc.key_space = key_space_of(sc.g);
    if (s == 0) sc.key_base = 0;
    desc_host_.resize(scenes_.size());
    SceneDesc<R>& d = desc_host_[s];
    std::memset(&d, 0, sizeof(d));
    d.shape_begin = (uint32_t)sc.shape_begin;
    d.shape_end = (uint32_t)(sc.shape_begin + sc.n_shapes);
    d.body_begin = (uint32_t)sc.body_begin;
    d.body_end = (uint32_t)(sc.body_begin + sc.n_bodies);
    d.gx = sc.g[0]; d.gy = sc.g[1]; d.gz = sc.g[2];
    d.rgx = recip_of((uint32_t)sc.g[0]);
    d.rgy = recip_of((uint32_t)sc.g[1]);
    d.key_base = (uint32_t)sc.key_base;
    selection_dirty_ = true;
    desc_dirty_ = true;
  }

  // Chunk table, per-scene chunk ranges and the step header for a set of
  // scenes; uploaded once per change of selection.
  void set_selection(const std::vector<uint32_t>& sel) {
    selection_ = sel;
    std::vector<Chunk> ch;
    for (auto& d : desc_host_) { d.chunk_begin = d.chunk_end = 0; }
    uint64_t ks_total = 0, nshapes = 0;
    for (uint32_t s : sel) {
      const HostScene& sc = scenes_[s];
      SceneDesc<R>& d = desc_host_[s];
      d.chunk_begin = (uint32_t)ch.size();
      for (uint64_t b = 0; b < sc.n_shapes; b += CHUNK) {
        Chunk c;
        c.scene = s;
        c.begin = (uint32_t)(sc.shape_begin + b);
        c.len = (uint32_t)std::min<uint64_t>(CHUNK, sc.n_shapes - b);
        c.pad = 0;
        ch.push_back(c);
      }
      d.chunk_end = (uint32_t)ch.size();
      // step-local key bases: the selected scenes' key ranges follow one
      // another from 0, so the keys a step sorts span only its own scenes
      d.key_base = (uint32_t)ks_total;
      ks_total += sc.key_space;
      CDX_REQUIRE(ks_total <= (1ull << 32), ERR_INVALID, "bin key space of a step exceeds 2^32: use fewer bins per axis");
      nshapes = std::max<uint64_t>(nshapes, sc.shape_begin + sc.n_shapes);
    }
    nchunks_ = ch.size();
    nslots_ = nchunks_ * CHUNK;
    nsel_ = (uint32_t)sel.size();
    for (const HostScene& sc : scenes_) nshapes = std::max<uint64_t>(nshapes, sc.shape_begin + sc.n_shapes);
    shape_bits_ = bits_for(nshapes ? nshapes - 1 : 0);
    if (nchunks_ > chunk_cap_ || !chunks_.p)
      grow_chunk_cap(std::max<uint64_t>(std::max<uint64_t>(nchunks_, 1), chunk_cap_ * 2));
    // Key width of the radix sort: a recorded step stays valid whichever
    // scenes a later step selects, so it covers any selection the buffers
    // hold: at most chunk_cap_ * CHUNK shapes, each scene holding at most
    // `ratio` keys per shape (and never more than every resident scene).
    {
      uint64_t ratio = 1, ks_all = 0;
      for (const HostScene& sc : scenes_) {
        ratio = std::max<uint64_t>(ratio, (sc.key_space + sc.n_shapes - 1) / std::max<uint64_t>(sc.n_shapes, 1));
        ks_all += sc.key_space;
      }
      const uint64_t bound = std::max(ks_total, std::min(ks_all, ratio * chunk_cap_ * CHUNK));
      key_bits_ = bits_for(bound ? bound - 1 : 0);
    }
    if (nsel_ > sel_cap_ || !sel_.p) grow_sel_cap(std::max<uint64_t>(std::max<uint64_t>(nsel_, 1), sel_cap_ * 2));
    if (desc_host_.size() > desc_cap_ || !desc_.p) {
      desc_cap_ = std::max<uint64_t>(std::max<uint64_t>(desc_host_.size(), 1), desc_cap_ * 2);
      desc_.reserve(desc_cap_ * sizeof(SceneDesc<R>));
      graph_valid_ = false;
      desc_dirty_ = true;
    }
    // Uploads through one pinned staging block, ordered on the scene stream:
    // the step header, the selection, the chunk table, and the selected
    // scenes' chunk ranges (applied to their descriptors by a small kernel;
    // the full descriptor table is uploaded only after scenes change).
    if (desc_dirty_) {
      CDX_HIP_CHECK(hipMemcpyAsync(desc_.p, desc_host_.data(), desc_host_.size() * sizeof(SceneDesc<R>),
                                   hipMemcpyHostToDevice, s_));
      CDX_HIP_CHECK(hipStreamSynchronize(s_));  // desc_host_ is pageable
      desc_dirty_ = false;
    }
    if (stage_pending_) {  // the previous selection's copies read the staging block
      CDX_HIP_CHECK(hipEventSynchronize(ev_stage_));
      stage_pending_ = false;
    }
    const size_t sel_bytes = ((nsel_ * 4 + 15) & ~size_t(15));
    const size_t rng_bytes = ((nsel_ * 16 + 15) & ~size_t(15));
    const size_t need = sizeof(StepHdr) + sel_bytes + rng_bytes + nchunks_ * sizeof(Chunk) + 64;
    if (need > stage_host_bytes_) {
      if (stage_host_) CDX_HIP_CHECK(hipHostFree(stage_host_));
      stage_host_bytes_ = std::max<size_t>(need * 2, 4096);
      CDX_HIP_CHECK(hipHostMalloc(&stage_host_, stage_host_bytes_, hipHostMallocDefault));
    }
    char* hs = static_cast<char*>(stage_host_);
    StepHdr* shdr = reinterpret_cast<StepHdr*>(hs);
    shdr->nslots = nslots_;
    shdr->nsel = nsel_;
    shdr->pad = 0;
    uint32_t* ssel = reinterpret_cast<uint32_t*>(hs + sizeof(StepHdr));
    uint32_t* srng = reinterpret_cast<uint32_t*>(hs + sizeof(StepHdr) + sel_bytes);
    Chunk* sch = reinterpret_cast<Chunk*>(hs + sizeof(StepHdr) + sel_bytes + rng_bytes);
    for (uint32_t j = 0; j < nsel_; ++j) {
      ssel[j] = sel[j];
      srng[4 * j] = sel[j];
      srng[4 * j + 1] = desc_host_[sel[j]].chunk_begin;
      srng[4 * j + 2] = desc_host_[sel[j]].chunk_end;
      srng[4 * j + 3] = desc_host_[sel[j]].key_base;
    }
    if (!ch.empty()) std::memcpy(sch, ch.data(), ch.size() * sizeof(Chunk));
    selrng_.reserve(std::max<size_t>(rng_bytes, 16));
    CDX_HIP_CHECK(hipMemcpyAsync(stephdr_.p, shdr, sizeof(StepHdr), hipMemcpyHostToDevice, s_));
    if (nsel_) {
      CDX_HIP_CHECK(hipMemcpyAsync(sel_.p, ssel, nsel_ * 4, hipMemcpyHostToDevice, s_));
      CDX_HIP_CHECK(hipMemcpyAsync(selrng_.p, srng, nsel_ * 16, hipMemcpyHostToDevice, s_));
      hipLaunchKernelGGL(k_apply_selection<R>, dim3(blocks_for(nsel_, 64)), dim3(64), 0, s_, (uint64_t)nsel_,
                         (const uint32_t*)selrng_.as<uint32_t>(), desc_.as<SceneDesc<R>>());
    }
    if (!ch.empty())
      CDX_HIP_CHECK(hipMemcpyAsync(chunks_.p, sch, ch.size() * sizeof(Chunk), hipMemcpyHostToDevice, s_));
    if (!ev_stage_) CDX_HIP_CHECK(hipEventCreateWithFlags(&ev_stage_, hipEventDisableTiming));
    CDX_HIP_CHECK(hipEventRecord(ev_stage_, s_));
    stage_pending_ = true;
    selection_dirty_ = false;
    if (key_bits_ != graph_key_bits_ || shape_bits_ != graph_shape_bits_) graph_valid_ = false;
  }

  void select_single() {
    if (selection_dirty_ || selection_.size() != 1 || selection_[0] != 0) {
      sync_internal();
      set_selection(std::vector<uint32_t>{0});
    }
  }

  // -------------------------------------------------------------- stages --
  DevStatus* st() { return status_.as<DevStatus>(); }

  uint64_t chunk_grid() const { return exact_ ? std::max<uint64_t>(nchunks_, 1) : std::max<uint64_t>(chunk_cap_, 1); }
  uint64_t sel_grid() const { return exact_ ? std::max<uint64_t>(nsel_, 1) : std::max<uint64_t>(sel_cap_, 1); }

  // The AABBs (or the scene grid) change: the bins of a cell-search
  // broadphase can no longer be counted or built.
  void aabbs_changed() {
    if (bins_pending_) {
      bins_pending_ = false;
      bins_lost_ = true;
    }
  }

  // Entries of the MPR queue's class histograms for up to P pairs.
  static uint64_t queue_hist_words(uint64_t P) { return (uint64_t)NQC * ((P + QUEUE_TILE - 1) / QUEUE_TILE); }

  // Exclusive scan on the step stream.
  void scan(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, const uint64_t* d_n) {
    scan_excl_u32(in, out, n, d_total, scan_tmp_.as<uint32_t>(), s_, d_n);
  }

  // Tile sums the producers of three scans add to (bin counts, pairs per
  // shape, filter selections), zeroed by k_stage_begin: tp_T_ tiles each for
  // the two per-slot scans, then tp_F_ for the per-pair one.
  uint64_t* tp_counts() { return tile_parts_.as<uint64_t>(); }
  uint64_t* tp_spairs() { return tile_parts_.as<uint64_t>() + tp_T_; }
  uint64_t* tp_filter() { return tile_parts_.as<uint64_t>() + 2 * tp_T_; }
  void reserve_tile_parts(uint64_t slots, uint64_t P) {
    const uint64_t T = scan_tile_count(slots), F = scan_tile_count(P);
    if (T <= tp_T_ && F <= tp_F_ && tile_parts_.p) return;
    tp_T_ = std::max(T, tp_T_);
    tp_F_ = std::max(F, tp_F_);
    tile_parts_.reserve((2 * tp_T_ + tp_F_) * 8);
    graph_valid_ = false;  // the layout moved
  }

  void stage_aabb() {
    aabbs_changed();
    if (!opened_)
      hipLaunchKernelGGL(k_stage_begin, dim3(1), dim3(64), 0, s_, (const StepHdr*)stephdr_.as<StepHdr>(), st(),
                       tp_counts(), (uint32_t)(2 * tp_T_), tp_filter(), (uint32_t)tp_F_,
                         (int)ZERO_AABB);
    mark(ST_AABB, true);
    hipLaunchKernelGGL(k_aabb<R>, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_, (const Chunk*)chunks_.as<Chunk>(),
                       (const DevStatus*)st(), NSC_, type_.as<uint8_t>(), body_.as<uint32_t>(), lpos_.as<R>(),
                       dims_.as<R>(), cpos_.as<R>(), lrot_.as<R>(), NBC_, bpos_.as<R>(), brot_.as<R>(), active_.as<uint8_t>(),
                       aabb_.as<R>(), rec_.as<ShapeRec<R>>(), hdr_.as<ShapeHdr>(), frec_.as<ShapeRecF>(),
                       bpart_.as<uint32_t>(), (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), st());
    mark(ST_AABB, false);
  }

  // Broadphase from the AABBs (exact mode reads totals back between steps).
  void stage_broadphase() {
    reserve_tile_parts(chunk_grid() * CHUNK, std::max<uint64_t>(pcap_, 1));
    if (!opened_)  // AABBs from the caller: no AABB stage cleared the error flags
      hipLaunchKernelGGL(k_stage_begin, dim3(1), dim3(64), 0, s_, (const StepHdr*)stephdr_.as<StepHdr>(), st(),
                       tp_counts(), (uint32_t)(2 * tp_T_), tp_filter(), (uint32_t)tp_F_,
                         (int)(ZERO_BP | (aabb_from_caller_ ? ZERO_AABB : 0)));
    mark(ST_BP_BOUNDS, true);
    if (aabb_from_caller_)
      hipLaunchKernelGGL(k_bounds<R>, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_,
                         (const Chunk*)chunks_.as<Chunk>(), (const DevStatus*)st(), NSC_, (const R*)aabb_.as<R>(),
                         bpart_.as<uint32_t>());
    hipLaunchKernelGGL(k_params<R>, dim3((unsigned)sel_grid()), dim3(PARAMS_BS), 0, s_, (const uint32_t*)sel_.as<uint32_t>(),
                       st(), (const uint32_t*)bpart_.as<uint32_t>(), desc_.as<SceneDesc<R>>());
    mark(ST_BP_BOUNDS, false);

    const uint64_t slots = chunk_grid() * CHUNK;
    scan_tmp_.reserve(std::max<size_t>(scan_tmp_words(slots), 16) * 4);
    cell_ = false;
    bins_pending_ = false;
    bins_lost_ = false;
    bins_counted_ = true;
    // The cell search needs only the bounds.  A scene whose last broadphase
    // took it (or that asks for it) decides right after them and counts
    // the reference grid's bins only when a caller reads bin totals or bins;
    // any other scene counts first, so the bin search waits for one status
    // read, not two.
    const bool cells_first = exact_ && (pair_search_ == PAIRS_CELLS || (pair_search_ == PAIRS_AUTO && last_cells_));
    if (cells_first) {
      read_status();
      if (plan_cells()) {
        bins_counted_ = false;
        M_ = 0;
        stage_cells();
        return;
      }
    }
    enqueue_bin_count();
    int key_bits = key_bits_;
    if (exact_) {
      read_status();
      take_bin_totals();
      key_bits = exact_key_bits_;
      if (!cells_first && pair_search_ != PAIRS_BINS && plan_cells()) {
        stage_cells();
        return;
      }
      last_cells_ = false;
    }
    emit_and_sort_bins(key_bits);
    meta_.reserve(mcap_ * sizeof(PairMeta));
    runs_part_.reserve((mcap_ / PAIR_BS + 1) * 4);
    mark(ST_BP_PAIRS, true);
    hipLaunchKernelGGL(k_pairs<R>, dim3(blocks_for(mcap_, PAIR_BS)), dim3(PAIR_BS), 0, s_, mcap_,
                       (const uint64_t*)keys_.as<uint64_t>(), (const BpRec<R>*)brec_.as<BpRec<R>>(),
                       (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), pcount_.as<uint32_t>(), meta_.as<PairMeta>(),
                       runs_part_.as<uint32_t>(), st());
    mark(ST_BP_PAIRS, false);
    mark(ST_BP_PAIRSCAN, true);
    hipLaunchKernelGGL(k_shape_pairs, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_,
                       (const Chunk*)chunks_.as<Chunk>(), st(),
                       (const uint32_t*)counts_buf_.as<uint32_t>(), (const uint32_t*)offs_.as<uint32_t>(),
                       (const uint32_t*)pcount_.as<uint32_t>(), mcap_, scount_.as<uint32_t>(), tp_spairs(),
                       (const uint32_t*)runs_part_.as<uint32_t>());
    scan_excl_u32_tiled(scount_.as<uint32_t>(), soff_.as<uint32_t>(), slots, &st()->P, tp_spairs(), s_,
                        &st()->nslots);
    mark(ST_BP_PAIRSCAN, false);
    if (exact_) {
      read_status();
      P_ = hstat_->P;
      A_ = hstat_->A;
      CDX_REQUIRE(P_ < (1ull << 32), ERR_CAPACITY, "more than 2^32 candidate pairs");
      pcap_ = std::max<uint64_t>(P_, 1);
    }
    pairs_.reserve(pcap_ * 8);
    const uint64_t long_cap = pcap_ / (GROUP_LONG + 1) + 1;
    longs_.reserve(long_cap * 4);
    mark(ST_BP_PAIREMIT, true);
    hipLaunchKernelGGL(k_pair_emit<R>, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_,
                       (const Chunk*)chunks_.as<Chunk>(), (const DevStatus*)st(),
                       (const uint32_t*)counts_buf_.as<uint32_t>(), (const uint32_t*)offs_.as<uint32_t>(),
                       (const uint32_t*)pcount_.as<uint32_t>(), (const PairMeta*)meta_.as<PairMeta>(),
                       (const uint32_t*)scount_.as<uint32_t>(), (const uint32_t*)soff_.as<uint32_t>(),
                       (const uint64_t*)keys_.as<uint64_t>(), (const BpRec<R>*)brec_.as<BpRec<R>>(),
                       (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), mcap_,
                       pcap_, pairs_.as<uint64_t>(), longs_.as<uint32_t>(), long_cap, st());
    mark(ST_BP_PAIREMIT, false);
    mark(ST_BP_PAIRSORT, true);
    hipLaunchKernelGGL(k_pair_groups_long, dim3(256), dim3(256), 0, s_, (const Chunk*)chunks_.as<Chunk>(),
                       (const DevStatus*)st(), (const uint32_t*)scount_.as<uint32_t>(),
                       (const uint32_t*)soff_.as<uint32_t>(), pairs_.as<uint64_t>(),
                       (const uint32_t*)longs_.as<uint32_t>(), long_cap);
    mark(ST_BP_PAIRSORT, false);
    if (exact_) {
      counts_.n = scenes_[0].n_shapes; counts_.m = M_; counts_.a = A_; counts_.p = P_;
      counts_.info = 0;
      counts_.key_max = hstat_->key_max;
      counts_.bad = hstat_->bad;
      SceneDesc<R> d;
      CDX_HIP_CHECK(hipMemcpy(&d, desc_.p, sizeof(d), hipMemcpyDeviceToHost));
      for (int c = 0; c < 3; ++c) { counts_.bmin[c] = ord2f(d.bmin_enc[c]); counts_.bmax[c] = ord2f(d.bmax_enc[c]); }
    }
  }

  // Bins spanned per AABB at the reference grid (Broadphase.cu:150-166),
  // the largest key, and the scan of the counts (bin entry total st->M).
  void enqueue_bin_count() {
    const uint64_t slots = chunk_grid() * CHUNK;
    // The key maximum: a step with a fixed key width leaves the chunk maxima
    // to the key emission's first workgroup; an exact step needs it before
    // the emission (for the key width), from atomics in small launches or
    // from k_reduce_keymax.
    const bool fold = exact_ && chunk_grid() <= FOLD_WGS;
    mark(ST_BP_COUNT, true);
    hipLaunchKernelGGL(k_count<R>, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_, (const Chunk*)chunks_.as<Chunk>(),
                       (const DevStatus*)st(), NSC_, (const R*)aabb_.as<R>(),
                       (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), counts_buf_.as<uint32_t>(),
                       fold ? nullptr : kpart_.as<uint32_t>(), tp_counts(), st());
    if (exact_ && !fold)
      hipLaunchKernelGGL(k_reduce_keymax, dim3(1), dim3(KEYMAX_BS), 0, s_, (const uint32_t*)kpart_.as<uint32_t>(), st());
    mark(ST_BP_COUNT, false);
    mark(ST_BP_SCAN, true);
    scan_excl_u32_tiled(counts_buf_.as<uint32_t>(), offs_.as<uint32_t>(), slots, &st()->M, tp_counts(), s_,
                        &st()->nslots);
    mark(ST_BP_SCAN, false);
  }

  // Bin entry total and key width from the status just read (exact mode).
  void take_bin_totals() {
    M_ = hstat_->M;
    CDX_REQUIRE(M_ < (1ull << 32), ERR_CAPACITY, "more than 2^32 bin entries: use fewer bins per axis");
    mcap_ = std::max<uint64_t>(M_, 1);
    exact_key_bits_ = bits_for(hstat_->key_max);
    bins_counted_ = true;
  }

  // The bin count of a cell-search broadphase that skipped it, run when a
  // caller reads bin totals or bins: the AABBs and the scene grid are that
  // broadphase's, so the counts are the ones a count before the search
  // would have given.
  void ensure_bin_counts() {
    if (bins_counted_) return;
    CDX_REQUIRE(!bins_lost_, ERR_NOT_READY,
                "the AABBs changed after a cell-search broadphase: run the broadphase again to read its bins");
    const bool prof = prof_;
    prof_ = false;  // the stage times stay those of the broadphase
    // the count adds to the scan's tile sums: clear them (a narrowphase may have moved them since)
    fill_u32(reinterpret_cast<uint32_t*>(tp_counts()), 0u, 2 * tp_T_, s_);
    enqueue_bin_count();
    prof_ = prof;
    read_status();
    take_bin_totals();
    counts_.m = M_;
    counts_.key_max = hstat_->key_max;
  }

  // (bin key << 32 | shape) entries (Broadphase.cu:172-209) sorted by key
  // (:428-432), plus each shape's broadphase record.
  void emit_and_sort_bins(int key_bits) {
    keys_.reserve(mcap_ * 8);
    keys_tmp_.reserve(mcap_ * 8);
    pcount_.reserve((mcap_ + 1) * 4);
    mark(ST_BP_EMIT, true);
    hipLaunchKernelGGL(k_emit_keys<R>, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_,
                       (const Chunk*)chunks_.as<Chunk>(), (const DevStatus*)st(), NSC_, (const R*)aabb_.as<R>(),
                       (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), (const uint32_t*)offs_.as<uint32_t>(),
                       (const uint32_t*)counts_buf_.as<uint32_t>(), mcap_, keys_.as<uint64_t>(),
                       pcount_.as<uint32_t>(), brec_.as<BpRec<R>>(),
                       exact_ ? (const uint32_t*)nullptr : (const uint32_t*)kpart_.as<uint32_t>(), st());
    mark(ST_BP_EMIT, false);
    radix_tmp_.reserve(std::max<size_t>(radix_tmp_words(mcap_, key_bits), 16) * 4);
    mark(ST_BP_SORT, true);
    if (radix_sort_u64(keys_.as<uint64_t>(), keys_tmp_.as<uint64_t>(), mcap_, 32, key_bits, radix_tmp_.as<uint32_t>(),
                       s_, &st()->M))
      swap_bufs(keys_, keys_tmp_);
    mark(ST_BP_SORT, false);
  }

  // Grid of the cell search for scene 0 (exact mode, after the bounds):
  // cells of edge (1 + 1e-3) * (W + 2^-20 C) along y and z, W the widest
  // AABB extent and C the largest coordinate magnitude (the allowance for
  // the rounding of the shift and of W itself; the search's rounding
  // argument needs the edge to bound the widest shifted extent), enlarged
  // until there are at most about two cells per shape, and CELL_KX times
  // finer along x.  Taken when the mean occupancy of the coarse cells stays
  // at most CELL_AUTO_DENSITY shapes (always, when the caller asks for the
  // cell search).
  bool plan_cells() {
    const uint64_t n = scenes_[0].n_shapes;
    if (n < 2) return false;
    SceneDesc<R> d;
    CDX_HIP_CHECK(hipMemcpy(&d, desc_.p, sizeof(d), hipMemcpyDeviceToHost));
    // The bin count, which reports AABBs the reference grid cannot bin, does
    // not run before the cell search.  With no inverted AABB, finite scene
    // bounds and finite, positive inverse bin sizes every AABB is finite and
    // maps into the grid (cells_ok), so none goes unreported; any other scene
    // counts first.
    if (hstat_->wmax & WIDTH_INVERTED) return false;
    float w;
    std::memcpy(&w, &hstat_->wmax, 4);
    if (!(w >= 0.f) || !std::isfinite(w)) return false;
    for (int c = 0; c < 3; ++c)
      if (!(d.inv[c] > (R)0) || !std::isfinite((double)d.inv[c])) return false;
    double ext[3], cmax = 0.0;
    for (int c = 0; c < 3; ++c) {
      // shifted coordinates lie in [0, bmax + origin] up to rounding (from
      // 2 bmin when bmin > 0: the origin is |bmin|, Broadphase.cu:377-381)
      const double lo = (double)ord2f(d.bmin_enc[c]), up = (double)ord2f(d.bmax_enc[c]);
      const double hi = up + (double)d.origin[c];
      if (!std::isfinite(hi) || !std::isfinite(lo)) return false;
      ext[c] = std::max(hi, 0.0);
      cmax = std::max(cmax, std::max(std::max(std::fabs(lo), std::fabs(up)), ext[c]));
    }
    const double target = 2.0 * (double)n + 64.0;
    // plus 8 of the smallest float denormals: the relative bounds end there
    double edge = ((double)w + std::ldexp(cmax, -20) + 8.0 * (double)std::numeric_limits<float>::denorm_min()) *
                  (1.0 + 1e-3);
    if (!(edge > 0.0)) edge = std::max(std::max(ext[0], ext[1]), std::max(ext[2], 1e-300));  // not reached
    double g[3], cells = 0;
    for (int it = 0; it < 64; ++it) {
      cells = 1;
      for (int c = 0; c < 3; ++c) {
        g[c] = std::floor(ext[c] / edge) + 1.0;
        cells *= g[c];
      }
      if (cells <= target) break;
      edge *= std::max(std::cbrt(cells / target), 1.0 + 1e-3);
    }
    if (!(cells <= target)) return false;
    if (pair_search_ == PAIRS_AUTO && (double)n / cells > CELL_AUTO_DENSITY) return false;
    const double gxf = std::floor(ext[0] * CELL_KX / edge) + 1.0;
    if (gxf * g[1] * g[2] >= 4294967295.0) return false;
    cg_.g[0] = (uint32_t)gxf;
    cg_.g[1] = (uint32_t)g[1];
    cg_.g[2] = (uint32_t)g[2];
    cg_.kx = (uint32_t)CELL_KX;
    cg_.inv[0] = CELL_KX / edge;
    cg_.inv[1] = cg_.inv[2] = 1.0 / edge;
    ncells_ = (uint64_t)cg_.g[0] * cg_.g[1] * cg_.g[2];
    return true;
  }

  // Pair search by cells (the reference grid's bin entries are counted,
  // emitted and sorted only when a caller reads bin totals or bins).
  void stage_cells() {
    const uint64_t n = scenes_[0].n_shapes;
    const uint64_t slots = chunk_grid() * CHUNK;
    const uint64_t ntab = ncells_ + 1;
    cell_tab_.reserve(ntab * 4);
    cell_rank_.reserve(n * 4);
    cell_lo_.reserve(n * sizeof(CellLo<R>));
    cell_hi_.reserve(n * sizeof(CellHi<R>));
    cell_slab_.reserve(n * CELL_SLAB * 4);
    cell_last_.reserve(ncells_ * 4);
    cell_wmax_.reserve(ncells_ * 4);
    cell_ovf_.reserve(n * 4);
    scan_tmp_.reserve(std::max<size_t>(std::max(scan_tmp_words(ntab), scan_tmp_words(slots)), 16) * 4);
    uint32_t* tab = cell_tab_.as<uint32_t>();
    uint32_t* rank = cell_rank_.as<uint32_t>();
    mark(ST_BP_EMIT, true);
    fill_u32(tab, 0u, ntab, s_);
    fill_u32(cell_last_.as<uint32_t>(), 0u, ncells_, s_);
    hipLaunchKernelGGL(k_cell_keys<R>, dim3((unsigned)chunk_grid()), dim3(CHUNK), 0, s_,
                       (const Chunk*)chunks_.as<Chunk>(), (const DevStatus*)st(), NSC_, (const R*)aabb_.as<R>(),
                       (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), cg_, tab, rank, cell_last_.as<uint32_t>());
    scan_excl_u32(tab, tab, ntab, &st()->cells, scan_tmp_.as<uint32_t>(), s_);
    mark(ST_BP_EMIT, false);
    mark(ST_BP_SORT, true);
    hipLaunchKernelGGL(k_cell_scatter<R>, dim3(blocks_for(n, 256)), dim3(256), 0, s_, n, (const uint32_t*)rank,
                       (const uint32_t*)tab, NSC_, (const R*)aabb_.as<R>(),
                       (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(), cg_, cell_lo_.as<CellLo<R>>(),
                       cell_hi_.as<CellHi<R>>());
    const CellLo<R>* clo = cell_lo_.as<CellLo<R>>();
    const CellHi<R>* chi = cell_hi_.as<CellHi<R>>();
    hipLaunchKernelGGL(k_cell_window, dim3(blocks_for(ncells_, 256)), dim3(256), 0, s_, ncells_, cg_,
                       (const uint32_t*)cell_last_.as<uint32_t>(), cell_wmax_.as<uint32_t>());
    mark(ST_BP_SORT, false);
    const uint32_t* wmx = cell_wmax_.as<uint32_t>();
    mark(ST_BP_PAIRS, true);
    // slots past the last shape take part in the scan: their counts are 0
    if (slots > n) fill_u32(scount_.as<uint32_t>() + n, 0u, slots - n, s_);
    hipLaunchKernelGGL(k_cell_find<R>, dim3(blocks_for(n, CELL_BS)), dim3(CELL_BS), 0, s_, n, clo, chi,
                       (const uint32_t*)tab, wmx, cg_, scount_.as<uint32_t>(), cell_slab_.as<uint32_t>(),
                       cell_ovf_.as<uint32_t>(), st());
    mark(ST_BP_PAIRS, false);
    mark(ST_BP_PAIRSCAN, true);
    scan_excl_u32(scount_.as<uint32_t>(), soff_.as<uint32_t>(), slots, &st()->P, scan_tmp_.as<uint32_t>(), s_,
                  &st()->nslots);
    mark(ST_BP_PAIRSCAN, false);
    read_status();
    P_ = hstat_->P;
    CDX_REQUIRE(hstat_->cells == n, ERR_HIP, "internal error: the cell table does not hold every shape");
    CDX_REQUIRE(P_ < (1ull << 32), ERR_CAPACITY, "more than 2^32 candidate pairs");
    pcap_ = std::max<uint64_t>(P_, 1);
    pairs_.reserve(pcap_ * 8);
    const uint64_t long_cap = pcap_ / (GROUP_LONG + 1) + 1;
    longs_.reserve(long_cap * 4);
    mark(ST_BP_PAIREMIT, true);
    hipLaunchKernelGGL(k_cell_place, dim3(blocks_for(n, CELL_BS)), dim3(CELL_BS), 0, s_, n,
                       (const uint32_t*)scount_.as<uint32_t>(), (const uint32_t*)soff_.as<uint32_t>(),
                       (const uint32_t*)cell_slab_.as<uint32_t>(), pcap_, pairs_.as<uint64_t>(), st());
    mark(ST_BP_PAIREMIT, false);
    mark(ST_BP_PAIRSORT, true);
    hipLaunchKernelGGL(k_cell_overflow<R>, dim3(256), dim3(CELL_BS), 0, s_, (const uint32_t*)cell_ovf_.as<uint32_t>(),
                       clo, chi, (const uint32_t*)tab, wmx, cg_, (const uint32_t*)scount_.as<uint32_t>(),
                       (const uint32_t*)soff_.as<uint32_t>(), pcap_, pairs_.as<uint64_t>(), longs_.as<uint32_t>(),
                       long_cap, st());
    hipLaunchKernelGGL(k_pair_groups_long, dim3(256), dim3(256), 0, s_, (const Chunk*)chunks_.as<Chunk>(),
                       (const DevStatus*)st(), (const uint32_t*)scount_.as<uint32_t>(),
                       (const uint32_t*)soff_.as<uint32_t>(), pairs_.as<uint64_t>(),
                       (const uint32_t*)longs_.as<uint32_t>(), long_cap);
    mark(ST_BP_PAIRSORT, false);
    cell_ = true;
    last_cells_ = true;
    bins_pending_ = true;
    A_ = 0;
    // bin totals: zero until counted (counts_full, read_bins)
    counts_.n = n; counts_.m = M_; counts_.a = 0; counts_.p = P_;
    counts_.info = INFO_CELL_SEARCH | INFO_NO_ACTIVE_BINS;
    counts_.key_max = bins_counted_ ? hstat_->key_max : 0u;
    counts_.bad = hstat_->bad;
    SceneDesc<R> d;
    CDX_HIP_CHECK(hipMemcpy(&d, desc_.p, sizeof(d), hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c) { counts_.bmin[c] = ord2f(d.bmin_enc[c]); counts_.bmax[c] = ord2f(d.bmax_enc[c]); }
  }

  // Bin entries of a cell-search broadphase, built on request (they
  // depend on the AABBs of that broadphase, so not after they change).
  void ensure_bins() {
    ensure_bin_counts();
    if (!bins_pending_) {
      CDX_REQUIRE(!bins_lost_, ERR_NOT_READY,
                  "the AABBs changed after a cell-search broadphase: run the broadphase again to read its bins");
      return;
    }
    const bool prof = prof_;
    prof_ = false;  // the stage times stay those of the broadphase
    emit_and_sort_bins(exact_key_bits_);
    prof_ = prof;
    bins_pending_ = false;
  }

  void stage_narrowphase() {
    reserve_tile_parts(chunk_grid() * CHUNK, std::max<uint64_t>(pcap_, 1));
    if (!opened_)
      hipLaunchKernelGGL(k_stage_begin, dim3(1), dim3(64), 0, s_, (const StepHdr*)stephdr_.as<StepHdr>(), st(),
                       tp_counts(), (uint32_t)(2 * tp_T_), tp_filter(), (uint32_t)tp_F_,
                         (int)ZERO_NP);
    const uint64_t P = std::max<uint64_t>(pcap_, 1);
    np_v_.reserve(P * sizeof(PairRec<R>));
    np_hit_.reserve(P * 4);
    np_pos_.reserve(std::max(queue_hist_words(P), P) * 4);
    queue_.reserve(P * 8);
    queue_tmp_.reserve(P * 8);
    qsel_.reserve(P * 4);
    qcls_.reserve(P);
    scan_tmp_.reserve(std::max<size_t>(scan_tmp_words(std::max(queue_hist_words(P), P)), 16) * 4);
    np_kind_.reserve(P);
    PairOut<R> o{np_v_.as<PairRec<R>>(), np_hit_.as<uint32_t>(), P, np_kind_.as<uint8_t>()};
    const uint64_t* dP = &st()->P;
    mark(ST_NP_MAIN, true);
    hipLaunchKernelGGL(k_np_front<R>, dim3(blocks_for(P, 128)), dim3(128), 0, s_, P, dP,
                       (const uint64_t*)pairs_.as<uint64_t>(), (const ShapeRec<R>*)rec_.as<ShapeRec<R>>(),
                       (const ShapeHdr*)hdr_.as<ShapeHdr>(), (const ShapeRecF*)frec_.as<ShapeRecF>(), tau_,
                       (int)(mode_ != NP_NATIVE), envelope_, o, qcls_.as<uint8_t>());
    // class histograms per tile -> first slot per (class, tile); the total is the queue length
    const unsigned qtiles = (unsigned)((P + QUEUE_TILE - 1) / QUEUE_TILE);
    hipLaunchKernelGGL(k_queue_hist, dim3(qtiles), dim3(QUEUE_BS), 0, s_, P, dP, (const uint8_t*)qcls_.as<uint8_t>(),
                       np_pos_.as<uint32_t>(), &st()->qhn);
    scan(np_pos_.as<uint32_t>(), np_pos_.as<uint32_t>(), queue_hist_words(P), &st()->Q, &st()->qhn);
    hipLaunchKernelGGL(k_queue_scatter, dim3(qtiles), dim3(QUEUE_BS), 0, s_, P, dP, (const uint8_t*)qcls_.as<uint8_t>(),
                       (const uint32_t*)np_pos_.as<uint32_t>(), queue_.as<uint64_t>());
    uint64_t q = P;
    if (exact_) {
      read_status();
      q = std::max<uint64_t>(hstat_->Q, 1);
    }
    mark(ST_NP_MAIN, false);
    mark(ST_NP_RESOLVE, true);
    const uint64_t* rq = queue_.as<uint64_t>();
    const uint64_t* dq = &st()->Q;
    if (mode_ != NP_NATIVE) {
      auto filt = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(blocks_for(q, 128)), dim3(128), 0, s_, q, dq, (const uint64_t*)queue_.as<uint64_t>(),
                           (const uint64_t*)pairs_.as<uint64_t>(), (const ShapeHdr*)hdr_.as<ShapeHdr>(),
                           (const ShapeRecF*)frec_.as<ShapeRecF>(), envelope_, tau_, o, qsel_.as<uint32_t>(),
                           tp_filter());
      };
      if (mode_ == NP_FAST) filt(k_np_filter<R, true>);
      else filt(k_np_filter<R, false>);
      scan_excl_u32_tiled(qsel_.as<uint32_t>(), np_pos_.as<uint32_t>(), q, &st()->Q2, tp_filter(), s_, dq);
      hipLaunchKernelGGL(k_queue_select, dim3(blocks_for(q, 256)), dim3(256), 0, s_, q, dq,
                         (const uint64_t*)queue_.as<uint64_t>(), (const uint32_t*)qsel_.as<uint32_t>(),
                         (const uint32_t*)np_pos_.as<uint32_t>(), queue_tmp_.as<uint64_t>());
      rq = queue_tmp_.as<uint64_t>();
      dq = &st()->Q2;
      if (exact_) {
        read_status();
        q = std::max<uint64_t>(hstat_->Q2, 1);
      }
    }
    mpr_in_tmp_ = mode_ != NP_NATIVE;  // the filter's selection went to queue_tmp_
    mpr_q_ = exact_ ? (mode_ != NP_NATIVE ? hstat_->Q2 : hstat_->Q) : 0;
    mark(ST_NP_MPR, true);
    // two launches over the class-sorted queue: the pairs of the five
    // filtered types, then those with a rectangle, cone or triangle (kept at
    // the queue's end), each kernel holding only the support maps it meets;
    // the second only when a resident scene holds such a shape
    auto resolve = [&](auto kern, unsigned grid) {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(128), 0, s_, q, dq, (const uint64_t*)pairs_.as<uint64_t>(),
                         (const ShapeRec<R>*)rec_.as<ShapeRec<R>>(), envelope_, o, rq);
    };
    const unsigned ext_grid = std::min<unsigned>(blocks_for(q, 128), EXT_RESOLVE_GRID);
    if (mode_ == NP_NATIVE) {
      resolve(k_np_resolve<R, R, false>, blocks_for(q, 128));
      if (has_ext_) resolve(k_np_resolve<R, R, true>, ext_grid);
    } else {
      resolve(k_np_resolve<R, double, false>, blocks_for(q, 128));
      if (has_ext_) resolve(k_np_resolve<R, double, true>, ext_grid);
    }
    mark(ST_NP_MPR, false);
    mark(ST_NP_RESOLVE, false);
    mark(ST_NP_COMPACT, true);
    scan(np_hit_.as<uint32_t>(), np_pos_.as<uint32_t>(), P, &st()->C, dP);
    // contacts are at most the pairs: the outputs are sized by the pair capacity
    out_.reserve(P * (10 * sizeof(R) + 8));
    perm_.reserve(P * 8);
    hipLaunchKernelGGL(k_np_compact<R>, dim3(blocks_for(P, COMPACT_BS)), dim3(COMPACT_BS), 0, s_, P, dP, o,
                       (const uint32_t*)np_pos_.as<uint32_t>(), (const uint64_t*)pairs_.as<uint64_t>(),
                       (const ShapeRec<R>*)rec_.as<ShapeRec<R>>(), envelope_, (const DevStatus*)st(), out_.as<R>(),
                       perm_.as<uint64_t>());
    mark(ST_NP_COMPACT, false);
    if (exact_) {
      read_status();
      C_ = hstat_->C;
      counts_.c = C_;
      counts_.resolved = mode_ == NP_NATIVE ? 0 : hstat_->Q2;
      counts_.p = P_;
    }
  }

  void stage_scene_ranges() {
    hipLaunchKernelGGL(k_scene_ranges<R>, dim3(blocks_for(sel_grid(), 64)), dim3(64), 0, s_,
                       (const uint32_t*)sel_.as<uint32_t>(), (const SceneDesc<R>*)desc_.as<SceneDesc<R>>(),
                       (const DevStatus*)st(), mcap_, pcap_, (const uint32_t*)offs_.as<uint32_t>(),
                       (const uint32_t*)soff_.as<uint32_t>(), (const uint32_t*)np_pos_.as<uint32_t>(),
                       souts_.as<SceneOut>());
  }

  // ------------------------------------------------------ capacity steps --
  // Capacities only grow (a recorded step stays valid while they hold):
  // entries and pairs per shape from the capacity factors, or from the
  // largest totals measured so far.
  void grow_chunk_cap(uint64_t cap) {
    chunk_cap_ = cap;
    chunks_.reserve(chunk_cap_ * sizeof(Chunk));
    bpart_.reserve(chunk_cap_ * BPART_WORDS * 4);
    kpart_.reserve(chunk_cap_ * 4);
    counts_buf_.reserve(chunk_cap_ * CHUNK * 4);
    offs_.reserve(chunk_cap_ * CHUNK * 4);
    scount_.reserve(chunk_cap_ * CHUNK * 4);
    soff_.reserve(chunk_cap_ * CHUNK * 4);
    graph_valid_ = false;
  }
  void grow_sel_cap(uint64_t cap) {
    sel_cap_ = cap;
    sel_.reserve(sel_cap_ * 4);
    souts_.reserve(sel_cap_ * sizeof(SceneOut));
    graph_valid_ = false;
  }

  void ensure_capacities() {
    const uint64_t n = total_selected_shapes();
    uint64_t need_m = std::max<uint64_t>((uint64_t)(cap_m_per_shape_ * n) + 4096, m_seen_ + m_seen_ / 4);
    uint64_t need_p = std::max<uint64_t>((uint64_t)(cap_p_per_shape_ * n) + 4096, p_seen_ + p_seen_ / 4);
    if (need_m > mcap_) { mcap_ = need_m; graph_valid_ = false; }
    if (need_p > pcap_) { pcap_ = need_p; graph_valid_ = false; }
  }

  void enqueue_capacity_step() {
    struct Restore {  // also on an exception
      Engine* e;
      bool prof;
      ~Restore() { e->prof_ = prof; e->opened_ = false; }
    } restore{this, prof_};
    prof_ = false;  // no event records inside a recorded step
    // one stage opening for the whole step (the stages' totals are disjoint)
    hipLaunchKernelGGL(k_stage_begin, dim3(1), dim3(64), 0, s_, (const StepHdr*)stephdr_.as<StepHdr>(), st(),
                       tp_counts(), (uint32_t)(2 * tp_T_), tp_filter(), (uint32_t)tp_F_,
                       (int)(ZERO_AABB | ZERO_BP | ZERO_NP));
    opened_ = true;
    stage_aabb();
    stage_broadphase();
    stage_narrowphase();
    stage_scene_ranges();
  }

  void launch_capacity_step(bool use_graph) {
    prepare_capacity_buffers();  // every buffer is sized before recording
    CDX_HIP_CHECK(hipEventRecord(ev_[2 * ST_COUNT], s_));
    if (use_graph) {
      if (!graph_valid_) {
        // Destroy, record and instantiate with no other library thread inside
        // HIP: other size classes launch and synchronise their own graphs
        // concurrently otherwise, and this stream's last launch of an old
        // graph must have ended before its executable is freed.
        GateExclusive gate;
        CDX_HIP_CHECK(hipStreamSynchronize(s_));
        if (graph_exec_) {
          CDX_HIP_CHECK(hipGraphExecDestroy(graph_exec_));
          graph_exec_ = nullptr;
        }
        hipGraph_t g = nullptr;
        CDX_HIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_capacity_step();
        } catch (...) {
          // close the capture so the stream stays usable, then report
          if (hipStreamEndCapture(s_, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
          throw;
        }
        CDX_HIP_CHECK(hipStreamEndCapture(s_, &g));
        CDX_HIP_CHECK(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
        CDX_HIP_CHECK(hipGraphDestroy(g));
        graph_valid_ = true;
        graph_key_bits_ = key_bits_;
        graph_shape_bits_ = shape_bits_;
      }
      CDX_HIP_CHECK(hipGraphLaunch(graph_exec_, s_));
    } else {
      enqueue_capacity_step();
    }
    CDX_HIP_CHECK(hipEventRecord(ev_[2 * ST_COUNT + 1], s_));
    enqueue_step_readback();
  }

  // Copies queued right behind a capacity step on its stream, so finish()
  // waits once: the totals, the per-scene outputs and (when the caller set
  // a prefetch target) the first rows of the packed contacts.
  void enqueue_step_readback() {
    CDX_HIP_CHECK(hipMemcpyAsync(hstat_, status_.p, sizeof(DevStatus), hipMemcpyDeviceToHost, s_));
    const size_t sb = selection_.size() * sizeof(SceneOut);
    if (sb) {
      if (sb > souts_pinned_bytes_) {
        CDX_HIP_CHECK(hipStreamSynchronize(s_));  // a previous copy may still target the old block
        if (souts_pinned_) CDX_HIP_CHECK(hipHostFree(souts_pinned_));
        souts_pinned_ = nullptr;
        souts_pinned_bytes_ = 0;
        CDX_HIP_CHECK(hipHostMalloc(&souts_pinned_, std::max<size_t>(sb * 2, 4096), hipHostMallocDefault));
        souts_pinned_bytes_ = std::max<size_t>(sb * 2, 4096);
      }
      CDX_HIP_CHECK(hipMemcpyAsync(souts_pinned_, souts_.p, sb, hipMemcpyDeviceToHost, s_));
    }
    if (pf_dst_ && pf_rows_) {
      const uint64_t bytes = std::min<uint64_t>(pf_rows_ * (10 * sizeof(R) + 8), out_.bytes);
      CDX_HIP_CHECK(hipMemcpyAsync(pf_dst_, out_.p, bytes, hipMemcpyDeviceToHost, s_));
    }
  }

  void prepare_capacity_buffers() {
    const uint64_t slots = std::max<uint64_t>(chunk_cap_, 1) * CHUNK;
    const uint64_t P = pcap_;
    reserve_tile_parts(slots, std::max<uint64_t>(P, 1));
    scan_tmp_.reserve(std::max<size_t>(std::max(scan_tmp_words(slots), scan_tmp_words(std::max(queue_hist_words(P), P))), 16) * 4);
    keys_.reserve(mcap_ * 8);
    keys_tmp_.reserve(mcap_ * 8);
    radix_tmp_.reserve(std::max<size_t>(radix_tmp_words(mcap_, key_bits_), 16) * 4);
    pcount_.reserve((mcap_ + 1) * 4);
    meta_.reserve(mcap_ * sizeof(PairMeta));
    runs_part_.reserve((mcap_ / PAIR_BS + 1) * 4);
    pairs_.reserve(P * 8);
    longs_.reserve((P / (GROUP_LONG + 1) + 1) * 4);
    np_v_.reserve(P * sizeof(PairRec<R>));
    np_hit_.reserve(P * 4);
    np_kind_.reserve(P);
    np_pos_.reserve(std::max(queue_hist_words(P), P) * 4);
    queue_.reserve(P * 8);
    queue_tmp_.reserve(P * 8);
    qsel_.reserve(P * 4);
    qcls_.reserve(P);
    out_.reserve(P * (10 * sizeof(R) + 8));
    perm_.reserve(P * 8);
  }

  // ------------------------------------------------------------- readback --
  void copy_contacts(uint64_t begin, uint64_t n, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids,
                     uint64_t body_base) {
    if (!n) return;
    // the step's contact arrays lie one after the other, each C_ rows long
    const R* o = out_.as<R>();
    const uint64_t C = C_;
    if (norm) stager_.d2h(norm, o + 3 * begin, n * 3 * sizeof(R), s_);
    if (cpta) stager_.d2h(cpta, o + 3 * C + 3 * begin, n * 3 * sizeof(R), s_);
    if (cptb) stager_.d2h(cptb, o + 6 * C + 3 * begin, n * 3 * sizeof(R), s_);
    if (dpth) stager_.d2h(dpth, o + 9 * C + begin, n * sizeof(R), s_);
    if (bids) {
      stager_.d2h(bids, reinterpret_cast<const int32_t*>(o + 10 * C) + 2 * begin, n * 8, s_);
      if (body_base)
        for (uint64_t i = 0; i < 2 * n; ++i) bids[i] -= (int32_t)body_base;
    }
  }

  void sync_internal() { CDX_HIP_CHECK(hipStreamSynchronize(s_)); }
  void read_status() {
    CDX_HIP_CHECK(hipMemcpyAsync(hstat_, status_.p, sizeof(DevStatus), hipMemcpyDeviceToHost, s_));
    CDX_HIP_CHECK(hipStreamSynchronize(s_));
    check_status();
  }
  void check_status() const {
    const uint32_t bad = hstat_->bad;
    CDX_REQUIRE(!(bad & 1u), ERR_UNSUPPORTED_SHAPE, "unsupported shape type");
    CDX_REQUIRE(!(bad & 2u), ERR_INVALID, "shape body id out of range");
    CDX_REQUIRE(!(bad & 4u), ERR_INVALID,
                "an AABB maps outside the grid: non-finite coordinates or a zero scene extent");
    CDX_REQUIRE(!(bad & 8u), ERR_INVALID, "internal error: a bin entry lies outside its shape's cell range");
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
  StepHdr* hhdr_ = nullptr;
  hipEvent_t ev_[2 * ST_COUNT + 2];
  bool used_[ST_COUNT] = {};
  bool prof_ = false;
  bool exact_ = true;

  // resident scenes
  std::vector<HostScene> scenes_;
  std::vector<SceneDesc<R>> desc_host_;
  uint64_t NSC_ = 0, NBC_ = 0;  // capacities = SoA strides
  bool shapes_set_ = false, bodies_set_ = false, records_valid_ = false;
  bool aabb_current_ = false, aabb_from_caller_ = false, pairs_given_ = false;
  uint64_t given_pairs_ = 0;

  // current selection
  std::vector<uint32_t> selection_;
  bool selection_dirty_ = true, explicit_selection_ = false;
  uint64_t nchunks_ = 0, nslots_ = 0, chunk_cap_ = 0, sel_cap_ = 0, desc_cap_ = 0;
  uint32_t nsel_ = 0;
  int key_bits_ = 0, shape_bits_ = 0;

  // capacities and totals
  uint64_t mcap_ = 0, pcap_ = 0, m_seen_ = 0, p_seen_ = 0;
  double cap_m_per_shape_ = 6.0, cap_p_per_shape_ = 4.0;
  uint64_t M_ = 0, A_ = 0, P_ = 0, C_ = 0;
  StepCounts counts_;
  std::vector<SceneOut> souts_host_;

  // graph of a capacity step
  hipGraphExec_t graph_exec_ = nullptr;
  bool graph_valid_ = false;
  bool has_ext_ = false;  // a resident scene holds a rectangle, cone or triangle (sticky)
  int graph_key_bits_ = -1, graph_shape_bits_ = -1;
  bool pending_step_ = false, pending_graph_ = false;

  R envelope_ = 0;
  int mode_ = NP_NATIVE;
  float tau_ = 1e-2f;

  DevBuf status_, stephdr_, stage_, upd_, tmp_out_, longs_, meta_, scount_, soff_, selrng_, bins_;
  void* stage_host_ = nullptr;
  size_t stage_host_bytes_ = 0;
  hipEvent_t ev_stage_ = nullptr;  // after the last copies out of stage_host_
  bool stage_pending_ = false;
  void* souts_pinned_ = nullptr;
  size_t souts_pinned_bytes_ = 0;
  void* pf_dst_ = nullptr;  // contacts prefetch target of the next step
  uint64_t pf_rows_ = 0;
  bool pf_ok_ = false;      // every contact of the last finished step is there
  bool desc_dirty_ = true;
  DevBuf type_, body_, lpos_, dims_, cpos_, lrot_, bpos_, brot_, active_;
  DevBuf aabb_, rec_, brec_, hdr_, frec_, bpart_, kpart_;
  DevBuf chunks_, sel_, desc_, souts_;
  DevBuf counts_buf_, offs_, keys_, keys_tmp_, pcount_, pairs_;
  DevBuf np_v_, np_hit_, np_kind_, np_pos_, queue_, queue_tmp_, qsel_, qcls_;
  DevBuf tile_parts_;
  uint64_t tp_T_ = 0, tp_F_ = 0;
  DevBuf out_, perm_;  // out_: the step's contact arrays back to back, C rows each
  HostStager stager_;  // large copies to and from caller memory
  DevBuf scan_tmp_, radix_tmp_, runs_part_;
  bool opened_ = false;  // the step's totals were zeroed by one k_stage_begin
  bool mpr_in_tmp_ = false;  // the last reference-order MPR launch read queue_tmp_ (else queue_)
  uint64_t mpr_q_ = 0;       // and this many pairs of it
  // cell search (exact mode)
  int pair_search_ = PAIRS_AUTO;
  bool cell_ = false;          // the last broadphase found its pairs by cells
  bool bins_pending_ = false;  // its bin entries are not built yet
  bool bins_lost_ = false;     // and can no longer be: the AABBs changed
  bool bins_counted_ = true;   // its bin counts, their scan and the key maximum are there
  bool last_cells_ = false;    // the last exact broadphase took the cell search
  int exact_key_bits_ = 0;
  CellGrid cg_{};
  uint64_t ncells_ = 0;
  DevBuf cell_tab_, cell_rank_, cell_lo_, cell_hi_, cell_slab_, cell_last_, cell_wmax_, cell_ovf_;
};

}  // namespace

std::unique_ptr<SceneBase> make_scene(int real_kind, int device) {
  if (real_kind == REAL_F64) return std::unique_ptr<SceneBase>(new Engine<double>(device));
  if (real_kind == REAL_F32) return std::unique_ptr<SceneBase>(new Engine<float>(device));
  throw Error(ERR_INVALID, "real kind must be CDX_REAL_F64 or CDX_REAL_F32");
}

std::unique_ptr<BatchBase> make_batch(int real_kind, int device) {
  if (real_kind == REAL_F64) return std::unique_ptr<BatchBase>(new Engine<double>(device));
  if (real_kind == REAL_F32) return std::unique_ptr<BatchBase>(new Engine<float>(device));
  throw Error(ERR_INVALID, "real kind must be CDX_REAL_F64 or CDX_REAL_F32");
}

}  // namespace cdx
