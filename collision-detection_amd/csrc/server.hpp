// Served request loop over device-resident scenes: continuous batching per
// scene-size class.  No reference counterpart (the reference API runs one
// synchronous call per scene); a request carries the pose inputs of
// AABBGenerator::GenerateAABB (AABBGenerator.cuh:21-30) for one scene and
// returns that scene's contacts as Narrowphase::DoNarrowphase does
// (Narrowphase.cuh:19-34).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "autotune.hpp"
#include "cdx_pipeline.hpp"

namespace cdx {

struct ServerConfig {
  int real_kind = REAL_F64;
  int np_mode = NP_MIXED;
  double tau = 1e-2;
  double envelope = 0.0;
  uint64_t hbm_budget = 0;          // bytes of resident scenes + step buffers; 0 = 90 % of free HBM
  uint64_t max_batch_shapes = 1u << 19;  // shapes per batched step (bounds step latency)
  uint64_t max_pending = 1u << 16;       // queued requests before submit refuses
  bool use_graph = true;
};

struct ServerStats {
  uint64_t submitted, completed, failed, batches, batched_requests;
  double p50_ms, p99_ms, mean_ms, max_ms;  // submit -> result ready, over the window
  double busy_ms;                          // summed device step time of the window
  double window_s;                         // wall time since the window started
  // host time of the workers' step phases, summed: pose uploads, selection,
  // launch, wait for the step, result block, contact readback
  double host_ms[6];
  // per size class
  uint64_t class_completed[4], class_steps[4];
  double class_p50_ms[4], class_p99_ms[4], class_busy_ms[4];
};

struct PoseBlock;
struct DevPools;

// Pinned host block holding the contacts of one batched step.
// The step's contact arrays lie back to back, each `rows` (the step's
// contact count) long: normals, points A, points B, depths, body id pairs.
struct ResultBlock {
  void* base = nullptr;
  uint64_t cap = 0;   // contacts
  uint64_t rows = 0;  // contacts of the step that filled it
  ~ResultBlock();
};

class Server {
 public:
  Server(int device, const ServerConfig& cfg);
  ~Server();

  // Resident scene (host arrays in the reference layouts; body ids local).
  // bins = 0,0,0 autotunes.  Refuses (ERR_CAPACITY) a scene whose estimated
  // footprint does not fit the HBM budget.
  uint32_t add_scene(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                     const uint32_t* ID,
                     uint64_t nb, const void* pos, const void* rot, const uint8_t* active, int bx, int by, int bz);
  uint64_t scene_count() const;
  void scene_info(uint32_t scene, uint64_t* n_shapes, uint64_t* n_bodies, int* size_class, int bins[3]) const;

  // Queues a step request for one scene with k pose updates (idx NULL:
  // bodies 0..k-1).  The inputs are copied before return.  Returns the
  // request id.
  uint64_t submit(uint32_t scene, uint64_t k, const uint32_t* idx, const void* pos, const void* rot);
  // Zero-copy submission: a pinned block for the poses of bodies 0..k-1 of
  // a scene, filled by the caller in place (k real3 positions, k real4
  // rotations), then queued by submit_buffer (which takes it over).
  uint64_t pose_buffer(uint32_t scene, uint64_t k, void** pos, void** rot);
  uint64_t submit_buffer(uint64_t handle);
  void drop_buffer(uint64_t handle);
  // 1 when the result is ready, 0 while queued or running; throws on a
  // failed request.
  int poll(uint64_t req);
  // Blocks until the request is done (timeout_ms < 0: no limit).  Returns
  // 1 when done, 0 on timeout.
  int wait(uint64_t req, double timeout_ms);
  void result_counts(uint64_t req, uint64_t* contacts, double* latency_ms);
  void read_contacts(uint64_t req, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids, uint64_t cap);
  void release(uint64_t req);

  ServerStats stats();
  void reset_stats();
  uint64_t device_bytes();
  uint64_t hbm_budget() const { return budget_; }
  uint64_t hbm_committed() const { return committed_; }

 private:
  struct Request {
    uint64_t id = 0;
    uint32_t scene = 0, local = 0;  // server scene id; index in its class engine
    uint64_t n = 0;                 // shapes of the scene
    uint64_t body_base = 0;         // global body id of the scene's body 0
    uint64_t k = 0;
    bool contiguous = false;  // updates cover bodies 0..k-1 (no id list)
    std::shared_ptr<PoseBlock> poses;  // pinned: global ids, positions, rotations (+ device copy)
    double t_submit = 0.0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    std::string err;
    std::shared_ptr<ResultBlock> block;
    uint64_t c_begin = 0, c_count = 0;
    double latency_ms = 0.0;
  };
  struct Class {
    uint64_t lo, hi;  // shape-count range [lo, hi)
    uint64_t max_shapes;  // shapes per step of this class
    uint64_t min_n = ~0ull;  // smallest resident scene (bounds the scenes per step)
    // largest bin entries / pairs per shape the autotuner measured in this
    // class (guarded by eng_mu, like the engine they size)
    double m_per_shape = 0.0, p_per_shape = 0.0;
    std::unique_ptr<BatchBase> eng;
    std::mutex eng_mu;  // the engine is used by the worker and add_scene
    std::deque<std::shared_ptr<Request>> queue;
    std::thread worker;
    std::vector<std::shared_ptr<ResultBlock>> blocks;  // pool, used by the worker only
    // contacts of each scene (by engine index) at its last step, ~0 before
    // its first (worker only, under eng_mu): sizes the contacts prefetch
    std::vector<uint64_t> last_contacts;
  };
  struct SceneRef {
    int cls;
    uint32_t local;
    uint64_t n, nb, body_begin;
    int bins[3];
  };

  uint64_t enqueue(std::shared_ptr<Request> req, int cls, double t_enter);
  void worker_loop(int c);
  void run_batch(int c, std::vector<std::shared_ptr<Request>>& batch);
  std::shared_ptr<ResultBlock> get_block(Class& k, uint64_t contacts);
  int class_of(uint64_t n) const;
  int scenes_cls(uint32_t s) const;
  static double now_ms();
  // Starts the copy of a request's poses to device memory on up_ (at submit,
  // so the copy overlaps the queueing and the steps before it).
  void upload_poses(Request& r, int cls);

  int dev_;
  // pose copies at submit, one stream per size class: a small scene's copy
  // does not wait behind a large scene's
  hipStream_t up_[4] = {nullptr, nullptr, nullptr, nullptr};
  std::shared_ptr<DevPools> pools_;  // device pose blocks and their copy events
  ServerConfig cfg_;
  uint64_t budget_ = 0, committed_ = 0;
  std::vector<std::unique_ptr<Class>> classes_;
  std::vector<SceneRef> scenes_;
  Autotuner tuner_;
  mutable std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  bool stop_ = false;
  uint64_t next_id_ = 1;
  uint64_t pending_ = 0;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs_;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> buffers_;  // filled by the caller, not yet queued
  uint64_t next_buffer_ = 1;
  // window statistics
  std::vector<double> lat_;
  std::vector<double> lat_cls_[4];
  uint64_t st_steps_cls_[4] = {0, 0, 0, 0};
  double st_busy_cls_[4] = {0, 0, 0, 0};
  uint64_t st_sub_ = 0, st_done_ = 0, st_fail_ = 0, st_batches_ = 0, st_breqs_ = 0;
  double st_busy_ = 0.0, st_t0_ = 0.0;
  double st_phase_[6] = {0, 0, 0, 0, 0, 0};
};

}  // namespace cdx
