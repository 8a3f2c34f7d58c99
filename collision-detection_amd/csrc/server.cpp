// Served request loop (declarations in server.hpp).
//
// Scenes live in one batch engine per size class.  Each class has a worker
// thread that forms a step from its queue as soon as the previous step of
// that class has finished (continuous batching): oldest requests first, at
// most one request per scene, up to the class's step size.  Classes step
// on their own HIP streams, so steps of different classes overlap on the
// device.  A request's pose inputs are copied at submit into a pinned
// block, and from there to a device block by an asynchronous copy started
// right away on the server's upload stream; the step waits for that copy
// and applies the block.  The worker reads every contact of the step back
// with one copy (the
// step's contact arrays lie back to back on the device) into a pinned
// result block that the step's requests share.
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <unordered_set>

#include "cdx_util.hpp"

namespace cdx {

namespace {

// Size classes by shape count (upper bounds; the last is open).  Each class
// steps on its own stream; four match the device's hardware queues.
constexpr uint64_t kClassHi[] = {8192, 49152, 131072, ~0ull};
constexpr int kNumClasses = 4;
static_assert(kNumClasses == 4, "the per-class statistics arrays hold four classes");
// A class's steps hold at most this many scenes of its largest size (and
// never more than max_batch_shapes shapes): step buffers, and so every
// launch grid of a recorded step, are sized by it, and a step of a few
// small scenes should not launch workgroups for half a million shapes.
constexpr uint64_t kScenesPerStep = 8;

// Pinned host blocks for request inputs, bucketed by power-of-two size.
class PinnedPool {
 public:
  ~PinnedPool() {
    for (auto& kv : free_)
      for (void* p : kv.second) (void)hipHostFree(p);
  }
  void* get(uint64_t bytes, uint64_t* bucket) {
    uint64_t b = 4096;
    while (b < bytes) b <<= 1;
    *bucket = b;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto& v = free_[b];
      if (!v.empty()) {
        void* p = v.back();
        v.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    GateShared g;
    CDX_HIP_CHECK(hipHostMalloc(&p, b, hipHostMallocDefault));
    return p;
  }
  void put(void* p, uint64_t bucket) {
    std::lock_guard<std::mutex> lk(mu_);
    free_[bucket].push_back(p);
  }

 private:
  std::mutex mu_;
  std::map<uint64_t, std::vector<void*>> free_;
};

PinnedPool& pinned_pool() {
  static PinnedPool* pool = new PinnedPool();  // lives until exit
  return *pool;
}

}  // namespace

ResultBlock::~ResultBlock() {
  if (base) {
    GateShared g;
    (void)hipHostFree(base);
  }
}

// Device pose blocks (bucketed by power-of-two size) and copy events,
// reused across requests; owned jointly by the server and the blocks in use.
struct DevPools {
  std::mutex mu;
  std::map<uint64_t, std::vector<void*>> blocks;
  std::vector<hipEvent_t> events;
  ~DevPools() {
    GateShared g;
    for (auto& kv : blocks)
      for (void* p : kv.second) (void)hipFree(p);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
  void* get_block(uint64_t bytes, uint64_t* bucket) {
    uint64_t b = 4096;
    while (b < bytes) b <<= 1;
    *bucket = b;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto& v = blocks[b];
      if (!v.empty()) {
        void* p = v.back();
        v.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    CDX_HIP_CHECK(hipMalloc(&p, b));
    return p;
  }
  hipEvent_t get_event() {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!events.empty()) {
        hipEvent_t e = events.back();
        events.pop_back();
        return e;
      }
    }
    hipEvent_t e = nullptr;
    CDX_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
  }
  void put(void* p, uint64_t bucket, hipEvent_t e) {
    std::lock_guard<std::mutex> lk(mu);
    if (p) blocks[bucket].push_back(p);
    if (e) events.push_back(e);
  }
};

struct PoseBlock {
  void* p = nullptr;  // pinned host block
  uint64_t bucket = 0;
  void* dev = nullptr;  // device copy, started at submit
  uint64_t dev_bucket = 0;
  hipEvent_t ev = nullptr;  // recorded after that copy
  std::shared_ptr<DevPools> pools;
  // No wait here: a request that ran has had its copy waited for by its
  // step; the copy of one that never ran may still be in flight, but every
  // later copy into the recycled device block is ordered after it on the
  // same stream, and what it reads from the recycled host block only ends
  // up in that device block.
  ~PoseBlock() {
    if (pools) pools->put(dev, dev_bucket, ev);
    if (p) pinned_pool().put(p, bucket);
  }
};

double Server::now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

Server::Server(int device, const ServerConfig& cfg) : dev_(device), cfg_(cfg), tuner_(device, cfg.real_kind) {
  GateShared gate;
  CDX_REQUIRE(cfg.real_kind == REAL_F64 || cfg.real_kind == REAL_F32, ERR_INVALID, "unknown real kind");
  CDX_REQUIRE(cfg.max_batch_shapes > 0 && cfg.max_pending > 0, ERR_INVALID, "batch and queue limits must be positive");
  CDX_HIP_CHECK(hipSetDevice(dev_));
  static_assert(kNumClasses == 4, "one upload stream per size class");
  for (hipStream_t& u : up_) CDX_HIP_CHECK(hipStreamCreateWithFlags(&u, hipStreamNonBlocking));
  pools_ = std::make_shared<DevPools>();
  size_t free_b = 0, total_b = 0;
  CDX_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  budget_ = cfg.hbm_budget ? cfg.hbm_budget : (uint64_t)(0.9 * (double)free_b);
  classes_.resize(kNumClasses);
  uint64_t lo = 0;
  for (int c = 0; c < kNumClasses; ++c) {
    classes_[c] = std::make_unique<Class>();
    Class& k = *classes_[c];
    k.lo = lo;
    k.hi = kClassHi[c];
    lo = k.hi;
    k.max_shapes = k.hi == ~0ull ? cfg.max_batch_shapes : std::min(cfg.max_batch_shapes, kScenesPerStep * k.hi);
    k.eng = make_batch(cfg.real_kind, dev_);
    k.eng->set_np_mode(cfg.np_mode, cfg.tau);
    k.eng->set_envelope(cfg.envelope);
  }
  st_t0_ = now_ms();
  for (int c = 0; c < kNumClasses; ++c) classes_[c]->worker = std::thread(&Server::worker_loop, this, c);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& k : classes_)
    if (k->worker.joinable()) k->worker.join();
  {
    GateShared g;
    for (hipStream_t u : up_) (void)hipStreamSynchronize(u);
  }
  reqs_.clear();
  buffers_.clear();
  pools_.reset();
  GateShared g;
  for (hipStream_t u : up_) (void)hipStreamDestroy(u);
}

int Server::class_of(uint64_t n) const {
  for (int c = 0; c < kNumClasses; ++c)
    if (n < classes_[c]->hi) return c;
  return kNumClasses - 1;
}

uint32_t Server::add_scene(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                           const uint32_t* ID, uint64_t nb, const void* pos, const void* rot, const uint8_t* active,
                           int bx, int by, int bz) {
  CDX_REQUIRE(n > 0 && nb > 0, ERR_INVALID, "a scene needs shapes and bodies");
  const int c = class_of(n);
  Class& k = *classes_[c];
  int g[3] = {bx, by, bz};
  std::lock_guard<std::mutex> elk(k.eng_mu);
  GateShared gate;  // after the engine lock: a worker holding it may need the gate exclusive
  const bool first_in_class = k.eng->scene_count() == 0;
  // footprint: the scene's resident arrays, plus the class's step buffers
  // the first time the class gets a scene
  uint64_t need = k.eng->resident_bytes_estimate(n, nb);
  if (first_in_class) need += k.eng->step_bytes_estimate(std::max<uint64_t>(k.max_shapes, n));
  {
    // check and reserve in one step, so concurrent admissions (of other
    // classes) cannot both pass the check; given back if admission fails
    std::lock_guard<std::mutex> lk(mu_);
    if (committed_ + need > budget_)
      throw Error(ERR_CAPACITY, "scene of " + std::to_string(n) + " shapes needs ~" + std::to_string(need >> 20) +
                                    " MiB of HBM; " + std::to_string((budget_ - committed_) >> 20) +
                                    " MiB of the budget are left");
    committed_ += need;
  }
  struct Reservation {
    Server* s;
    uint64_t bytes;
    ~Reservation() {
      if (bytes) {
        std::lock_guard<std::mutex> lk(s->mu_);
        s->committed_ -= bytes;
      }
    }
  } reservation{this, need};
  if (bx == 0 && by == 0 && bz == 0) {
    TuneChoice t = tuner_.choose(n, T, A, B, C, R, ID, nb, pos, rot, active);
    for (int i = 0; i < 3; ++i) g[i] = t.grid[i];
    if (t.m_per_shape > k.m_per_shape || t.p_per_shape > k.p_per_shape) {
      k.m_per_shape = std::max(k.m_per_shape, t.m_per_shape);
      k.p_per_shape = std::max(k.p_per_shape, t.p_per_shape);
      k.eng->set_capacity_factors(k.m_per_shape * 1.25 + 1.0, k.p_per_shape * 1.25 + 1.0);
    }
  }
  const uint32_t local = k.eng->add_scene(n, T, A, B, C, R, ID, nb, pos, rot, active, g[0], g[1], g[2]);
  // step buffers for the largest step this class can form, sized now
  const uint64_t nsc = k.eng->scene_count();
  k.min_n = std::min(k.min_n, n);  // a step holds at most max_shapes / min_n scenes
  const uint64_t per_step = std::min<uint64_t>(nsc, std::max<uint64_t>(1, k.max_shapes / k.min_n));
  k.eng->reserve_step(std::max<uint64_t>(k.max_shapes, n), per_step);
  reservation.bytes = 0;  // admitted: the footprint stays committed
  std::lock_guard<std::mutex> lk(mu_);
  SceneRef r;
  r.cls = c;
  r.local = local;
  r.n = n;
  r.nb = nb;
  r.body_begin = k.eng->scene_body_begin(local);
  for (int i = 0; i < 3; ++i) r.bins[i] = g[i];
  scenes_.push_back(r);
  return (uint32_t)(scenes_.size() - 1);
}

uint64_t Server::scene_count() const {
  std::lock_guard<std::mutex> lk(mu_);
  return scenes_.size();
}

void Server::scene_info(uint32_t s, uint64_t* n, uint64_t* nb, int* cls, int bins[3]) const {
  std::lock_guard<std::mutex> lk(mu_);
  CDX_REQUIRE(s < scenes_.size(), ERR_INVALID, "scene index out of range");
  const SceneRef& r = scenes_[s];
  if (n) *n = r.n;
  if (nb) *nb = r.nb;
  if (cls) *cls = r.cls;
  if (bins)
    for (int i = 0; i < 3; ++i) bins[i] = r.bins[i];
}

uint64_t Server::submit(uint32_t s, uint64_t k, const uint32_t* idx, const void* pos, const void* rot) {
  const double t_enter = now_ms();
  SceneRef ref;
  {
    std::lock_guard<std::mutex> lk(mu_);
    CDX_REQUIRE(s < scenes_.size(), ERR_INVALID, "scene index out of range");
    CDX_REQUIRE(pending_ < cfg_.max_pending, ERR_CAPACITY, "request queue is full");
    ref = scenes_[s];
  }
  CDX_REQUIRE(k == 0 || (pos && rot), ERR_INVALID, "null pose input");
  CDX_REQUIRE(idx || k <= ref.nb, ERR_INVALID, "more pose updates than bodies");
  const int kind = cfg_.real_kind;
  const uint64_t rs = kind == REAL_F64 ? 8 : 4;
  auto req = std::make_shared<Request>();
  req->scene = s;
  req->local = ref.local;
  req->n = ref.n;
  req->body_base = ref.body_begin;
  req->k = k;
  if (k) {
    auto blk = std::make_shared<PoseBlock>();
    blk->p = pinned_pool().get(pose_block_bytes(k, kind), &blk->bucket);
    if (idx) {
      uint32_t* gi = static_cast<uint32_t*>(blk->p);
      for (uint64_t j = 0; j < k; ++j) {
        CDX_REQUIRE(idx[j] < ref.nb, ERR_INVALID, "body index out of range in a pose update");
        gi[j] = (uint32_t)(ref.body_begin + idx[j]);
      }
    }
    req->contiguous = idx == nullptr;
    char* base = static_cast<char*>(blk->p);
    std::memcpy(base + pose_block_pos_offset(k), pos, k * 3 * rs);
    std::memcpy(base + pose_block_rot_offset(k, kind), rot, k * 4 * rs);
    req->poses = std::move(blk);
  }
  return enqueue(std::move(req), ref.cls, t_enter);
}

void Server::upload_poses(Request& r, int cls) {
  if (!r.k || !r.poses) return;
  PoseBlock& pb = *r.poses;
  const uint64_t bytes = pose_block_bytes(r.k, cfg_.real_kind);
  const uint64_t from = r.contiguous ? pose_block_pos_offset(r.k) : 0;  // contiguous: no id list
  GateShared g;
  pb.pools = pools_;
  pb.dev = pools_->get_block(bytes, &pb.dev_bucket);
  pb.ev = pools_->get_event();
  CDX_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(pb.dev) + from, static_cast<const char*>(pb.p) + from,
                               bytes - from, hipMemcpyHostToDevice, up_[cls]));
  CDX_HIP_CHECK(hipEventRecord(pb.ev, up_[cls]));
}

uint64_t Server::enqueue(std::shared_ptr<Request> req, int cls, double t_enter) {
  upload_poses(*req, cls);
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    CDX_REQUIRE(pending_ < cfg_.max_pending, ERR_CAPACITY, "request queue is full");
    id = next_id_++;
    req->id = id;
    req->t_submit = t_enter;  // latency includes the input copy, if any
    reqs_[id] = req;
    classes_[cls]->queue.push_back(req);
    ++pending_;
    ++st_sub_;
  }
  cv_work_.notify_all();
  return id;
}

uint64_t Server::pose_buffer(uint32_t s, uint64_t k, void** pos, void** rot) {
  SceneRef ref;
  {
    std::lock_guard<std::mutex> lk(mu_);
    CDX_REQUIRE(s < scenes_.size(), ERR_INVALID, "scene index out of range");
    ref = scenes_[s];
  }
  CDX_REQUIRE(k > 0 && k <= ref.nb, ERR_INVALID, "pose buffer size must be in 1..bodies of the scene");
  CDX_REQUIRE(pos && rot, ERR_INVALID, "null output");
  const int kind = cfg_.real_kind;
  auto req = std::make_shared<Request>();
  req->scene = s;
  req->local = ref.local;
  req->n = ref.n;
  req->body_base = ref.body_begin;
  req->k = k;
  req->contiguous = true;
  auto blk = std::make_shared<PoseBlock>();
  blk->p = pinned_pool().get(pose_block_bytes(k, kind), &blk->bucket);
  char* base = static_cast<char*>(blk->p);
  *pos = base + pose_block_pos_offset(k);
  *rot = base + pose_block_rot_offset(k, kind);
  req->poses = std::move(blk);
  std::lock_guard<std::mutex> lk(mu_);
  const uint64_t h = next_buffer_++;
  buffers_[h] = std::move(req);
  return h;
}

uint64_t Server::submit_buffer(uint64_t h) {
  const double t_enter = now_ms();
  std::shared_ptr<Request> req;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = buffers_.find(h);
    CDX_REQUIRE(it != buffers_.end(), ERR_INVALID, "unknown pose buffer");
    CDX_REQUIRE(pending_ < cfg_.max_pending, ERR_CAPACITY, "request queue is full");
    req = std::move(it->second);
    buffers_.erase(it);
  }
  const int cls = scenes_cls(req->scene);
  return enqueue(std::move(req), cls, t_enter);
}

void Server::drop_buffer(uint64_t h) {
  std::shared_ptr<Request> req;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = buffers_.find(h);
  CDX_REQUIRE(it != buffers_.end(), ERR_INVALID, "unknown pose buffer");
  req = std::move(it->second);
  buffers_.erase(it);
}

int Server::scenes_cls(uint32_t s) const {
  std::lock_guard<std::mutex> lk(mu_);
  return scenes_[s].cls;
}

void Server::worker_loop(int c) {
  {
    GateShared gate;
    (void)hipSetDevice(dev_);
  }
  Class& k = *classes_[c];
  std::vector<std::shared_ptr<Request>> batch;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !k.queue.empty(); });
      if (stop_) {
        for (auto& r : k.queue) {
          r->state = 3;
          r->err = "server stopped";
        }
        k.queue.clear();
        cv_done_.notify_all();
        return;
      }
      uint64_t shapes = 0;
      std::unordered_set<uint32_t> taken;
      for (auto it = k.queue.begin(); it != k.queue.end();) {
        const auto& r = *it;
        if (taken.count(r->scene)) { ++it; continue; }  // the scene's next step waits for this one
        const uint64_t n = r->n;
        if (!batch.empty() && shapes + n > k.max_shapes) break;
        shapes += n;
        taken.insert(r->scene);
        r->state = 1;
        batch.push_back(r);
        it = k.queue.erase(it);
      }
    }
    run_batch(c, batch);
  }
}

// Result blocks come in power-of-two capacities of at least 4,096 contacts.
// A step takes an idle block whose capacity is its need rounded up to a
// power of two, or up to twice that: blocks stay near the size of the steps
// they serve, and a new one is pinned (about a millisecond per 50 MB, with
// the worker waiting) only while every block of the size is held by results
// the caller has not released.  The pool keeps what it allocated.
std::shared_ptr<ResultBlock> Server::get_block(Class& k, uint64_t contacts) {
  uint64_t cap = 4096;
  while (cap < contacts) cap <<= 1;
  for (auto& b : k.blocks)
    if (b.use_count() == 1 && b->cap >= cap && b->cap <= 2 * cap) return b;
  auto b = std::make_shared<ResultBlock>();
  const uint64_t rs = cfg_.real_kind == REAL_F64 ? 8 : 4;
  {
    GateShared g;
    CDX_HIP_CHECK(hipHostMalloc(&b->base, cap * (10 * rs + 8), hipHostMallocDefault));
  }
  b->cap = cap;
  k.blocks.push_back(b);
  return b;
}

void Server::run_batch(int c, std::vector<std::shared_ptr<Request>>& batch) {
  Class& k = *classes_[c];
  std::string err;
  double step_ms = 0.0;
  double ph[6] = {0, 0, 0, 0, 0, 0};
  try {
    std::lock_guard<std::mutex> elk(k.eng_mu);
    GateShared gate;
    BatchBase& e = *k.eng;
    std::vector<uint32_t> sel;
    sel.reserve(batch.size());
    double t0 = now_ms();
    for (auto& r : batch) {
      const int64_t first = r->contiguous ? (int64_t)r->body_base : -1;
      if (r->k && r->poses->dev) e.update_bodies_device(r->k, r->poses->dev, first, r->poses->ev);
      else if (r->k) e.update_bodies_pinned(r->k, r->poses->p, first);
      sel.push_back(r->local);
    }
    double t1 = now_ms();
    e.select(sel.data(), sel.size());
    // Contacts expected: each scene's count at its previous step, plus a
    // quarter.  When every scene of the step has one, the contacts are
    // copied right behind the step into a block of that size, and a step
    // that fits needs no copy of its own after finish().
    std::shared_ptr<ResultBlock> blk;
    uint64_t expect = 0;
    bool known = true;
    for (auto& r : batch) {
      if (r->local >= k.last_contacts.size() || k.last_contacts[r->local] == ~0ull) {
        known = false;
        break;
      }
      expect += k.last_contacts[r->local];
    }
    if (known) {
      const uint64_t rows = expect + expect / 4 + 64;
      blk = get_block(k, rows);
      e.prefetch_contacts(blk->base, rows);
    }
    double t2 = now_ms();
    e.step(cfg_.use_graph);
    double t3 = now_ms();
    e.finish();
    double t4 = now_ms();
    step_ms = e.last_step_ms();
    const uint64_t C = e.totals().c;
    double t5 = t4;
    if (!(blk && e.contacts_prefetched())) {
      blk.reset();
      blk = get_block(k, C);
      t5 = now_ms();
      e.read_all_contacts_packed(blk->base);  // one copy: the arrays are C rows each
    }
    blk->rows = C;
    double t6 = now_ms();
    ph[0] = t1 - t0; ph[1] = t2 - t1; ph[2] = t3 - t2; ph[3] = t4 - t3; ph[4] = t5 - t4; ph[5] = t6 - t5;
    for (uint64_t j = 0; j < batch.size(); ++j) {
      uint64_t b0, b1;
      e.scene_contact_range(j, &b0, &b1);
      batch[j]->block = blk;
      batch[j]->c_begin = b0;
      batch[j]->c_count = b1 - b0;
      const uint32_t loc = batch[j]->local;
      if (loc >= k.last_contacts.size()) k.last_contacts.resize(loc + 1, ~0ull);
      k.last_contacts[loc] = b1 - b0;
    }
  } catch (const std::exception& ex) {
    err = ex.what();
  }
  const double t = now_ms();
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& r : batch) {
      r->poses.reset();
      --pending_;
      if (err.empty()) {
        r->state = 2;
        r->latency_ms = t - r->t_submit;
        lat_.push_back(r->latency_ms);
        lat_cls_[c].push_back(r->latency_ms);
        ++st_done_;
      } else {
        r->state = 3;
        r->err = err;
        ++st_fail_;
      }
    }
    ++st_batches_;
    ++st_steps_cls_[c];
    st_breqs_ += batch.size();
    st_busy_ += step_ms;
    st_busy_cls_[c] += step_ms;
    for (int q = 0; q < 6; ++q) st_phase_[q] += ph[q];
  }
  cv_done_.notify_all();
}

int Server::poll(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  CDX_REQUIRE(it != reqs_.end(), ERR_INVALID, "unknown request id");
  if (it->second->state == 3) throw Error(ERR_INVALID, "request failed: " + it->second->err);
  return it->second->state == 2 ? 1 : 0;
}

int Server::wait(uint64_t id, double timeout_ms) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  CDX_REQUIRE(it != reqs_.end(), ERR_INVALID, "unknown request id");
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->state >= 2; };
  if (timeout_ms < 0) cv_done_.wait(lk, ready);
  else if (!cv_done_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) return 0;
  if (r->state == 3) throw Error(ERR_INVALID, "request failed: " + r->err);
  return 1;
}

void Server::result_counts(uint64_t id, uint64_t* contacts, double* latency_ms) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  CDX_REQUIRE(it != reqs_.end(), ERR_INVALID, "unknown request id");
  const Request& r = *it->second;
  CDX_REQUIRE(r.state == 2, ERR_NOT_READY, "request not finished");
  if (contacts) *contacts = r.c_count;
  if (latency_ms) *latency_ms = r.latency_ms;
}

void Server::read_contacts(uint64_t id, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids, uint64_t cap) {
  std::shared_ptr<Request> r;
  uint64_t body_base;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    CDX_REQUIRE(it != reqs_.end(), ERR_INVALID, "unknown request id");
    r = it->second;
    CDX_REQUIRE(r->state == 2, ERR_NOT_READY, "request not finished");
    body_base = r->body_base;
  }
  const uint64_t n = r->c_count;
  CDX_REQUIRE(cap >= n, ERR_CAPACITY, "contact buffers too small: need " + std::to_string(n));
  if (!n) return;
  const uint64_t rs = cfg_.real_kind == REAL_F64 ? 8 : 4;
  const char* base = static_cast<const char*>(r->block->base);
  const uint64_t bc = r->block->rows, b0 = r->c_begin;
  if (norm) std::memcpy(norm, base + 3 * rs * b0, 3 * rs * n);
  if (cpta) std::memcpy(cpta, base + 3 * rs * (bc + b0), 3 * rs * n);
  if (cptb) std::memcpy(cptb, base + 3 * rs * (2 * bc + b0), 3 * rs * n);
  if (dpth) std::memcpy(dpth, base + 9 * rs * bc + rs * b0, rs * n);
  if (bids) {
    const int32_t* ids = reinterpret_cast<const int32_t*>(base + 10 * rs * bc) + 2 * b0;
    for (uint64_t i = 0; i < 2 * n; ++i) bids[i] = ids[i] - (int32_t)body_base;
  }
}

void Server::release(uint64_t id) {
  std::shared_ptr<Request> r;  // destroyed after the lock: it may free a pinned block
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  CDX_REQUIRE(it != reqs_.end(), ERR_INVALID, "unknown request id");
  CDX_REQUIRE(it->second->state >= 2, ERR_NOT_READY, "request still queued or running");
  r = std::move(it->second);
  reqs_.erase(it);
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  ServerStats s{};
  s.submitted = st_sub_;
  s.completed = st_done_;
  s.failed = st_fail_;
  s.batches = st_batches_;
  s.batched_requests = st_breqs_;
  s.busy_ms = st_busy_;
  for (int q = 0; q < 6; ++q) s.host_ms[q] = st_phase_[q];
  s.window_s = (now_ms() - st_t0_) / 1000.0;
  // nearest-rank quantile of a sorted sample
  auto q = [](const std::vector<double>& v, double f) {
    size_t i = (size_t)std::ceil(f * (double)v.size()) - 1;
    return v[std::min(i, v.size() - 1)];
  };
  if (!lat_.empty()) {
    std::vector<double> v = lat_;
    std::sort(v.begin(), v.end());
    s.p50_ms = q(v, 0.50);
    s.p99_ms = q(v, 0.99);
    s.max_ms = v.back();
    double sum = 0.0;
    for (double x : v) sum += x;
    s.mean_ms = sum / (double)v.size();
  }
  for (int c = 0; c < kNumClasses; ++c) {
    s.class_completed[c] = lat_cls_[c].size();
    s.class_steps[c] = st_steps_cls_[c];
    s.class_busy_ms[c] = st_busy_cls_[c];
    s.class_p50_ms[c] = s.class_p99_ms[c] = 0.0;
    if (!lat_cls_[c].empty()) {
      std::vector<double> v = lat_cls_[c];
      std::sort(v.begin(), v.end());
      s.class_p50_ms[c] = q(v, 0.50);
      s.class_p99_ms[c] = q(v, 0.99);
    }
  }
  return s;
}

void Server::reset_stats() {
  std::lock_guard<std::mutex> lk(mu_);
  lat_.clear();
  for (int c = 0; c < kNumClasses; ++c) {
    lat_cls_[c].clear();
    st_steps_cls_[c] = 0;
    st_busy_cls_[c] = 0.0;
  }
  st_sub_ = st_done_ = st_fail_ = st_batches_ = st_breqs_ = 0;
  st_busy_ = 0.0;
  for (int q = 0; q < 6; ++q) st_phase_[q] = 0.0;
  st_t0_ = now_ms();
}

uint64_t Server::device_bytes() {
  uint64_t t = 0;
  for (auto& k : classes_) {
    std::lock_guard<std::mutex> elk(k->eng_mu);
    t += k->eng->device_bytes();
  }
  return t;
}

}  // namespace cdx
