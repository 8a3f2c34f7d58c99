// extern "C" layer of libcdx.so (declarations in include/cdx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cdx.h"
#include "autotune.hpp"
#include "cdx_noalloc.hpp"
#include "cdx_pipeline.hpp"
#include "server.hpp"
#include "cdx_util.hpp"

struct cdx_ctx {
  int device = 0;
  std::string err;
  std::unique_ptr<cdx::SceneBase> aabb_scene[2];
  std::unique_ptr<cdx::SceneBase> bp_scene[2];
  std::unique_ptr<cdx::SceneBase> np_scene[2];
  int last_bp_kind = -1, last_np_kind = -1;
};

struct cdx_scene {
  cdx_ctx* ctx = nullptr;
  std::unique_ptr<cdx::SceneBase> s;
  std::string err;
};

struct cdx_batch {
  cdx_ctx* ctx = nullptr;
  int kind = 0;
  std::unique_ptr<cdx::BatchBase> b;
  std::unique_ptr<cdx::Autotuner> tuner;
  std::vector<std::array<int, 3>> bins;
  double m_per_shape = 0.0, p_per_shape = 0.0;  // largest measured by the autotuner
  std::string err;
};

struct cdx_server {
  std::unique_ptr<cdx::Server> srv;
};

namespace {

thread_local std::string t_server_err;

// Runs f with the HIP gate held shared (gate = true) and turns exceptions
// into status codes.  Server calls pass gate = false: they may block on
// other threads' work and take the gate themselves around their HIP calls.
template <class F> int guarded(std::string* err, F&& f, bool gate = true) {
  try {
    if (gate) {
      cdx::GateShared g;
      f();
    } else {
      f();
    }
    if (err) err->clear();
    return CDX_OK;
  } catch (const cdx::Error& e) {
    if (err) *err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (err) *err = "host allocation failed";
    return CDX_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return CDX_ERR_INVALID;
  }
}

void fill_counts(const cdx::StepCounts& c, cdx_counts* out) {
  out->n_shapes = c.n; out->bin_entries = c.m; out->active_bins = c.a; out->pairs = c.p;
  out->contacts = c.c; out->resolved = c.resolved; out->key_max = c.key_max; out->flags = c.bad | c.info;
  for (int i = 0; i < 3; ++i) { out->bmin[i] = c.bmin[i]; out->bmax[i] = c.bmax[i]; }
}

void check_kind(int k) {
  CDX_REQUIRE(k == CDX_REAL_F64 || k == CDX_REAL_F32, CDX_ERR_INVALID, "real_kind must be CDX_REAL_F64 or CDX_REAL_F32");
}

cdx::SceneBase& lazy(std::unique_ptr<cdx::SceneBase>& p, int kind, int device) {
  if (!p) p = cdx::make_scene(kind, device);
  return *p;
}

}  // namespace

extern "C" {

int cdx_version(void) { return CDX_VERSION; }

const char* cdx_status_string(int s) {
  switch (s) {
    case CDX_OK: return "ok";
    case CDX_ERR_INVALID: return "invalid argument";
    case CDX_ERR_HIP: return "HIP runtime error";
    case CDX_ERR_UNSUPPORTED_SHAPE: return "unsupported shape type";
    case CDX_ERR_CAPACITY: return "capacity exceeded";
    case CDX_ERR_NO_DEVICE: return "no GPU device";
    case CDX_ERR_OUT_OF_MEMORY: return "out of memory";
    case CDX_ERR_NOT_READY: return "stage inputs not ready";
    default: return "unknown status";
  }
}

int cdx_device_count(int* count) {
  return guarded(nullptr, [&] {
    CDX_REQUIRE(count, CDX_ERR_INVALID, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
  });
}

int cdx_ctx_create(int device, cdx_ctx** out) {
  return guarded(nullptr, [&] {
    CDX_REQUIRE(out, CDX_ERR_INVALID, "null output");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) throw cdx::Error(CDX_ERR_NO_DEVICE, "no HIP device visible");
    CDX_REQUIRE(device >= 0 && device < c, CDX_ERR_INVALID, "device index out of range");
    CDX_HIP_CHECK(hipSetDevice(device));
    auto* x = new cdx_ctx();
    x->device = device;
    *out = x;
  });
}

int cdx_ctx_destroy(cdx_ctx* ctx) {
  delete ctx;
  return CDX_OK;
}

const char* cdx_ctx_last_error(const cdx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cdx_generate_aabb(cdx_ctx* ctx, int real_kind, uint64_t n, const int32_t* T, const void* A, const void* B,
                      const void* C, const void* R, const uint32_t* ID, uint64_t nb, const void* body_pos,
                      const void* body_rot, void* aabb_out) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    if (n == 0) return;
    CDX_REQUIRE(T && A && B && R && ID && body_pos && body_rot && aabb_out, CDX_ERR_INVALID, "null input");
    cdx::SceneBase& s = lazy(ctx->aabb_scene[real_kind], real_kind, ctx->device);
    s.set_shapes(n, T, A, B, C, R, ID);
    s.set_bodies(nb, body_pos, body_rot, nullptr);
    s.run_aabb();
    s.sync();
    s.read_aabb(aabb_out, false);
  });
}

int cdx_detect_possible_collisions(cdx_ctx* ctx, int real_kind, void* aabb_inout, uint64_t n_aabb, int bx, int by,
                                   int bz, uint64_t* n_pairs) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    CDX_REQUIRE(n_pairs, CDX_ERR_INVALID, "null n_pairs");
    cdx::SceneBase& s = lazy(ctx->bp_scene[real_kind], real_kind, ctx->device);
    ctx->last_bp_kind = real_kind;
    s.set_grid(bx, by, bz);
    CDX_REQUIRE(n_aabb == 0 || aabb_inout, CDX_ERR_INVALID, "null aabb_data");
    s.set_aabbs(n_aabb, aabb_inout);
    s.run_broadphase();
    s.sync();
    if (n_aabb) s.read_aabb(aabb_inout, true);
    *n_pairs = s.counts().p;
  });
}

int cdx_detect_possible_collisions_noalloc(cdx_ctx* ctx, int real_kind, void* aabb_inout, uint64_t n_aabb, int bx,
                                           int by, int bz, int64_t* pairs_out, uint64_t pair_capacity,
                                           uint64_t* n_pairs, void* stream) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    CDX_REQUIRE(n_pairs, CDX_ERR_INVALID, "null n_pairs");
    *n_pairs = 0;
    *n_pairs = cdx::noalloc_broadphase(ctx->device, real_kind, aabb_inout, n_aabb, bx, by, bz, pairs_out,
                                       pair_capacity, stream);
  });
}

int cdx_read_possible_collisions(cdx_ctx* ctx, int64_t* out, uint64_t capacity) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    CDX_REQUIRE(ctx->last_bp_kind >= 0, CDX_ERR_NOT_READY, "no broadphase has run on this context");
    cdx::SceneBase& s = *ctx->bp_scene[ctx->last_bp_kind];
    uint64_t p = s.counts().p;
    CDX_REQUIRE(capacity >= p, CDX_ERR_CAPACITY, "pair buffer too small: need " + std::to_string(p));
    CDX_REQUIRE(p == 0 || out, CDX_ERR_INVALID, "null output");
    s.read_pairs(out);
  });
}

int cdx_do_narrowphase(cdx_ctx* ctx, int real_kind, int np_mode, double envelope, uint64_t n, const int32_t* T,
                       const void* A, const void* B, const void* C, const void* R, const uint32_t* ID, uint64_t nb,
                       const uint8_t* obj_active, const void* body_pos, const void* body_rot, int64_t* pairs_inout,
                       uint64_t n_pairs, uint64_t* n_contacts) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    CDX_REQUIRE(n_contacts, CDX_ERR_INVALID, "null n_contacts");
    cdx::SceneBase& s = lazy(ctx->np_scene[real_kind], real_kind, ctx->device);
    ctx->last_np_kind = real_kind;
    s.set_np_mode(np_mode, 1e-2);
    s.set_envelope(envelope);
    s.set_shapes(n, T, A, B, C, R, ID);
    s.set_bodies(nb, body_pos, body_rot, obj_active);
    s.set_pairs(n_pairs, pairs_inout);
    s.run_narrowphase();
    s.sync();
    if (n_pairs) s.read_pairs_permuted(pairs_inout);
    *n_contacts = s.counts().c;
  });
}

int cdx_mpr_store(cdx_ctx* ctx, int real_kind, int np_mode, double envelope, uint64_t n, const int32_t* T,
                  const void* A, const void* B, const void* C, const void* R, const uint32_t* ID, uint64_t nb,
                  const uint8_t* obj_active, const void* body_pos, const void* body_rot, const int64_t* pairs,
                  uint64_t n_pairs, uint32_t* contact_active, void* norm, void* ptA, void* ptB, void* depth,
                  int32_t* ids) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    if (n_pairs == 0) return;
    CDX_REQUIRE(pairs, CDX_ERR_INVALID, "null pairs");
    cdx::SceneBase& s = lazy(ctx->np_scene[real_kind], real_kind, ctx->device);
    ctx->last_np_kind = real_kind;
    s.set_np_mode(np_mode, 1e-2);
    s.set_envelope(envelope);
    s.set_shapes(n, T, A, B, C, R, ID);
    s.set_bodies(nb, body_pos, body_rot, obj_active);
    s.set_pairs(n_pairs, pairs);
    s.run_narrowphase();
    s.sync();
    // contacts come compacted in pair order; put each back at its pair
    std::vector<uint32_t> hit(n_pairs);
    s.read_pair_hits(hit.data());
    const uint64_t c = s.counts().c;
    const size_t rs = real_kind == CDX_REAL_F64 ? 8 : 4;
    std::vector<char> nrm(c * 3 * rs), pa(c * 3 * rs), pb(c * 3 * rs), dp(c * rs);
    std::vector<int32_t> bi(2 * c);
    s.read_contacts(nrm.data(), pa.data(), pb.data(), dp.data(), bi.data());
    uint64_t j = 0;
    for (uint64_t i = 0; i < n_pairs; ++i) {
      if (!hit[i]) continue;
      CDX_REQUIRE(j < c, CDX_ERR_HIP, "internal error: more hit flags than contacts");
      if (norm) std::memcpy(static_cast<char*>(norm) + i * 3 * rs, nrm.data() + j * 3 * rs, 3 * rs);
      if (ptA) std::memcpy(static_cast<char*>(ptA) + i * 3 * rs, pa.data() + j * 3 * rs, 3 * rs);
      if (ptB) std::memcpy(static_cast<char*>(ptB) + i * 3 * rs, pb.data() + j * 3 * rs, 3 * rs);
      if (depth) std::memcpy(static_cast<char*>(depth) + i * rs, dp.data() + j * rs, rs);
      if (ids) { ids[2 * i] = bi[2 * j]; ids[2 * i + 1] = bi[2 * j + 1]; }
      if (contact_active) contact_active[i] = 0;
      ++j;
    }
    CDX_REQUIRE(j == c, CDX_ERR_HIP, "internal error: hit flags and contact count disagree");
  });
}

int cdx_read_contacts(cdx_ctx* ctx, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids,
                      uint64_t capacity) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    CDX_REQUIRE(ctx->last_np_kind >= 0, CDX_ERR_NOT_READY, "no narrowphase has run on this context");
    cdx::SceneBase& s = *ctx->np_scene[ctx->last_np_kind];
    uint64_t c = s.counts().c;
    CDX_REQUIRE(capacity >= c, CDX_ERR_CAPACITY, "contact buffers too small: need " + std::to_string(c));
    s.read_contacts(norm, cpta, cptb, dpth, bids);
  });
}

// ---- scenes ---------------------------------------------------------------

int cdx_scene_create(cdx_ctx* ctx, int real_kind, cdx_scene** out) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    CDX_REQUIRE(out, CDX_ERR_INVALID, "null output");
    auto* sc = new cdx_scene();
    sc->ctx = ctx;
    try {
      sc->s = cdx::make_scene(real_kind, ctx->device);
    } catch (...) {
      delete sc;
      throw;
    }
    *out = sc;
  });
}

int cdx_scene_destroy(cdx_scene* scene) {
  delete scene;
  return CDX_OK;
}

const char* cdx_scene_last_error(const cdx_scene* scene) { return scene ? scene->err.c_str() : "null scene"; }

#define SCENE_CALL(body)                                     \
  if (!scene) return CDX_ERR_INVALID;                        \
  return guarded(&scene->err, [&] { cdx::SceneBase& s = *scene->s; body; })

int cdx_scene_set_shapes(cdx_scene* scene, uint64_t n, const int32_t* T, const void* A, const void* B, const void* C,
                         const void* R, const uint32_t* ID) {
  SCENE_CALL(CDX_REQUIRE(n == 0 || (T && A && B && R && ID), CDX_ERR_INVALID, "null input");
             s.set_shapes(n, T, A, B, C, R, ID));
}
int cdx_scene_set_bodies(cdx_scene* scene, uint64_t nb, const void* pos, const void* rot, const uint8_t* active) {
  SCENE_CALL(CDX_REQUIRE(nb == 0 || (pos && rot), CDX_ERR_INVALID, "null input"); s.set_bodies(nb, pos, rot, active));
}
int cdx_scene_update_bodies(cdx_scene* scene, uint64_t k, const uint32_t* idx, const void* pos, const void* rot) {
  SCENE_CALL(CDX_REQUIRE(k == 0 || (idx && pos && rot), CDX_ERR_INVALID, "null input");
             s.update_bodies(k, idx, pos, rot));
}
int cdx_scene_set_aabbs(cdx_scene* scene, uint64_t n, const void* aabb) {
  SCENE_CALL(CDX_REQUIRE(n == 0 || aabb, CDX_ERR_INVALID, "null input"); s.set_aabbs(n, aabb));
}
int cdx_scene_set_pairs(cdx_scene* scene, uint64_t p, const int64_t* pairs) {
  SCENE_CALL(CDX_REQUIRE(p == 0 || pairs, CDX_ERR_INVALID, "null input"); s.set_pairs(p, pairs));
}
int cdx_scene_set_bins(cdx_scene* scene, int bx, int by, int bz) { SCENE_CALL(s.set_grid(bx, by, bz)); }
int cdx_scene_set_envelope(cdx_scene* scene, double e) { SCENE_CALL(s.set_envelope(e)); }
int cdx_scene_set_narrowphase_mode(cdx_scene* scene, int mode, double tau) {
  SCENE_CALL(CDX_REQUIRE(tau >= 0.0, CDX_ERR_INVALID, "tau must be non-negative"); s.set_np_mode(mode, tau));
}
int cdx_scene_set_profiling(cdx_scene* scene, int on) { SCENE_CALL(s.set_profiling(on != 0)); }
int cdx_scene_set_pair_search(cdx_scene* scene, int mode) { SCENE_CALL(s.set_pair_search(mode)); }
int cdx_scene_aabb(cdx_scene* scene) { SCENE_CALL(s.run_aabb()); }
int cdx_scene_broadphase(cdx_scene* scene) { SCENE_CALL(s.run_broadphase()); }
int cdx_scene_narrowphase(cdx_scene* scene) { SCENE_CALL(s.run_narrowphase()); }
int cdx_scene_step(cdx_scene* scene) { SCENE_CALL(s.run_aabb(); s.run_broadphase(); s.run_narrowphase()); }
int cdx_scene_sync(cdx_scene* scene) { SCENE_CALL(s.sync()); }
int cdx_scene_counts(cdx_scene* scene, cdx_counts* out) {
  SCENE_CALL(CDX_REQUIRE(out, CDX_ERR_INVALID, "null output"); fill_counts(s.counts_full(), out));
}
int cdx_scene_read_aabb(cdx_scene* scene, int shifted, void* out) {
  SCENE_CALL(CDX_REQUIRE(out, CDX_ERR_INVALID, "null output"); s.read_aabb(out, shifted != 0));
}
int cdx_scene_read_bins(cdx_scene* scene, uint32_t* bin_counts, uint32_t* sorted_keys, uint32_t* sorted_bodies,
                        uint32_t* active_keys, uint32_t* active_counts) {
  SCENE_CALL(s.read_bins(bin_counts, sorted_keys, sorted_bodies, active_keys, active_counts));
}
int cdx_scene_read_pairs(cdx_scene* scene, int64_t* out) { SCENE_CALL(s.read_pairs(out)); }
int cdx_scene_read_pairs_permuted(cdx_scene* scene, int64_t* out) { SCENE_CALL(s.read_pairs_permuted(out)); }
int cdx_scene_read_contacts(cdx_scene* scene, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids) {
  SCENE_CALL(s.read_contacts(norm, cpta, cptb, dpth, bids));
}
int cdx_scene_read_mpr_queue(cdx_scene* scene, uint32_t* pair_index, uint64_t capacity, uint64_t* count) {
  SCENE_CALL(CDX_REQUIRE(count, CDX_ERR_INVALID, "null count"); *count = s.read_mpr_queue(pair_index, capacity));
}
static_assert(cdx::ST_COUNT == CDX_STAGE_COUNT, "stage list of cdx.h and the pipeline differ");
static_assert(cdx::PAIRS_AUTO == CDX_PAIRS_AUTO && cdx::PAIRS_BINS == CDX_PAIRS_BINS && cdx::PAIRS_CELLS == CDX_PAIRS_CELLS,
              "pair search modes");
static_assert(cdx::INFO_CELL_SEARCH == CDX_FLAG_CELL_SEARCH && cdx::INFO_NO_ACTIVE_BINS == CDX_FLAG_NO_ACTIVE_BINS,
              "count flags");
static_assert(cdx::ST_NP_MPR == CDX_STAGE_NP_MPR && cdx::ST_BP_PAIREMIT == CDX_STAGE_BP_PAIREMIT, "stage numbers");

int cdx_scene_stage_ms(cdx_scene* scene, double* ms) {
  SCENE_CALL(CDX_REQUIRE(ms, CDX_ERR_INVALID, "null output"); s.stage_ms(ms));
}
int cdx_scene_device_bytes(cdx_scene* scene, uint64_t* bytes) {
  SCENE_CALL(CDX_REQUIRE(bytes, CDX_ERR_INVALID, "null output"); *bytes = s.device_bytes());
}

// ---- batches --------------------------------------------------------------

int cdx_batch_create(cdx_ctx* ctx, int real_kind, cdx_batch** out) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    check_kind(real_kind);
    CDX_REQUIRE(out, CDX_ERR_INVALID, "null output");
    auto bt = std::make_unique<cdx_batch>();
    bt->ctx = ctx;
    bt->kind = real_kind;
    bt->b = cdx::make_batch(real_kind, ctx->device);
    *out = bt.release();
  });
}

int cdx_batch_destroy(cdx_batch* batch) {
  delete batch;
  return CDX_OK;
}

const char* cdx_batch_last_error(const cdx_batch* batch) { return batch ? batch->err.c_str() : "null batch"; }

#define BATCH_CALL(...)                                      \
  if (!batch) return CDX_ERR_INVALID;                        \
  return guarded(&batch->err, [&] { cdx::BatchBase& b = *batch->b; __VA_ARGS__; })

int cdx_batch_add_scene(cdx_batch* batch, uint64_t n, const int32_t* T, const void* A, const void* B, const void* C,
                        const void* R, const uint32_t* ID, uint64_t nb, const void* body_pos, const void* body_rot,
                        const uint8_t* obj_active, int bins_x, int bins_y, int bins_z, uint32_t* scene_out) {
  BATCH_CALL({
    CDX_REQUIRE(n > 0 && nb > 0, CDX_ERR_INVALID, "a scene needs shapes and bodies");
    CDX_REQUIRE(T && A && B && R && ID && body_pos && body_rot, CDX_ERR_INVALID, "null input");
    int g[3] = {bins_x, bins_y, bins_z};
    if (bins_x == 0 && bins_y == 0 && bins_z == 0) {
      if (!batch->tuner) batch->tuner = std::make_unique<cdx::Autotuner>(batch->ctx->device, batch->kind);
      cdx::TuneChoice t = batch->tuner->choose(n, T, A, B, C, R, ID, nb, body_pos, body_rot, obj_active);
      for (int k = 0; k < 3; ++k) g[k] = t.grid[k];
      if (t.m_per_shape > batch->m_per_shape || t.p_per_shape > batch->p_per_shape) {
        batch->m_per_shape = std::max(batch->m_per_shape, t.m_per_shape);
        batch->p_per_shape = std::max(batch->p_per_shape, t.p_per_shape);
        b.set_capacity_factors(batch->m_per_shape * 1.25 + 1.0, batch->p_per_shape * 1.25 + 1.0);
      }
    }
    uint32_t s = b.add_scene(n, T, A, B, C, R, ID, nb, body_pos, body_rot, obj_active, g[0], g[1], g[2]);
    if (batch->bins.size() <= s) batch->bins.resize(s + 1);
    batch->bins[s] = {g[0], g[1], g[2]};
    if (scene_out) *scene_out = s;
  });
}

int cdx_batch_scene_bins(cdx_batch* batch, uint32_t scene, int* bins_xyz) {
  BATCH_CALL((void)b; CDX_REQUIRE(bins_xyz, CDX_ERR_INVALID, "null output");
             CDX_REQUIRE(scene < batch->bins.size(), CDX_ERR_INVALID, "scene index out of range");
             for (int k = 0; k < 3; ++k) bins_xyz[k] = batch->bins[scene][k]);
}
int cdx_batch_scene_count(cdx_batch* batch, uint64_t* count) {
  BATCH_CALL(CDX_REQUIRE(count, CDX_ERR_INVALID, "null output"); *count = b.scene_count());
}
int cdx_batch_update_bodies(cdx_batch* batch, uint32_t scene, uint64_t k, const uint32_t* idx, const void* body_pos,
                            const void* body_rot) {
  BATCH_CALL(CDX_REQUIRE(k == 0 || (idx && body_pos && body_rot), CDX_ERR_INVALID, "null input");
             b.update_scene_bodies(scene, k, idx, body_pos, body_rot));
}
int cdx_batch_set_envelope(cdx_batch* batch, double envelope) { BATCH_CALL(b.set_envelope(envelope)); }
int cdx_batch_set_narrowphase_mode(cdx_batch* batch, int np_mode, double tau) {
  BATCH_CALL(CDX_REQUIRE(tau >= 0.0, CDX_ERR_INVALID, "tau must be non-negative"); b.set_np_mode(np_mode, tau));
}
int cdx_batch_select(cdx_batch* batch, const uint32_t* scenes, uint64_t k) {
  BATCH_CALL(CDX_REQUIRE(scenes || k == 0, CDX_ERR_INVALID, "null scene list"); b.select(scenes, k));
}
int cdx_batch_step(cdx_batch* batch, int use_graph) { BATCH_CALL(b.step(use_graph != 0)); }
int cdx_batch_finish(cdx_batch* batch, cdx_counts* totals) {
  BATCH_CALL(b.finish(); if (totals) fill_counts(b.totals(), totals));
}
int cdx_batch_scene_counts(cdx_batch* batch, uint64_t j, cdx_counts* out) {
  BATCH_CALL(CDX_REQUIRE(out, CDX_ERR_INVALID, "null output"); cdx::StepCounts c; b.scene_counts(j, &c);
             fill_counts(c, out));
}
int cdx_batch_read_pairs(cdx_batch* batch, uint64_t j, int64_t* pairs_out, uint64_t capacity) {
  BATCH_CALL(cdx::StepCounts c; b.scene_counts(j, &c);
             CDX_REQUIRE(capacity >= c.p, CDX_ERR_CAPACITY, "pair buffer too small: need " + std::to_string(c.p));
             CDX_REQUIRE(c.p == 0 || pairs_out, CDX_ERR_INVALID, "null output"); b.read_scene_pairs(j, pairs_out));
}
int cdx_batch_read_contacts(cdx_batch* batch, uint64_t j, void* norm, void* cpta, void* cptb, void* dpth,
                            int32_t* bids, uint64_t capacity) {
  BATCH_CALL(cdx::StepCounts c; b.scene_counts(j, &c);
             CDX_REQUIRE(capacity >= c.c, CDX_ERR_CAPACITY, "contact buffers too small: need " + std::to_string(c.c));
             b.read_scene_contacts(j, norm, cpta, cptb, dpth, bids));
}
int cdx_batch_step_ms(cdx_batch* batch, double* ms) {
  BATCH_CALL(CDX_REQUIRE(ms, CDX_ERR_INVALID, "null output"); *ms = b.last_step_ms());
}
int cdx_batch_device_bytes(cdx_batch* batch, uint64_t* bytes) {
  BATCH_CALL(CDX_REQUIRE(bytes, CDX_ERR_INVALID, "null output"); *bytes = b.device_bytes());
}
int cdx_batch_scene_bytes(cdx_batch* batch, uint64_t n, uint64_t nb, uint64_t* bytes) {
  BATCH_CALL(CDX_REQUIRE(bytes, CDX_ERR_INVALID, "null output"); *bytes = b.scene_bytes_estimate(n, nb));
}

// ---- served loop ----------------------------------------------------------

int cdx_server_create(cdx_ctx* ctx, const cdx_server_config* cfg, cdx_server** out) {
  if (!ctx) return CDX_ERR_INVALID;
  return guarded(&ctx->err, [&] {
    CDX_REQUIRE(cfg && out, CDX_ERR_INVALID, "null argument");
    check_kind(cfg->real_kind);
    CDX_REQUIRE(cfg->np_mode == CDX_NP_NATIVE || cfg->np_mode == CDX_NP_MIXED || cfg->np_mode == CDX_NP_FAST,
                CDX_ERR_INVALID, "unknown narrowphase mode");
    CDX_REQUIRE(cfg->tau >= 0.0, CDX_ERR_INVALID, "tau must be non-negative");
    cdx::ServerConfig c;
    c.real_kind = cfg->real_kind;
    c.np_mode = cfg->np_mode;
    c.tau = cfg->tau;
    c.envelope = cfg->envelope;
    c.hbm_budget = cfg->hbm_budget;
    if (cfg->max_batch_shapes) c.max_batch_shapes = cfg->max_batch_shapes;
    if (cfg->max_pending) c.max_pending = cfg->max_pending;
    c.use_graph = cfg->use_graph != 0;
    auto sv = std::make_unique<cdx_server>();
    sv->srv = std::make_unique<cdx::Server>(ctx->device, c);
    *out = sv.release();
  });
}

int cdx_server_destroy(cdx_server* server) {
  delete server;
  return CDX_OK;
}

const char* cdx_server_last_error(const cdx_server* server) {
  return server ? t_server_err.c_str() : "null server";
}

#define SERVER_CALL(...)                                                 \
  if (!server) return CDX_ERR_INVALID;                                   \
  return guarded(&t_server_err, [&] { cdx::Server& sv = *server->srv; __VA_ARGS__; }, false)

int cdx_server_add_scene(cdx_server* server, uint64_t n, const int32_t* T, const void* A, const void* B,
                         const void* C, const void* R, const uint32_t* ID, uint64_t nb, const void* body_pos,
                         const void* body_rot, const uint8_t* obj_active, int bins_x, int bins_y, int bins_z,
                         uint32_t* scene_out) {
  SERVER_CALL(CDX_REQUIRE(T && A && B && R && ID && body_pos && body_rot, CDX_ERR_INVALID, "null input");
              uint32_t s = sv.add_scene(n, T, A, B, C, R, ID, nb, body_pos, body_rot, obj_active, bins_x, bins_y,
                                        bins_z);
              if (scene_out) *scene_out = s);
}
int cdx_server_scene_info(cdx_server* server, uint32_t scene, uint64_t* n_shapes, uint64_t* n_bodies,
                          int* size_class, int* bins_xyz) {
  SERVER_CALL(sv.scene_info(scene, n_shapes, n_bodies, size_class, bins_xyz));
}
int cdx_server_submit(cdx_server* server, uint32_t scene, uint64_t k, const uint32_t* idx, const void* body_pos,
                      const void* body_rot, uint64_t* request_out) {
  SERVER_CALL(CDX_REQUIRE(request_out, CDX_ERR_INVALID, "null output");
              *request_out = sv.submit(scene, k, idx, body_pos, body_rot));
}
int cdx_server_pose_buffer(cdx_server* server, uint32_t scene, uint64_t k, void** body_pos, void** body_rot,
                           uint64_t* handle) {
  SERVER_CALL(CDX_REQUIRE(handle, CDX_ERR_INVALID, "null output");
              *handle = sv.pose_buffer(scene, k, body_pos, body_rot));
}
int cdx_server_submit_buffer(cdx_server* server, uint64_t handle, uint64_t* request_out) {
  SERVER_CALL(CDX_REQUIRE(request_out, CDX_ERR_INVALID, "null output"); *request_out = sv.submit_buffer(handle));
}
int cdx_server_drop_buffer(cdx_server* server, uint64_t handle) { SERVER_CALL(sv.drop_buffer(handle)); }
int cdx_server_poll(cdx_server* server, uint64_t request, int* done) {
  SERVER_CALL(CDX_REQUIRE(done, CDX_ERR_INVALID, "null output"); *done = sv.poll(request));
}
int cdx_server_wait(cdx_server* server, uint64_t request, double timeout_ms, int* done) {
  SERVER_CALL(int d = sv.wait(request, timeout_ms); if (done) *done = d);
}
int cdx_server_result(cdx_server* server, uint64_t request, uint64_t* n_contacts, double* latency_ms) {
  SERVER_CALL(sv.result_counts(request, n_contacts, latency_ms));
}
int cdx_server_read_contacts(cdx_server* server, uint64_t request, void* norm, void* cpta, void* cptb, void* dpth,
                             int32_t* bids, uint64_t capacity) {
  SERVER_CALL(sv.read_contacts(request, norm, cpta, cptb, dpth, bids, capacity));
}
int cdx_server_release(cdx_server* server, uint64_t request) { SERVER_CALL(sv.release(request)); }
int cdx_server_stats(cdx_server* server, cdx_server_stats_t* out, int reset) {
  SERVER_CALL(CDX_REQUIRE(out, CDX_ERR_INVALID, "null output"); cdx::ServerStats s = sv.stats();
              out->submitted = s.submitted; out->completed = s.completed; out->failed = s.failed;
              out->steps = s.batches; out->stepped_requests = s.batched_requests; out->p50_ms = s.p50_ms;
              out->p99_ms = s.p99_ms; out->mean_ms = s.mean_ms; out->max_ms = s.max_ms; out->busy_ms = s.busy_ms;
              out->window_s = s.window_s; for (int q = 0; q < 6; ++q) out->host_ms[q] = s.host_ms[q];
              for (int c = 0; c < 4; ++c) {
                out->class_completed[c] = s.class_completed[c]; out->class_steps[c] = s.class_steps[c];
                out->class_p50_ms[c] = s.class_p50_ms[c]; out->class_p99_ms[c] = s.class_p99_ms[c];
                out->class_busy_ms[c] = s.class_busy_ms[c];
              }
              if (reset) sv.reset_stats());
}
int cdx_server_memory(cdx_server* server, uint64_t* hbm_budget, uint64_t* hbm_committed, uint64_t* device_bytes) {
  SERVER_CALL(if (hbm_budget) *hbm_budget = sv.hbm_budget(); if (hbm_committed) *hbm_committed = sv.hbm_committed();
              if (device_bytes) *device_bytes = sv.device_bytes());
}

}  // extern "C"
