/* cdx — MI355X (gfx950) collision detection: the C ABI of libcdx.so.
 *
 * Drop-in boundary for the hot path of uwsbel/collision-detection
 * (Chrono::Collide): shape/pose upload -> AABB generation -> uniform-grid
 * broadphase -> narrowphase (sphere-sphere + MPR) -> contact readback.
 * Plain pointers and sizes only; no C++ or HIP types cross this boundary.
 *
 * Layouts follow the reference API's containers:
 *   real3  = 3 packed reals (x,y,z); real4 = quaternion (w,x,y,z);
 *   shape types are the reference enum (src/collision_detection.h:57)
 *   SPHERE=0 ELLIPSOID=1 BOX=2 CYLINDER=3 RECT=4 CONE=5 TRIANGLEMESH=6 plus
 *   CDX_CAPSULE=7 (no reference counterpart); other values are rejected with
 *   CDX_ERR_UNSUPPORTED_SHAPE.  A triangle's world vertices are
 *   R (A + p), R (B + p), R (C + p) with p the body position and R the body
 *   rotation times the local rotation (Narrowphase.cu:363-367); RECT and
 *   CONE are centred at their local position A, rotated by R, without the
 *   body position (Narrowphase.cu:359-377 applies none to them).  Their
 *   AABBs, which the reference leaves unwritten (RECT, TRIANGLEMESH) or
 *   computes by the triangle formula (CONE, AABBGenerator.cu:103-112), are
 *   the bounds of the shape as the narrowphase places it;
 *   aabb_data = 2n real3, the n minimum corners followed by the n maximum
 *   corners (AABBGenerator.cu:114-115);
 *   pairs = int64 (a << 32 | b), a < b shape indices (Broadphase.cu:302-308);
 *   obj_active = one byte per body (custom_vector<bool>).
 * real is double (CDX_REAL_F64) or float (CDX_REAL_F32), chosen per call or
 * per scene, the reference's compile-time `real` typedef made a parameter.
 *
 * Every function returns a status (CDX_OK = 0).  The message of the last
 * failure is kept per context / per scene.  A context or scene must not be
 * used from two threads at once; distinct contexts and scenes are
 * independent (each scene owns its own HIP stream and device buffers).
 */
#ifndef CDX_H
#define CDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDX_VERSION 1

#define CDX_OK 0
#define CDX_ERR_INVALID 1
#define CDX_ERR_HIP 2
#define CDX_ERR_UNSUPPORTED_SHAPE 3
#define CDX_ERR_CAPACITY 4
#define CDX_ERR_NO_DEVICE 5
#define CDX_ERR_OUT_OF_MEMORY 6
#define CDX_ERR_NOT_READY 7

#define CDX_REAL_F64 0
#define CDX_REAL_F32 1

#define CDX_CAPSULE 7

/* Narrowphase precision.
 * NATIVE: the MPR in the scene's real type with the reference's evaluation
 *   order (Narrowphase.cu:122-423).
 * MIXED: an FP32 MPR in the frame of the first shape decides clear misses;
 *   every hit, and every pair whose FP32 verdict lies within
 *   tau x (sum of the two bounding radii) of a decision boundary (near-zero
 *   separation, degenerate portal, coincident centres, iteration cap), is
 *   re-solved in FP64 on the reference path, so contacts equal NATIVE FP64.
 * FAST: like MIXED, but FP32 contacts are kept for clear hits (depth beyond
 *   tau); only uncertain pairs are re-solved in FP64.  FP32 resolves the
 *   MPR portal only to about sqrt(FLT_EPSILON) on curved surfaces, so
 *   normals differ from FP64 by ~1e-4 (median) to ~1e-3 (p99). */
#define CDX_NP_NATIVE 0
#define CDX_NP_MIXED 1
#define CDX_NP_FAST 2

/* Stages reported by cdx_scene_stage_ms (HIP-event times of the last run). */
#define CDX_STAGE_AABB 0
#define CDX_STAGE_BP_BOUNDS 1
#define CDX_STAGE_BP_COUNT 2
#define CDX_STAGE_BP_SCAN 3
#define CDX_STAGE_BP_EMIT 4
#define CDX_STAGE_BP_SORT 5
#define CDX_STAGE_BP_RUNS 6 /* 0: the active bins are summed inside CDX_STAGE_BP_PAIRSCAN */
#define CDX_STAGE_BP_PAIRS 7 /* pair counting pass; the scan and the emission pass are 12 and 13 */
#define CDX_STAGE_BP_PAIRSORT 8
#define CDX_STAGE_NP_MAIN 9
#define CDX_STAGE_NP_RESOLVE 10
#define CDX_STAGE_NP_COMPACT 11
#define CDX_STAGE_BP_PAIRSCAN 12
#define CDX_STAGE_BP_PAIREMIT 13
#define CDX_STAGE_NP_MPR 14 /* the reference-order MPR launch alone (part of NP_RESOLVE) */
#define CDX_STAGE_COUNT 15

/* cdx_counts.flags */
#define CDX_FLAG_UNSUPPORTED_SHAPE 0x1u
#define CDX_FLAG_BAD_BODY_ID 0x2u
#define CDX_FLAG_CELL_SEARCH 0x100u    /* the broadphase found its pairs by the cell search */
#define CDX_FLAG_NO_ACTIVE_BINS 0x200u /* bin_entries, active_bins and key_max not computed (0):
                                          the AABBs or the bins per axis changed after a
                                          cell-search broadphase before its counts were read */

/* Pair search of a scene's broadphase (cdx_scene_set_pair_search). */
#define CDX_PAIRS_AUTO 0  /* cell search for near-uniform shape sizes, else bins (default) */
#define CDX_PAIRS_BINS 1  /* the reference's bin decomposition always */
#define CDX_PAIRS_CELLS 2 /* cell search wherever its grid fits */

typedef struct cdx_ctx cdx_ctx;
typedef struct cdx_scene cdx_scene;

typedef struct cdx_counts {
  uint64_t n_shapes;
  uint64_t bin_entries;   /* number_of_bin_intersections (Broadphase.cu:404) */
  uint64_t active_bins;   /* last_active_bin (Broadphase.cu:440) */
  uint64_t pairs;         /* number_of_contacts_possible after unique (:528) */
  uint64_t contacts;      /* number_of_contacts (Narrowphase.cu:580) */
  uint64_t resolved;      /* pairs re-solved in FP64 (CDX_NP_MIXED) */
  uint32_t key_max;       /* largest bin key emitted */
  uint32_t flags;         /* CDX_FLAG_* bits */
  float bmin[3], bmax[3]; /* scene bounds, float (Broadphase.cu:368-376) */
} cdx_counts;

int cdx_version(void);
const char* cdx_status_string(int status);
int cdx_device_count(int* count);

/* ---- contexts --------------------------------------------------------- */
int cdx_ctx_create(int device, cdx_ctx** out);
int cdx_ctx_destroy(cdx_ctx* ctx);
const char* cdx_ctx_last_error(const cdx_ctx* ctx);

/* ---- drop-in calls: one per reference method, host buffers ------------ */

/* Replaces AABBGenerator::GenerateAABB (AABBGenerator.cuh:21-30,
 * AABBGenerator.cu:135-188).  aabb_out receives 2*n_shapes real3.
 * C (obj_data_C) holds the triangles' third vertices (NULL: zeros).  Types
 * outside the enum fail with CDX_ERR_UNSUPPORTED_SHAPE (the reference
 * leaves their AABB slots unwritten). */
int cdx_generate_aabb(cdx_ctx* ctx, int real_kind, uint64_t n_shapes, const int32_t* T, const void* A,
                      const void* B, const void* C, const void* R, const uint32_t* ID, uint64_t n_bodies,
                      const void* body_pos, const void* body_rot, void* aabb_out);

/* Replaces Broadphase::detectPossibleCollisions (Broadphase.cuh:17,
 * Broadphase.cu:344-545) with setBinsPerAxis's grid (Broadphase.cu:57-60).
 * aabb_inout (2*n_aabb real3) is shifted in place by the scene origin, as
 * the reference does (Broadphase.cu:382-386).  *n_pairs receives the
 * number of unique candidate pairs; fetch them with
 * cdx_read_possible_collisions.  An empty input yields 0 pairs (the
 * reference reads aabb_data[0] there). */
int cdx_detect_possible_collisions(cdx_ctx* ctx, int real_kind, void* aabb_inout, uint64_t n_aabb, int bins_x,
                                   int bins_y, int bins_z, uint64_t* n_pairs);
int cdx_read_possible_collisions(cdx_ctx* ctx, int64_t* pairs_out, uint64_t capacity);

/* Replaces Narrowphase::DoNarrowphase (Narrowphase.cuh:19-34,
 * Narrowphase.cu:511-595) and SetCollisionEnvelope (Narrowphase.cuh:53-55).
 * pairs_inout is permuted in place, contacts first in pair order, as the
 * reference's sort_by_key leaves it (Narrowphase.cu:584-588).
 * *n_contacts receives number_of_contacts; fetch the contact arrays with
 * cdx_read_contacts. */
int cdx_do_narrowphase(cdx_ctx* ctx, int real_kind, int np_mode, double envelope, uint64_t n_shapes,
                       const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                       const uint32_t* ID, uint64_t n_bodies, const uint8_t* obj_active, const void* body_pos,
                       const void* body_rot, int64_t* pairs_inout, uint64_t n_pairs, uint64_t* n_contacts);
/* norm, cpta, cptb: real3[n]; dpth: real[n]; bids: int2[n] (body ids).
 * Any output pointer may be NULL. */
int cdx_read_contacts(cdx_ctx* ctx, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids,
                      uint64_t capacity);

/* Replaces Narrowphase::host_MPR_Store (Narrowphase.cuh:36-52, body
 * Narrowphase.cu:464-504 over function_MPR_Store :316-423): the per-pair
 * store without compaction.  For each pair i that makes a contact, norm[i],
 * ptA[i], ptB[i] (real3), depth[i] (real) and ids[2i..2i+1] (body ids) are
 * written and contact_active[i] is set to 0; for every other pair nothing
 * is written, so the caller's values stay (the reference leaves them, and
 * DoNarrowphase pre-fills contact_active with 1).  pairs is not modified.
 * Any output pointer may be NULL. */
int cdx_mpr_store(cdx_ctx* ctx, int real_kind, int np_mode, double envelope, uint64_t n_shapes, const int32_t* T,
                  const void* A, const void* B, const void* C, const void* R, const uint32_t* ID, uint64_t n_bodies,
                  const uint8_t* obj_active, const void* body_pos, const void* body_rot, const int64_t* pairs,
                  uint64_t n_pairs, uint32_t* contact_active, void* norm, void* ptA, void* ptB, void* depth,
                  int32_t* ids);

/* No counterpart in the reference: detectPossibleCollisions for callers
 * that cannot allocate (code under HIP graph capture set-up, tight serving
 * loops, harnesses that forbid allocation).  The same kernels as
 * cdx_detect_possible_collisions, on DEVICE buffers: aabb_inout (2*n_aabb
 * real3 in device memory) is shifted in place, the sorted unique pairs go to
 * pairs_out (device memory, pair_capacity int64), their count to *n_pairs
 * (host).  Every intermediate lives in one fixed device array of the library
 * (48 MiB, taken once per device and held until the process ends); no
 * device or pinned memory is allocated.  Limits: n_aabb <= 65536, and the
 * bin entries must fit what the array holds at that size (about 1.2M for
 * FP64 at the limit): beyond either, or with more pairs than pair_capacity,
 * the call fails with CDX_ERR_CAPACITY.  Synchronous on `stream` (a
 * hipStream_t; NULL = the default stream).  Calls on one device are
 * serialised (they share its array). */
int cdx_detect_possible_collisions_noalloc(cdx_ctx* ctx, int real_kind, void* aabb_inout, uint64_t n_aabb,
                                           int bins_x, int bins_y, int bins_z, int64_t* pairs_out,
                                           uint64_t pair_capacity, uint64_t* n_pairs, void* stream);

/* ---- device-resident scenes (no reference counterpart: the reference
 * re-uploads through thrust containers on every call) ------------------- */
int cdx_scene_create(cdx_ctx* ctx, int real_kind, cdx_scene** out);
int cdx_scene_destroy(cdx_scene* scene);
const char* cdx_scene_last_error(const cdx_scene* scene);

int cdx_scene_set_shapes(cdx_scene* scene, uint64_t n_shapes, const int32_t* T, const void* A, const void* B,
                         const void* C, const void* R, const uint32_t* ID);
int cdx_scene_set_bodies(cdx_scene* scene, uint64_t n_bodies, const void* body_pos, const void* body_rot,
                         const uint8_t* obj_active);
/* Incremental pose update of k bodies (indices + new pos/rot). */
int cdx_scene_update_bodies(cdx_scene* scene, uint64_t k, const uint32_t* body_idx, const void* body_pos,
                            const void* body_rot);
/* Caller-supplied AABBs for a broadphase-only scene. */
int cdx_scene_set_aabbs(cdx_scene* scene, uint64_t n_aabb, const void* aabb_data);
/* Caller-supplied candidate pairs for a narrowphase-only scene. */
int cdx_scene_set_pairs(cdx_scene* scene, uint64_t n_pairs, const int64_t* pairs);

int cdx_scene_set_bins(cdx_scene* scene, int bins_x, int bins_y, int bins_z);
int cdx_scene_set_envelope(cdx_scene* scene, double envelope);
int cdx_scene_set_narrowphase_mode(cdx_scene* scene, int np_mode, double tau);
int cdx_scene_set_profiling(cdx_scene* scene, int on);
/* How the broadphase finds candidate pairs; the pair list is the same
 * either way (every pair of overlapping shifted AABBs, ascending, as
 * Broadphase.cu:484-535 leaves it).  The bins (Broadphase.cu:393-480) walk
 * every bin each AABB spans.  The cell search keys each shape once, by the
 * cell of its AABB's minimum corner in an internal grid of cells at least as
 * wide as the widest AABB, and tests the cells around it.  After it, the
 * reference grid's bin counts, entry total, sorted entries, active bins and
 * occupancy are computed when cdx_scene_counts or cdx_scene_read_bins asks
 * for them, as long as the AABBs and the bins per axis have not changed. */
int cdx_scene_set_pair_search(cdx_scene* scene, int mode);

int cdx_scene_aabb(cdx_scene* scene);
int cdx_scene_broadphase(cdx_scene* scene);
int cdx_scene_narrowphase(cdx_scene* scene);
/* AABB + broadphase + narrowphase. */
int cdx_scene_step(cdx_scene* scene);
int cdx_scene_sync(cdx_scene* scene);
int cdx_scene_counts(cdx_scene* scene, cdx_counts* out);

/* Readback.  shifted = 1 returns the AABBs as the reference leaves them
 * after detectPossibleCollisions (plus the scene origin). */
int cdx_scene_read_aabb(cdx_scene* scene, int shifted, void* aabb_out);
/* Broadphase intermediates the reference keeps internal: per-AABB bin
 * counts (Bins_Intersected before its scan, Broadphase.cu:393-402), the
 * sorted (bin, body) entries (after sort_by_key, :428-432) and the active
 * bins with their occupancy (reduce_by_key, :439-443).  NULL skips one. */
int cdx_scene_read_bins(cdx_scene* scene, uint32_t* bin_counts, uint32_t* sorted_keys, uint32_t* sorted_bodies,
                        uint32_t* active_keys, uint32_t* active_counts);
int cdx_scene_read_pairs(cdx_scene* scene, int64_t* pairs_out);
int cdx_scene_read_pairs_permuted(cdx_scene* scene, int64_t* pairs_out);
/* The pairs (indices into the cdx_scene_read_pairs order) that the
 * reference-order MPR launch of the last narrowphase ran on: every queued
 * pair in NATIVE, the FP64 re-solves in MIXED / FAST.  *count receives how
 * many; at most capacity indices are written (pair_index may be NULL). */
int cdx_scene_read_mpr_queue(cdx_scene* scene, uint32_t* pair_index, uint64_t capacity, uint64_t* count);
int cdx_scene_read_contacts(cdx_scene* scene, void* norm, void* cpta, void* cptb, void* dpth, int32_t* bids);
/* Milliseconds per stage of the last run (profiling on), CDX_STAGE_COUNT values. */
int cdx_scene_stage_ms(cdx_scene* scene, double* ms_out);
int cdx_scene_device_bytes(cdx_scene* scene, uint64_t* bytes);

/* ---- batches: many resident scenes stepped by one launch sequence -------
 * No reference counterpart: the reference steps one scene per
 * detectPossibleCollisions / DoNarrowphase call (Broadphase.cu:364-455,
 * Narrowphase.cu:426-489).  A batch holds any number of scenes resident in
 * HBM, each with its own grid, bounds and key range; cdx_batch_step runs
 * AABB + broadphase + narrowphase over a selection of them as one sequence
 * of launches (recorded once as a HIP graph and replayed while the
 * selection, settings and capacities are unchanged).  Step buffers are
 * sized by capacities that only grow; a step whose totals overflow them is
 * re-run at the larger size inside cdx_batch_finish.  Each scene's pairs
 * and contacts equal what a one-scene call on the same scene returns
 * (shape and body ids local to the scene). */
typedef struct cdx_batch cdx_batch;

int cdx_batch_create(cdx_ctx* ctx, int real_kind, cdx_batch** out);
int cdx_batch_destroy(cdx_batch* batch);
const char* cdx_batch_last_error(const cdx_batch* batch);
/* Appends a scene (arrays as for cdx_do_narrowphase; obj_active may be NULL
 * for all active).  bins_x = bins_y = bins_z = 0 lets the autotuner pick the
 * grid: the first scene of each size class (octaves of 4 shapes) is timed
 * at several cell sizes, later scenes of the class reuse the fastest. */
int cdx_batch_add_scene(cdx_batch* batch, uint64_t n, const int32_t* T, const void* A, const void* B, const void* C,
                        const void* R, const uint32_t* ID, uint64_t nb, const void* body_pos, const void* body_rot,
                        const uint8_t* obj_active, int bins_x, int bins_y, int bins_z, uint32_t* scene_out);
int cdx_batch_scene_bins(cdx_batch* batch, uint32_t scene, int* bins_xyz);
int cdx_batch_scene_count(cdx_batch* batch, uint64_t* count);
/* Replaces the poses of bodies idx[0..k) of one scene (local body ids). */
int cdx_batch_update_bodies(cdx_batch* batch, uint32_t scene, uint64_t k, const uint32_t* idx, const void* body_pos,
                            const void* body_rot);
int cdx_batch_set_envelope(cdx_batch* batch, double envelope);
int cdx_batch_set_narrowphase_mode(cdx_batch* batch, int np_mode, double tau);
/* Scenes of the following steps (scenes = NULL: every resident scene). */
int cdx_batch_select(cdx_batch* batch, const uint32_t* scenes, uint64_t k);
/* Enqueues one step on the batch's stream (use_graph: record / replay). */
int cdx_batch_step(cdx_batch* batch, int use_graph);
/* Waits for the step, re-runs it if a capacity overflowed, reads totals. */
int cdx_batch_finish(cdx_batch* batch, cdx_counts* totals);
/* Per selected scene j (its position in the selection), after finish. */
int cdx_batch_scene_counts(cdx_batch* batch, uint64_t j, cdx_counts* out);
int cdx_batch_read_pairs(cdx_batch* batch, uint64_t j, int64_t* pairs_out, uint64_t capacity);
int cdx_batch_read_contacts(cdx_batch* batch, uint64_t j, void* norm, void* cpta, void* cptb, void* dpth,
                            int32_t* bids, uint64_t capacity);
/* Device milliseconds of the last step (HIP events around its launches). */
int cdx_batch_step_ms(cdx_batch* batch, double* ms);
int cdx_batch_device_bytes(cdx_batch* batch, uint64_t* bytes);
/* Estimated HBM bytes a scene of n shapes and nb bodies adds when resident. */
int cdx_batch_scene_bytes(cdx_batch* batch, uint64_t n, uint64_t nb, uint64_t* bytes);

/* ---- served request loop ------------------------------------------------
 * No reference counterpart.  A server keeps scenes resident in HBM (one
 * batch per size class: < 8192, < 49152, < 131072 and >= 131072 shapes),
 * admits a new
 * scene only if its estimated footprint fits the HBM budget, and runs one
 * worker thread per size class that keeps stepping: each step takes the
 * oldest queued requests of its class (at most one per scene, up to
 * max_batch_shapes shapes), so requests that arrive while a step runs join
 * the next one.  Steps of different classes run on their own streams and
 * overlap on the device.  A request carries pose updates for one scene
 * (the body_pos / body_rot inputs of AABBGenerator::GenerateAABB,
 * AABBGenerator.cuh:21-30) and yields that scene's contacts (the outputs of
 * Narrowphase::DoNarrowphase, Narrowphase.cuh:28-34), equal to what the
 * scene gives stepped alone.  Requests for one scene run in submit order.
 * Thread safety: submit / poll / wait / result / read / release / stats may
 * be called from any threads; add_scene from one thread at a time.  The
 * last-error text is kept per calling thread. */
typedef struct cdx_server cdx_server;

typedef struct cdx_server_config {
  int real_kind;             /* CDX_REAL_F64 or CDX_REAL_F32 */
  int np_mode;               /* CDX_NP_* */
  double tau;                /* narrowphase re-solve band, as cdx_scene_set_narrowphase_mode */
  double envelope;           /* Narrowphase::SetCollisionEnvelope */
  uint64_t hbm_budget;       /* bytes for resident scenes + step buffers; 0 = 90 % of free HBM */
  uint64_t max_batch_shapes; /* shapes per step and class (bounds step latency); 0 = 524288 */
  uint64_t max_pending;      /* queued requests before submit returns CDX_ERR_CAPACITY; 0 = 65536 */
  int use_graph;             /* replay recorded HIP graphs */
} cdx_server_config;

typedef struct cdx_server_stats_t {
  uint64_t submitted, completed, failed, steps, stepped_requests;
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> result ready */
  double busy_ms;                         /* summed device time of the steps */
  double window_s;                        /* wall time of the statistics window */
  double host_ms[6]; /* workers' host time per phase: pose upload, selection, launch, wait, result block, readback */
  /* per size class (< 8192, < 49152, < 131072, >= 131072 shapes) */
  uint64_t class_completed[4], class_steps[4];
  double class_p50_ms[4], class_p99_ms[4]; /* submit -> result ready */
  double class_busy_ms[4];                 /* summed device time of the class's steps */
} cdx_server_stats_t;

int cdx_server_create(cdx_ctx* ctx, const cdx_server_config* cfg, cdx_server** out);
/* Stops the workers (queued requests fail) and frees the scenes. */
int cdx_server_destroy(cdx_server* server);
const char* cdx_server_last_error(const cdx_server* server);
/* bins 0,0,0 autotunes, as cdx_batch_add_scene.  CDX_ERR_CAPACITY when the
 * scene does not fit the HBM budget. */
int cdx_server_add_scene(cdx_server* server, uint64_t n, const int32_t* T, const void* A, const void* B,
                         const void* C, const void* R, const uint32_t* ID, uint64_t nb, const void* body_pos,
                         const void* body_rot, const uint8_t* obj_active, int bins_x, int bins_y, int bins_z,
                         uint32_t* scene_out);
int cdx_server_scene_info(cdx_server* server, uint32_t scene, uint64_t* n_shapes, uint64_t* n_bodies,
                          int* size_class, int* bins_xyz);
/* Queues a step of one scene after replacing the poses of bodies idx[0..k)
 * (idx NULL: bodies 0..k-1).  Inputs are copied before return.
 * CDX_ERR_CAPACITY when max_pending requests are queued. */
int cdx_server_submit(cdx_server* server, uint32_t scene, uint64_t k, const uint32_t* idx, const void* body_pos,
                      const void* body_rot, uint64_t* request_out);
/* Zero-copy submission: a pinned block for the poses of bodies 0..k-1 of
 * one scene (k real3 positions at *body_pos, k real4 rotations at
 * *body_rot) that the caller fills in place; cdx_server_submit_buffer
 * queues it (the server takes the block over), cdx_server_drop_buffer
 * discards it unsubmitted. */
int cdx_server_pose_buffer(cdx_server* server, uint32_t scene, uint64_t k, void** body_pos, void** body_rot,
                           uint64_t* handle);
int cdx_server_submit_buffer(cdx_server* server, uint64_t handle, uint64_t* request_out);
int cdx_server_drop_buffer(cdx_server* server, uint64_t handle);
int cdx_server_poll(cdx_server* server, uint64_t request, int* done);
/* timeout_ms < 0 waits without limit; *done = 0 on timeout. */
int cdx_server_wait(cdx_server* server, uint64_t request, double timeout_ms, int* done);
int cdx_server_result(cdx_server* server, uint64_t request, uint64_t* n_contacts, double* latency_ms);
int cdx_server_read_contacts(cdx_server* server, uint64_t request, void* norm, void* cpta, void* cptb, void* dpth,
                             int32_t* bids, uint64_t capacity);
/* Frees a finished request and its share of the pinned result block. */
int cdx_server_release(cdx_server* server, uint64_t request);
/* reset != 0 starts a new statistics window after reading. */
int cdx_server_stats(cdx_server* server, cdx_server_stats_t* out, int reset);
int cdx_server_memory(cdx_server* server, uint64_t* hbm_budget, uint64_t* hbm_committed, uint64_t* device_bytes);

#ifdef __cplusplus
}
#endif

#endif /* CDX_H */
