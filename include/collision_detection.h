// collision_detection.h — C++ drop-in for the Chrono::Collide API
// (uwsbel/collision-detection), backed by libcdx.so on an MI355X.
//
// Same class names, method names, argument meaning and side effects as:
//   AABBGenerator   src/aabb/AABBGenerator.cuh:13-49
//   Broadphase      src/broadphase/Broadphase.cuh:14-82
//   Narrowphase     src/narrowphase/Narrowphase.cuh:13-63
//   types / enum    src/collision_detection.h:57-69 and gpu_math (real, real3, ...)
// The classes are header-only wrappers over the C ABI in cdx.h: no thrust,
// CUDA or HIP type appears here.  Containers are host vectors
// (custom_vector<T> = std::vector<T>, with custom_vector<bool> stored one
// byte per element as thrust::host_vector<bool> does).
//
// `real` is double unless CDX_REAL_FLOAT is defined, mirroring the
// reference's compile-time choice in gpu_math.
//
// Error behaviour: the reference never reports errors (every int return is
// 0).  Here a failure of the GPU path throws cdx_error (derived from
// std::runtime_error) carrying the library's message; methods that return
// int still return 0 on success.
#ifndef CDX_COLLISION_DETECTION_H
#define CDX_COLLISION_DETECTION_H

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cdx.h"

#ifdef CDX_REAL_FLOAT
typedef float real;
#define CDX_REAL_KIND CDX_REAL_F32
#else
typedef double real;
#define CDX_REAL_KIND CDX_REAL_F64
#endif

typedef unsigned int uint;

struct real3 { real x, y, z; };
struct real4 { real w, x, y, z; };  // quaternion, scalar first
struct int2 { int x, y; };
struct int3 { int x, y, z; };
static_assert(sizeof(real3) == 3 * sizeof(real), "real3 must be three packed reals");
static_assert(sizeof(real4) == 4 * sizeof(real), "real4 must be four packed reals");
static_assert(sizeof(int2) == 2 * sizeof(int), "int2 must be two packed ints");

inline real3 make_real3(real x, real y, real z) { return real3{x, y, z}; }
inline real3 R3(real x, real y, real z) { return real3{x, y, z}; }
inline real3 R3(real v) { return real3{v, v, v}; }
inline real4 R4(real w, real x, real y, real z) { return real4{w, x, y, z}; }
inline int2 I2(int x, int y) { return int2{x, y}; }
inline int3 I3(int x, int y, int z) { return int3{x, y, z}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }

// src/collision_detection.h:57, plus CAPSULE (no reference counterpart).
enum shape_type { SPHERE, ELLIPSOID, BOX, CYLINDER, RECT, CONE, TRIANGLEMESH, CAPSULE };
static_assert(sizeof(shape_type) == sizeof(int32_t), "shape_type must be 32-bit like the reference enum");

namespace cdx_detail {
template <class T> struct vec { typedef std::vector<T> type; };
template <> struct vec<bool> { typedef std::vector<unsigned char> type; };
}  // namespace cdx_detail
template <class T> using custom_vector = typename cdx_detail::vec<T>::type;

class cdx_error : public std::runtime_error {
 public:
  cdx_error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

namespace cdx_detail {
// One library context per process, on device 0 (CDX_DEVICE overrides).
inline cdx_ctx* context() {
  static cdx_ctx* ctx = nullptr;
  static std::once_flag once;
  static int status = CDX_OK;
  std::call_once(once, [] {
    int dev = 0;
    if (const char* e = std::getenv("CDX_DEVICE")) dev = std::atoi(e);
    status = cdx_ctx_create(dev, &ctx);
  });
  if (status != CDX_OK) throw cdx_error(status, std::string("cdx: cannot create a context: ") + cdx_status_string(status));
  return ctx;
}
inline void check(int status) {
  if (status != CDX_OK) throw cdx_error(status, std::string("cdx: ") + cdx_ctx_last_error(context()));
}
}  // namespace cdx_detail

// AABBGenerator.cuh:13-49
class AABBGenerator {
 public:
  AABBGenerator() {}
  // AABBGenerator.cu:135-188: aabb_data is resized to 2N (N mins, N maxes).
  void GenerateAABB(const custom_vector<shape_type>& obj_data_T, const custom_vector<real3>& obj_data_A,
                    const custom_vector<real3>& obj_data_B, const custom_vector<real3>& obj_data_C,
                    const custom_vector<real4>& obj_data_R, const custom_vector<uint>& obj_data_ID,
                    const custom_vector<real3>& body_pos, const custom_vector<real4>& body_rot,
                    custom_vector<real3>& aabb_data) {
    numAABB = (uint)obj_data_T.size();
    aabb_data.resize(2 * (size_t)numAABB);
    if (numAABB == 0) return;
    cdx_detail::check(cdx_generate_aabb(cdx_detail::context(), CDX_REAL_KIND, numAABB,
                                        reinterpret_cast<const int32_t*>(obj_data_T.data()), obj_data_A.data(),
                                        obj_data_B.data(), obj_data_C.empty() ? nullptr : obj_data_C.data(),
                                        obj_data_R.data(), obj_data_ID.data(), body_pos.size(), body_pos.data(),
                                        body_rot.data(), aabb_data.data()));
  }

 private:
  uint numAABB = 0;
};

// Broadphase.cuh:14-82
class Broadphase {
 public:
  Broadphase() { setParallelConfiguration(1, 0, 1, 0, 0, 0, 0, 0, 1, 0); }

  // Broadphase.cu:344-545.  aabb_data is shifted in place by the scene
  // origin, as the reference does; potentialCollisions receives the sorted
  // unique (a << 32 | b) list.
  int detectPossibleCollisions(custom_vector<real3>& aabb_data, custom_vector<long long>& potentialCollisions) {
    potentialCollisions.clear();
    uint64_t np = 0;
    cdx_detail::check(cdx_detect_possible_collisions(cdx_detail::context(), CDX_REAL_KIND, aabb_data.data(),
                                                     aabb_data.size() / 2, grid_size.x, grid_size.y, grid_size.z,
                                                     &np));
    potentialCollisions.resize(np);
    cdx_detail::check(cdx_read_possible_collisions(cdx_detail::context(),
                                                   reinterpret_cast<int64_t*>(potentialCollisions.data()), np));
    number_of_contacts_possible = (uint)np;
    return 0;
  }
  int setBinsPerAxis(int3 binsPerAxis) {
    grid_size = binsPerAxis;
    return 0;
  }
  int3 getBinsPerAxis() { return grid_size; }
  // Kept for API compatibility: the reference stores these and never reads
  // them (its adaptive grid is commented out, Broadphase.cu:458-468).
  int setBodyPerBin(int max, int min) {
    max_body_per_bin = max;
    min_body_per_bin = min;
    return 0;
  }
  int getMaxBodyPerBin() { return max_body_per_bin; }
  int getMinBodyPerBin() { return min_body_per_bin; }
  uint getNumPossibleContacts() { return number_of_contacts_possible; }
  // Broadphase.cu:69-90 chooses OpenMP vs sequential thrust policies; the
  // GPU path has no such choice, so the flags are stored and reported only.
  void setParallelConfiguration(bool a, bool b, bool c, bool d, bool e, bool f, bool g, bool h, bool i, bool j) {
    bool v[10] = {a, b, c, d, e, f, g, h, i, j};
    for (int k = 0; k < 10; ++k) par_[k] = v[k];
  }
  void getParallelConfiguration() {
    for (int k = 0; k < 10; ++k) std::cout << par_[k] << " ";
  }

 private:
  int3 grid_size{16, 16, 16};
  int min_body_per_bin = 25, max_body_per_bin = 50;
  uint number_of_contacts_possible = 0;
  bool par_[10];
};

// Narrowphase.cuh:13-63
class Narrowphase {
 public:
  Narrowphase() {}
  // Narrowphase.cu:511-595.  Outputs are resized to the number of contacts,
  // in ascending pair order; potentialCollisions is permuted contacts first.
  void DoNarrowphase(const custom_vector<shape_type>& obj_data_T, const custom_vector<real3>& obj_data_A,
                     const custom_vector<real3>& obj_data_B, const custom_vector<real3>& obj_data_C,
                     const custom_vector<real4>& obj_data_R, const custom_vector<uint>& obj_data_ID,
                     const custom_vector<bool>& obj_active, const custom_vector<real3>& body_pos,
                     const custom_vector<real4>& body_rot, custom_vector<long long>& potentialCollisions,
                     custom_vector<real3>& norm_data, custom_vector<real3>& cpta_data, custom_vector<real3>& cptb_data,
                     custom_vector<real>& dpth_data, custom_vector<int2>& bids_data, uint& number_of_contacts) {
    uint64_t nc = 0;
    cdx_detail::check(cdx_do_narrowphase(
        cdx_detail::context(), CDX_REAL_KIND, mode_, (double)collision_envelope, obj_data_T.size(),
        reinterpret_cast<const int32_t*>(obj_data_T.data()), obj_data_A.data(), obj_data_B.data(),
        obj_data_C.empty() ? nullptr : obj_data_C.data(), obj_data_R.data(), obj_data_ID.data(), body_pos.size(),
        obj_active.data(), body_pos.data(), body_rot.data(), reinterpret_cast<int64_t*>(potentialCollisions.data()),
        potentialCollisions.size(), &nc));
    norm_data.resize(nc);
    cpta_data.resize(nc);
    cptb_data.resize(nc);
    dpth_data.resize(nc);
    bids_data.resize(nc);
    cdx_detail::check(cdx_read_contacts(cdx_detail::context(), norm_data.data(), cpta_data.data(), cptb_data.data(),
                                        dpth_data.data(), reinterpret_cast<int32_t*>(bids_data.data()), nc));
    number_of_contacts = (uint)nc;
    total_possible_contacts = (uint)potentialCollisions.size();
  }
  // Narrowphase.cuh:36-52, Narrowphase.cu:464-504: function_MPR_Store over
  // the first total_possible_contacts pairs (the count of the last
  // DoNarrowphase call, 0 before one, as the reference's member), without
  // compaction: a contact's outputs go to its pair index and clear
  // contact_active there; other entries are left as they are.  The shape
  // and body counts the GPU upload needs are the largest index the pairs
  // reach, as the reference reads nothing beyond them.
  void host_MPR_Store(const shape_type* obj_data_T, const real3* obj_data_A, const real3* obj_data_B,
                      const real3* obj_data_C, const real4* obj_data_R, const uint* obj_data_ID,
                      const bool* obj_active, const real3* body_pos, const real4* body_rot, long long* contact_pair,
                      uint* contact_active, real3* norm, real3* ptA, real3* ptB, real* contactDepth, int2* ids) {
    const uint64_t np = total_possible_contacts;
    if (np == 0) return;
    uint64_t ns = 0, nb = 0;
    for (uint64_t i = 0; i < np; ++i) {
      const uint64_t a = (uint64_t)contact_pair[i] >> 32, b = (uint64_t)contact_pair[i] & 0xffffffffu;
      ns = std::max<uint64_t>(ns, std::max(a, b) + 1);
    }
    for (uint64_t i = 0; i < ns; ++i) nb = std::max<uint64_t>(nb, (uint64_t)obj_data_ID[i] + 1);
    static_assert(sizeof(bool) == 1, "obj_active is read as one byte per body");
    cdx_detail::check(cdx_mpr_store(
        cdx_detail::context(), CDX_REAL_KIND, mode_, (double)collision_envelope, ns,
        reinterpret_cast<const int32_t*>(obj_data_T), obj_data_A, obj_data_B, obj_data_C, obj_data_R, obj_data_ID,
        nb, reinterpret_cast<const uint8_t*>(obj_active), body_pos, body_rot,
        reinterpret_cast<const int64_t*>(contact_pair), np, contact_active, norm, ptA, ptB, contactDepth,
        reinterpret_cast<int32_t*>(ids)));
  }
  // custom_vector<bool> holds bytes here: its data() comes as unsigned char.
  void host_MPR_Store(const shape_type* obj_data_T, const real3* obj_data_A, const real3* obj_data_B,
                      const real3* obj_data_C, const real4* obj_data_R, const uint* obj_data_ID,
                      const unsigned char* obj_active, const real3* body_pos, const real4* body_rot,
                      long long* contact_pair, uint* contact_active, real3* norm, real3* ptA, real3* ptB,
                      real* contactDepth, int2* ids) {
    host_MPR_Store(obj_data_T, obj_data_A, obj_data_B, obj_data_C, obj_data_R, obj_data_ID,
                   reinterpret_cast<const bool*>(obj_active), body_pos, body_rot, contact_pair, contact_active, norm,
                   ptA, ptB, contactDepth, ids);
  }
  void SetCollisionEnvelope(const real& envelope) { collision_envelope = envelope; }
  real GetCollisionEnvelope() { return collision_envelope; }
  // Precision of the MPR (CDX_NP_NATIVE, CDX_NP_MIXED or CDX_NP_FAST);
  // no reference counterpart.  NATIVE reproduces the reference's arithmetic.
  void SetPrecisionMode(int mode) { mode_ = mode; }

 private:
  real collision_envelope = 0;
  int mode_ = CDX_NP_NATIVE;
  uint total_possible_contacts = 0;
};

#endif  // CDX_COLLISION_DETECTION_H
