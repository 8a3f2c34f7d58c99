// Grid autotuner: picks the broadphase cell size (bins per axis) per scene
// size class, and the bin-entry and pair capacities per shape that batched
// steps are sized by.  The reference fixes bins per axis by hand
// (setBinsPerAxis, Broadphase.cu:57-60) and leaves its body-per-bin
// adaptivity commented out (Broadphase.cu:458-468).
#pragma once
#include <cstdint>
#include <map>
#include <mutex>

namespace cdx {

struct TuneChoice {
  int grid[3];
  double cell_factor;   // cell edge / mean AABB edge
  double m_per_shape;   // bin entries per shape measured at that grid
  double p_per_shape;   // candidate pairs per shape measured at that grid
  bool measured;        // this scene was the one timed for its class
};

class Autotuner {
 public:
  Autotuner(int device, int real_kind) : dev_(device), kind_(real_kind) {}
  // Size class of a scene: octaves of 4 in shape count.
  static int size_class(uint64_t n);
  // Bins per axis covering [lo, hi] with cubic cells of edge `cell`, with
  // at most max_bins bins in all.
  static void grid_for(const double lo[3], const double hi[3], double cell, uint64_t max_bins, int out[3]);
  // Grid for a scene (host arrays in the reference's layouts).  The AABBs are
  // computed on the GPU; the first scene of a size class is measured: the
  // broadphase runs at each candidate cell factor and the fastest wins; the
  // class's later scenes reuse that factor with their own bounds.
  TuneChoice choose(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                    const uint32_t* ID, uint64_t nb, const void* pos, const void* rot, const uint8_t* active);

 private:
  struct ClassInfo {
    double cell_factor;
    double m_per_shape, p_per_shape;
  };
  int dev_, kind_;
  std::mutex mu_;
  std::map<int, ClassInfo> classes_;
};

}  // namespace cdx
