// Grid autotuner: picks the broadphase cell size (bins per axis) per scene
// size class, and the bin-entry and pair capacities per shape that batched
// steps are sized by.  The reference fixes bins per axis by hand
// (setBinsPerAxis, Broadphase.cu:57-60) and leaves its body-per-bin
// adaptivity commented out (Broadphase.cu:458-468).
#pragma once
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>

namespace cdx {

struct TuneChoice {
  int grid[3];
  double cell_factor;   // cell edge / mean AABB edge
  double m_per_shape;   // bin entries per shape measured at that grid
  double p_per_shape;   // candidate pairs per shape measured at that grid
  bool measured;        // this scene was the one timed for its class
  int size_class;       // octave of 4 in shape count
  int density;          // density bucket: round(log2 of the estimated neighbours per shape)
};

class Autotuner {
 public:
  Autotuner(int device, int real_kind) : dev_(device), kind_(real_kind) {}
  // Size class of a scene: octaves of 4 in shape count.
  static int size_class(uint64_t n);
  // Bins per axis covering [lo, hi] with cubic cells of edge `cell`, with
  // at most max_bins bins in all.
  static void grid_for(const double lo[3], const double hi[3], double cell, uint64_t max_bins, int out[3]);
  // Density bucket of a scene from its AABBs: neighbours per shape are
  // about 8 x (summed AABB volume / volume of the bounds) for shapes spread
  // uniformly (two boxes overlap when their centres are closer than the
  // Minkowski sum, 8 box volumes for equal cubes), bucketed by rounded log2.
  static int density_bucket(double aabb_volume_sum, const double lo[3], const double hi[3]);
  // Grid for a scene (host arrays in the reference's layouts).  The AABBs are
  // computed on the GPU.  The first scene of each (size class, density
  // bucket) is measured: its broadphase runs at each candidate cell factor,
  // then at factors around the fastest, and the fastest wins; later scenes
  // of that class and bucket reuse the factor with their own bounds.  The
  // narrowphase is not timed: the pair list, and so its work, is the same at
  // every grid.
  TuneChoice choose(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                    const uint32_t* ID, uint64_t nb, const void* pos, const void* rot, const uint8_t* active);

 private:
  struct ClassInfo {
    double cell_factor;
    double m_per_shape, p_per_shape;
  };
  int dev_, kind_;
  std::mutex mu_;
  std::map<std::pair<int, int>, ClassInfo> classes_;  // by (size class, density bucket)
};

}  // namespace cdx
