// Grid autotuner (declarations in autotune.hpp).
#include "autotune.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "cdx_pipeline.hpp"
#include "cdx_util.hpp"

namespace cdx {

namespace {

// Cell edge / mean AABB edge candidates.  Below ~1 a shape spans many bins
// (entries grow as the cube of the span); above ~3 the bins hold many
// shapes and the per-bin pair search grows quadratically.
constexpr double kFactors[] = {0.75, 1.0, 1.5, 2.0, 3.0};
constexpr double kRefine[] = {0.85, 1.18};  // around the coarse winner
constexpr int kReps = 3;

struct Bounds {
  double lo[3], hi[3], mean_width, volume_sum;
};

template <class R> Bounds bounds_of(const std::vector<R>& ab, uint64_t n) {
  Bounds b;
  for (int k = 0; k < 3; ++k) { b.lo[k] = INFINITY; b.hi[k] = -INFINITY; }
  double w = 0.0, vs = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    double v = 1.0;
    for (int k = 0; k < 3; ++k) {
      double lo = (double)ab[3 * i + k], hi = (double)ab[3 * (n + i) + k];
      b.lo[k] = std::min(b.lo[k], lo);
      b.hi[k] = std::max(b.hi[k], hi);
      w += hi - lo;
      v *= hi - lo;
    }
    vs += v;
  }
  b.volume_sum = vs;
  b.mean_width = n ? w / (3.0 * (double)n) : 1.0;
  if (!(b.mean_width > 0.0)) b.mean_width = 1.0;
  return b;
}

}  // namespace

int Autotuner::density_bucket(double aabb_volume_sum, const double lo[3], const double hi[3]) {
  double vol = 1.0;
  for (int k = 0; k < 3; ++k) vol *= std::max(hi[k] - lo[k], 0.0);
  if (!(vol > 0.0) || !(aabb_volume_sum > 0.0) || !std::isfinite(aabb_volume_sum / vol)) return 0;
  const double neighbours = 8.0 * aabb_volume_sum / vol;
  return (int)std::lround(std::min(std::max(std::log2(neighbours), -8.0), 16.0));
}

int Autotuner::size_class(uint64_t n) {
  int c = 0;
  while (n >= 4) { n >>= 2; ++c; }
  return c;
}

void Autotuner::grid_for(const double lo[3], const double hi[3], double cell, uint64_t max_bins, int out[3]) {
  for (;;) {
    uint64_t tot = 1;
    for (int k = 0; k < 3; ++k) {
      double span = hi[k] - lo[k];
      double g = std::ceil(span / cell);
      if (!(g >= 1.0)) g = 1.0;
      if (g > 2048.0) g = 2048.0;
      out[k] = (int)g;
      tot *= (uint64_t)out[k];
    }
    if (tot <= max_bins) return;
    cell *= 1.25;
  }
}

TuneChoice Autotuner::choose(uint64_t n, const int32_t* T, const void* A, const void* B, const void* C, const void* R,
                             const uint32_t* ID, uint64_t nb, const void* pos, const void* rot,
                             const uint8_t* active) {
  CDX_REQUIRE(n > 0, ERR_INVALID, "a scene needs shapes");
  std::unique_ptr<SceneBase> s = make_scene(kind_, dev_);
  s->set_shapes(n, T, A, B, C, R, ID);
  s->set_bodies(nb, pos, rot, active);
  s->run_aabb();
  s->sync();
  Bounds bd;
  if (kind_ == REAL_F64) {
    std::vector<double> ab(6 * n);
    s->read_aabb(ab.data(), false);
    bd = bounds_of(ab, n);
  } else {
    std::vector<float> ab(6 * n);
    s->read_aabb(ab.data(), false);
    bd = bounds_of(ab, n);
  }
  // More bins than a few per shape only adds empty key space.
  const uint64_t max_bins = std::max<uint64_t>(64, std::min<uint64_t>(4 * n, 1ull << 24));

  const int cls = size_class(n);
  const int dens = density_bucket(bd.volume_sum, bd.lo, bd.hi);
  const std::pair<int, int> key{cls, dens};
  TuneChoice out{};
  out.size_class = cls;
  out.density = dens;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = classes_.find(key);
    if (it != classes_.end()) {
      out.cell_factor = it->second.cell_factor;
      out.m_per_shape = it->second.m_per_shape;
      out.p_per_shape = it->second.p_per_shape;
      out.measured = false;
      grid_for(bd.lo, bd.hi, out.cell_factor * bd.mean_width, max_bins, out.grid);
      return out;
    }
  }

  double best_ms = INFINITY;
  s->set_profiling(true);
  s->set_pair_search(PAIRS_BINS);  // batched steps search by bins
  std::vector<double> ms(ST_COUNT);
  auto trial = [&](double f) {
    int g[3];
    grid_for(bd.lo, bd.hi, f * bd.mean_width, max_bins, g);
    s->set_grid(g[0], g[1], g[2]);
    double t = INFINITY;
    uint64_t m = 0, p = 0;
    for (int r = 0; r < kReps; ++r) {
      s->run_aabb();  // the broadphase shifts the AABBs in place
      s->run_broadphase();
      s->stage_ms(ms.data());
      double bp = 0.0;
      for (int k = ST_BP_BOUNDS; k <= ST_BP_PAIRSORT; ++k) bp += ms[k];
      bp += ms[ST_BP_PAIRSCAN] + ms[ST_BP_PAIREMIT];
      t = std::min(t, bp);
      m = s->counts().m;
      p = s->counts().p;
    }
    if (t < best_ms) {
      best_ms = t;
      out.cell_factor = f;
      out.m_per_shape = (double)m / (double)n;
      out.p_per_shape = (double)p / (double)n;
      for (int k = 0; k < 3; ++k) out.grid[k] = g[k];
    }
  };
  for (double f : kFactors) trial(f);
  // second round: between the winner and its neighbours on the coarse list
  const double f0 = out.cell_factor;
  for (double r : kRefine) trial(f0 * r);
  out.measured = true;
  std::lock_guard<std::mutex> lk(mu_);
  classes_[key] = ClassInfo{out.cell_factor, out.m_per_shape, out.p_per_shape};
  return out;
}

}  // namespace cdx
