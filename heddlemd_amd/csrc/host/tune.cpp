// Launch-shape autotuner: the pair grid, the FFT pass grids, the SPME
// brick and whether SPME runs on its own stream are timed on the force field's own system and the fastest kept.
// None of these shapes is visible in the reference, whose launches are
// fixed; the candidates are the ones the shapes' costs trade between
// (occupancy against tail waves, LDS footprint against halo overhead).
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "forcefield.h"

namespace hmd {

static const uint32_t kPairCandidates[] = {8, 16, 24, 32, 48, 64};
static const uint32_t kFftCandidates[] = {1, 2, 3, 4, 6, 8};

double ForceField::time_compute(uint32_t n_evals) {
  compute(0);   // warm-up (and the list build after a shape change)
  check(hipStreamSynchronize(c.stream), "sync");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  // best of three runs of n_evals evaluations: clock ramps and other work
  // on the device only ever add time
  float ms = 0.0f;
  try {
    for (int rep = 0; rep < 3; ++rep) {
      check(hipEventRecord(e0, c.stream), "hipEventRecord");
      for (uint32_t i = 0; i < n_evals; ++i) compute(0);
      check(hipEventRecord(e1, c.stream), "hipEventRecord");
      check(hipEventSynchronize(e1), "hipEventSynchronize");
      float t = 0.0f;
      check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
      ms = rep == 0 ? t : std::min(ms, t);
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_status();
  return 1000.0 * ms / n_evals;
}

void ForceField::get_tuning(hmd_tuning* t) const {
  *t = hmd_tuning{};
  t->pair_blocks_per_cu = pair_blocks_per_cu;
  if (spme_on) {
    t->fft_blocks_per_cu = fft_blocks_per_cu;
    for (int k = 0; k < 3; ++k) t->brick[k] = brick[k];
    t->spme_stream = spme_stream;
    t->fft_slab = fft_slab_ok ? 1u : 0u;
  }
}

void ForceField::set_tuning(const hmd_tuning& t) {
  if (t.pair_blocks_per_cu == 0 || t.pair_blocks_per_cu > 256)
    throw Error(HMD_ERR_INVALID, "pair_blocks_per_cu must be 1..256");
  if (spme_on) {
    if (t.fft_blocks_per_cu == 0 || t.fft_blocks_per_cu > 64)
      throw Error(HMD_ERR_INVALID, "fft_blocks_per_cu must be 1..64");
    check_brick(t.brick);
    if (t.spme_stream > 1) throw Error(HMD_ERR_INVALID, "spme_stream must be 0 or 1");
    if (t.fft_slab > 1) throw Error(HMD_ERR_INVALID, "fft_slab must be 0 or 1");
  }
  check(hipStreamSynchronize(c.stream), "sync");   // nothing in flight uses the old buffers
  pair_blocks_per_cu = t.pair_blocks_per_cu;
  if (spme_on) {
    spme_stream = t.spme_stream;
    make_side_stream();
    if (t.fft_blocks_per_cu != fft_blocks_per_cu || (t.fft_slab != 0) != fft_slab_ok) {
      fft_blocks_per_cu = t.fft_blocks_per_cu;
      fft_slab_ok = t.fft_slab != 0;
      plan_fft();
    }
    if (t.brick[0] != brick[0] || t.brick[1] != brick[1] || t.brick[2] != brick[2]) {
      for (int k = 0; k < 3; ++k) brick[k] = t.brick[k];
      alloc_bins();
    }
    check(hipStreamSynchronize(c.stream), "sync");
  }
  graph_invalid = true;
}

// Brick shapes: per axis the divisors of the mesh in [order, 32], combined
// with b0 >= b1 >= b2 and 512..4096 points, that fit the LDS of the spread
// (i64 per point) and of the gather (f32 per halo point).
static std::vector<std::array<uint32_t, 3>> brick_candidates(const ForceField& f) {
  std::vector<uint32_t> ax[3];
  for (int k = 0; k < 3; ++k)
    for (uint32_t d = f.order; d <= 32; ++d)
      if (f.grid[k] % d == 0) ax[k].push_back(d);
  std::vector<std::array<uint32_t, 3>> out;
  for (uint32_t a : ax[0])
    for (uint32_t b : ax[1])
      for (uint32_t c : ax[2]) {
        const uint32_t pts = a * b * c;
        if (a < b || b < c || pts < 512 || pts > 4096) continue;
        const uint32_t br[3] = {a, b, c};
        try {
          f.check_brick(br);
        } catch (const Error&) {
          continue;
        }
        out.push_back({a, b, c});
      }
  // nearest to 2,048 points first; at most 8
  std::stable_sort(out.begin(), out.end(), [](const auto& x, const auto& y) {
    return std::fabs(std::log2((double)x[0] * x[1] * x[2]) - 11.0) <
           std::fabs(std::log2((double)y[0] * y[1] * y[2]) - 11.0);
  });
  if (out.size() > 8) out.resize(8);
  return out;
}

void ForceField::autotune(uint32_t n_evals, hmd_tuning* out) {
  if (n_evals == 0) n_evals = 20;
  hmd_tuning best;
  get_tuning(&best);
  uint32_t count = 1;
  const double t_initial = time_compute(n_evals);
  double t_best = t_initial;
  auto trial = [&](const hmd_tuning& t) {
    set_tuning(t);
    const double us = time_compute(n_evals);
    ++count;
    if (us < t_best) {
      t_best = us;
      best = t;
    }
  };
  for (uint32_t v : kPairCandidates) {
    if (v == best.pair_blocks_per_cu) continue;
    hmd_tuning t = best;
    t.pair_blocks_per_cu = v;
    trial(t);
  }
  if (spme_on) {
    for (uint32_t v : kFftCandidates) {
      if (v == best.fft_blocks_per_cu) continue;
      hmd_tuning t = best;
      t.fft_blocks_per_cu = v;
      trial(t);
    }
    {
      hmd_tuning t = best;
      t.spme_stream = best.spme_stream ? 0u : 1u;
      trial(t);
    }
    for (const auto& b : brick_candidates(*this)) {
      if (b[0] == best.brick[0] && b[1] == best.brick[1] && b[2] == best.brick[2]) continue;
      hmd_tuning t = best;
      for (int k = 0; k < 3; ++k) t.brick[k] = b[k];
      trial(t);
    }
  }
  set_tuning(best);
  if (out) {
    *out = best;
    out->n_candidates = count;
    out->us_initial = t_initial;
    out->us_tuned = t_best;
  }
}

}  // namespace hmd

using namespace hmd;

extern "C" {

hmd_status hmd_forcefield_autotune(hmd_forcefield* f, uint32_t n_evals, hmd_tuning* out) {
  HMD_GUARD({
    if (!f) throw Error(HMD_ERR_INVALID, "null force field");
    f->ff->autotune(n_evals, out);
  });
}

hmd_status hmd_forcefield_set_tuning(hmd_forcefield* f, const hmd_tuning* t) {
  HMD_GUARD({
    if (!f || !t) throw Error(HMD_ERR_INVALID, "null argument");
    f->ff->set_tuning(*t);
  });
}

hmd_status hmd_forcefield_get_tuning(const hmd_forcefield* f, hmd_tuning* out) {
  HMD_GUARD({
    if (!f || !out) throw Error(HMD_ERR_INVALID, "null argument");
    f->ff->get_tuning(out);
  });
}

}  // extern "C"
