// ForceField host: the potential registry, the neighbour list, the fused
// pair slot and the SPME reciprocal slot.
//
// Mirrors src/forces/mod.rs (ForceField, PotentialRegistry with builtins in
// evaluation order), src/forces/neighbor_list.rs (NeighborListState) and
// src/forces/spme.rs (SpmeReciprocalGrid / SpmeReciprocalState).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>

#include <cstring>

#include "forcefield.h"

namespace hmd {

// ------------------------------------------------------------ registry ---
// PotentialRegistry::with_builtins (src/forces/mod.rs:363-398): builders in
// evaluation order; each claims its part of the descriptor.
struct PotentialBuilder {
  const char* label;
  std::function<bool(const hmd_forcefield_desc&)> claims;
};

static const std::vector<PotentialBuilder>& builtin_registry() {
  static const std::vector<PotentialBuilder> reg = {
      {"lennard-jones", [](const hmd_forcefield_desc& d) { return d.lj_on != 0; }},
      {"spme-real", [](const hmd_forcefield_desc& d) { return d.coulomb_on != 0; }},
      {"spme-reciprocal", [](const hmd_forcefield_desc& d) { return d.spme_on != 0; }},
  };
  return reg;
}

// Radix plan for the LDS FFT: 4s first, then 2, 3, 5.
static std::vector<uint32_t> radix_plan(uint32_t n) {
  std::vector<uint32_t> r;
  uint32_t m = n;
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  while (m % 2 == 0) { r.push_back(2); m /= 2; }
  while (m % 3 == 0) { r.push_back(3); m /= 3; }
  while (m % 5 == 0) { r.push_back(5); m /= 5; }
  if (m != 1 || r.size() > 8)
    throw Error(HMD_ERR_INVALID, "SPME grid size " + std::to_string(n) +
                                     " must factor into 2, 3 and 5");
  return r;
}

static void twiddles(DevBuf<float>& buf, uint32_t n, hipStream_t s) {
  std::vector<float> t(2 * (size_t)n);
  for (uint32_t m = 0; m < n; ++m) {
    double a = -2.0 * M_PI * (double)m / (double)n;
    t[2 * m] = (float)std::cos(a);
    t[2 * m + 1] = (float)std::sin(a);
  }
  buf.alloc(t.size(), false);
  buf.upload(t.data(), t.size(), s);
}

// |b(k)|^2 per axis with near-zero neighbour averaging
// (src/forces/spme.rs:524-585), host f64 -> f32.
static double cardinal_bspline(int p, double x) {
  std::vector<double> v(p);
  for (int i = 0; i < p; ++i) {
    double xi = x - i;
    v[i] = (xi >= 0.0 && xi < 1.0) ? 1.0 : 0.0;
  }
  for (int order = 2; order <= p; ++order) {
    double inv = 1.0 / (order - 1.0);
    for (int i = 0; i < p - order + 1; ++i) {
      double xi = x - i;
      v[i] = xi * inv * v[i] + (order - xi) * inv * v[i + 1];
    }
  }
  return v[0];
}

static std::vector<double> b_factors(uint32_t n, uint32_t p) {
  std::vector<double> mod(n);
  for (uint32_t k = 0; k < n; ++k) {
    double th = 2.0 * M_PI * k / n, re = 0.0, im = 0.0;
    for (uint32_t j = 0; j + 1 < p; ++j) {
      double m = cardinal_bspline((int)p, j + 1.0);
      re += m * std::cos(th * j);
      im += m * std::sin(th * j);
    }
    mod[k] = re * re + im * im;
  }
  for (uint32_t k = 0; k < n; ++k)
    if (mod[k] < 1.0e-7) mod[k] = 0.5 * (mod[(k + n - 1) % n] + mod[(k + 1) % n]);
  std::vector<double> out(n);
  for (uint32_t k = 0; k < n; ++k) out[k] = mod[k] > 0.0 ? 1.0 / mod[k] : 0.0;
  return out;
}

static uint32_t pick_brick(uint32_t n, uint32_t p) {
  const uint32_t prefs[] = {16, 8, 12, 10, 6};
  const uint32_t want = n >= 128 ? 16 : 8;
  if (n % want == 0 && want >= p) return want;
  for (uint32_t b : prefs)
    if (b >= p && n % b == 0) return b;
  for (uint32_t b = p; b <= n; ++b)
    if (n % b == 0) return b;
  return n;
}

static uint32_t lines_per_block(uint32_t n) {
  uint32_t L = 40960u / (16u * (n + 1));
  return std::max(1u, std::min(64u, L));
}

// Power-of-two line lengths with a compile-time kernel: index 0..4 for
// 16, 32, 64, 128, 256 (else -1), and the second-stage radix N2.
static int fft_p2_index(uint32_t n) {
  switch (n) {
    case 16: return 0;
    case 32: return 1;
    case 64: return 2;
    case 128: return 3;
    case 256: return 4;
    default: return -1;
  }
}
static const uint32_t kP2N2[5] = {4, 4, 8, 8, 16};
static const KernelId kP2Axis[5] = {K_hmd_fft_axis_16, K_hmd_fft_axis_32, K_hmd_fft_axis_64,
                                    K_hmd_fft_axis_128, K_hmd_fft_axis_256};
static const KernelId kP2R2c[5] = {K_hmd_fft_r2c_16, K_hmd_fft_r2c_32, K_hmd_fft_r2c_64,
                                   K_hmd_fft_r2c_128, K_hmd_fft_r2c_256};
static const KernelId kP2C2r[5] = {K_hmd_fft_c2r_16, K_hmd_fft_c2r_32, K_hmd_fft_c2r_64,
                                   K_hmd_fft_c2r_128, K_hmd_fft_c2r_256};

// Strided axis of length n over n_outer slabs of `cols` adjacent lines:
// column tiles of width W (stage-1 capacity 2 x 256 / N2, 64 KiB of LDS),
// at least `target` blocks in all (a few per compute unit).
// `tiles` LDS tiles of n x W complex values per workgroup (the a-axis pass
// stages the influence table beside the data), at most 64 KiB in all.
static ForceField::FftPass plan_axis(uint32_t n, uint32_t n_outer, uint32_t cols,
                                     uint32_t target, uint32_t tiles_lds = 1) {
  ForceField::FftPass ps;
  const int ix = fft_p2_index(n);
  if (ix < 0) return ps;
  const uint32_t wmax = std::min(512u / kP2N2[ix], 8192u / (n * tiles_lds));
  uint32_t tiles = std::max(ceil_div(cols, wmax), ceil_div(target, n_outer));
  tiles = std::min(tiles, cols);
  ps.p2 = true;
  ps.k = kP2Axis[ix];
  ps.W = ceil_div(cols, tiles);
  ps.grid = n_outer * ceil_div(cols, ps.W);
  ps.shmem = (size_t)n * ps.W * 8 * tiles_lds;
  return ps;
}

// Contiguous lines (the c axis, N = n2 / 2 complex points each).
static ForceField::FftPass plan_lines(uint32_t N, uint32_t n_lines, uint32_t target, bool r2c) {
  ForceField::FftPass ps;
  const int ix = fft_p2_index(N);
  if (ix < 0) return ps;
  const uint32_t wmax = std::min(512u / kP2N2[ix], 8192u / N);
  ps.p2 = true;
  ps.k = r2c ? kP2R2c[ix] : kP2C2r[ix];
  ps.W = std::max(1u, std::min(wmax, ceil_div(n_lines, target)));
  ps.grid = ceil_div(n_lines, ps.W);
  ps.shmem = (size_t)N * ps.W * 8;
  return ps;
}

// The reference's A&S 7.1.26 erfc (src/forces/spme.rs:712-760), in FP64.
static double erfc_as(double x) {
  const double t = 1.0 / (1.0 + 0.3275911 * x);
  const double poly =
      t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 +
                                                                        t * 1.061405429))));
  return poly * std::exp(-x * x);
}

// Real-space Ewald kernel in the octave table of r^2 (hmd_coul_bin): the
// table spans [2^(e - HMD_COUL_OCTAVES), 2^e) with 2^e the first power of two
// above r_c^2, each octave split into 2^HMD_COUL_BITS bins of equal width in
// r^2, so a bin's relative width is at most 2^-HMD_COUL_BITS everywhere.  Per
// bin the cubic Hermite interpolant, in the bin's own coordinate s in [0, 1),
// of the force factor G(r) = [erfc(a r)/r + 2a/sqrt(pi) exp(-a^2 r^2)]/r^2
// and of the energy E(r) = erfc(a r)/r, from FP64 values and derivatives in
// r^2 at both bin ends (derivatives by a fourth-order central difference),
// stored as the same cubic in t = 1 + s (the kernels read t off the bits of
// r^2 and evaluate it by Horner's rule; the basis change is done in FP64).
// Distances below the table's first r (between r_c/16 and r_c/11 with eight
// octaves: under 0.9 angstrom for a 1 nm cutoff) clamp to its first value;
// no non-excluded pair comes that close.
void ForceField::build_coulomb_table(double a, double rc, hipStream_t s) {
  const uint32_t nb = HMD_COUL_BINS, per = 1u << HMD_COUL_BITS;
  const int e_hi = (int)std::floor(std::log2(rc * rc)) + 1;   // 2^e_hi > rc^2
  const int e_lo = e_hi - (int)HMD_COUL_OCTAVES;
  auto G = [a](double r2) {
    const double r = std::sqrt(r2);
    return (erfc_as(a * r) / r + 2.0 * a / std::sqrt(M_PI) * std::exp(-a * a * r2)) / r2;
  };
  auto E = [a](double r2) { return erfc_as(a * std::sqrt(r2)) / std::sqrt(r2); };
  std::vector<float> tab(2 * 4 * nb);
  for (int f = 0; f < 2; ++f) {
    auto F = [&](double r2) { return f == 0 ? G(r2) : E(r2); };
    for (uint32_t k = 0; k < nb; ++k) {
      const double w = std::ldexp(1.0, e_lo + (int)(k / per)) / per;   // bin width in r^2
      const double x0 = std::ldexp(1.0, e_lo + (int)(k / per)) + (k % per) * w;
      double v[2], dv[2];
      for (int end = 0; end < 2; ++end) {
        const double x = x0 + end * w, d = 1e-3 * w;
        v[end] = F(x);
        dv[end] = w * (F(x - 2 * d) - 8 * F(x - d) + 8 * F(x + d) - F(x + 2 * d)) / (12 * d);
      }
      // Hermite cubic in s, then the same cubic in t = 1 + s
      const double a0 = v[0], a1 = dv[0];
      const double a2 = 3.0 * (v[1] - v[0]) - 2.0 * dv[0] - dv[1];
      const double a3 = 2.0 * (v[0] - v[1]) + dv[0] + dv[1];
      float* c = &tab[4 * (f * nb + k)];
      c[0] = (float)(a0 - a1 + a2 - a3);
      c[1] = (float)(a1 - 2.0 * a2 + 3.0 * a3);
      c[2] = (float)(a2 - 3.0 * a3);
      c[3] = (float)a3;
    }
  }
  coul_table.alloc(tab.size(), false);
  coul_table.upload(tab.data(), tab.size(), s);
  const float lo = std::ldexp(1.0f, e_lo), hi = std::ldexp(1.0f, e_hi);
  std::memcpy(&coul_u_lo, &lo, 4);
  std::memcpy(&coul_u_hi, &hi, 4);
  coul_u_hi -= 1u;   // the largest f32 below 2^e_hi
}

// ---------------------------------------------------------- construction --
ForceField::ForceField(Context& ctx, Particles& part, const hmd_forcefield_desc& d)
    : c(ctx), p(part) {
  hipStream_t s = c.stream;
  n = p.n;
  for (const auto& b : builtin_registry())
    if (b.claims(d)) slots.push_back(b.label);
  if (slots.empty()) throw Error(HMD_ERR_INVALID, "force field has no active slot");
  lj_on = d.lj_on != 0;
  coul_on = d.coulomb_on != 0;
  spme_on = d.spme_on != 0;
  if (spme_on && !coul_on)
    throw Error(HMD_ERR_INVALID, "spme-reciprocal requires the spme-real slot");

  // ---- pair slot parameters (src/forces/lj.rs, src/forces/spme.rs)
  n_types = std::max(1u, d.n_types);
  if (n_types > 16) throw Error(HMD_ERR_UNSUPPORTED, "at most 16 particle types");
  if (p.max_type >= n_types) throw Error(HMD_ERR_INVALID, "particle type index >= n_types");
  double cut = 0.0;
  std::vector<float> c6c12(2 * n_types * n_types, 0.0f);
  double lj_cut = -1.0, lj_sw = -1.0;
  if (lj_on) {
    for (uint32_t k = 0; k < n_types * n_types; ++k) {
      double sg = d.lj_sigma[k], ep = d.lj_epsilon[k];
      double s6 = std::pow(sg, 6.0);
      c6c12[2 * k] = (float)(4.0 * ep * s6);
      c6c12[2 * k + 1] = (float)(4.0 * ep * s6 * s6);
      if (c6c12[2 * k] != 0.0f || c6c12[2 * k + 1] != 0.0f)
        lj_type_mask |= (1u << (k % n_types)) | (1u << (k / n_types));
      if (lj_cut < 0.0) { lj_cut = d.lj_cutoff[k]; lj_sw = d.lj_switch[k]; }
      if (d.lj_cutoff[k] != lj_cut || d.lj_switch[k] != lj_sw)
        throw Error(HMD_ERR_UNSUPPORTED,
                    "per-type-pair LJ cutoffs / switches differ (uniform only)");
    }
    cut = lj_cut;
  }
  if (coul_on) {
    if (lj_on && d.r_cut_real != lj_cut)
      throw Error(HMD_ERR_UNSUPPORTED, "LJ and real-space cutoffs must match");
    cut = d.r_cut_real;
    alpha = (float)d.alpha;
    k_coulomb = (float)d.k_coulomb;
  }
  r_cut = (float)cut;
  rc2 = (float)(cut * cut);
  if (lj_on && lj_sw < lj_cut) {
    lj_rs2 = (float)(lj_sw * lj_sw);
    lj_inv_delta = (float)(1.0 / (lj_cut * lj_cut - lj_sw * lj_sw));
  } else {
    lj_rs2 = rc2;
    lj_inv_delta = 0.0f;
  }
  lj_table.alloc(c6c12.size(), false);
  lj_table.upload(c6c12.data(), c6c12.size(), s);
  if (coul_on) build_coulomb_table(d.alpha, cut, s);

  // ---- exclusions
  std::vector<uint32_t> eoff(n + 1, 0u), epar;
  if (d.excl_offsets) {
    eoff.assign(d.excl_offsets, d.excl_offsets + n + 1);
    for (uint32_t i = 0; i < n; ++i) {
      if (eoff[i + 1] < eoff[i]) throw Error(HMD_ERR_INVALID, "exclusion offsets not monotonic");
      if (eoff[i + 1] - eoff[i] > 8)
        throw Error(HMD_ERR_UNSUPPORTED, "more than 8 exclusion partners for one atom");
    }
    epar.assign(d.excl_partners, d.excl_partners + eoff[n]);
    for (uint32_t m = 0; m < eoff[n]; ++m)
      if (epar[m] >= n) throw Error(HMD_ERR_INVALID, "exclusion partner out of range");
    h_eoff = eoff;
    h_epar = epar;
    h_slj.assign(eoff[n], 0.0f);
    h_sco.assign(eoff[n], 0.0f);
    for (uint32_t m = 0; m < eoff[n]; ++m) {
      const float a = d.excl_lj_scale ? d.excl_lj_scale[m] : 0.0f;
      const float b = d.excl_coul_scale ? d.excl_coul_scale[m] : 0.0f;
      if (!std::isfinite(a) || !std::isfinite(b))
        throw Error(HMD_ERR_INVALID, "exclusion scale not finite");
      h_slj[m] = lj_on ? a : 0.0f;
      h_sco[m] = coul_on ? b : 0.0f;
    }
    build_scaled_list(s);
  }
  if (epar.empty()) epar.push_back(0u);
  excl_off.alloc(eoff.size(), false);
  excl_off.upload(eoff.data(), eoff.size(), s);
  excl_par.alloc(epar.size(), false);
  excl_par.upload(epar.data(), epar.size(), s);

  init_neighbor_list(d);
  if (spme_on) init_spme(d);
  check(hipStreamSynchronize(s), "sync");
  probe_capacity();
  // SPME beside the list build and pair kernel pays for large systems
  // (196,608 atoms: 0.543 -> 0.469 ms/step); for small ones the extra graph
  // branches cost more than they overlap (64 boxes of 648-12,288 atoms in one
  // graph: 4.9 -> 7.5 ms per batched step), so it starts on at 16,384 atoms.
  spme_stream = (n >= 16384 && c.owns_stream) ? 1u : 0u;
  make_side_stream();
  p.observers.push_back(this);   // box changes reach this force field
}

// The SPME stream and its fork / join events, made only when the fork is on
// (streams share the hardware queues round-robin: an idle one still takes a
// queue slot from the streams of other force fields) and never during a
// capture (the set-up and set_tuning paths call this).
void ForceField::make_side_stream() {
  if (!spme_on || !spme_stream || side) return;
  check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
  check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
}

// The box requirement (r_c below half of every perpendicular width) and
// the orthorhombic flags, for the current lattice.
void ForceField::check_box() {
  const HmdLattice& L = p.lattice;
  // perpendicular widths of the lower-triangular cell
  double wa = L.lx, wb = L.ly, wc = L.lz;  // exact for orthorhombic
  if (L.xy != 0.0f || L.xz != 0.0f || L.yz != 0.0f) {
    double ax = L.lx, bx = L.xy, by = L.ly, cx = L.xz, cy = L.yz, cz = L.lz;
    double vol = ax * by * cz;
    // |b x c|, |c x a|, |a x b| for a=(ax,0,0), b=(bx,by,0), c=(cx,cy,cz)
    double bxc = std::sqrt(std::pow(by * cz, 2) + std::pow(bx * cz, 2) +
                           std::pow(bx * cy - by * cx, 2));
    double cxa = std::sqrt(std::pow(cz * ax, 2) + std::pow(cy * ax, 2));
    double axb = ax * by;
    wa = vol / bxc; wb = vol / cxa; wc = vol / axb;
  }
  const double widths[3] = {wa, wb, wc};
  for (int k = 0; k < 3; ++k)
    if (!(widths[k] > 2.0 * (double)r_cut)) {
      char msg[160];
      std::snprintf(msg, sizeof msg,
                    "box too small: perpendicular width %.4f <= 2 * r_cut = %.4f",
                    widths[k], 2.0 * (double)r_cut);
      throw Error(HMD_ERR_BOX_TOO_SMALL, msg);
    }
  orthorhombic = (L.xy == 0.0f && L.xz == 0.0f && L.yz == 0.0f);
  ortho_fast = orthorhombic && 0.5 * L.lx > 1.001 * r_search &&
               0.5 * L.ly > 1.001 * r_search && 0.5 * L.lz > 1.001 * r_search;

}

// compute_cell_layout analogue (src/forces/neighbor_list.rs:1563-1583).
// The reference's cell list needs >= 3 cells of edge r_search per axis
// (BoxTooSmallForCells otherwise).  Nothing here depends on cells: the fine
// Hilbert grid only orders atoms and every distance test is a minimum-image
// one.  The minimum-image distance is a metric on the periodic cell, so the
// skin argument (a pair within r_c now was within r_c + r_skin at the
// rebuild when neither atom moved more than r_skin / 2) holds for any box,
// and the in-cutoff image of a pair is unique while r_c is below half of
// every perpendicular width.  That is the only box requirement; small
// replicas (edges down to ~1.9 nm with r_c = 0.8 nm) rely on it.
void ForceField::init_neighbor_list(const hmd_forcefield_desc& d) {
  r_skin = (float)d.r_skin;
  r_search = r_cut + r_skin;
  r_search_sq = r_search * r_search;
  check_box();
  const HmdLattice& L = p.lattice;

  // Fine ordering grid of 2^bits cells per axis (the Hilbert curve needs a
  // power of two), at least as fine as ~1.6 atoms per cell.
  const double vol = (double)L.lx * L.ly * L.lz;
  const double cell = std::cbrt(1.6 * vol / n);
  uint32_t gmax = 1;
  for (int k = 0; k < 3; ++k) {
    double w = k == 0 ? L.lx : (k == 1 ? L.ly : L.lz);
    gmax = std::max(gmax, (uint32_t)std::max(1.0, std::min(1024.0, std::floor(w / cell))));
  }
  grid_bits = 1;
  while ((1u << grid_bits) < gmax) ++grid_bits;
  for (int k = 0; k < 3; ++k) g[k] = 1u << grid_bits;
  const uint32_t bits = grid_bits;
  n_keys = 1u << (3 * bits);
  if (n_keys > (1u << 24)) throw Error(HMD_ERR_UNSUPPORTED, "fine grid too large");

  n_clusters = ceil_div(n, HMD_CLUSTER);
  n_pad = n_clusters * HMD_CLUSTER;
  if (n_clusters > 65536u)   // the build's candidate masks (20 B per 64 clusters) share its LDS
    throw Error(HMD_ERR_UNSUPPORTED, "more than 65,536 clusters (4,194,304 atoms)");
  status.alloc(4);
  atom_key.alloc(n);
  cell_count.alloc(n_keys);
  cell_start.alloc(n_keys + 1);
  sorted_id.alloc(n_pad);
  slot_of.alloc(n);
  type_sorted.alloc(n_pad);
  posq_sorted.alloc(4 * (size_t)n_pad);
  ref_pos.alloc(4 * (size_t)n);
  image_sorted.alloc(4 * (size_t)n_pad);
  bbox.alloc(8 * (size_t)n_clusters);
  sub_bbox.alloc(64 * (size_t)n_clusters);
  excl_slot.alloc(8 * (size_t)n_pad, false);
  entry_count.alloc(n_clusters);
  excl_entries.alloc(n_clusters);
  entry_offset.alloc(n_clusters + 1);
  n_work.alloc(1);
  acc.alloc(5 * (size_t)n_pad);
  cell_scan.init(n_keys);
  entry_scan.init(n_clusters);
  alloc_entries(8);
}

// Entries per cluster: at most n_clusters (a cluster lists every atom above
// it; all_pairs_tile_capacity, src/forces/neighbor_list.rs:276-287), and the
// work-list indices c * cap + k stay below 2^32.
void ForceField::alloc_entries(uint32_t cap) {
  cap = std::max(1u, std::min(cap, n_clusters));
  if ((uint64_t)n_clusters * cap >= (1ull << 32))
    throw Error(HMD_ERR_UNSUPPORTED,
                "neighbour list needs " + std::to_string(cap) + " entries per cluster for " +
                    std::to_string(n_clusters) + " clusters (work indices exceed 32 bits)");
  cap_entries = cap;
  const size_t slots_total = (size_t)n_clusters * cap * HMD_ENTRY;
  entry_j.alloc(slots_total, false);
  entry_mask.alloc((size_t)n_clusters * cap, false);
  entry_excl.alloc(slots_total, false);
  work.alloc((size_t)n_clusters * cap);
}

void ForceField::init_spme(const hmd_forcefield_desc& d) {
  order = d.spline_order;
  if (order != 4 && order != 5 && order != 6 && order != 8)
    throw Error(HMD_ERR_UNSUPPORTED, "spline order must be 4, 5, 6 or 8");
  for (int k = 0; k < 3; ++k) {
    grid[k] = d.grid[k];
    if (grid[k] < 2 * order)  // SpmeError::InvalidGrid (spme.rs:236-248)
      throw Error(HMD_ERR_INVALID, "SPME grid must be at least 2 * spline_order per axis");
    brick[k] = d.brick[k] ? d.brick[k] : pick_brick(grid[k], order);
  }
  // Default 16^3 bricks (large meshes) are halved along c: 2,048 points per
  // workgroup is where the brick spread and gather ran fastest at 128^3
  // with 196,608 atoms (spread 61 -> 49-50 us, gather 26 -> 23-24 us; 1,024
  // points 72 us, 512 points 100 us).
  if (!d.brick[0] && !d.brick[1] && !d.brick[2] && brick[0] == 16 && brick[1] == 16 &&
      brick[2] == 16 && 8 >= order)
    brick[2] = 8;
  check_brick(brick);
  if (grid[2] % 2 != 0) throw Error(HMD_ERR_UNSUPPORTED, "SPME grid[2] must be even");
  hipStream_t s = c.stream;
  const uint32_t n0 = grid[0], n1 = grid[1], n2 = grid[2];
  nch = n2 / 2 + 1;
  M = (size_t)n0 * n1 * n2;
  alloc_bins();
  rho.alloc(M);
  V.alloc(M);
  mesh_c.alloc(2 * (size_t)n0 * n1 * nch);
  slow_f.alloc(3 * (size_t)n);
  slow_e.alloc(n);
  slow_w.alloc(n);
  spme_out.alloc(2);
  auto up = [&](DevBuf<float>& b, const std::vector<float>& v) {
    b.alloc(v.size(), false);
    b.upload(v.data(), v.size(), s);
  };
  for (int d = 0; d < 3; ++d) {
    const std::vector<double> b = b_factors(grid[d], order);
    bfac[d].alloc(b.size(), false);
    bfac[d].upload(b.data(), b.size(), s);
  }
  twiddles(tw_a, n0, s);
  twiddles(tw_b, n1, s);
  twiddles(tw_c, n2 / 2, s);
  twiddles(tw_r2c, n2, s);
  rad_a = radix_plan(n0);
  rad_b = radix_plan(n1);
  rad_c = radix_plan(n2 / 2);
  // per-atom self energy k_C alpha/sqrt(pi) q^2 (spme.rs:808-818)
  std::vector<float> posq(4 * (size_t)n);
  p.posq.download(posq.data(), posq.size(), s);
  std::vector<float> us(n);
  u_self_total = 0.0;
  const double pref = (double)d.k_coulomb * d.alpha / std::sqrt(M_PI);
  for (uint32_t i = 0; i < n; ++i) {
    double q = posq[4 * (size_t)i + 3];
    double u = pref * q * q;
    us[i] = (float)u;
    u_self_total += u;
  }
  up(u_self, us);
  plan_fft();
}

void ForceField::check_brick(const uint32_t b[3]) const {
  for (int k = 0; k < 3; ++k)
    if (b[k] == 0 || grid[k] % b[k] != 0 || b[k] < order)
      throw Error(HMD_ERR_INVALID, "SPME brick must divide the grid and be >= order");
  if ((size_t)b[0] * b[1] * b[2] * 8 > 65536 ||
      (size_t)(b[0] + order - 1) * (b[1] + order - 1) * (b[2] + order - 1) * 4 > 65536)
    throw Error(HMD_ERR_INVALID, "SPME brick exceeds 64 KiB of LDS");
}

// Atom bins of the brick spread / gather for the current brick shape.
void ForceField::alloc_bins() {
  for (int k = 0; k < 3; ++k) nb[k] = grid[k] / brick[k];
  n_bins = nb[0] * nb[1] * nb[2];
  atom_bin.alloc(n);
  bin_count.alloc(n_bins);
  bin_start.alloc(n_bins + 1);
  bin_atoms.alloc(n);
  bin_scan.init(n_bins);
}

// Launch shapes of the five FFT passes (fft_blocks_per_cu workgroups per
// CU as the target) and the a-axis pass's energy partials.
void ForceField::plan_fft() {
  const uint32_t n0 = grid[0], n1 = grid[1], n2 = grid[2];
  const uint32_t target = (uint32_t)c.num_cus * fft_blocks_per_cu;
  pass_r2c = plan_lines(n2 / 2, n0 * n1, target, true);
  pass_c2r = plan_lines(n2 / 2, n0 * n1, target, false);
  pass_b = plan_axis(n1, n0, nch, target);
  pass_a = plan_axis(n0, n1, nch, target, 2);
  slab = FftSlab{};
  if (n0 == n1 && n1 == n2 && (n0 == 128 || n0 == 64)) {
    const size_t nc = n2 / 2;
    const size_t lds_fwd = 8 * (nc + 1) * (size_t)(n1 + 1);
    if (lds_fwd <= c.max_lds && fft_slab_ok) {
      slab.on = true;
      slab.fwd = n0 == 128 ? K_hmd_fft_slab_fwd_128 : K_hmd_fft_slab_fwd_64;
      slab.solve = n0 == 128 ? K_hmd_fft_slab_solve_128 : K_hmd_fft_slab_solve_64;
      slab.inv = n0 == 128 ? K_hmd_fft_slab_inv_128 : K_hmd_fft_slab_inv_64;
      slab.lds_fwd = lds_fwd;
      slab.lds_solve = 8 * (size_t)n0 * nch;
    }
  }
  if (slab.on) {
    partial_blocks = n1;
  } else if (pass_a.p2) {
    partial_blocks = pass_a.grid;
  } else {
    const uint32_t La = lines_per_block(n0);
    partial_blocks = ceil_div((size_t)n1 * nch, La);
  }
  partials.alloc(2 * (size_t)partial_blocks);
  if (pass_a.p2 || slab.on) gv.alloc(2 * (size_t)n0 * n1 * nch, false);
  compute_gv_table();
}

// (G, vf) influence table of the power-of-two a-axis pass for the current
// box (rebuilt when the box changes).
void ForceField::compute_gv_table() {
  if (pass_a.p2 || slab.on) {
    const uint32_t n0 = grid[0], n1 = grid[1], n2 = grid[2], nch = n2 / 2 + 1;
    const size_t m_c = (size_t)n0 * n1 * nch;
    FftParams g{};
    g.L = p.lattice;
    g.alpha = alpha;
    g.k_coulomb = k_coulomb;
    g.n0 = n0; g.n1 = n1; g.n2 = n2; g.nch = nch;
    g.bfac0 = bfac[0].ptr; g.bfac1 = bfac[1].ptr; g.bfac2 = bfac[2].ptr;
    g.gv = gv.ptr;
    launch(c, K_hmd_spme_gv_table, (uint32_t)ceil_div(m_c, (size_t)256), 256, 0, g);
  }
}

// Capacity for a build that needed `need` entries in its fullest cluster:
// 1.5x headroom, clamped to the all-pairs ceiling (alloc_entries).
uint32_t ForceField::grown_capacity(uint32_t need) const {
  const uint64_t want = (uint64_t)std::ceil((double)need * kNblGrowth) + 1;
  return (uint32_t)std::max<uint64_t>(4, std::min<uint64_t>(want, n_clusters));
}

// Probe rebuild: size the per-cluster entry capacity from the measured high
// water (the reference's probe sizing + x1.5 growth,
// src/forces/neighbor_list.rs:1151-1221, 421-445).
void ForceField::probe_capacity() {
  request_rebuild();
  for (int attempt = 0; attempt < 3; ++attempt) {
    rebuild_chain();
    uint32_t st[4];
    status.download(st, 4, c.stream);
    const uint32_t need = st[2];
    const bool fits = !(st[0] & HMD_STATUS_ENTRY_OVERFLOW) && need <= cap_entries;
    if (fits && (cap_entries >= grown_capacity(need) || cap_entries == n_clusters)) break;
    alloc_entries(grown_capacity(need));
    uint32_t clear[4] = {HMD_STATUS_DISPLACED, st[1], 0u, 0u};
    status.upload(clear, 4, c.stream);
  }
  check(hipStreamSynchronize(c.stream), "sync");
}

// A new box (hmd_particles_set_box): re-check it, refresh what depends on
// it, and have the next evaluation rebuild the list and re-capture graphs.
void ForceField::on_box_change() {
  check_box();
  if (spme_on) compute_gv_table();
  request_rebuild();
  graph_invalid = true;
}

ForceField::~ForceField() {
  auto& obs = p.observers;
  obs.erase(std::remove(obs.begin(), obs.end(), this), obs.end());
  if (side) (void)hipStreamSynchronize(side);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
}

void ForceField::request_rebuild() {
  uint32_t st[4];
  status.download(st, 4, c.stream);
  st[0] |= HMD_STATUS_DISPLACED;
  status.upload(st, 4, c.stream);
  check(hipStreamSynchronize(c.stream), "sync");
}

// Status for a batch re-run from restored positions: any overflow and the
// high-water mark of the abandoned batch are dropped (the capacity has
// already grown) and the next evaluation rebuilds the list.
void ForceField::restart_list() {
  uint32_t st[4];
  status.download(st, 4, c.stream);
  uint32_t reset[4] = {HMD_STATUS_DISPLACED, st[1], 0u, 0u};
  status.upload(reset, 4, c.stream);
  check(hipStreamSynchronize(c.stream), "sync");
}

// ----------------------------------------------------------------- step ---
void ForceField::rebuild_chain() {
  rebuild_sort();
  rebuild_list();
}

// Cell keys and the cell sort (gated on a rebuild), then the sorted
// positions (every step).
void ForceField::rebuild_sort() {
  const uint32_t* gate = status.ptr;
  const HmdLattice L = p.lattice;
  NblKeyParams kp{p.posq.ptr, n, L, g[0], g[1], g[2], grid_bits, atom_key.ptr, cell_count.ptr,
                  cell_start.ptr, sorted_id.ptr, gate};
  launch(c, K_hmd_nbl_keys, ceil_div(n, 256), 256, 0, kp);
  cell_scan.run(c, cell_count.ptr, cell_start.ptr, gate, HMD_STATUS_DISPLACED);
  launch(c, K_hmd_nbl_scatter, ceil_div(n, 256), 256, 0, kp);
  NblSortParams sp{cell_start.ptr, n_keys, sorted_id.ptr, gate};
  launch(c, K_hmd_nbl_sort_cells, ceil_div(n_keys, 256), 256, 0, sp);
  NblGatherParams gp{p.posq.ptr, p.type.ptr, n, n_pad, sorted_id.ptr, slot_of.ptr,
                     posq_sorted.ptr, type_sorted.ptr, ref_pos.ptr, gate, p.lattice,
                     image_sorted.ptr};
  launch(c, K_hmd_nbl_gather, ceil_div(n_pad, 256), 256, 0, gp);
}

// Cluster boxes, the cluster-pair list and the work list (gated on a
// rebuild).
void ForceField::rebuild_list() {
  const uint32_t* gate = status.ptr;
  const HmdLattice L = p.lattice;
  NblBboxParams bp{posq_sorted.ptr, n, n_clusters, bbox.ptr, sub_bbox.ptr, gate,
                   sorted_id.ptr, slot_of.ptr, excl_off.ptr, excl_par.ptr,
                   excl_slot.ptr, n_pad};
  launch(c, K_hmd_nbl_bbox, ceil_div(n_clusters, 4), 256, 0, bp);
  NblBuildParams bd{};
  bd.posq_sorted = posq_sorted.ptr;
  bd.sorted_id = sorted_id.ptr;
  bd.slot_of = slot_of.ptr;
  bd.bbox = bbox.ptr;
  bd.sub_bbox = sub_bbox.ptr;
  bd.L = L;
  bd.n = n;
  bd.n_clusters = n_clusters;
  bd.r_search_sq = r_search_sq;
  bd.r_search = r_search;
  bd.orthorhombic_fast = ortho_fast ? 1u : 0u;
  bd.cap_entries = cap_entries;
  bd.max_chunks = ceil_div(n_clusters, 64);   // one ballot per 64 j-clusters
  const size_t chunk_lds = 16 * (size_t)bd.max_chunks + 4 * (size_t)((bd.max_chunks + 2) & ~1u);
  uint32_t lds_max = (uint32_t)((kNblBuildLds - kNblBuildStaticLds - chunk_lds) / (HMD_ENTRY * 4));
  if (lds_entries) lds_max = std::min(lds_max, lds_entries);
  bd.lds_slots = std::min(cap_entries, lds_max) * HMD_ENTRY;
  bd.entry_j = entry_j.ptr;
  bd.entry_mask = entry_mask.ptr;
  bd.entry_excl = entry_excl.ptr;
  bd.entry_count = entry_count.ptr;
  bd.excl_entries = excl_entries.ptr;
  bd.excl_slot = excl_slot.ptr;
  bd.n_pad = n_pad;
  bd.status = status.ptr;
#ifndef HMD_NBL_WAVES
#define HMD_NBL_WAVES 4u
#endif
  launch(c, K_hmd_nbl_build, n_clusters, 64 * HMD_NBL_WAVES, chunk_lds + (size_t)bd.lds_slots * 4,
         bd);
  entry_scan.run(c, entry_count.ptr, entry_offset.ptr, gate, HMD_STATUS_DISPLACED);
  NblWorkParams wp{entry_count.ptr, entry_offset.ptr, n_clusters, cap_entries,
                   work.ptr, n_work.ptr, status.ptr};
  launch(c, K_hmd_nbl_worklist, ceil_div(n_clusters, 4u), 256, 0, wp);
  launch(c, K_hmd_nbl_close, 1, 64, 0, wp);
}

NblDispParams ForceField::disp_params() const {
  return NblDispParams{p.posq.ptr, ref_pos.ptr, n, p.lattice, 0.25f * r_skin * r_skin,
                       status.ptr};
}

void ForceField::neighbor_pre_step(const std::function<void()>& after_sort,
                                   bool displacement_checked) {
  if (!displacement_checked) {
    const NblDispParams dp = disp_params();
    launch(c, K_hmd_nbl_displacement, ceil_div(n, 256), 256, 0, dp);
    c.mark("nbl_displacement");
  }
  rebuild_sort();
  if (after_sort) after_sort();
  rebuild_list();
  c.mark("nbl_rebuild");
}

void ForceField::pair_forces(bool fev) {
  PairParams pp{};
  pp.posq_sorted = posq_sorted.ptr;
  pp.type_sorted = type_sorted.ptr;
  pp.work = work.ptr;
  pp.n_work = n_work.ptr;
  pp.entry_j = entry_j.ptr;
  pp.entry_mask = entry_mask.ptr;
  pp.entry_excl = entry_excl.ptr;
  pp.excl_entries = excl_entries.ptr;
  pp.bbox = bbox.ptr;
  pp.L = p.lattice;
  pp.n = n;
  pp.cap_entries = cap_entries;
  pp.n_types = n_types;
  pp.lj_c6c12 = lj_table.ptr;
  pp.lj_type_mask = lj_type_mask;
  pp.lj_single_type = __builtin_popcount(lj_type_mask) == 1 ? (uint32_t)__builtin_ctz(lj_type_mask)
                                                            : n_types;
  pp.rc2 = rc2;
  pp.lj_rs2 = lj_rs2;
  pp.lj_inv_delta = lj_inv_delta;
  pp.alpha = alpha;
  pp.k_coulomb = k_coulomb;
  pp.coul_table = coul_table.ptr;
  pp.coul_u_lo = coul_u_lo;
  pp.coul_u_hi = coul_u_hi;
  pp.shift_margin = r_cut + 0.5f * r_skin;
  pp.acc = acc.ptr;
  pp.acc_stride = n_pad;
  // four waves per block, and never more waves than the list has entries
  // (small systems would otherwise launch mostly idle blocks)
  const uint32_t entry_blocks = ceil_div((uint64_t)n_clusters * cap_entries, 4);
  const uint32_t blocks =
      std::max(1u, std::min((uint32_t)c.num_cus * pair_blocks_per_cu, entry_blocks));
  KernelId k = lj_on && coul_on ? (fev ? K_hmd_pair_force_fev : K_hmd_pair_force_f)
               : lj_on ? (fev ? K_hmd_pair_force_lj_fev : K_hmd_pair_force_lj_f)
                       : (fev ? K_hmd_pair_force_coul_fev : K_hmd_pair_force_coul_f);
  pp.user_params = user_params.ptr;
  if (jit && c.rec) {
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&pp);
    c.rec->push_back(LaunchRec{k, pair_function(k), nullptr, blocks, 256, 0,
                               std::vector<unsigned char>(b, b + sizeof pp)});
  } else if (jit) {
    void* args[] = {&pp};
    check(hipModuleLaunchKernel(pair_function(k), blocks, 1, 1, 256, 1, 1, 0, c.stream, args,
                                nullptr),
          "composed pair kernel");
  } else {
    launch(c, k, blocks, 256, 0, pp);
  }
  if (n_scaled) {
    PairScaledParams q{};
    q.posq = p.posq.ptr;
    q.type = p.type.ptr;
    q.slot_of = slot_of.ptr;
    q.pair_i = sc_i.ptr;
    q.pair_j = sc_j.ptr;
    q.s_lj = sc_lj.ptr;
    q.s_coul = sc_coul.ptr;
    q.n_pairs = n_scaled;
    q.n_types = n_types;
    q.L = p.lattice;
    q.lj_c6c12 = lj_table.ptr;
    q.rc2 = rc2;
    q.lj_rs2 = lj_rs2;
    q.lj_inv_delta = lj_inv_delta;
    q.k_coulomb = k_coulomb;
    q.coul_table = coul_table.ptr;
    q.coul_u_lo = coul_u_lo;
    q.coul_u_hi = coul_u_hi;
    q.lj_on = lj_on ? 1u : 0u;
    q.coul_on = coul_on ? 1u : 0u;
    q.fev = fev ? 1u : 0u;
    q.acc = acc.ptr;
    q.acc_stride = n_pad;
    q.s_user = jit && sc_user.count == n_scaled ? sc_user.ptr : nullptr;
    q.user_params = user_params.ptr;
    const uint32_t sb = ceil_div(n_scaled, 256);
    if (jit && c.rec) {
      const unsigned char* b = reinterpret_cast<const unsigned char*>(&q);
      c.rec->push_back(LaunchRec{K_hmd_pair_scaled, pair_function(K_hmd_pair_scaled), nullptr,
                                 sb, 256, 0, std::vector<unsigned char>(b, b + sizeof q)});
    } else if (jit) {
      void* args[] = {&q};
      check(hipModuleLaunchKernel(pair_function(K_hmd_pair_scaled), sb, 1, 1, 256, 1, 1, 0,
                                  c.stream, args, nullptr),
            "composed scaled-pair kernel");
    } else {
      launch(c, K_hmd_pair_scaled, sb, 256, 0, q);
    }
  }
  c.mark("pair_force");
}

void ForceField::spme_forces(bool fev) {
  const HmdLattice L = p.lattice;
  const uint32_t n0 = grid[0], n1 = grid[1], n2 = grid[2];
  SpmeBinParams bp{p.posq.ptr, sorted_id.ptr, n, L, n0, n1, n2, brick[0], brick[1], brick[2],
                   nb[0], nb[1], nb[2], atom_bin.ptr, bin_count.ptr, bin_start.ptr,
                   bin_atoms.ptr};
  launch(c, K_hmd_spme_bin, ceil_div(n, 256), 256, 0, bp);
  bin_scan.run(c, bin_count.ptr, bin_start.ptr, nullptr, 0);
  launch(c, K_hmd_spme_bin_scatter, ceil_div(n, 256), 256, 0, bp);
  c.mark("spme_bin");

  SpmeSpreadParams sp{p.posq.ptr, n, L, n0, n1, n2, brick[0], brick[1], brick[2],
                      nb[0], nb[1], nb[2], bin_start.ptr, bin_atoms.ptr, rho.ptr};
  const size_t sh_spread = (size_t)brick[0] * brick[1] * brick[2] * 8;
  KernelId ks = order == 4 ? K_hmd_spme_spread_p4
                : order == 5 ? K_hmd_spme_spread_p5
                : order == 6 ? K_hmd_spme_spread_p6 : K_hmd_spme_spread_p8;
  launch(c, ks, n_bins, 256, sh_spread, sp);
  c.mark("spme_spread");

  FftParams f{};
  f.L = L;
  f.alpha = alpha;
  f.k_coulomb = k_coulomb;
  f.n0 = n0; f.n1 = n1; f.n2 = n2; f.nch = nch;
  f.bfac0 = bfac[0].ptr; f.bfac1 = bfac[1].ptr; f.bfac2 = bfac[2].ptr;
  f.gv = gv.ptr;
  auto set_radix = [&](const std::vector<uint32_t>& r) {
    f.n_stages = (uint32_t)r.size();
    for (size_t k = 0; k < r.size(); ++k) f.radix[k] = r[k];
  };
  if (slab.on) {
    // slab path: c-axis R2C + b forward per a; a forward * G * a inverse per
    // b; b inverse + c-axis C2R per a
    FftParams q = f;
    q.twiddle = tw_c.ptr;
    q.twiddle_r2c = tw_r2c.ptr;
    q.twiddle2 = tw_b.ptr;
    q.data = rho.ptr;
    q.out = mesh_c.ptr;
    launch(c, slab.fwd, n0, 1024, slab.lds_fwd, q);
    FftParams qa = f;
    qa.data = mesh_c.ptr;
    qa.twiddle = tw_a.ptr;
    qa.partials = partials.ptr;
    launch(c, slab.solve, n1, 1024, slab.lds_solve, qa);
    q.data = mesh_c.ptr;
    q.out = V.ptr;
    launch(c, slab.inv, n0, 1024, slab.lds_fwd, q);
  } else {
  // 1. c axis, R2C
  {
    FftParams q = f;
    q.data = rho.ptr;
    q.out = mesh_c.ptr;
    q.n = n2 / 2;
    q.n_lines = n0 * n1;
    q.lines_per_block = lines_per_block(q.n);
    q.twiddle = tw_c.ptr;
    q.twiddle_r2c = tw_r2c.ptr;
    q.sign = -1;
    FftParams& r = q;
    r.n_stages = (uint32_t)rad_c.size();
    for (size_t k = 0; k < rad_c.size(); ++k) r.radix[k] = rad_c[k];
    if (pass_r2c.p2) {
      q.lines_per_block = pass_r2c.W;
      launch(c, pass_r2c.k, pass_r2c.grid, 256, pass_r2c.shmem, q);
    } else {
      launch(c, K_hmd_fft_r2c_c, ceil_div(q.n_lines, q.lines_per_block), 256,
             2 * (size_t)q.lines_per_block * (q.n + 1) * 8, q);
    }
  }
  // 2-4. b forward, a forward * G * a inverse, b inverse
  FftParams fb = f;
  fb.data = mesh_c.ptr;
  fb.n = n1;
  fb.n_lines = n0 * nch;
  fb.lines_per_block = lines_per_block(n1);
  fb.elem_stride = nch;
  fb.line_minor = nch;
  fb.major_stride = (uint64_t)n1 * nch;
  fb.minor_stride = 1;
  fb.twiddle = tw_b.ptr;
  fb.sign = -1;
  fb.n_stages = (uint32_t)rad_b.size();
  for (size_t k = 0; k < rad_b.size(); ++k) fb.radix[k] = rad_b[k];
  const size_t sh_b = 2 * (size_t)fb.lines_per_block * (n1 + 1) * 8;
  auto launch_b = [&]() {
    if (pass_b.p2) {
      FftParams q = fb;
      q.lines_per_block = pass_b.W;
      launch(c, pass_b.k, pass_b.grid, 256, pass_b.shmem, q);
    } else {
      launch(c, K_hmd_fft_axis, ceil_div(fb.n_lines, fb.lines_per_block), 256, sh_b, fb);
    }
  };
  launch_b();
  FftParams fa = f;
  fa.data = mesh_c.ptr;
  fa.n = n0;
  fa.n_lines = n1 * nch;
  fa.lines_per_block = lines_per_block(n0);
  fa.elem_stride = n1 * nch;
  fa.line_minor = nch;
  fa.major_stride = nch;
  fa.minor_stride = 1;
  fa.twiddle = tw_a.ptr;
  fa.sign = -1;
  fa.apply_influence = 1;
  fa.partials = partials.ptr;
  set_radix(rad_a);
  fa.n_stages = f.n_stages;
  for (int k = 0; k < 8; ++k) fa.radix[k] = f.radix[k];
  if (pass_a.p2) {
    fa.lines_per_block = pass_a.W;
    launch(c, pass_a.k, pass_a.grid, 256, pass_a.shmem, fa);   // data + (G, vf) tiles
  } else {
    launch(c, K_hmd_fft_axis, ceil_div(fa.n_lines, fa.lines_per_block), 256,
           2 * (size_t)fa.lines_per_block * (n0 + 1) * 8, fa);
  }
  fb.sign = +1;
  launch_b();
  // 5. c axis, C2R
  {
    FftParams q = f;
    q.data = mesh_c.ptr;
    q.out = V.ptr;
    q.n = n2 / 2;
    q.n_lines = n0 * n1;
    q.lines_per_block = lines_per_block(q.n);
    q.twiddle = tw_c.ptr;
    q.twiddle_r2c = tw_r2c.ptr;
    q.sign = +1;
    q.n_stages = (uint32_t)rad_c.size();
    for (size_t k = 0; k < rad_c.size(); ++k) q.radix[k] = rad_c[k];
    if (pass_c2r.p2) {
      q.lines_per_block = pass_c2r.W;
      launch(c, pass_c2r.k, pass_c2r.grid, 256, pass_c2r.shmem, q);
    } else {
      launch(c, K_hmd_fft_c2r_c, ceil_div(q.n_lines, q.lines_per_block), 256,
             2 * (size_t)q.lines_per_block * (q.n + 1) * 8, q);
    }
  }
  }
  c.mark("spme_fft_solve");
  SpmeReduceParams rp{partials.ptr, partial_blocks, spme_out.ptr};
  launch(c, K_hmd_spme_reduce, 1, 256, 0, rp);

  SpmeGatherParams gp{p.posq.ptr, n, L, n0, n1, n2, brick[0], brick[1], brick[2],
                      nb[0], nb[1], nb[2], bin_start.ptr, bin_atoms.ptr, V.ptr,
                      u_self.ptr, spme_out.ptr, slow_f.ptr, slow_e.ptr, slow_w.ptr,
                      fev ? 1u : 0u};
  const size_t sh_gather = (size_t)(brick[0] + order - 1) * (brick[1] + order - 1) *
                           (brick[2] + order - 1) * 4;
  KernelId kg = order == 4 ? K_hmd_spme_gather_p4
                : order == 5 ? K_hmd_spme_gather_p5
                : order == 6 ? K_hmd_spme_gather_p6 : K_hmd_spme_gather_p8;
  launch(c, kg, n_bins, 256, sh_gather, gp);
  c.mark("spme_gather");
}

void ForceField::compute(int level, bool displacement_checked) {
  const bool fev = level != 0;
  PairFinalizeParams fp{};
  // SPME on a second stream: forked once the sorted positions are written
  // (it reads nothing else of the list), so that on a rebuild step it runs
  // beside the list build, and on every step beside the pair kernel.  The
  // pair kernel is VALU-bound, the list build latency-bound, the spread
  // LDS-bound and the FFT passes latency-bound, so they share the CUs.  The
  // two sides write disjoint buffers (list and fixed-point pair
  // accumulators; slow-class forces, energies and mesh), which the finalize
  // combines after the join: the result is the same bytes either way.  Not
  // while recording a batched replica step or marking per-stage events.
  const bool fork = spme_on && spme_stream && side && !c.rec && !c.profiling;
  auto fork_spme = [&]() {
    check(hipEventRecord(ev_fork, c.stream), "hipEventRecord");
    check(hipStreamWaitEvent(side, ev_fork, 0), "hipStreamWaitEvent");
    std::swap(c.stream, side);
    try {
      spme_forces(fev);
    } catch (...) {
      std::swap(c.stream, side);
      throw;
    }
    std::swap(c.stream, side);
    check(hipEventRecord(ev_join, side), "hipEventRecord");
  };
  if (fork) neighbor_pre_step(fork_spme, displacement_checked);
  else neighbor_pre_step({}, displacement_checked);
  pair_forces(fev);
  bonded_forces(fev);
  if (spme_on && !fork) spme_forces(fev);
  if (fork) check(hipStreamWaitEvent(c.stream, ev_join, 0), "hipStreamWaitEvent");
  if (mlp) mlp_correct(*mlp, *this);
  fp.acc = acc.ptr;
  fp.acc_stride = n_pad;
  fp.slot_of = slot_of.ptr;
  fp.slow_f = spme_on ? slow_f.ptr : nullptr;
  fp.slow_e = spme_on ? slow_e.ptr : nullptr;
  fp.slow_w = spme_on ? slow_w.ptr : nullptr;
  fp.f = p.force.ptr;
  fp.e = p.energy.ptr;
  fp.w = p.virial.ptr;
  fp.n = n;
  fp.fev = fev ? 1u : 0u;
  fp.acc_clear = acc.ptr;
  launch(c, K_hmd_pair_finalize, ceil_div(n, 256), 256, 0, fp);
  c.mark("finalize");
}

// Listed pairs with a non-zero scale for any term (taken from i's list for
// i < j), added back after the fused kernel by hmd_pair_scaled.
void ForceField::build_scaled_list(hipStream_t s) {
  std::vector<uint32_t> si, sj;
  std::vector<float> slj, sco, sus;
  for (uint32_t i = 0; i + 1 < (uint32_t)h_eoff.size(); ++i)
    for (uint32_t m = h_eoff[i]; m < h_eoff[i + 1]; ++m) {
      const float a = h_slj[m], b = h_sco[m], u = h_suser.empty() ? 0.0f : h_suser[m];
      if (h_epar[m] <= i || (a == 0.0f && b == 0.0f && u == 0.0f)) continue;
      si.push_back(i);
      sj.push_back(h_epar[m]);
      slj.push_back(a);
      sco.push_back(b);
      sus.push_back(u);
    }
  n_scaled = (uint32_t)si.size();
  if (n_scaled) {
    sc_i.alloc(n_scaled, false);
    sc_j.alloc(n_scaled, false);
    sc_lj.alloc(n_scaled, false);
    sc_coul.alloc(n_scaled, false);
    sc_i.upload(si.data(), n_scaled, s);
    sc_j.upload(sj.data(), n_scaled, s);
    sc_lj.upload(slj.data(), n_scaled, s);
    sc_coul.upload(sco.data(), n_scaled, s);
    if (!h_suser.empty()) {
      sc_user.alloc(n_scaled, false);
      sc_user.upload(sus.data(), n_scaled, s);
    }
  }
  graph_invalid = true;
}

// The composed user term's exclusion scale for each listed partner (CSR
// entry m of the force field's exclusion lists): the per-fragment
// exclusion_scale(i, j) of src/forces/lj.rs:451 and jit_composed.rs:939-972.
// Without it a composed term has scale 0 for listed partners.
void ForceField::set_fragment_exclusion_scales(const float* scales, uint32_t n_entries) {
  if (!jit) throw Error(HMD_ERR_INVALID, "no composed pair term on this force field");
  const uint32_t want = h_eoff.empty() ? 0u : h_eoff.back();
  if (n_entries != want)
    throw Error(HMD_ERR_INVALID, "expected one scale per listed exclusion partner (" +
                                     std::to_string(want) + ")");
  for (uint32_t m = 0; m < n_entries; ++m)
    if (!std::isfinite(scales[m])) throw Error(HMD_ERR_INVALID, "exclusion scale not finite");
  h_suser.assign(scales, scales + n_entries);
  build_scaled_list(c.stream);
  check(hipStreamSynchronize(c.stream), "sync");
}

// Bonded terms (src/forces/harmonic_bond.rs, morse.rs, angle.rs,
// dihedral.rs; validate_params of each): terms [n][arity + 1] (canonical
// atoms, then the type), params [n_types][3].  The step graphs are
// re-captured.
void ForceField::add_bonded(uint32_t kind, uint32_t n_terms, const uint32_t* terms,
                            uint32_t n_types, const double* params) {
  uint32_t arity = 0;
  switch (kind) {
    case HMD_BONDED_HARMONIC_BOND: case HMD_BONDED_MORSE_BOND: arity = 2; break;
    case HMD_BONDED_HARMONIC_ANGLE: arity = 3; break;
    case HMD_BONDED_PERIODIC_DIHEDRAL: arity = 4; break;
    default: throw Error(HMD_ERR_INVALID, "unknown bonded kind");
  }
  if (n_terms && (!terms || !n_types || !params)) throw Error(HMD_ERR_INVALID, "null argument");
  const double pi = 3.14159265358979323846;
  std::vector<float> ty(4 * (size_t)n_types, 0.0f);
  for (uint32_t t = 0; t < n_types; ++t) {
    const double* q = params + 3 * (size_t)t;
    auto pos = [](double v) { return std::isfinite(v) && v > 0.0; };
    bool ok = true;
    if (kind == HMD_BONDED_HARMONIC_BOND) ok = pos(q[0]) && pos(q[1]);
    else if (kind == HMD_BONDED_MORSE_BOND) ok = pos(q[0]) && pos(q[1]) && pos(q[2]);
    else if (kind == HMD_BONDED_HARMONIC_ANGLE)
      ok = pos(q[0]) && std::isfinite(q[1]) && q[1] >= 0.0 && q[1] <= pi;
    else
      ok = std::isfinite(q[0]) && std::isfinite(q[1]) && q[2] >= 0.0 && q[2] <= 64.0 &&
           q[2] == std::floor(q[2]);
    if (!ok) throw Error(HMD_ERR_INVALID, "bonded type " + std::to_string(t) + ": invalid parameters");
    for (int k = 0; k < 3; ++k) ty[4 * (size_t)t + k] = (float)q[k];
  }
  const uint32_t stride = arity + 1;
  for (uint32_t m = 0; m < n_terms; ++m) {
    const uint32_t* q = terms + (size_t)stride * m;
    for (uint32_t a = 0; a < arity; ++a) {
      if (q[a] >= n) throw Error(HMD_ERR_INVALID, "bonded term atom out of range");
      for (uint32_t b = 0; b < a; ++b)
        if (q[a] == q[b]) throw Error(HMD_ERR_INVALID, "bonded term repeats an atom");
    }
    if (q[arity] >= n_types) throw Error(HMD_ERR_INVALID, "bonded term type out of range");
  }
  auto set = std::make_unique<BondedSet>();
  set->kind = kind;
  set->n_terms = n_terms;
  if (n_terms) {
    set->terms.alloc((size_t)stride * n_terms, false);
    set->terms.upload(terms, (size_t)stride * n_terms, c.stream);
    set->types.alloc(ty.size(), false);
    set->types.upload(ty.data(), ty.size(), c.stream);
  }
  bonded.push_back(std::move(set));
  graph_invalid = true;
}

// Every bonded set adds into the pair accumulators (enqueue only).
void ForceField::bonded_forces(bool fev) {
  for (const auto& b : bonded) {
    if (!b->n_terms) continue;
    BondedParams q{};
    q.posq = p.posq.ptr;
    q.terms = b->terms.ptr;
    q.types = reinterpret_cast<const float4*>(b->types.ptr);
    q.slot_of = slot_of.ptr;
    q.acc = acc.ptr;
    q.acc_stride = n_pad;
    q.L = p.lattice;
    q.n_terms = b->n_terms;
    q.kind = b->kind;
    q.fev = fev ? 1u : 0u;
    launch(c, K_hmd_bonded_terms, ceil_div(b->n_terms, 256), 256, 0, q);
  }
  if (!bonded.empty()) c.mark("bonded");
}

// Host-side status check at a batch boundary: the analogue of the 4-byte
// neighbour status read in pre_step (neighbor_list.rs:1308-1320).  The
// high-water word holds the most entries any cluster needed since the
// capacity was last set.  An overflow (entries dropped) grows the capacity and is reported:
// the batch is to be re-run.  A build past the fill threshold grows the
// capacity ahead of any dropped entry (neighbor_list.rs:401-413): nothing
// was lost, so the batch stands.  Either way the list is rebuilt at once
// from the current positions into the new buffers, so that it is whole for
// any reader before the next step, and step graphs are re-captured.
void ForceField::check_status() {
  uint32_t st[4];
  status.download(st, 4, c.stream);
  const uint32_t need = st[2];
  const bool overflow = (st[0] & HMD_STATUS_ENTRY_OVERFLOW) != 0;
  const bool near_full = need > (uint32_t)((double)cap_entries * kNblFillThreshold) &&
                         cap_entries < n_clusters;
  if (!overflow && !near_full) return;
  const uint32_t old_cap = cap_entries;
  alloc_entries(std::max(grown_capacity(need), cap_entries));
  uint32_t reset[4] = {HMD_STATUS_DISPLACED, st[1], 0u, 0u};
  status.upload(reset, 4, c.stream);
  rebuild_chain();
  check(hipStreamSynchronize(c.stream), "sync");
  graph_invalid = true;
  if (overflow)
    throw Error(HMD_ERR_NBL_OVERFLOW,
                "packed neighbour list overflowed: a cluster needed " + std::to_string(need) +
                    " entries of 64 j-atoms (capacity " + std::to_string(old_cap) +
                    "); capacity grown to " + std::to_string(cap_entries) + ", re-run the batch");
}

}  // namespace hmd

using namespace hmd;

extern "C" {

hmd_status hmd_forcefield_create(hmd_ctx* h, hmd_particles* hp,
                                 const hmd_forcefield_desc* d, hmd_forcefield** out) {
  HMD_GUARD({
    if (!d) throw Error(HMD_ERR_INVALID, "null descriptor");
    if (!hp || hp->p.n == 0) throw Error(HMD_ERR_INVALID, "force field: empty particle set");
    auto* f = new hmd_forcefield();
    try {
      f->ff = new ForceField(h->c, hp->p, *d);
    } catch (...) {
      delete f;
      throw;
    }
    *out = f;
  });
}

void hmd_forcefield_destroy(hmd_forcefield* f) {
  if (!f) return;
  delete f->ff;
  delete f;
}

hmd_status hmd_forcefield_compute(hmd_forcefield* f, int32_t level) {
  HMD_GUARD({
    f->ff->compute(level);
    check(hipStreamSynchronize(f->ff->c.stream), "sync");
    f->ff->check_status();
  });
}

// ---- slot level: the reference's Potential::compute per slot, writing +=
// into caller-owned per-atom outputs (SlotOutputView), enqueue only.
static void check_slot_output(const ForceField& F, int32_t level, const hmd_slot_output* o) {
  if (!o || !o->force_x || !o->force_y || !o->force_z)
    throw Error(HMD_ERR_INVALID, "slot output: null force arrays");
  (void)F;
  (void)level;
}

hmd_status hmd_nbl_pre_step(hmd_forcefield* f) {
  HMD_GUARD({
    if (!f) throw Error(HMD_ERR_INVALID, "null force field");
    f->ff->neighbor_pre_step();
  });
}

hmd_status hmd_pair_force(hmd_forcefield* f, int32_t level, const hmd_slot_output* out) {
  HMD_GUARD({
    if (!f) throw Error(HMD_ERR_INVALID, "null force field");
    ForceField& F = *f->ff;
    check_slot_output(F, level, out);
    const bool fev = level != 0;
    F.pair_forces(fev);
    SlotAddParams q{};
    q.acc = F.acc.ptr;
    q.acc_stride = F.n_pad;
    q.slot_of = F.slot_of.ptr;
    q.acc_clear = F.acc.ptr;
    q.out_x = out->force_x; q.out_y = out->force_y; q.out_z = out->force_z;
    q.out_e = out->energy; q.out_w = out->virial;
    q.n = F.n;
    q.fev = fev ? 1u : 0u;
    launch(F.c, K_hmd_slot_add, ceil_div(F.n, 256), 256, 0, q);
  });
}

hmd_status hmd_forcefield_set_fragment_exclusion_scales(hmd_forcefield* f, const float* scales,
                                                        uint32_t n_entries) {
  HMD_GUARD({
    if (!f || (n_entries && !scales)) throw Error(HMD_ERR_INVALID, "null argument");
    f->ff->set_fragment_exclusion_scales(scales, n_entries);
  });
}

hmd_status hmd_forcefield_add_bonded(hmd_forcefield* f, const hmd_bonded_desc* d) {
  HMD_GUARD({
    if (!f || !d) throw Error(HMD_ERR_INVALID, "null argument");
    f->ff->add_bonded(d->kind, d->n_terms, d->terms, d->n_types, d->params);
  });
}

hmd_status hmd_bonded_compute(hmd_forcefield* f, int32_t level, const hmd_slot_output* out) {
  HMD_GUARD({
    if (!f) throw Error(HMD_ERR_INVALID, "null force field");
    ForceField& F = *f->ff;
    check_slot_output(F, level, out);
    const bool fev = level != 0;
    F.bonded_forces(fev);
    SlotAddParams q{};
    q.acc = F.acc.ptr;
    q.acc_stride = F.n_pad;
    q.slot_of = F.slot_of.ptr;
    q.acc_clear = F.acc.ptr;
    q.out_x = out->force_x; q.out_y = out->force_y; q.out_z = out->force_z;
    q.out_e = out->energy; q.out_w = out->virial;
    q.n = F.n;
    q.fev = fev ? 1u : 0u;
    launch(F.c, K_hmd_slot_add, ceil_div(F.n, 256), 256, 0, q);
  });
}

hmd_status hmd_spme_compute(hmd_forcefield* f, int32_t level, const hmd_slot_output* out) {
  HMD_GUARD({
    if (!f) throw Error(HMD_ERR_INVALID, "null force field");
    ForceField& F = *f->ff;
    if (!F.spme_on) throw Error(HMD_ERR_INVALID, "no spme-reciprocal slot");
    check_slot_output(F, level, out);
    const bool fev = level != 0;
    F.spme_forces(fev);
    if (F.mlp) mlp_correct(*F.mlp, F);
    SlotAddParams q{};
    q.src_f = F.slow_f.ptr; q.src_e = F.slow_e.ptr; q.src_w = F.slow_w.ptr;
    q.out_x = out->force_x; q.out_y = out->force_y; q.out_z = out->force_z;
    q.out_e = out->energy; q.out_w = out->virial;
    q.n = F.n;
    q.fev = fev ? 1u : 0u;
    launch(F.c, K_hmd_slot_add, ceil_div(F.n, 256), 256, 0, q);
  });
}

hmd_status hmd_forcefield_check(hmd_forcefield* f) {
  HMD_GUARD({
    if (!f) throw Error(HMD_ERR_INVALID, "null force field");
    check(hipStreamSynchronize(f->ff->c.stream), "sync");
    f->ff->check_status();
  });
}

hmd_status hmd_forcefield_slots(hmd_forcefield* f, char* buf, int32_t len) {
  HMD_GUARD({
    std::string s;
    for (auto& l : f->ff->slots) s += l + "\n";
    if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", s.c_str());
  });
}

hmd_status hmd_spme_scalars(hmd_forcefield* f, double out[3]) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    if (!F.spme_on) throw Error(HMD_ERR_INVALID, "no spme-reciprocal slot");
    double v[2];
    F.spme_out.download(v, 2, F.c.stream);
    out[0] = v[0];
    out[1] = v[1];
    out[2] = F.u_self_total;
  });
}

hmd_status hmd_spme_grids(hmd_forcefield* f, float* rho, float* pot) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    if (!F.spme_on) throw Error(HMD_ERR_INVALID, "no spme-reciprocal slot");
    if (rho) F.rho.download(rho, F.M, F.c.stream);
    if (pot) F.V.download(pot, F.M, F.c.stream);
  });
}

hmd_status hmd_spme_forces(hmd_forcefield* f, float* force) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    if (!F.spme_on) throw Error(HMD_ERR_INVALID, "no spme-reciprocal slot");
    std::vector<float> soa(3 * (size_t)F.n);
    F.slow_f.download(soa.data(), soa.size(), F.c.stream);
    for (uint32_t i = 0; i < F.n; ++i)
      for (int d = 0; d < 3; ++d) force[3 * (size_t)i + d] = soa[(size_t)d * F.n + i];
  });
}

hmd_status hmd_nbl_request_rebuild(hmd_forcefield* f) {
  HMD_GUARD(f->ff->request_rebuild());
}

hmd_status hmd_nbl_pair_snapshot(hmd_forcefield* f, uint32_t* pairs, uint64_t cap,
                                 uint64_t* n_pairs) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    Context& c = F.c;
    DevBuf<uint32_t> dpairs;
    dpairs.alloc(std::max<uint64_t>(2, 2 * cap));
    DevBuf<unsigned long long> cnt;
    cnt.alloc(1);
    uint32_t nw = 0;
    F.n_work.download(&nw, 1, c.stream);
    NblPairsParams pp{F.posq_sorted.ptr, F.sorted_id.ptr, F.work.ptr, F.n_work.ptr,
                      F.entry_j.ptr, F.p.lattice, F.n, F.cap_entries, F.r_search_sq,
                      dpairs.ptr, cnt.ptr, cap, nullptr};
    launch(c, K_hmd_nbl_pairs, ceil_div(nw, 4), 256, 0, pp);
    unsigned long long total = 0;
    cnt.download(&total, 1, c.stream);
    *n_pairs = total;
    uint64_t k = std::min<uint64_t>(total, cap);
    if (k) dpairs.download(pairs, 2 * k, c.stream);
  });
}

hmd_status hmd_nbl_count_pairs(hmd_forcefield* f, double radius, uint64_t* n_pairs) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    Context& c = F.c;
    const float r = (float)radius;
    if (!(r > 0.0f) || r > F.r_search) throw Error(HMD_ERR_INVALID, "radius must be in (0, r_search]");
    DevBuf<unsigned long long> cnt;
    cnt.alloc(1);
    uint32_t nw = 0;
    F.n_work.download(&nw, 1, c.stream);
    NblPairsParams pp{F.posq_sorted.ptr, F.sorted_id.ptr, F.work.ptr, F.n_work.ptr,
                      F.entry_j.ptr, F.p.lattice, F.n, F.cap_entries, r * r,
                      nullptr, cnt.ptr, 0ull, nullptr};
    launch(c, K_hmd_nbl_pairs, ceil_div(nw, 4), 256, 0, pp);
    unsigned long long total = 0;
    cnt.download(&total, 1, c.stream);
    *n_pairs = total;
  });
}

hmd_status hmd_nbl_set_capacity(hmd_forcefield* f, uint32_t entries, uint32_t lds_entries) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    if (entries == 0) throw Error(HMD_ERR_INVALID, "capacity must be at least one entry");
    F.lds_entries = lds_entries;
    F.alloc_entries(entries);
    uint32_t st[4];
    F.status.download(st, 4, F.c.stream);
    uint32_t reset[4] = {st[0] | HMD_STATUS_DISPLACED, st[1], 0u, 0u};
    F.status.upload(reset, 4, F.c.stream);
    check(hipStreamSynchronize(F.c.stream), "sync");
    F.graph_invalid = true;
  });
}

hmd_status hmd_nbl_stats(hmd_forcefield* f, uint32_t out[8]) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    uint32_t st[4], nw = 0;
    F.status.download(st, 4, F.c.stream);
    F.n_work.download(&nw, 1, F.c.stream);
    out[0] = nw;
    out[1] = st[1];
    out[2] = st[0];
    out[3] = st[2];
    out[4] = F.cap_entries;
    out[5] = F.n_clusters;
    out[6] = F.g[0];
    // i-sub-cluster visits of the pair kernel (8 steps of 64 pairs each):
    // the listed entries' sub-cluster masks, summed
    std::vector<uint32_t> cnt(F.n_clusters), mask((size_t)F.n_clusters * F.cap_entries);
    F.entry_count.download(cnt.data(), cnt.size(), F.c.stream);
    F.entry_mask.download(mask.data(), mask.size(), F.c.stream);
    uint64_t visits = 0;
    for (uint32_t c = 0; c < F.n_clusters; ++c)
      for (uint32_t e = 0; e < cnt[c] && e < F.cap_entries; ++e)
        visits += (uint64_t)__builtin_popcount(mask[(size_t)c * F.cap_entries + e] & 0xFFu);
    out[7] = (uint32_t)std::min<uint64_t>(visits, 0xFFFFFFFFu);
  });
}

}  // extern "C"
