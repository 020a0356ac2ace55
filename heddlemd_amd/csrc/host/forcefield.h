// ForceField host object: potential slots, neighbour list, SPME mesh.
#pragma once
#include <functional>
#include <memory>

#include "hmd_internal.h"

namespace hmd {

struct Mlp;

// Pair kernels recompiled at run time with a user pair term (compose.cpp).
struct JitPairModule {
  hipModule_t module = nullptr;
  KernelId kernel[7] = {};
  hipFunction_t fn[7] = {};
  JitPairModule() = default;
  JitPairModule(const JitPairModule&) = delete;
  JitPairModule& operator=(const JitPairModule&) = delete;
  ~JitPairModule();
};

// LDS of the list build: 64 KiB per workgroup, less its static arrays (at
// most 24 KiB) and its candidate-mask arrays (20 bytes per 64 clusters); the
// rest stages the head of a cluster's j list (list positions past it are
// written to the entry array directly).
constexpr size_t kNblBuildLds = 65536;
constexpr size_t kNblBuildStaticLds = 24576;
// Capacity policy of the packed list (src/forces/neighbor_list.rs:235-244,
// 415-445): grow by 1.5x, and grow ahead of any dropped entry once a build
// fills more than 0.8 of the capacity.
constexpr double kNblGrowth = 1.5;
constexpr double kNblFillThreshold = 0.8;

struct ForceField {
  Context& c;
  Particles& p;
  std::vector<std::string> slots;   // labels in evaluation order
  uint32_t n = 0;

  // ---- fused pair slot
  bool lj_on = false, coul_on = false, spme_on = false;
  uint32_t n_types = 1;
  uint32_t lj_type_mask = 0;
  float r_cut = 0, rc2 = 0, lj_rs2 = 0, lj_inv_delta = 0;
  float alpha = 0, k_coulomb = 1;
  DevBuf<float> lj_table;
  DevBuf<float> coul_table;    // float4[2][HMD_COUL_BINS]
  uint32_t coul_u_lo = 0, coul_u_hi = 0;   // table bounds as f32 bits
  void build_coulomb_table(double alpha, double r_cut, hipStream_t s);
  DevBuf<uint32_t> excl_off, excl_par;
  DevBuf<uint32_t> excl_slot;   // [8][n_pad] partners' sorted slots (each rebuild)
  // listed pairs with non-zero exclusion scales (i < j)
  uint32_t n_scaled = 0;
  DevBuf<uint32_t> sc_i, sc_j;
  DevBuf<float> sc_lj, sc_coul, sc_user;
  // the exclusion lists and their scales on the host (CSR, one entry per
  // listed partner), to rebuild the scaled list when a composed term brings
  // scales of its own
  std::vector<uint32_t> h_eoff, h_epar;
  std::vector<float> h_slj, h_sco, h_suser;
  void build_scaled_list(hipStream_t s);
  void set_fragment_exclusion_scales(const float* scales, uint32_t n_entries);
#ifndef HMD_PAIR_BLOCKS_PER_CU
#define HMD_PAIR_BLOCKS_PER_CU 32
#endif
  uint32_t pair_blocks_per_cu = HMD_PAIR_BLOCKS_PER_CU;
// Workgroups per CU the FFT passes aim for (several, so that one group's
// loads overlap another's transform).
#ifndef HMD_FFT_BLOCKS_PER_CU
#define HMD_FFT_BLOCKS_PER_CU 3
#endif
  uint32_t fft_blocks_per_cu = HMD_FFT_BLOCKS_PER_CU;
  // the slab FFT path where the mesh allows it (off by default: faster alone,
  // but its one-workgroup-per-CU kernels co-run worse beside the pair kernel)
  bool fft_slab_ok = false;
  // SPME enqueued on a second stream beside the pair kernel (fork / join)
  uint32_t spme_stream = 1;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void make_side_stream();
  // Launch-shape autotuner: times candidate pair grids, FFT grids and SPME
  // bricks on this system and keeps the fastest; set_tuning applies a
  // recorded choice (so that a run can repeat another's bytes).
  void autotune(uint32_t n_evals, hmd_tuning* out);
  void set_tuning(const hmd_tuning& t);
  void get_tuning(hmd_tuning* t) const;
  double time_compute(uint32_t n_evals);
  void check_brick(const uint32_t b[3]) const;
  void alloc_bins();
  void plan_fft();
  // run-time composed pair term (null: the built-in kernels)
  std::unique_ptr<JitPairModule> jit;
  DevBuf<float> user_params;
  void set_pair_fragment(const std::string& name, const std::string& source,
                         const float* params, uint32_t n_params);
  hipFunction_t pair_function(KernelId k) const;

  // ---- neighbour list
  float r_skin = 0, r_search = 0, r_search_sq = 0;
  bool orthorhombic = true, ortho_fast = false;
  uint32_t g[3] = {1, 1, 1};
  uint32_t grid_bits = 1;
  uint32_t n_keys = 1, n_clusters = 0, n_pad = 0, cap_entries = 0;
  uint32_t lds_entries = 0;         // list entries staged in LDS (0: as many as fit)
  DevBuf<uint32_t> status;          // [0] flags [1] generation [2] high water
  DevBuf<uint32_t> atom_key, cell_count, cell_start, sorted_id, slot_of, type_sorted;
  DevBuf<float> posq_sorted, ref_pos, bbox, sub_bbox;
  DevBuf<int32_t> image_sorted;   // int4[n_pad]
  DevBuf<uint32_t> entry_j, entry_mask, entry_count, entry_offset, work, n_work;
  DevBuf<uint64_t> entry_excl;
  DevBuf<uint32_t> excl_entries;   // [n_clusters]
  DevBuf<long long> acc;            // [5][n_pad] fixed-point accumulators
  Scanner cell_scan, entry_scan;
  bool graph_invalid = false;

  // ---- SPME reciprocal slot
  uint32_t order = 4, grid[3] = {0, 0, 0}, brick[3] = {0, 0, 0}, nb[3] = {0, 0, 0};
  uint32_t nch = 0, n_bins = 0, partial_blocks = 0;
  size_t M = 0;
  DevBuf<uint32_t> atom_bin, bin_count, bin_start, bin_atoms;
  Scanner bin_scan;
  DevBuf<float> rho, V, mesh_c, tw_a, tw_b, tw_c, tw_r2c, u_self;
  DevBuf<double> bfac[3];
  DevBuf<float> gv;            // float2 (G, vf) per complex mesh point (p2 a-axis pass)
  DevBuf<float> slow_f, slow_e, slow_w;
  DevBuf<double> partials, spme_out;
  std::vector<uint32_t> rad_a, rad_b, rad_c;
  // Launch shape of one FFT pass: compile-time power-of-two kernel (p2) or
  // the generic mixed-radix one.
  struct FftPass {
    bool p2 = false;
    KernelId k = K_COUNT;
    uint32_t W = 0, grid = 0;
    size_t shmem = 0;
  };
  FftPass pass_r2c, pass_c2r, pass_b, pass_a;
  // Slab path (128^3 and 64^3 meshes): three kernels in place of the five
  // passes when a (b, c) slab fits in one workgroup's LDS.
  struct FftSlab {
    bool on = false;
    KernelId fwd = K_COUNT, solve = K_COUNT, inv = K_COUNT;
    size_t lds_fwd = 0, lds_solve = 0;
  };
  FftSlab slab;
  double u_self_total = 0.0;

  // ---- learned correction of the reciprocal forces (optional)
  Mlp* mlp = nullptr;
  // bonded terms (one set per hmd_forcefield_add_bonded call)
  struct BondedSet {
    uint32_t kind = 0, n_terms = 0;
    DevBuf<uint32_t> terms;
    DevBuf<float> types;   // [n_types][4]
  };
  std::vector<std::unique_ptr<BondedSet>> bonded;
  void add_bonded(uint32_t kind, uint32_t n_terms, const uint32_t* terms, uint32_t n_types,
                  const double* params);
  void bonded_forces(bool fev);

  ForceField(Context& ctx, Particles& part, const hmd_forcefield_desc& d);

  void init_neighbor_list(const hmd_forcefield_desc& d);
  void init_spme(const hmd_forcefield_desc& d);
  void alloc_entries(uint32_t cap);
  uint32_t grown_capacity(uint32_t need) const;
  void probe_capacity();
  void request_rebuild();
  void restart_list();
  void check_box();
  void compute_gv_table();
  void on_box_change();
  ~ForceField();
  void rebuild_chain();
  void rebuild_sort();
  void rebuild_list();
  void neighbor_pre_step(const std::function<void()>& after_sort = {},
                         bool displacement_checked = false);
  NblDispParams disp_params() const;
  void pair_forces(bool fev);
  void spme_forces(bool fev);
  // One force evaluation, enqueued only (no host synchronisation): safe to
  // capture into a graph.
  // displacement_checked: the caller's kernel has already run the skin
  // check on the current positions (the fused first integrator kernel)
  void compute(int level, bool displacement_checked = false);
  void check_status();
};

// Adds the network's correction to the slow-class forces (enqueue only).
void mlp_correct(Mlp& m, ForceField& ff);

}  // namespace hmd

struct hmd_forcefield {
  hmd::ForceField* ff = nullptr;
};
