// Kernel parameter blocks shared by the host library and the gfx950 code
// object.  Every kernel takes exactly one of these structs by value, so the
// kernarg segment is the struct itself and the host and device agree on the
// argument layout by construction (this replaces the reference's positional
// argument tuples and their KernelArgSchema guard,
// src/forces/jit_composed.rs:457-639).
//
// Plain C layout only: pointers, 32/64-bit integers, float, double.
#pragma once
#ifndef __HIPCC_RTC__
#include <stdint.h>
#else   // hiprtc: its runtime header keeps the fixed-width types in a namespace
using __hip_internal::int32_t;
using __hip_internal::uint16_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

#define HMD_CLUSTER 64u             // atoms per cluster = lanes per wave
#define HMD_ENTRY 64u               // j-atom slots per neighbour entry
#define HMD_SENTINEL 0xFFFFFFFFu    // empty j slot
#define HMD_FIXED_SHIFT 40          // pair accumulators: value * 2^40 in i64
// real-space Ewald table: HMD_COUL_OCTAVES octaves of r^2 below the first
// power of two above r_cut^2, 2^HMD_COUL_BITS bins per octave
#define HMD_COUL_OCTAVES 8u
#define HMD_COUL_BITS 7u
#define HMD_COUL_BINS (HMD_COUL_OCTAVES << HMD_COUL_BITS)
#define HMD_SPREAD_SHIFT 32         // mesh accumulators: value * 2^32 in i64

// neighbour status word bits (kernels set them, the host reads them)
#define HMD_STATUS_DISPLACED 1u     // an atom moved > r_skin/2 since rebuild
#define HMD_STATUS_ENTRY_OVERFLOW 2u
#define HMD_STATUS_CELL_OVERFLOW 4u
#define HMD_STATUS_REBUILT 8u       // a rebuild ran since the host cleared it

// A batched launch (one launch over the blocks of several instances of a
// kernel): instance r's parameter struct sits at params + r * stride and its
// blocks are block_start[r] .. block_start[r + 1] - 1 (n + 1 entries).
struct HmdBatch {
  const unsigned char* params;
  const uint32_t* block_start;
  uint32_t n;
  uint32_t stride;
};

struct HmdLattice {
  float lx, ly, lz, xy, xz, yz;
};

// ---------------------------------------------------------------- scan ---
struct ScanParams {
  const uint32_t* in;       // may alias out
  uint32_t* out;
  uint32_t* block_sums;     // [num_blocks]
  uint32_t len;
  const uint32_t* gate;     // optional: skip when (*gate & gate_mask) == 0
  uint32_t gate_mask;
  uint32_t total_slot;      // write out[len] = total when != 0 (single block)
};

// ----------------------------------------------------- neighbour list ---
struct NblKeyParams {
  const float* posq;        // canonical float4[n]
  uint32_t n;
  HmdLattice L;
  uint32_t g0, g1, g2;      // fine cells per axis (2^bits each)
  uint32_t bits;
  uint32_t* atom_key;       // [n]
  uint32_t* cell_count;     // [n_keys]  (zero on entry, zero again after scatter)
  const uint32_t* cell_start;  // [n_keys+1] (scatter only)
  uint32_t* sorted_id;      // [n_pad]   (scatter only)
  const uint32_t* status;   // rebuild gate
};

struct NblSortParams {
  const uint32_t* cell_start;
  uint32_t n_keys;
  uint32_t* sorted_id;
  const uint32_t* status;
};

struct NblGatherParams {
  const float* posq;        // canonical float4[n]
  const uint32_t* type;     // canonical [n]
  uint32_t n, n_pad;
  const uint32_t* sorted_id;   // [n_pad] (pad entries written here)
  uint32_t* slot_of;        // [n]
  float* posq_sorted;       // float4[n_pad]
  uint32_t* type_sorted;    // [n_pad]
  float* ref_pos;           // float4[n] canonical positions at rebuild
  const uint32_t* status;   // rebuild extras gated on the DISPLACED bit
  HmdLattice L;
  int32_t* image;           // int4[n_pad]: lattice images removed at the rebuild
};

struct NblBboxParams {
  const float* posq_sorted;
  uint32_t n;
  uint32_t n_clusters;
  float* bbox;              // float4[2*n_clusters]: (cx,cy,cz,-), (hx,hy,hz,-)
  float* sub_bbox;          // float4[16*n_clusters]: 8 sub-cluster boxes
  const uint32_t* status;
  // exclusion partners as sorted slots, for the list build: slot s's m-th
  // partner (canonical CSR order, at most 8) at excl_slot[m * n_pad + s]
  const uint32_t* sorted_id;
  const uint32_t* slot_of;
  const uint32_t* excl_offsets;   // canonical CSR [n+1]
  const uint32_t* excl_partners;
  uint32_t* excl_slot;            // [8][n_pad], sentinel-filled
  uint32_t n_pad;
};

struct NblBuildParams {
  const float* posq_sorted;
  const uint32_t* sorted_id;
  const uint32_t* slot_of;
  const float* bbox;
  const float* sub_bbox;
  HmdLattice L;
  uint32_t n, n_clusters;
  float r_search_sq;
  float r_search;
  uint32_t orthorhombic_fast;  // 1: exact fast min-image valid (see kernel)
  uint32_t cap_entries;        // entries per cluster
  uint32_t lds_slots;          // list slots staged in LDS (later ones go to entry_j directly)
  uint32_t max_chunks;         // ceil(n_clusters / 64): LDS candidate-mask arrays
  uint32_t* entry_j;           // [n_clusters*cap_entries*64]
  uint32_t* entry_mask;        // [n_clusters*cap_entries] i-sub-clusters any j reaches
  uint64_t* entry_excl;        // [n_clusters*cap_entries*64]
  uint32_t* entry_count;       // [n_clusters] (entries written / needed)
  uint32_t* excl_entries;      // [n_clusters] leading entries with masks
  const uint32_t* excl_slot;   // [8][n_pad] partners' sorted slots (hmd_nbl_bbox)
  uint32_t n_pad;
  uint32_t* status;
};

struct NblWorkParams {
  const uint32_t* entry_count;   // [n_clusters]
  const uint32_t* entry_offset;  // [n_clusters+1] exclusive scan
  uint32_t n_clusters, cap_entries;
  uint32_t* work;                // [total]
  uint32_t* n_work;              // [1]
  uint32_t* status;              // [0] flags, [1] generation, [2] high water
};

struct NblDispParams {
  const float* posq;
  const float* ref_pos;
  uint32_t n;
  HmdLattice L;
  float threshold_sq;
  uint32_t* status;
};

struct NblPairsParams {        // decode the list into exact pair set
  const float* posq_sorted;
  const uint32_t* sorted_id;
  const uint32_t* work;
  const uint32_t* n_work;
  const uint32_t* entry_j;
  HmdLattice L;
  uint32_t n;
  uint32_t cap_entries;
  float r_search_sq;
  uint32_t* pairs;             // [cap*2] canonical (min,max)
  unsigned long long* count;   // [1]
  unsigned long long cap;
  uint32_t* visits;            // optional [n*?] unused
};

// ------------------------------------------------------------- pair ----
struct PairParams {
  const float* posq_sorted;    // float4[n_pad]
  const uint32_t* type_sorted; // [n_pad]
  const uint32_t* work;
  const uint32_t* n_work;
  const uint32_t* entry_j;
  const uint32_t* entry_mask;  // per entry: 8-atom i-sub-clusters any of its j-atoms reaches
  const uint64_t* entry_excl;
  const uint32_t* excl_entries;  // [n_clusters]: masks exist for entries below
  const float* user_params;    // parameters of a run-time composed pair term
  const float* bbox;           // cluster centres + half extents
  HmdLattice L;
  uint32_t n, cap_entries;
  uint32_t n_types;
  const float* lj_c6c12;       // float2[n_types^2]: (4 eps s^6, 4 eps s^12)
  uint32_t lj_type_mask;       // bit t: type t has LJ with some type
  uint32_t lj_single_type;     // the one type in lj_type_mask, or n_types if several
  float rc2;                   // max cutoff squared (mask)
  float lj_rs2;                // switch start squared
  float lj_inv_delta;          // 1/(rc^2 - rs^2); 0 disables the switch
  float alpha, k_coulomb;
  // real-space Ewald kernel, tabulated: float4[2][HMD_COUL_BINS] cubic
  // coefficients per bin of the octave table in r^2 (hmd_coul_bin), force
  // factor G(r) = [erfc(a r)/r + 2a/sqrt(pi) exp(-a^2 r^2)] / r^2 then
  // energy erfc(a r)/r
  const float* coul_table;
  uint32_t coul_u_lo, coul_u_hi;   // f32 bits of the table's first r^2 and of its last
  float shift_margin;          // r_cut + r_skin/2: per-cluster shift eligibility
  long long* acc;              // [5][n_pad] fixed-point F_x,F_y,F_z,E,W (sorted)
  uint32_t acc_stride;         // n_pad
};

// Listed partner pairs with a non-zero exclusion scale (1-4 pairs and the
// like): the fused kernel skips every listed partner, and this pass adds
// s_lj x LJ + s_coul x real-space Coulomb for each scaled pair back into the
// same fixed-point accumulators.
struct PairScaledParams {
  const float* posq;            // float4[n], canonical order
  const uint32_t* type;         // [n]
  const uint32_t* slot_of;      // canonical -> sorted slot
  const uint32_t* pair_i;       // [n_pairs], i < j
  const uint32_t* pair_j;
  const float* s_lj;            // [n_pairs]
  const float* s_coul;
  uint32_t n_pairs;
  uint32_t n_types;
  HmdLattice L;
  const float* lj_c6c12;        // float2 per type pair
  float rc2, lj_rs2, lj_inv_delta;
  float k_coulomb;
  const float* coul_table;      // float4 cubic coefficients, force then energy
  uint32_t coul_u_lo, coul_u_hi;
  uint32_t lj_on, coul_on, fev;
  long long* acc;
  uint32_t acc_stride;
  const float* s_user;          // [n_pairs] scale of a composed user term, or null
  const float* user_params;
};

// Bonded terms (harmonic_bond.rs, morse.rs, angle.rs, dihedral.rs), added
// into the pair accumulators.
#define HMD_BONDED_HARMONIC_BOND 1u
#define HMD_BONDED_MORSE_BOND 2u
#define HMD_BONDED_HARMONIC_ANGLE 3u
#define HMD_BONDED_PERIODIC_DIHEDRAL 4u
struct BondedParams {
  const float* posq;           // [n] float4, canonical
  const uint32_t* terms;       // [n_terms][arity + 1]: canonical atoms, then the type
  const float4* types;         // per type (k, r0) / (De, a, r0) / (k, theta0) / (k, phi0, n)
  const uint32_t* slot_of;     // canonical -> sorted
  long long* acc;              // [5][acc_stride] fixed point (2^-40)
  uint32_t acc_stride;
  HmdLattice L;
  uint32_t n_terms;
  uint32_t kind;
  uint32_t fev;
};

struct PairFinalizeParams {
  const long long* acc;
  uint32_t acc_stride;
  const uint32_t* slot_of;     // canonical -> sorted
  const float* slow_f;         // [3][n] canonical SPME forces (may be null)
  const float* slow_e;         // [n]
  const float* slow_w;         // [n]
  float* f;                    // [3][n] canonical total forces
  float* e;                    // [n]  (fev only)
  float* w;                    // [n]
  uint32_t n;
  uint32_t fev;
  long long* acc_clear;        // accumulators zeroed after reading
};

// One slot's per-atom results added (+=) into caller-owned outputs, the
// SlotOutputView of the reference's Potential::compute: either the pair
// slots' fixed-point accumulators (sorted order, cleared as read) or the
// reciprocal slot's f32 arrays (canonical order).
struct SlotAddParams {
  const long long* acc;        // [5][acc_stride] fixed point, or null
  uint32_t acc_stride;
  const uint32_t* slot_of;     // canonical -> sorted (with acc)
  long long* acc_clear;
  const float* src_f;          // [3][n] canonical, or null (with acc null)
  const float* src_e;          // [n]
  const float* src_w;          // [n]
  float* out_x;                // [n] each, canonical; +=
  float* out_y;
  float* out_z;
  float* out_e;                // energy level only (may be null)
  float* out_w;
  uint32_t n;
  uint32_t fev;
};

// ------------------------------------------------------------- spme ----
struct SpmeBinParams {
  const float* posq;           // canonical order (what the spread and gather read)
  const uint32_t* sorted_id;   // sorted slot -> canonical id (visiting order)
  uint32_t n;
  HmdLattice L;
  uint32_t n0, n1, n2;         // mesh
  uint32_t b0, b1, b2;         // brick dims (mesh points)
  uint32_t nb0, nb1, nb2;      // bricks per axis
  uint32_t* atom_bin;          // [n] bin of each sorted slot
  uint32_t* bin_count;         // [n_bins] zero on entry, zero after scatter
  const uint32_t* bin_start;   // [n_bins+1]
  uint32_t* bin_atoms;         // [n]
};

struct SpmeSpreadParams {
  const float* posq;
  uint32_t n;
  HmdLattice L;
  uint32_t n0, n1, n2;
  uint32_t b0, b1, b2;
  uint32_t nb0, nb1, nb2;
  const uint32_t* bin_start;
  const uint32_t* bin_atoms;
  float* rho;                  // real mesh [n0*n1*n2]
};

struct SpmeGatherParams {
  const float* posq;
  uint32_t n;
  HmdLattice L;
  uint32_t n0, n1, n2;
  uint32_t b0, b1, b2;
  uint32_t nb0, nb1, nb2;
  const uint32_t* bin_start;
  const uint32_t* bin_atoms;
  const float* V;              // real mesh
  const float* u_self;         // [n]
  const double* w_recip;       // [1] scalar reciprocal virial (total)
  float* f;                    // [3][n] canonical, overwritten
  float* e;                    // [n]
  float* w;                    // [n]
  uint32_t fev;
};

// one axis of the 3D FFT, LDS resident
struct FftParams {
  float* data;                 // complex mesh (float2), or real mesh for c-axis
  float* out;                  // c-axis passes: the other array
  uint32_t n;                  // transform length along the axis
  uint32_t n_lines;            // lines in the pass
  uint32_t lines_per_block;
  uint32_t elem_stride;        // complex elements between successive points
  uint32_t line_major, line_minor;  // line l -> (l / line_minor, l % line_minor)
  uint64_t major_stride, minor_stride;  // complex elements
  const float* twiddle;        // float2[n] e^{-2 pi i m/n}
  const float* twiddle_r2c;    // float2[nc] e^{-2 pi i m/nc} (c-axis only)
  const float* twiddle2;       // slab passes: float2[n1] e^{-2 pi i m/n1} (b axis)
  uint32_t radix[8];
  uint32_t n_stages;
  int32_t sign;                // -1 forward, +1 inverse
  // influence (a-axis pass only)
  uint32_t apply_influence;
  uint32_t n0, n1, n2, nch;
  HmdLattice L;
  float alpha, k_coulomb;
  const double* bfac0;         // |b(k)|^2 moduli per axis, FP64
  const double* bfac1;
  const double* bfac2;
  double* partials;            // [grid blocks][2]: virial, energy
  const float* gv;             // float2[n0][n1][nch]: (G, vf) in f32, or null
};

struct SpmeReduceParams {
  const double* partials;
  uint32_t n_partials;
  double* out;                 // [2]: W_recip, U_recip (half sums applied)
};

// ------------------------------------------------------- integrators ---
struct VvParams {
  float* posq;                 // canonical float4
  float* v;                    // [3][n]
  const float* f;              // [3][n]
  const float* m;              // [n] masses
  int32_t* images;             // [3][n]
  double* pos_lo;              // [3][n] compensated mode
  double* vel_lo;              // [3][n]
  HmdLattice L;
  float dt;
  uint32_t n;
  uint32_t compensated;
};

struct SettleParams {
  float* posq;
  float* v;                    // [3][n]
  double* pos_lo;              // compensated (may be null)
  double* vel_lo;
  float* snapshot;             // float4 per atom slot [3*n_groups]
  const uint32_t* group_atoms; // [3*n_groups] O, H1, H2
  uint32_t n_groups;
  uint32_t n;
  float ra, rb, rc, m_o, m_h;
  HmdLattice L;
  float dt;
  float* constraint_virial;    // [3*n_groups]
  float* particle_virial;      // [n]: the constraint virial is added here (may be null)
};

struct LangevinParams {
  float* posq;
  float* v;
  const float* m;              // [n] masses
  int32_t* images;
  HmdLattice L;
  float dt;
  float a;                     // exp(-gamma dt)
  float kt;                    // k_B T (Hartree)
  uint32_t seed_lo, seed_hi;
  const unsigned long long* step;  // device step counter
  uint32_t n;
};

// Fused integrator sub-steps (hmd_step_*): the per-op parameter blocks,
// the free (unconstrained) atoms and the Langevin draw counter.
struct StepParams {
  VvParams vv;
  SettleParams st;
  LangevinParams lan;
  const uint32_t* free_atoms;  // [n_free]; null: atoms 0..n_free-1
  uint32_t n_free;
  float dt;
  unsigned long long* counter; // Langevin draw counter (null for VV)
  NblDispParams disp;          // the force field's skin check on the new
                               // positions (disp.status null: not done here)
};

struct CounterParams {
  unsigned long long* counter;
};

// Stochastic velocity rescaling (CSVR; kernels/nose_hoover.cu:181-360,
// src/integrator/csvr.rs:164-226).
struct CsvrParams {
  const double* ke;            // [1] FP64 kinetic energy of the velocities
  double* partials;            // [n_partials] sums of xi^2
  uint32_t n_partials;
  uint32_t g_dof;              // thermal degrees of freedom 3N - n_constraints - 3
  unsigned long long* counter; // [1] Philox draw counter, +1 per application
  uint32_t seed_lo, seed_hi;
  double c;                    // exp(-dt / tau)
  double one_minus_c;
  double kt_half;              // k_target / N_f = k_B T / 2
  float* factor;               // [1] rescale factor
  double* injection;           // [1] accumulated k_new - k_old
  float* v;                    // [3][n]
  double* vel_lo;              // compensated low parts (may be null)
  uint32_t n;
};

// C-rescale barostat (kernels/barostat.cu:169-227,
// src/integrator/c_rescale_barostat.rs:172-232).
struct CRescaleParams {
  const double* ke;            // [1] FP64 kinetic energy after the step's last kick
  const double* w;             // [1] FP64 sum of the per-atom virials
  double v_pre;                // box volume before the move
  unsigned long long* counter; // [1] Philox draw counter, +1 per application
  uint32_t seed_lo, seed_hi;
  double p_target, tau, beta, kt, dt, mu3_min;
  double* state;               // [4] mu, P, V after the move, accumulated P_target dV
  float* posq;                 // [n] float4, scaled by (float) mu
  double* pos_lo;              // compensated low parts (may be null)
  uint32_t n;
};

struct ReduceParams {
  const float* x;              // [comps][n] or null
  const float* v;              // velocities [3][n] (KE mode)
  const float* m;              // masses (KE mode)
  uint32_t n;
  uint32_t mode;               // 0 = sum x, 1 = kinetic energy
  double* partials;            // [blocks]
  double* out;                 // [1]
  uint32_t n_blocks;
};

struct SdParams {              // capped steepest-descent move (setup only)
  float* posq;
  const float* f;              // [3][n]
  uint32_t n;
  HmdLattice L;
  float h;                     // step per unit force
  float max_step;              // cap on |dx| per atom
};

// ------------------------------------------- learned mesh correction ----
#define HMD_NN_RBF 16u               // radial basis functions
#define HMD_NN_HID 64u               // hidden width
#define HMD_NN_IN 51u                // 16 rbf + 3 one-hot + 2 x 16 descriptors
#define HMD_NN_PARAMS 7553u          // W1 64x51, b1, W2 64x64, b2, w3, b3
#define HMD_NN_DESC_SHIFT 32         // descriptor accumulators: value * 2^32
#define HMD_NN_PARTIAL 11328u        // per-wave gradient partial (floats)
#define HMD_NN_OPERAND_BYTES 8192u   // one packed 64x64 bf16 operand set

// Packed network (written by hmd_nn_pack from the f32 master parameters):
// three matrix-core operand sets of 4 x 2 x 64 lanes x 8 bf16, then f32
// b1[64], b2[64], w3[64] (bf16-rounded values), b3.
struct NnPackParams {
  const float* master;         // [HMD_NN_PARAMS] canonical
  uint16_t* ops;               // [3][4][2][64][8] bf16 bits: W1, W2, W2^T
  float* vec;                  // [64*3 + 1]
};

// Shared by the descriptor, forward and backward passes: the packed
// neighbour list of the force field plus the network's per-atom tables.
struct NnPassParams {
  const float* posq_sorted;    // float4[n_pad]
  const uint32_t* type_sorted; // [n_pad]
  const uint32_t* work;
  const uint32_t* n_work;
  const uint32_t* entry_j;
  const uint32_t* entry_mask;  // per entry: 8-atom i-sub-clusters any of its j-atoms reaches
  const float* bbox;           // cluster centres + half extents
  HmdLattice L;
  uint32_t n, cap_entries, n_pad;
  float shift_margin;          // r_cut + r_skin/2: single-image shift eligibility
  uint32_t species_mask;       // bit t = species (0/1) of type t
  float r_cut2, inv_r_cut;     // r_nn^2, 1/r_nn
  float force_scale;           // s_out
  long long* desc_acc;         // [16][n_pad] fixed point (descriptor pass)
  const float* desc;           // [n_pad][16] sorted (forward/backward)
  const uint16_t* ops;         // packed operands (forward/backward)
  const float* vec;            // b1, b2, w3, b3
  long long* force_acc;        // [3][n_pad] fixed point dF (forward)
  const float* resid;          // [3][n_pad] dL/d(dF), sorted (backward)
  float* partials;             // [n_waves][HMD_NN_PARTIAL] (backward)
};

struct NnDescFinalizeParams {
  long long* desc_acc;         // [16][n_pad], cleared after reading
  uint32_t n_pad, n;
  const uint32_t* sorted_id;   // sorted slot -> canonical id
  float* desc;                 // [n_pad][16] sorted
  float* desc_canon;           // [n][16] canonical (optional)
};

struct NnFinalizeParams {
  long long* force_acc;        // [3][n_pad], cleared after reading
  uint32_t n_pad, n;
  const uint32_t* slot_of;     // canonical -> sorted
  float* df;                   // [3][n] canonical
  float* add_to;               // [3][n] canonical (optional): += dF
};

struct NnResidualParams {
  const float* df;             // [3][n] canonical
  const float* target;         // [3][n] canonical
  const uint32_t* slot_of;
  uint32_t n, n_pad;
  float scale;                 // 2 / (3 n s_out^2 B)
  double loss_scale;           // 1 / (3 n s_out^2 B)
  float* resid;                // [3][n_pad] sorted
  double* partials;            // [blocks]
};

struct NnGradReduceParams {
  const float* partials;       // [n_waves][HMD_NN_PARTIAL]
  uint32_t n_waves;
  double* grad;                // [HMD_NN_PARAMS] accumulated (+=)
};

struct NnAdamParams {
  float* master;               // [HMD_NN_PARAMS]
  float* m;
  float* v;
  double* grad;                // cleared after the update
  float lr, beta1, beta2, eps;
  float bc1, bc2;              // 1 - beta^t
};

struct NnRecordParams {
  const float* posq;           // canonical float4[n]
  const float* fine_f;         // [3][n] fine-mesh reciprocal forces
  const float* coarse_f;       // [3][n] coarse-mesh reciprocal forces
  uint32_t n;
  float* frame_posq;           // float4[n]
  float* frame_target;         // [3][n]
};

struct NnDiffParams {          // sum |a - b|^2 and |b|^2 (FP64 partials)
  const float* a;              // [3][n]
  const float* b;              // [3][n]
  uint32_t n;
  double* partials;            // [2][blocks]: |a - b|^2, then |b|^2
};
