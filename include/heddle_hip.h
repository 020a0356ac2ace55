/*
 * heddle_hip.h -- C ABI of the MI355X-native HeddleMD nonbonded hot path.
 *
 * One shared library, libheddle_hip.so, built from C++17 host code and a
 * gfx950 code object (embedded, loaded with hipModuleLoadData).  Every entry
 * point takes plain pointers and sizes; no C++ or torch types cross this
 * boundary.  Each declaration names the reference interface it replaces
 * (taylor-a-barnes/heddlemd, Rust host + CUDA kernels).
 *
 * Conventions
 *  - Return value: hmd_status, 0 on success.  hmd_last_error() holds a
 *    message for the calling thread's last failure.
 *  - Units: Hartree atomic units (Bohr, Hartree, electron mass, e,
 *    hbar/E_h), as the reference uses internally (src/units/).
 *  - Host arrays are row-major: positions+charges posq[n][4], velocities and
 *    forces [n][3], images [n][3].
 *  - Lattice: {lx, ly, lz, xy, xz, yz}, lower triangular (kernels/pbc.cuh).
 *  - Handles own all device memory.  Every call enqueues on the context's
 *    single HIP stream; calls on one context are not thread safe.
 */
#ifndef HEDDLE_HIP_H
#define HEDDLE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hmd_status;
enum {
  HMD_OK = 0,
  HMD_ERR_GPU = 1,            /* GpuError(DriverError), src/gpu/device.rs:9 */
  HMD_ERR_INVALID = 2,        /* ForceFieldError / SpmeError validation */
  HMD_ERR_NBL_OVERFLOW = 3,   /* NeighborListError::PackedNeighborOverflow */
  HMD_ERR_BOX_TOO_SMALL = 4,  /* NeighborListError::BoxTooSmallForCells */
  HMD_ERR_UNSUPPORTED = 5,
  HMD_ERR_NO_DEVICE = 6
};

typedef struct hmd_ctx hmd_ctx;
typedef struct hmd_particles hmd_particles;
typedef struct hmd_forcefield hmd_forcefield;
typedef struct hmd_integrator hmd_integrator;

/* ---- errors / build info ------------------------------------------------ */
const char* hmd_last_error(void);
const char* hmd_version(void);

/* ---- device context ------------------------------------------------------
 * Replaces init_device + the define_kernels! manifest
 * (src/gpu/device.rs:34-74, src/gpu/kernel_macros.rs:63-113): loads the
 * embedded gfx950 code object and resolves every kernel by name. */
hmd_status hmd_ctx_create(int device_ordinal, hmd_ctx** out);
void hmd_ctx_destroy(hmd_ctx* ctx);
hmd_status hmd_ctx_synchronize(hmd_ctx* ctx);
hmd_status hmd_ctx_device_info(hmd_ctx* ctx, int* compute_units, char* name,
                               int name_len);
/* Kernel manifest (no GPU needed): names of every kernel in the code object. */
int hmd_kernel_count(void);
const char* hmd_kernel_name(int index);

/* ---- particle buffers ----------------------------------------------------
 * ParticleBuffers (src/gpu/buffers.rs:11-178): SoA velocities/forces,
 * posq float4, masses, type indices, image counters; lo parts for the
 * compensated integrator (src/gpu/lossless_buffers.rs). */
hmd_status hmd_particles_create(hmd_ctx* ctx, uint32_t n, const float* posq,
                                const float* vel, const float* mass,
                                const uint32_t* type, const float lattice[6],
                                hmd_particles** out);
void hmd_particles_destroy(hmd_particles* p);
/* Any output pointer may be NULL. */
hmd_status hmd_particles_download(hmd_particles* p, float* posq, float* vel,
                                  float* force, int32_t* images);
hmd_status hmd_particles_upload(hmd_particles* p, const float* posq,
                                const float* vel);
/* Per-atom potential-energy and virial shares of the last energy-level
 * evaluation (SlotOutputView energy/virial, src/forces/mod.rs:206-212). */
hmd_status hmd_particles_download_energies(hmd_particles* p, float* energy,
                                           float* virial);
/* f64 low parts of the compensated integrator ([n][3] each; zeros when the
 * integrator is not compensated). */
hmd_status hmd_particles_download_lo(hmd_particles* p, double* pos_lo,
                                     double* vel_lo);
/* Deterministic FP64 reductions (kernels/nose_hoover.cu:22-104,
 * kernels/barostat.cu:11-85 restated in FP64). */
hmd_status hmd_particles_kinetic_energy(hmd_particles* p, double* ke);
hmd_status hmd_particles_energy_sums(hmd_particles* p, double* pe,
                                     double* virial);

/* ---- force field ---------------------------------------------------------
 * ForceField::new_with_combining + the builtin PotentialRegistry
 * (src/forces/mod.rs:363-398, :514): slots are built in registry order
 * "lennard-jones", "spme-real", "spme-reciprocal".  The pair slots are fused
 * into one kernel (the reference's JIT composer, src/forces/jit_composed.rs);
 * the reciprocal slot is the SPME mesh pipeline (src/forces/spme.rs). */
typedef struct {
  uint32_t n_types;
  /* "lennard-jones" slot: per type pair [n_types*n_types] tables
   * (src/forces/lj.rs:58-75; r_switch == cutoff disables the switch). */
  int32_t lj_on;
  const double* lj_sigma;
  const double* lj_epsilon;
  const double* lj_cutoff;
  const double* lj_switch;
  /* "spme-real" slot (src/forces/spme.rs:686-777) */
  int32_t coulomb_on;
  double alpha;
  double r_cut_real;
  double k_coulomb;
  /* "spme-reciprocal" slot (src/forces/spme.rs:82-159, :858-914) */
  int32_t spme_on;
  uint32_t grid[3];
  uint32_t spline_order;  /* 4, 5, 6 or 8 */
  uint32_t brick[3];      /* owner-computes brick; 0 = choose */
  /* [neighbor_list] (src/io/config.rs): r_skin */
  double r_skin;
  /* exclusion CSR over canonical atom ids, symmetric, with per-slot scales
   * (src/forces/topology.rs:312-430).  Listed partners are fully excluded
   * from the fused pair kernel; scales other than 0 are added back by an
   * exact correction (not yet supported: HMD_ERR_UNSUPPORTED). */
  const uint32_t* excl_offsets;  /* [n+1] or NULL */
  const uint32_t* excl_partners;
  const float* excl_lj_scale;
  const float* excl_coul_scale;
} hmd_forcefield_desc;

hmd_status hmd_forcefield_create(hmd_ctx* ctx, hmd_particles* particles,
                                 const hmd_forcefield_desc* desc,
                                 hmd_forcefield** out);
void hmd_forcefield_destroy(hmd_forcefield* ff);
/* ForceField::run (src/forces/mod.rs:944-1432).  level 0 = forces,
 * 1 = forces + per-atom energies and virials (AggregateLevel). */
hmd_status hmd_forcefield_compute(hmd_forcefield* ff, int32_t level);
/* Slot names in evaluation order, '\n' separated. */
hmd_status hmd_forcefield_slots(hmd_forcefield* ff, char* buf, int32_t len);
/* Reciprocal scalars of the last evaluation: out[0] = W_recip,
 * out[1] = U_recip (without self energy), out[2] = U_self. */
hmd_status hmd_spme_scalars(hmd_forcefield* ff, double out[3]);
/* Mesh charge density and potential of the last evaluation (f32[M]). */
hmd_status hmd_spme_grids(hmd_forcefield* ff, float* rho, float* potential);
/* Reciprocal-only forces of the last evaluation ([n][3]). */
hmd_status hmd_spme_forces(hmd_forcefield* ff, float* force);

/* ---- neighbour list ------------------------------------------------------
 * NeighborListState (src/forces/neighbor_list.rs). */
/* Force a rebuild at the next evaluation (rebuild request). */
hmd_status hmd_nbl_request_rebuild(hmd_forcefield* ff);
/* PairSnapshot (neighbor_list.rs:1358-1445): every listed pair whose exact
 * f32 min-image distance is <= r_c + r_skin, as canonical (min, max) ids
 * (unsorted).  *n_pairs receives the total even when it exceeds cap. */
hmd_status hmd_nbl_pair_snapshot(hmd_forcefield* ff, uint32_t* pairs,
                                 uint64_t cap, uint64_t* n_pairs);
/* Number of listed pairs whose exact f32 min-image distance is <= radius
 * (radius <= r_c + r_skin); with radius = r_c this is the number of pairs
 * the pair kernel evaluates inside the cutoff. */
hmd_status hmd_nbl_count_pairs(hmd_forcefield* ff, double radius,
                               uint64_t* n_pairs);
/* out[0] entries in the work list, [1] rebuild generation, [2] status bits,
 * [3] high-water entries per cluster, [4] capacity per cluster,
 * [5] clusters, [6] fine cells per axis, [7] pad. */
hmd_status hmd_nbl_stats(hmd_forcefield* ff, uint32_t out[8]);

/* ---- integrators ---------------------------------------------------------
 * Integrator plans (src/integrator/mod.rs:420, velocity_verlet.rs:67-110,
 * langevin_baoab.rs:86-111) with SETTLE constraint hooks
 * (src/integrator/settle.rs:310-404). */
typedef struct {
  const char* kind;     /* "velocity-verlet" or "langevin-baoab" */
  double dt;
  int32_t compensated;  /* (f32, f64) compensated kick/drift */
  double temperature;   /* k_B T in Hartree (Langevin) */
  double gamma;         /* friction, 1/time (Langevin) */
  uint64_t seed;        /* Philox key (Langevin) */
  /* SETTLE groups: [n_groups][3] canonical ids (O, H1, H2) */
  uint32_t n_groups;
  const uint32_t* group_atoms;
  double d_oh, d_hh, m_o, m_h;
} hmd_integrator_desc;

hmd_status hmd_integrator_create(hmd_ctx* ctx, hmd_particles* particles,
                                 hmd_forcefield* ff,
                                 const hmd_integrator_desc* desc,
                                 hmd_integrator** out);
void hmd_integrator_destroy(hmd_integrator* it);
/* run_step (src/integrator/mod.rs:523-672), eager launches. */
hmd_status hmd_integrator_step(hmd_integrator* it, uint32_t n_steps);
/* Capture one step into a hipGraph (runner.rs:2321-2470) and replay it. */
hmd_status hmd_integrator_capture(hmd_integrator* it);
hmd_status hmd_integrator_replay(hmd_integrator* it, uint32_t n_steps);
/* Steps taken so far. */
uint64_t hmd_integrator_steps(hmd_integrator* it);
/* Per-stage GPU time (HIP events) of n eager steps: writes up to
 * max_stages averages in microseconds and '\n' separated stage names. */
hmd_status hmd_integrator_profile(hmd_integrator* it, uint32_t n_steps,
                                  double* stage_us, int32_t max_stages,
                                  int32_t* n_stages, char* names,
                                  int32_t names_len);
/* Steepest-descent relaxation with SETTLE position projection (setup only;
 * the reference's minimizer, src/minimizer/). */
hmd_status hmd_integrator_minimize(hmd_integrator* it, uint32_t n_iter,
                                   double max_step);

/* ---- setup helpers (host only, no GPU) -----------------------------------
 * generate_velocities (src/runner.rs:3108-3200): ChaCha8 seeded with
 * seed_from_u64, Box-Muller cos branch, COM removal, rescale to
 * (3N - n_constraints - 3)/2 k_B T.  vel_out is [n][3]. */
hmd_status hmd_generate_velocities(uint32_t n, uint32_t n_constraints,
                                   double temperature, uint64_t seed,
                                   const double* masses, float* vel_out);
/* ChaCha block function on a raw 16-word state (test hook). */
void hmd_chacha_block(const uint32_t state[16], uint32_t rounds,
                      uint32_t out[16]);

#ifdef __cplusplus
}
#endif
#endif /* HEDDLE_HIP_H */
