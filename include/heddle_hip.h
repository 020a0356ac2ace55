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
  HMD_ERR_NO_DEVICE = 6,
  HMD_ERR_OUTPUT_EXISTS = 7,  /* TrajectoryWriterError/LogWriterError::OutputExists */
  HMD_ERR_IO = 8              /* TrajectoryWriterError/LogWriterError::Io */
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
/* A context that enqueues on the caller's HIP stream (a hipStream_t, passed
 * as void*) instead of a stream of its own, so that a host's stream capture
 * (the runner's graph capture, src/runner.rs:2402-2408) records the path's
 * launches; the stream must outlive the context, which does not destroy
 * it.  Force fields on such a context run SPME on that stream (no second
 * stream unless hmd_forcefield_set_tuning asks for one). */
hmd_status hmd_ctx_create_on_stream(int device_ordinal, void* stream, hmd_ctx** out);
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
/* Uploading positions (velocities) also clears their compensated low
 * parts: the uploaded values are the whole new state. */
hmd_status hmd_particles_upload(hmd_particles* p, const float* posq,
                                const float* vel);
/* New box {lx, ly, lz, xy, xz, yz} (SimulationBox update with its generation
 * counter, src/pbc.rs; a barostat's box change).  Positions are not
 * touched (scale them with hmd_particles_upload).  Every force field built
 * on these particles re-checks the box (HMD_ERR_BOX_TOO_SMALL leaves the
 * old box in place), refreshes its SPME influence table, rebuilds its list
 * at the next evaluation and re-captures its step graphs. */
hmd_status hmd_particles_set_box(hmd_particles* p, const float lattice[6]);
/* The current box {lx, ly, lz, xy, xz, yz} (SimulationBox::lattice,
 * src/pbc.rs; after a barostat move, the moved one). */
hmd_status hmd_particles_get_box(hmd_particles* p, float lattice[6]);
/* Image counters [n][3] (the images of an init state,
 * src/io/init_state.rs InitImages). */
hmd_status hmd_particles_upload_images(hmd_particles* p, const int32_t* images);
/* Forces [n][3] into the force buffer (the forces_{x,y,z} part of
 * ParticleBuffers::upload, src/gpu/buffers.rs:252-310); the next force
 * evaluation overwrites them. */
hmd_status hmd_particles_upload_forces(hmd_particles* p, const float* force);
/* v <- -v, both the f32 velocities and their f64 low parts: the velocity
 * reversal of a time-reversal check (the negate_velocities helper of
 * tests/integration_lossless.rs). */
hmd_status hmd_particles_negate_velocities(hmd_particles* p);
/* One velocity-Verlet half: v += F/m dt/2, x += v dt, wrap + image count
 * (vv_kick_drift, src/gpu/kernels.rs:34-58; compensated != 0:
 * vv_kick_drift_lossless, :2292-2330, on the f64 low parts), with the
 * forces currently in the buffer. */
hmd_status hmd_vv_kick_drift(hmd_particles* p, float dt, int32_t compensated);
/* v += F/m dt/2 (vv_kick, src/gpu/kernels.rs:128-143; vv_kick_lossless,
 * :2416-2450). */
hmd_status hmd_vv_kick(hmd_particles* p, float dt, int32_t compensated);
/* Langevin BAOAB A step: x += v dt/2, wrap + image count (lan_drift_half,
 * src/gpu/kernels.rs:2339-2372; kernels/langevin.cu:13-50). */
hmd_status hmd_lan_drift_half(hmd_particles* p, float dt);
/* Langevin O step v = alpha v + sqrt((1 - alpha^2) kT / m) xi, xi the
 * Philox-4x32-10 Box-Muller normal keyed by seed with counter
 * (draw_counter lo, hi, particle, axis) (lan_ou_step,
 * src/gpu/kernels.rs:2374-2410; kernels/langevin.cu:52-80).  The counter is
 * passed by value; the reference's device counter increment is the
 * caller's. */
hmd_status hmd_lan_ou_step(hmd_particles* p, uint64_t seed, float alpha, float kt,
                           uint64_t draw_counter);
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
   * (src/forces/topology.rs:312-430; NULL scales = 0).  Listed partners are
   * fully excluded from the fused pair kernel; a pair with a non-zero scale
   * (read from the lower atom's list) gets scale x its LJ / real-space
   * Coulomb term added back by a correction pass over the scaled pairs,
   * under the same cutoff (jit_composed.rs:939-972). */
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
/* ---- bonded terms ---------------------------------------------------------
 * The intramolecular potentials the reference composes into one module per
 * arity (src/forces/jit_composed.rs:1739-2406): HMD_BONDED_HARMONIC_BOND
 * (harmonic_bond.rs:316-360; params k, r0), HMD_BONDED_MORSE_BOND
 * (morse.rs:330-360; De, a, r0), HMD_BONDED_HARMONIC_ANGLE
 * (angle.rs:350-410; k_theta, theta0 in radians, angle i-j-k at j) and
 * HMD_BONDED_PERIODIC_DIHEDRAL (dihedral.rs:440-530; k_phi, phi0, n).
 * terms: [n_terms][arity + 1] canonical atom ids then the type index;
 * params: [n_types][3] in atomic units (unused entries ignored).  Each call
 * adds one set (a slot) to the force field; hmd_forcefield_compute adds
 * every set's forces, and at level 1 the energy and virial shares (1/2,
 * 1/3, 1/4 of the term's per endpoint), to the fast-class totals.  Invalid
 * parameters (k, r0, De, a not finite and positive; theta0 outside [0, pi];
 * a dihedral n not an integer in [0, 64]), atoms out of range or repeated
 * within a term, and types out of range are HMD_ERR_INVALID.  Exclusions of
 * bonded partners from the pair terms are the caller's, through the
 * force-field descriptor's exclusion lists, as the reference's topology
 * supplies them. */
typedef struct {
  uint32_t kind;
  uint32_t n_terms;
  const uint32_t* terms;
  uint32_t n_types;
  const double* params;
} hmd_bonded_desc;
#define HMD_BONDED_HARMONIC_BOND 1u
#define HMD_BONDED_MORSE_BOND 2u
#define HMD_BONDED_HARMONIC_ANGLE 3u
#define HMD_BONDED_PERIODIC_DIHEDRAL 4u
hmd_status hmd_forcefield_add_bonded(hmd_forcefield* ff, const hmd_bonded_desc* desc);

/* ---- slot level (Potential::compute with a SlotOutputView,
 * src/forces/mod.rs:86-133, 206-212) ------------------------------------
 * The same evaluation one slot at a time, each adding (+=) its per-atom
 * results into caller-owned device arrays (canonical atom order), so that
 * a host can keep slots of its own (bonded terms, a Morse term) beside
 * these and sum them into one view.  Enqueue only, on the context's stream
 * (capturable into a caller's graph, see hmd_ctx_create_on_stream): per
 * evaluation call hmd_nbl_pre_step first (NeighborListState::pre_step,
 * neighbor_list.rs:1265-1354: skin check, device-gated rebuild, sorted
 * positions), then hmd_pair_force and hmd_spme_compute in either order, and
 * at a batch boundary hmd_forcefield_check (the status read: an overflow is
 * reported there, as hmd_forcefield_compute reports it).  Outputs the
 * caller zeroes once per evaluation give the bytes hmd_forcefield_compute
 * writes. */
typedef struct {
  float* force_x;   /* [n] each; += */
  float* force_y;
  float* force_z;
  float* energy;    /* level 1: per-atom energy shares (may be NULL) */
  float* virial;    /* level 1: per-atom virial shares (may be NULL) */
} hmd_slot_output;
hmd_status hmd_nbl_pre_step(hmd_forcefield* ff);
/* The fast-class pair slots ("lennard-jones", "spme-real", with scaled
 * exclusions): the fused kernel's fixed-point sums added into out. */
hmd_status hmd_pair_force(hmd_forcefield* ff, int32_t level, const hmd_slot_output* out);
/* The "spme-reciprocal" slot (spme.rs:858-914), with an attached network's
 * correction when there is one, added into out. */
hmd_status hmd_spme_compute(hmd_forcefield* ff, int32_t level, const hmd_slot_output* out);
/* Every bonded set of the force field as one slot (the composed bonded,
 * angle and dihedral modules' launches, jit_composed.rs:1759-2406), += into
 * out; enqueue only. */
hmd_status hmd_bonded_compute(hmd_forcefield* ff, int32_t level, const hmd_slot_output* out);
/* Synchronize the context's stream and read the neighbour status (the
 * per-batch 4-byte read of pre_step, neighbor_list.rs:1308-1320). */
hmd_status hmd_forcefield_check(hmd_forcefield* ff);
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
/* hmd_nbl_pair_snapshot without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code near
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The pairs and their counter are staged in a fixed device array of the
 * library (2,097,152 pairs, 16 MiB + 256 B per GPU, held until the process
 * ends); cap above 2,097,152 is refused with HMD_ERR_INVALID before any
 * device work, and *n_pairs still receives the full total.  Calls on one
 * GPU share that array: calls that overlap on one GPU race. */
hmd_status hmd_nbl_pair_snapshot_noalloc(hmd_forcefield* ff, uint32_t* pairs,
                                         uint64_t cap, uint64_t* n_pairs);
/* Number of listed pairs whose exact f32 min-image distance is <= radius
 * (radius <= r_c + r_skin); with radius = r_c this is the number of pairs
 * the pair kernel evaluates inside the cutoff. */
hmd_status hmd_nbl_count_pairs(hmd_forcefield* ff, double radius,
                               uint64_t* n_pairs);
/* Set the list capacity (entries of 64 j-atoms per cluster; clamped to the
 * number of clusters, the all-pairs ceiling) and rebuild at the next
 * evaluation: the capacity field of PackedNeighborData and its growth
 * (neighbor_list.rs:415-445).  The library sizes the capacity itself (probe
 * at creation, 1.5x growth on overflow and past 0.8 fill); a smaller seed
 * exercises that growth.  lds_entries caps the entries the build stages in
 * LDS (0: as many as fit beside its other arrays, about 150); list
 * positions past them are written to the entry array directly, with the
 * same result. */
hmd_status hmd_nbl_set_capacity(hmd_forcefield* ff, uint32_t entries, uint32_t lds_entries);
/* out[0] entries in the work list, [1] rebuild generation, [2] status bits,
 * [3] high-water entries per cluster, [4] capacity per cluster,
 * [5] clusters, [6] fine cells per axis, [7] i-sub-cluster visits of the
 * pair kernel over the list (each 8 steps of 64 pairs). */
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
  /* 0: the sub-steps between two force evaluations run as one fused
   * launch per group of sub-steps (same per-atom arithmetic and order);
   * HMD_INTEGRATOR_UNFUSED: one launch per sub-step, as the reference's
   * plan executes them.  HMD_INTEGRATOR_SCALARS: every step's force
   * evaluation is at the energy level (per-atom energies and virials,
   * AggregateLevel) and, with SETTLE, the constraint virial of the step
   * (position-level after the drift plus velocity-level after the kick) is
   * added to each atom's virial after the last kick, as
   * settle_virial_scatter does (kernels/settle.cu:498-510, called from
   * apply_after_kick, src/integrator/settle.rs:371-401). */
  uint32_t flags;
  /* Thermostat applied after every step's last kick, before the SETTLE
   * velocity projection (apply_post, src/integrator/mod.rs:500-517;
   * velocity Verlet only):
   * HMD_THERMOSTAT_CSVR is the stochastic velocity rescaling of
   * src/integrator/csvr.rs:164-226 (kernels/nose_hoover.cu:181-360) with
   * temperature = k_B T, tau = coupling time and seed = Philox key; N_f =
   * 3N - 3 n_groups - 3.  The kinetic energy is the FP64 reduction. */
  uint32_t thermostat;
  double tau;
} hmd_integrator_desc;
#define HMD_INTEGRATOR_UNFUSED 1u
#define HMD_INTEGRATOR_SCALARS 2u
#define HMD_THERMOSTAT_NONE 0u
#define HMD_THERMOSTAT_CSVR 1u

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
/* Switch the integrator to another force field on the same particles: the
 * forces are re-evaluated with it at the current positions and the step
 * graph is re-captured (the revert of a judged run; no counterpart in the
 * reference, whose ForceField is fixed for a run). */
hmd_status hmd_integrator_set_forcefield(hmd_integrator* it, hmd_forcefield* ff);
/* The learned leg's guarded run: n_steps graph-replayed steps of an
 * integrator whose force field carries a network (hmd_mlp_attach), in
 * batches of judge_every steps; after each batch the fallback judge
 * (hmd_mlp_judge) compares its forces with fine_ff's at the same positions,
 * and the first relative RMS error above threshold switches the integrator
 * to fine_ff (hmd_integrator_set_forcefield) for the rest of the run and
 * after it.  judges[] receives up to max_judges judge values and *n_judges
 * their count (may be NULL); *reverted_at the integrator's step count at the
 * switch, or UINT64_MAX.  The learned leg's fallback policy (plain fine-mesh
 * SPME as the judge every 50 steps, revert above 1e-3); the reference's nearest policy is its
 * runner's fallback from graph replay to eager steps (src/runner.rs:
 * 1650-1656). */
hmd_status hmd_integrator_run_judged(hmd_integrator* it, hmd_forcefield* fine_ff,
                                     uint32_t n_steps, uint32_t judge_every, double threshold,
                                     double* judges, uint32_t max_judges, uint32_t* n_judges,
                                     uint64_t* reverted_at);
/* Device copy of everything a batch of steps changes (positions,
 * velocities, images, forces, the compensated low parts, the random-stream
 * and thermostat counters, the step count), and its restore with a list
 * rebuild at the next evaluation: the re-run of a batch that reported
 * HMD_ERR_NBL_OVERFLOW starts from exactly the batch's start.  No
 * counterpart in the reference, whose runner halts on a packed-list
 * overflow (NeighborListError::PackedNeighborOverflow,
 * src/forces/neighbor_list.rs:1308-1320).  Rollback without a checkpoint is
 * HMD_ERR_INVALID; a checkpoint stays valid until the next one. */
hmd_status hmd_integrator_checkpoint(hmd_integrator* it);
hmd_status hmd_integrator_rollback(hmd_integrator* it);
/* CSVR state (flush_pending_injection, csvr.rs:134-160): out[0] the energy
 * the thermostat has injected, sum of K' - K (the conserved quantity is
 * KE + PE - out[0]), out[1] the Philox draws so far. */
hmd_status hmd_integrator_thermostat_state(hmd_integrator* it, double out[2]);
/* C-rescale barostat (src/integrator/c_rescale_barostat.rs:21-28
 * CRescaleBarostatParams, :172-232 apply; kernels/barostat.cu:169-227
 * c_rescale_compute_mu, :93-107 rescale_positions_device_factor): one
 * isotropic move per step at the plan's BarostatPoint, after the last kick
 * and the thermostat and before the SETTLE velocity projection
 * (src/integrator/velocity_verlet.rs:97-109, mod.rs:500-517).  pressure =
 * target (Hartree/Bohr^3), temperature = k_B T of the noise term, tau the
 * coupling time, compressibility in 1/pressure, seed the Philox key (draw
 * counter from 0).  With K and W the FP64 kinetic-energy and per-atom-virial
 * sums at that point and V the volume, P = (2K + W)/(3V), mu^3 = 1 -
 * beta (dt/tau)(P_target - P) + sqrt(2 beta kT dt/(tau V)) r, floored at
 * 1e-18; positions and all six lattice components are multiplied by
 * (float) mu.  Validation as validate_params (:282-290): pressure finite,
 * the others finite and positive, else HMD_ERR_INVALID.  Turns on
 * energy-level evaluations (HMD_INTEGRATOR_SCALARS), as the reference's
 * runner does for a per-step barostat (src/runner.rs:2521-2539).  The box
 * is a kernel argument here, so barostat steps run eagerly:
 * hmd_integrator_replay steps eagerly, hmd_integrator_capture and
 * hmd_replicas_step return HMD_ERR_UNSUPPORTED.  kind HMD_BAROSTAT_NONE
 * removes the barostat (energy-level evaluations stay on). */
typedef struct {
  uint32_t kind;
  double pressure;
  double temperature;
  double tau;
  double compressibility;
  uint64_t seed;
} hmd_barostat_desc;
#define HMD_BAROSTAT_NONE 0u
#define HMD_BAROSTAT_C_RESCALE 1u
hmd_status hmd_integrator_set_barostat(hmd_integrator* it, const hmd_barostat_desc* desc);
/* Barostat state (flush_pending_injection, c_rescale_barostat.rs:141-170):
 * out[0] the pressure of the latest move, out[1] the volume after it,
 * out[2] the accumulated P_target (V_post - V_pre) (the conserved quantity
 * is KE + PE + P_target V - out[2], log_column_values :240-252), out[3] the
 * Philox draws so far, out[4] the latest mu.  Zeros before the first
 * move. */
hmd_status hmd_integrator_barostat_state(hmd_integrator* it, double out[5]);
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
/* Per-stage kernel timings of n eager steps (HIP events between the
 * stage marks on the context stream): the Timings collector of
 * src/timings.rs:228-495 (count, total, min, max per stage, in plan
 * order).  Arrays hold max_stages entries; names are '\n'-separated. */
hmd_status hmd_integrator_timings(hmd_integrator* it, uint32_t n_steps, int32_t max_stages,
                                  int32_t* n_stages, char* names, int32_t names_len,
                                  uint64_t* counts, uint64_t* total_ns, uint64_t* min_ns,
                                  uint64_t* max_ns);
hmd_status hmd_integrator_minimize(hmd_integrator* it, uint32_t n_iter,
                                   double max_step);

/* The .timings report (write_timings_file, src/timings.rs:513-553): a
 * fixed-width table of the stages with a nonzero count; an existing path
 * is HMD_ERR_OUTPUT_EXISTS. */
hmd_status hmd_timings_write(const char* path, uint32_t n_stages, const char* const* names,
                             const uint64_t* counts, const uint64_t* total_ns,
                             const uint64_t* min_ns, const uint64_t* max_ns);

/* ---- run-time composed pair terms -------------------------------------------
 * A user pair potential compiled into the force kernels with hiprtc (the
 * PairForceFragment composer of src/forces/jit_composed.rs:112-151,
 * :697-720, :1162-1643).  `functor_source` defines
 *   struct <functor_name> {
 *     __device__ void operator()(const float* params, uint32_t type_i,
 *         uint32_t type_j, float r2, float inv_r,
 *         float& force_over_r, float& energy) const;
 *   };
 * called for every listed pair (before the cutoff and exclusion masks,
 * which then apply to it as to the built-in terms); force_over_r * (x_i -
 * x_j) is added to atom i and subtracted from atom j, energy is split half
 * and half.  `params` (n_params floats) is copied to the device.  The
 * force field's pair kernels are replaced by the composed ones (captured
 * graphs are re-captured).  A source that does not compile returns
 * HMD_ERR_INVALID with the compiler log in hmd_last_error(). */
hmd_status hmd_forcefield_set_pair_fragment(hmd_forcefield* ff, const char* functor_name,
                                            const char* functor_source, const float* params,
                                            uint32_t n_params);
/* The composed term's exclusion_scale(i, j) (src/forces/lj.rs:451, the
 * per-fragment scale the composer multiplies into each fragment's result,
 * jit_composed.rs:939-972): one scale per entry of the force field's
 * exclusion lists (excl_offsets[n] entries, in CSR order; the pair (i, j) is
 * taken from i's list for i < j).  Listed pairs within the cutoff then get
 * scale x the composed term, beside the built-in terms' own scales; without
 * this call the composed term has scale 0 for listed partners.  Needs a
 * composed term (HMD_ERR_INVALID otherwise), the exact entry count and
 * finite scales. */
hmd_status hmd_forcefield_set_fragment_exclusion_scales(hmd_forcefield* ff, const float* scales,
                                                        uint32_t n_entries);
/* Compile a fragment without a device (validation); the compiler log goes
 * to `log` (log_len bytes, may be 0). */
hmd_status hmd_pair_fragment_compile(const char* functor_name, const char* functor_source,
                                     char* log, int32_t log_len);

/* ---- SETTLE as operators ---------------------------------------------------
 * The Constraint hooks of the reference's SETTLE (src/integrator/settle.rs:
 * 310-404): before_drift snapshots the group positions, after_drift resets
 * each group to the rigid geometry (Miyamoto & Kollman 1992) and adds the
 * correction / dt to the velocities, after_kick projects out the bond
 * velocities; f32 state (the integrators use the same kernels, and the
 * compensated forms, internally).  Groups are (O, H1, H2) atom ids. */
typedef struct hmd_settle hmd_settle;
hmd_status hmd_settle_create(hmd_particles* p, uint32_t n_groups, const uint32_t* group_atoms,
                             double d_oh, double d_hh, double m_o, double m_h, hmd_settle** out);
void hmd_settle_destroy(hmd_settle* s);
hmd_status hmd_settle_before_drift(hmd_settle* s);
hmd_status hmd_settle_after_drift(hmd_settle* s, float dt);
hmd_status hmd_settle_after_kick(hmd_settle* s, float dt);

/* ---- launch-shape autotuner ----------------------------------------------
 * No counterpart in the reference (its launch shapes are fixed).  The
 * shapes a force field launches with:
 *   pair_blocks_per_cu  pair-kernel workgroups per CU (capped by the list's
 *                       entries).  The f32 per-wave force sums follow the
 *                       work split, so each value gives reproducible results
 *                       that differ from another value's in the last bits;
 *   fft_blocks_per_cu   FFT-pass workgroups per CU (forces unchanged);
 *   brick[3]            SPME spread / gather brick (forces unchanged: the
 *                       mesh sums are exact integers);
 *   spme_stream         1: SPME runs on a second stream beside the pair
 *                       kernel (fork / join in the step graph), 0: after it
 *                       on one stream (results unchanged).
 * hmd_forcefield_autotune times candidates of each (coordinate descent;
 * per candidate the best of three runs of n_evals eager force evaluations on
 * the current positions)
 * and keeps the fastest; hmd_forcefield_set_tuning applies a recorded
 * choice, so that a later run repeats a tuned run's bytes.  Both invalidate
 * captured step graphs. */
typedef struct hmd_tuning {
  uint32_t pair_blocks_per_cu;
  uint32_t fft_blocks_per_cu;  /* 0 without SPME */
  uint32_t brick[3];           /* 0 without SPME */
  uint32_t spme_stream;        /* 0 without SPME */
  uint32_t fft_slab;           /* 1: three-kernel slab FFT where the mesh allows
                                  (128^3, 64^3); 0 (default): five line passes */
  uint32_t n_candidates;       /* autotune: configurations timed */
  double us_initial;           /* autotune: us per evaluation with the shapes at entry */
  double us_tuned;             /* autotune: us per evaluation with the chosen shapes */
} hmd_tuning;
hmd_status hmd_forcefield_autotune(hmd_forcefield* f, uint32_t n_evals, hmd_tuning* out);
hmd_status hmd_forcefield_set_tuning(hmd_forcefield* f, const hmd_tuning* t);
hmd_status hmd_forcefield_get_tuning(const hmd_forcefield* f, hmd_tuning* out);

/* ---- replica batches -----------------------------------------------------
 * Independent systems stepped together (the batched-replica workload; no
 * counterpart in the reference, which runs one system per process).  Each
 * integrator must come from its own context (hmd_ctx) on the same device.
 * Default: all replicas' steps are captured into one graph (one launch per
 * batched step, a fork per replica); HMD_REPLICAS_PER_REPLICA_GRAPHS keeps
 * one step graph per replica, each launched on its own stream;
 * HMD_REPLICAS_BATCHED merges the replicas' steps into one launch per stage
 * (each kernel's batched twin over the blocks of every replica, with the
 * replicas' parameter blocks in device memory), captured into one graph --
 * the same arithmetic per replica, so the results are those of stepping
 * each replica alone.  Status is
 * checked for every replica after each call; an overflow grows that
 * replica's list and is reported as HMD_ERR_NBL_OVERFLOW (re-run the
 * batch).  The integrators must outlive the set. */
typedef struct hmd_replicas hmd_replicas;
#define HMD_REPLICAS_PER_REPLICA_GRAPHS 1u
#define HMD_REPLICAS_BATCHED 2u
/* with HMD_REPLICAS_BATCHED: issue the merged launches directly every step
 * instead of through a graph (for kernel tracers; same results) */
#define HMD_REPLICAS_EAGER 4u
hmd_status hmd_replicas_create(hmd_integrator* const* integrators, uint32_t n, uint32_t flags,
                               hmd_replicas** out);
void hmd_replicas_destroy(hmd_replicas* set);
hmd_status hmd_replicas_step(hmd_replicas* set, uint32_t n_steps);
uint64_t hmd_replicas_steps(hmd_replicas* set);
/* The merged plan of a HMD_REPLICAS_BATCHED set (after its first step):
 * batched launches and single launches per batched step (zeros in the other
 * modes or before the first step). */
hmd_status hmd_replicas_plan(hmd_replicas* set, uint32_t* n_batched, uint32_t* n_single);

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

/* ---- learned coarse-mesh correction --------------------------------------
 * No counterpart in the reference: this is the hmd_mlp_{create, forward,
 * backward, adam_step} part of the hot path's C ABI, for the pair-scalar
 * network that corrects a coarse SPME mesh toward a fine one.  Model: for
 * every pair within r_cut, x_ij = [16 Gaussian radial basis values of
 * r_ij/r_cut, one-hot(species_i + species_j), D_i, D_j] with the per-atom
 * charge-weighted radial descriptor D_i = sum_j q_j phi(r_ij) fc(r_ij);
 * N = 51 -> 64 -> 64 -> 1 with SiLU (bf16 weights and activations on the
 * matrix cores, f32 accumulation); g_ij = force_scale fc(r_ij)
 * (N(x_ij) + N(x_ji))/2 and dF_i = sum_j g_ij (r_i - r_j)/r_ij.
 * Parameters (hmd_mlp_param_count() = 7553 floats): W1[64][51], b1[64],
 * W2[64][64], b2[64], w3[64], b3. */
typedef struct hmd_mlp hmd_mlp;
typedef struct {
  double r_cut;                     /* pair range (Bohr), <= r_c + r_skin */
  double force_scale;               /* output scale (Hartree/Bohr) */
  uint32_t n_types;                 /* <= 32 */
  const uint32_t* species_of_type;  /* 0 or 1 per type */
  uint64_t seed;                    /* parameter init and frame order */
  double lr, beta1, beta2, eps;     /* Adam; 0 = 1e-3, 0.9, 0.999, 1e-8 */
} hmd_mlp_desc;

uint32_t hmd_mlp_param_count(void);
hmd_status hmd_mlp_create(hmd_ctx* ctx, const hmd_mlp_desc* desc,
                          hmd_mlp** out);
void hmd_mlp_destroy(hmd_mlp* mlp);
hmd_status hmd_mlp_set_params(hmd_mlp* mlp, const float* params);
hmd_status hmd_mlp_get_params(hmd_mlp* mlp, float* params);
hmd_status hmd_mlp_set_force_scale(hmd_mlp* mlp, double force_scale);
/* Descriptors ([n][16]) of ff's particles at their current positions. */
hmd_status hmd_mlp_descriptors(hmd_mlp* mlp, hmd_forcefield* ff, float* out);
/* Force correction ([n][3]) at the current positions. */
hmd_status hmd_mlp_forward(hmd_mlp* mlp, hmd_forcefield* ff, float* dforce);
/* Loss sum_i |dF_i - T_i|^2 / (3 n force_scale^2) of one frame (the current
 * positions, target [n][3]) and its gradient (param_count floats). */
hmd_status hmd_mlp_loss_grad(hmd_mlp* mlp, hmd_forcefield* ff,
                             const float* target, double* loss, float* grad);
/* One Adam step (bias-corrected) with the given gradient. */
hmd_status hmd_mlp_adam_step(hmd_mlp* mlp, const float* grad);
/* Add the correction to ff's reciprocal (slow-class) forces in every
 * hmd_forcefield_compute, eager or graph-replayed; NULL detaches. */
hmd_status hmd_mlp_attach(hmd_forcefield* ff, hmd_mlp* mlp);
/* Training frames (device resident): positions and the target
 * F_recip(fine) - F_recip(coarse) of two force fields on the same
 * particles at their current positions.  A neighbour list that overflows
 * at these positions is grown and the frame evaluated again;
 * HMD_ERR_NBL_OVERFLOW is returned only if it overflows twice. */
hmd_status hmd_mlp_dataset_reserve(hmd_mlp* mlp, uint32_t n_atoms,
                                   uint32_t n_frames);
hmd_status hmd_mlp_record_frame(hmd_mlp* mlp, hmd_forcefield* fine,
                                hmd_forcefield* coarse);
/* out[0] frames, [1] RMS target component, [2] RMS fine reciprocal
 * force component. */
hmd_status hmd_mlp_dataset_stats(hmd_mlp* mlp, double out[3]);
/* n_steps Adam steps over batches of recorded frames (seeded shuffle per
 * epoch), forward and backward passes on the GPU; ff (the coarse force
 * field) is the scratch system: its particle positions are overwritten.
 * losses[n_steps] receives the batch losses (may be NULL). */
hmd_status hmd_mlp_train(hmd_mlp* mlp, hmd_forcefield* ff, uint32_t n_steps,
                         uint32_t batch_frames, double* losses);
/* Fallback judge: total forces of nn_ff (with its network) and fine_ff at
 * the current positions; *rel_rms = ||F_nn - F_fine|| / ||F_fine||.  The
 * particle forces are left as nn_ff's.  Overflow is handled as in
 * hmd_mlp_record_frame. */
hmd_status hmd_mlp_judge(hmd_forcefield* nn_ff, hmd_forcefield* fine_ff,
                         double* rel_rms);

/* ---- output formats ------------------------------------------------------
 * Extended-XYZ trajectory (src/io/trajectory.rs:39-176) and CSV log
 * (src/io/log_output.rs:36-110), byte for byte.  Values are atomic units;
 * the writer converts them to the chosen unit system (src/units/mod.rs:
 * 27-127) and formats Real columns with HMD_REAL_FMT_DIGITS fractional
 * digits (src/precision.rs:72, the f32 build).  Files are created new:
 * an existing path is HMD_ERR_OUTPUT_EXISTS. */
#define HMD_REAL_FMT_DIGITS 9
enum { HMD_UNITS_SI = 0, HMD_UNITS_ATOMIC = 1 };   /* UnitSystem */
enum {                                             /* Dimension */
  HMD_DIM_DIMENSIONLESS = 0, HMD_DIM_LENGTH, HMD_DIM_INVERSE_LENGTH, HMD_DIM_MASS,
  HMD_DIM_CHARGE, HMD_DIM_ENERGY, HMD_DIM_TIME, HMD_DIM_INVERSE_TIME, HMD_DIM_FORCE,
  HMD_DIM_PRESSURE, HMD_DIM_INVERSE_PRESSURE, HMD_DIM_TEMPERATURE, HMD_DIM_VELOCITY,
  HMD_DIM_STIFFNESS
};
typedef struct hmd_traj hmd_traj;
typedef struct hmd_log hmd_log;
/* TrajectoryWriter::open (trajectory.rs:39-62) */
hmd_status hmd_traj_open(const char* path, uint32_t units, int32_t include_velocities,
                         int32_t include_images, const char* const* type_names,
                         uint32_t n_type_names, hmd_traj** out);
/* TrajectoryWriter::write_frame (trajectory.rs:64-168); lattice
 * {lx, ly, lz, xy, xz, yz}; velocities / images may be NULL when the
 * writer was opened without them. */
hmd_status hmd_traj_write_frame(hmd_traj* t, uint64_t step, double dt, const float lattice[6],
                                uint32_t n, const uint32_t* type_idx, const float* x,
                                const float* y, const float* z, const float* vx, const float* vy,
                                const float* vz, const int32_t* ix, const int32_t* iy,
                                const int32_t* iz);
hmd_status hmd_traj_flush(hmd_traj* t);         /* TrajectoryWriter::flush */
void hmd_traj_close(hmd_traj* t);               /* flushes (Drop) */
/* LogWriter::open (log_output.rs:39-66): extra columns with their
 * dimensions (HMD_DIM_*) */
hmd_status hmd_log_open(const char* path, uint32_t units, const char* const* extra_names,
                        const uint32_t* extra_dims, uint32_t n_extra, hmd_log** out);
/* LogWriter::write_row (log_output.rs:74-101) */
hmd_status hmd_log_write_row(hmd_log* l, uint64_t step, double time, double kinetic_energy,
                             double temperature, const double* extras);
hmd_status hmd_log_flush(hmd_log* l);           /* LogWriter::flush */
void hmd_log_close(hmd_log* l);                 /* flushes (Drop) */

#ifdef __cplusplus
}
#endif
#endif /* HEDDLE_HIP_H */
