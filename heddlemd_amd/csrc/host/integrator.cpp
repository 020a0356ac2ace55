// Integrators (velocity Verlet f32 / compensated, Langevin BAOAB), SETTLE
// constraint hooks, hipGraph capture and replay, per-stage profiling, and
// setup helpers (steepest descent, velocity generation).
//
// Mirrors src/integrator/mod.rs (StepPlan, run_step), velocity_verlet.rs
// (:67-110), langevin_baoab.rs (:86-111), settle.rs (:310-404) and the
// runner's graph capture / batched replay (src/runner.rs:2321-2470).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>

#include "forcefield.h"

namespace hmd {

enum class Kind { VelocityVerlet, LangevinBaoab };

struct Integrator {
  Context& c;
  Particles& p;
  ForceField* ff;   // the force field the steps evaluate (a judged run may switch it)
  Kind kind = Kind::VelocityVerlet;
  float dt = 0;
  bool compensated = false;
  // SETTLE
  uint32_t n_groups = 0;
  float ra = 0, rb = 0, rc = 0, m_o = 0, m_h = 0;
  DevBuf<uint32_t> group_atoms;
  DevBuf<float> snapshot, constraint_virial;
  // Langevin
  float lan_a = 1.0f;
  float lan_kt = 0.0f;
  bool fused = true;
  // every step's evaluation at the energy level, SETTLE virial folded in
  bool scalars = false;
  DevBuf<uint32_t> free_atoms;   // atoms outside SETTLE groups (fused plan)
  uint32_t n_free = 0;
  uint64_t seed = 0;
  DevBuf<unsigned long long> step_counter;
  // CSVR thermostat
  bool csvr = false;
  double csvr_c = 1.0, csvr_kt_half = 0.0;
  uint32_t g_dof = 1;
  static constexpr uint32_t kCsvrPartials = 128;
  DevBuf<double> csvr_ke, csvr_partials, csvr_injection;
  DevBuf<float> csvr_factor;
  DevBuf<unsigned long long> csvr_counter;
  // c-rescale barostat (a per-step move after the last kick)
  bool baro = false;
  double baro_p = 0.0, baro_tau = 0.0, baro_beta = 0.0, baro_kt = 0.0;
  uint64_t baro_seed = 0;
  DevBuf<double> baro_ke, baro_w, baro_state;
  DevBuf<unsigned long long> baro_counter;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t steps = 0;
  bool primed = false;

  Integrator(Context& cc, Particles& pp, ForceField& f, const hmd_integrator_desc& d)
      : c(cc), p(pp), ff(&f) {
    std::string k = d.kind ? d.kind : "velocity-verlet";
    if (k == "velocity-verlet") kind = Kind::VelocityVerlet;
    else if (k == "langevin-baoab") kind = Kind::LangevinBaoab;
    else throw Error(HMD_ERR_INVALID, "unknown integrator kind '" + k + "'");
    dt = (float)d.dt;
    if (!(dt > 0.0f)) throw Error(HMD_ERR_INVALID, "dt must be positive");
    compensated = d.compensated != 0;
    if (compensated && kind != Kind::VelocityVerlet)
      throw Error(HMD_ERR_UNSUPPORTED, "compensated mode is velocity-Verlet only");
    hipStream_t s = c.stream;
    n_groups = d.n_groups;
    if (n_groups) {
      // canonical geometry (src/integrator/settle.rs:83-93)
      double total = d.m_o + 2.0 * d.m_h;
      double h = std::sqrt(d.d_oh * d.d_oh - 0.25 * d.d_hh * d.d_hh);
      ra = (float)(2.0 * d.m_h / total * h);
      rb = (float)(d.m_o / total * h);
      rc = (float)(0.5 * d.d_hh);
      m_o = (float)d.m_o;
      m_h = (float)d.m_h;
      for (uint32_t k2 = 0; k2 < 3 * n_groups; ++k2)
        if (d.group_atoms[k2] >= p.n) throw Error(HMD_ERR_INVALID, "SETTLE atom out of range");
      group_atoms.alloc(3 * (size_t)n_groups, false);
      group_atoms.upload(d.group_atoms, 3 * (size_t)n_groups, s);
      snapshot.alloc(12 * (size_t)n_groups);
      constraint_virial.alloc(3 * (size_t)n_groups);
    }
    fused = (d.flags & HMD_INTEGRATOR_UNFUSED) == 0;
    scalars = (d.flags & HMD_INTEGRATOR_SCALARS) != 0;
    if (n_groups) {
      std::vector<uint8_t> in_group(p.n, 0);
      for (uint32_t k2 = 0; k2 < 3 * n_groups; ++k2) {
        if (in_group[d.group_atoms[k2]]) throw Error(HMD_ERR_INVALID, "atom in two SETTLE groups");
        in_group[d.group_atoms[k2]] = 1;
      }
      std::vector<uint32_t> fr;
      for (uint32_t i = 0; i < p.n; ++i)
        if (!in_group[i]) fr.push_back(i);
      n_free = (uint32_t)fr.size();
      if (n_free) {
        free_atoms.alloc(n_free, false);
        free_atoms.upload(fr.data(), n_free, s);
      }
    } else {
      n_free = p.n;   // identity
    }
    if (kind == Kind::LangevinBaoab) {
      // alpha = exp(-gamma dt) in f32 (src/integrator/langevin_baoab.rs:140)
      lan_a = std::exp(-(float)d.gamma * (float)d.dt);
      lan_kt = (float)d.temperature;
      seed = d.seed;
      step_counter.alloc(1);
    }
    if (d.thermostat == HMD_THERMOSTAT_CSVR) {
      if (kind != Kind::VelocityVerlet)
        throw Error(HMD_ERR_INVALID, "the CSVR thermostat goes with velocity Verlet");
      if (!(d.temperature > 0.0) || !(d.tau > 0.0) || !std::isfinite(d.tau))
        throw Error(HMD_ERR_INVALID, "CSVR needs a positive temperature and tau");
      csvr = true;
      seed = d.seed;
      // thermal degrees of freedom 3N - n_constraints - 3, floored at 1
      // (csvr.rs:96-106); SETTLE fixes three distances per group
      const int64_t g = 3 * (int64_t)p.n - 3 * (int64_t)n_groups - 3;
      g_dof = (uint32_t)std::max<int64_t>(1, g);
      // from the f32 step, widened (csvr.rs:190)
      csvr_c = std::exp(-(double)dt / d.tau);
      csvr_kt_half = 0.5 * d.temperature;
      csvr_ke.alloc(1);
      csvr_partials.alloc(kCsvrPartials);
      csvr_injection.alloc(1);
      csvr_factor.alloc(1);
      csvr_counter.alloc(1);
    } else if (d.thermostat != HMD_THERMOSTAT_NONE) {
      throw Error(HMD_ERR_INVALID, "unknown thermostat");
    }
    check(hipMemsetAsync(p.pos_lo.ptr, 0, p.pos_lo.count * sizeof(double), s), "memset");
    check(hipMemsetAsync(p.vel_lo.ptr, 0, p.vel_lo.count * sizeof(double), s), "memset");
    check(hipStreamSynchronize(s), "sync");
  }

  // The thermostat after the step's last kick (apply_post, csvr.rs:164-206;
  // before the barostat and the SETTLE velocity projection): FP64 kinetic
  // energy into device memory, the CSVR chain, the rescale; enqueue only.
  void thermostat_post() {
    if (!csvr) return;
    ReduceParams r{};
    r.v = p.vel.ptr;
    r.m = p.mass.ptr;
    r.n = p.n;
    r.mode = 1;
    r.partials = p.red_partials.ptr;
    r.out = csvr_ke.ptr;
    r.n_blocks = (uint32_t)p.red_partials.count;
    launch(c, K_hmd_reduce_partial, r.n_blocks, 256, 0, r);
    launch(c, K_hmd_reduce_final, 1, 256, 0, r);
    CsvrParams q{};
    q.ke = csvr_ke.ptr;
    q.partials = csvr_partials.ptr;
    q.n_partials = kCsvrPartials;
    q.g_dof = g_dof;
    q.counter = csvr_counter.ptr;
    q.seed_lo = (uint32_t)(seed & 0xFFFFFFFFull);
    q.seed_hi = (uint32_t)(seed >> 32);
    q.c = csvr_c;
    q.one_minus_c = 1.0 - csvr_c;
    q.kt_half = csvr_kt_half;
    q.factor = csvr_factor.ptr;
    q.injection = csvr_injection.ptr;
    q.v = p.vel.ptr;
    q.vel_lo = compensated ? p.vel_lo.ptr : nullptr;
    q.n = p.n;
    launch(c, K_hmd_csvr_sample, kCsvrPartials, 256, 0, q);
    launch(c, K_hmd_csvr_finish, 1, 256, 0, q);
    launch(c, K_hmd_csvr_rescale, ceil_div(p.n, 256), 256, 0, q);
    c.mark("thermostat");
  }

  // c_rescale_barostat.rs:282-300 (validate_params / build): pressure
  // finite, temperature, tau and compressibility finite and positive.  A
  // per-step barostat reads the virial, so every force evaluation is at the
  // energy level from here on (runner.rs:2521-2539), with the SETTLE virial
  // folded in after the last kick.
  void set_barostat(const hmd_barostat_desc& d) {
    if (d.kind == HMD_BAROSTAT_NONE) {
      baro = false;
      drop_graph();
      return;
    }
    if (d.kind != HMD_BAROSTAT_C_RESCALE) throw Error(HMD_ERR_INVALID, "unknown barostat");
    auto finite_pos = [](double v) { return std::isfinite(v) && v > 0.0; };
    if (!std::isfinite(d.pressure)) throw Error(HMD_ERR_INVALID, "barostat pressure must be finite");
    if (!finite_pos(d.temperature) || !finite_pos(d.tau) || !finite_pos(d.compressibility))
      throw Error(HMD_ERR_INVALID,
                  "barostat temperature, tau and compressibility must be finite and positive");
    baro = true;
    baro_p = d.pressure;
    baro_kt = d.temperature;
    baro_tau = d.tau;
    baro_beta = d.compressibility;
    baro_seed = d.seed;
    scalars = true;
    if (!baro_state.count) {
      baro_ke.alloc(1);
      baro_w.alloc(1);
      baro_state.alloc(4);
      baro_counter.alloc(1);
    }
    drop_graph();
  }

  void drop_graph() {
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  }

  // BarostatPoint (c_rescale_barostat.rs:172-232): FP64 kinetic energy and
  // virial sums, the move on the device, then the new box.  The box is a
  // kernel argument here (the reference keeps it in device memory), so the
  // host reads mu back, scales the lattice by (float) mu as the device
  // lattice is scaled, and the position rescale and the SETTLE velocity
  // projection that follow see the new box.  Not capturable.
  void barostat_apply() {
    if (!baro) return;
    ReduceParams r{};
    r.v = p.vel.ptr;
    r.m = p.mass.ptr;
    r.n = p.n;
    r.mode = 1;
    r.partials = p.red_partials.ptr;
    r.out = baro_ke.ptr;
    r.n_blocks = (uint32_t)p.red_partials.count;
    launch(c, K_hmd_reduce_partial, r.n_blocks, 256, 0, r);
    launch(c, K_hmd_reduce_final, 1, 256, 0, r);
    r.mode = 0;
    r.x = p.virial.ptr;
    r.out = baro_w.ptr;
    launch(c, K_hmd_reduce_partial, r.n_blocks, 256, 0, r);
    launch(c, K_hmd_reduce_final, 1, 256, 0, r);
    const HmdLattice L = p.lattice;
    CRescaleParams q{};
    q.ke = baro_ke.ptr;
    q.w = baro_w.ptr;
    q.v_pre = (double)L.lx * (double)L.ly * (double)L.lz;
    q.counter = baro_counter.ptr;
    q.seed_lo = (uint32_t)(baro_seed & 0xFFFFFFFFull);
    q.seed_hi = (uint32_t)(baro_seed >> 32);
    q.p_target = baro_p;
    q.tau = baro_tau;
    q.beta = baro_beta;
    q.kt = baro_kt;
    q.dt = (double)dt;
    q.mu3_min = 1e-18;   // MU_MIN^3, MU_MIN = 1e-6 (c_rescale_barostat.rs:58)
    q.state = baro_state.ptr;
    q.posq = p.posq.ptr;
    q.pos_lo = compensated ? p.pos_lo.ptr : nullptr;
    q.n = p.n;
    launch(c, K_hmd_crescale_mu, 1, 64, 0, q);
    double st[4];
    baro_state.download(st, 4, c.stream);
    const float mu = (float)st[0];
    p.lattice = HmdLattice{L.lx * mu, L.ly * mu, L.lz * mu, L.xy * mu, L.xz * mu, L.yz * mu};
    launch(c, K_hmd_crescale_positions, ceil_div(p.n, 256), 256, 0, q);
    c.mark("barostat");
    box_moved = true;
  }
  bool box_moved = false;

  // The force fields on these particles follow a box the barostat moved:
  // a new SPME influence table and a list rebuild at the next evaluation.
  void publish_box() {
    if (!box_moved) return;
    box_moved = false;
    for (ForceField* f : p.observers) f->on_box_change();
  }

  ~Integrator() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }

  VvParams vv() const {
    VvParams v{};
    v.posq = p.posq.ptr;
    v.v = p.vel.ptr;
    v.f = p.force.ptr;
    v.m = p.mass.ptr;
    v.images = p.images.ptr;
    v.pos_lo = p.pos_lo.ptr;
    v.vel_lo = p.vel_lo.ptr;
    v.L = p.lattice;
    v.dt = dt;
    v.n = p.n;
    v.compensated = compensated ? 1u : 0u;
    return v;
  }

  SettleParams settle(float step_dt) const {
    SettleParams s{};
    s.posq = p.posq.ptr;
    s.v = p.vel.ptr;
    s.pos_lo = compensated ? p.pos_lo.ptr : nullptr;
    s.vel_lo = compensated ? p.vel_lo.ptr : nullptr;
    s.snapshot = snapshot.ptr;
    s.group_atoms = group_atoms.ptr;
    s.n_groups = n_groups;
    s.n = p.n;
    s.ra = ra; s.rb = rb; s.rc = rc; s.m_o = m_o; s.m_h = m_h;
    s.L = p.lattice;
    s.dt = step_dt;
    s.constraint_virial = constraint_virial.ptr;
    s.particle_virial = scalars ? p.virial.ptr : nullptr;
    return s;
  }

  LangevinParams lan(float step_dt) const {
    LangevinParams l{};
    l.posq = p.posq.ptr;
    l.v = p.vel.ptr;
    l.m = p.mass.ptr;
    l.kt = lan_kt;
    l.images = p.images.ptr;
    l.L = p.lattice;
    l.dt = step_dt;
    l.a = lan_a;
    l.seed_lo = (uint32_t)(seed & 0xFFFFFFFFull);
    l.seed_hi = (uint32_t)(seed >> 32);
    l.step = step_counter.ptr;
    l.n = p.n;
    return l;
  }

  void settle_snapshot() {
    if (!n_groups) return;
    launch(c, K_hmd_settle_snapshot, ceil_div(n_groups, 256), 256, 0, settle(dt));
  }
  void settle_positions(float step_dt) {
    if (!n_groups) return;
    launch(c, K_hmd_settle_positions, ceil_div(n_groups, 256), 256, 0, settle(step_dt));
  }
  void settle_velocities() {
    if (!n_groups) return;
    launch(c, K_hmd_settle_velocities, ceil_div(n_groups, 256), 256, 0, settle(dt));
  }
  // AfterKick of the step's last kick: the constraint virial into the
  // particle virials (energy-level steps only)
  void settle_virial() {
    if (!n_groups || !scalars) return;
    launch(c, K_hmd_settle_virial, ceil_div(n_groups, 256), 256, 0, settle(dt));
  }

  StepParams fused_params() const {
    StepParams sp{};
    sp.vv = vv();
    sp.st = settle(dt);
    sp.lan = lan(dt);
    sp.free_atoms = n_groups ? free_atoms.ptr : nullptr;
    sp.n_free = n_free;
    sp.dt = dt;
    sp.counter = kind == Kind::LangevinBaoab ? step_counter.ptr : nullptr;
    // the first fused kernel runs the force field's skin check
    sp.disp = ff->disp_params();
    return sp;
  }

  int level() const { return scalars ? 1 : 0; }

  // The trailing kick and the post-force markers in the reference's order
  // (src/integrator/mod.rs:500-517): the thermostat's apply_post after the
  // kick, then the barostat, then the SETTLE velocity projection
  // (velocity_verlet.rs:92-109).
  void post_force_tail(uint32_t nb) {
    launch(c, K_hmd_vv_kick, nb, 256, 0, vv());               // KickHalf
    thermostat_post();                                        // apply_post
    barostat_apply();                                         // BarostatPoint
    settle_velocities();                                      // AfterKick
    settle_virial();
    if (kind == Kind::LangevinBaoab) {
      CounterParams cp{step_counter.ptr};
      launch(c, K_hmd_counter_increment, 1, 64, 0, cp);
    }
    c.mark("kick_settle");
  }

  // One step, enqueued only (capturable).
  void enqueue_step() {
    const uint32_t nb = ceil_div(p.n, 256);
    if (fused) {
      const StepParams sp = fused_params();
      const uint32_t nu = ceil_div((uint64_t)n_groups + n_free, 256);
      launch(c, kind == Kind::VelocityVerlet ? K_hmd_step_vv_first : K_hmd_step_baoa, nu, 256,
             0, sp);
      c.mark(kind == Kind::VelocityVerlet ? "kick_drift_settle" : "baoa");
      ff->compute(level(), true);
      if (csvr || baro) {
        post_force_tail(nb);
        return;
      }
      launch(c, K_hmd_step_kick_last, nu, 256, 0, sp);
      c.mark("kick_settle");
      return;
    }
    if (kind == Kind::VelocityVerlet) {
      settle_snapshot();                                        // BeforeDrift
      launch(c, K_hmd_vv_kick_drift, nb, 256, 0, vv());         // KickDrift
      settle_positions(dt);                                     // AfterDrift
      c.mark("kick_drift_settle");
      ff->compute(level());                                      // ForceEval
      post_force_tail(nb);
    } else {
      // B A O A (force) B, with SETTLE after each position and velocity
      // update (the reference's BAOAB plan has no constraint points).
      VvParams k = vv();
      launch(c, K_hmd_vv_kick, nb, 256, 0, k);
      settle_velocities();
      settle_snapshot();
      launch(c, K_hmd_lan_drift_half, nb, 256, 0, lan(dt));
      settle_positions(0.5f * dt);
      launch(c, K_hmd_lan_ou, nb, 256, 0, lan(dt));
      settle_velocities();
      settle_snapshot();
      launch(c, K_hmd_lan_drift_half, nb, 256, 0, lan(dt));
      settle_positions(0.5f * dt);
      c.mark("baoa");
      ff->compute(level());
      post_force_tail(nb);
    }
  }

  void prime() {
    if (primed) return;
    ff->compute(level());
    check(hipStreamSynchronize(c.stream), "sync");
    ff->check_status();
    primed = true;
  }

  void step(uint32_t n) {
    prime();
    for (uint32_t i = 0; i < n; ++i) {
      enqueue_step();
      publish_box();
    }
    check(hipStreamSynchronize(c.stream), "sync");
    steps += n;
    ff->check_status();
  }

  void capture() {
    if (baro)
      throw Error(HMD_ERR_UNSUPPORTED,
                  "a per-step barostat moves the box every step: its steps run eagerly");
    prime();
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    check(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      enqueue_step();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(c.stream, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    ff->graph_invalid = false;
  }

  // Steps from here on evaluate forces with `f` (same particles): the
  // forces in the buffer are re-evaluated with it at the current positions
  // (the next kick uses them) and the step graph is captured anew.
  void set_forcefield(ForceField& f) {
    if (&f.p != &p) throw Error(HMD_ERR_INVALID, "force field is on other particles");
    ff = &f;
    ff->compute(level());
    check(hipStreamSynchronize(c.stream), "sync");
    ff->check_status();
    primed = true;
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  }

  void replay(uint32_t n) {
    if (baro) {   // the box is a kernel argument: barostat steps are eager
      step(n);
      return;
    }
    if (!exec || ff->graph_invalid) capture();
    for (uint32_t i = 0; i < n; ++i) check(hipGraphLaunch(exec, c.stream), "hipGraphLaunch");
    check(hipStreamSynchronize(c.stream), "sync");
    steps += n;
    ff->check_status();
  }
};

// ------------------------------------------------------- replica batches ---
// Independent systems, each with its own context (stream, code object),
// particles, force field and integrator, stepped as one batch.  In the fused
// mode every replica's step is captured into ONE graph: the master stream
// forks to each replica's stream and joins them back, so a batched step is
// a single graph launch whose branches the runtime may run side by side on
// its hardware queues.  In the per-replica mode each replica keeps its own
// step graph, launched on its own stream, all replicas before any wait.  In
// the batched mode every replica's step is recorded (not launched), and the
// records are merged into one launch per stage: the kernel's _batch twin over
// the blocks of all replicas that reach that stage together, each block
// reading its replica's parameter block from device memory.  The merged
// sequence is captured into one graph on the master stream.
struct ReplicaSet {
  enum class Mode { Fused, PerReplica, Batched };
  std::vector<Integrator*> its;
  Mode mode = Mode::Fused;
  // batched mode: parameter blocks and block tables of the merged launches
  DevBuf<unsigned char> batch_params;
  DevBuf<uint32_t> batch_starts;
  uint32_t n_batched = 0, n_single = 0;   // launches per step of the merged plan
  // the merged plan as launches (kernel, shape, argument bytes); eager:
  // issued directly every step instead of through a graph (kernel tracers)
  struct Planned {
    hipFunction_t fn;
    uint32_t grid, block;
    size_t shmem;
    std::vector<unsigned char> args;
    std::string name;
  };
  std::vector<Planned> planned;
  bool eager = false;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t steps = 0;

  ReplicaSet(std::vector<Integrator*> v, Mode m) : its(std::move(v)), mode(m) {
    if (its.empty()) throw Error(HMD_ERR_INVALID, "empty replica set");
    const int dev = its[0]->c.device;
    for (auto* it : its) {
      if (it->c.device != dev) throw Error(HMD_ERR_INVALID, "replicas on different devices");
      for (auto* jt : its)
        if (jt != it && jt->c.stream == it->c.stream)
          throw Error(HMD_ERR_INVALID, "each replica needs its own context");
    }
    check(hipSetDevice(dev), "hipSetDevice");
    check(hipStreamCreateWithFlags(&master, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
    joins.resize(its.size());
    for (auto& e : joins) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  }

  ~ReplicaSet() {
    release();
    for (auto& e : joins) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (master) (void)hipStreamDestroy(master);
  }

  void release() {
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
  }

  void capture() {
    for (auto* it : its) it->prime();
    release();
    check(hipStreamBeginCapture(master, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      check(hipEventRecord(fork, master), "hipEventRecord");
      for (size_t r = 0; r < its.size(); ++r) {
        Integrator& I = *its[r];
        check(hipStreamWaitEvent(I.c.stream, fork, 0), "hipStreamWaitEvent");
        I.enqueue_step();
        check(hipEventRecord(joins[r], I.c.stream), "hipEventRecord");
        check(hipStreamWaitEvent(master, joins[r], 0), "hipStreamWaitEvent");
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(master, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(master, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    for (auto* it : its) it->ff->graph_invalid = false;
  }

  // Two recorded launches can share a batched launch: the same kernel of
  // the shared code object (each replica's context loads its own copy, so
  // the function handles differ) at the same block size.  Kernels without a
  // batched twin (composed pair kernels) pair up only with themselves.
  static bool same_stage(const LaunchRec& x, const LaunchRec& y) {
    if (x.block != y.block) return false;
    if (x.fn_batch && y.fn_batch) return x.k == y.k;
    return x.fn == y.fn;
  }

  // Batched mode: record each replica's step, merge, capture one graph.
  // Merging is greedy over the replicas' next unissued launches: the
  // (kernel, block size) at the head of the replica with the most launches
  // left is issued next, as one batched launch over every replica whose
  // next launch it is (ties go to the head shared by more replicas).
  // Replicas whose step has extra launches (a multi-tile scan) thus catch
  // up before the others move on, and each stage stays one launch.  Each
  // replica's launches keep their order, so every replica sees exactly its
  // own step.
  void capture_batched(bool capture_graph = true) {
    for (auto* it : its) it->prime();
    release();
    const size_t R = its.size();
    std::vector<std::vector<LaunchRec>> recs(R);
    for (size_t r = 0; r < R; ++r) {
      Context& c = its[r]->c;
      c.rec = &recs[r];
      try {
        its[r]->enqueue_step();
      } catch (...) {
        c.rec = nullptr;
        throw;
      }
      c.rec = nullptr;
    }
    struct Item {
      const LaunchRec* first;                // kernel, block size
      std::vector<const LaunchRec*> members;
      size_t param_off = 0, start_off = 0;   // into the device tables
      uint32_t blocks = 0;
      size_t shmem = 0;
    };
    std::vector<Item> plan;
    std::vector<size_t> head(R, 0);
    for (;;) {
      int best = -1;
      size_t best_count = 0, best_left = 0;
      for (size_t r = 0; r < R; ++r) {
        if (head[r] >= recs[r].size()) continue;
        const LaunchRec& a = recs[r][head[r]];
        size_t count = 0;
        for (size_t q = 0; q < R; ++q)
          if (head[q] < recs[q].size() && same_stage(recs[q][head[q]], a)) ++count;
        const size_t left = recs[r].size() - head[r];
        if (left > best_left || (left == best_left && count > best_count)) {
          best = (int)r;
          best_count = count;
          best_left = left;
        }
      }
      if (best < 0) break;
      const LaunchRec& a = recs[best][head[best]];
      Item item;
      item.first = &a;
      for (size_t q = 0; q < R; ++q)
        if (head[q] < recs[q].size() && same_stage(recs[q][head[q]], a))
          item.members.push_back(&recs[q][head[q]++]);
      plan.push_back(std::move(item));
    }
    // device tables for the batched items (parameter blocks 256-B aligned)
    std::vector<unsigned char> hp;
    std::vector<uint32_t> hs;
    n_batched = n_single = 0;
    for (Item& it : plan) {
      if (it.members.size() < 2 || !it.first->fn_batch) {
        n_single += (uint32_t)it.members.size();
        continue;
      }
      ++n_batched;
      const size_t stride = it.first->params.size();
      it.param_off = (hp.size() + 255) / 256 * 256;
      hp.resize(it.param_off + stride * it.members.size());
      it.start_off = hs.size();
      uint64_t total = 0;
      for (size_t m = 0; m < it.members.size(); ++m) {
        const LaunchRec* L = it.members[m];
        if (L->params.size() != stride) throw Error(HMD_ERR_INVALID, "batched parameter size");
        std::memcpy(&hp[it.param_off + m * stride], L->params.data(), stride);
        hs.push_back((uint32_t)total);
        total += L->grid;
        it.shmem = std::max(it.shmem, L->shmem);
      }
      if (total > 0x7FFFFFFFull) throw Error(HMD_ERR_INVALID, "batched launch too large");
      hs.push_back((uint32_t)total);
      it.blocks = (uint32_t)total;
    }
    batch_params.alloc(std::max<size_t>(hp.size(), 1), false);
    batch_starts.alloc(std::max<size_t>(hs.size(), 1), false);
    if (!hp.empty())
      check(hipMemcpy(batch_params.ptr, hp.data(), hp.size(), hipMemcpyHostToDevice), "H2D");
    if (!hs.empty())
      check(hipMemcpy(batch_starts.ptr, hs.data(), hs.size() * 4, hipMemcpyHostToDevice), "H2D");
    // null-stream copies: landed before the replicas' non-blocking streams read them
    check(hipStreamSynchronize(nullptr), "sync");
    planned.clear();
    for (const Item& it : plan) {
      if (it.blocks) {
        HmdBatch hb{batch_params.ptr + it.param_off, batch_starts.ptr + it.start_off,
                    (uint32_t)it.members.size(), (uint32_t)it.first->params.size()};
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&hb);
        planned.push_back(Planned{it.first->fn_batch, it.blocks, it.first->block, it.shmem,
                                  std::vector<unsigned char>(b, b + sizeof hb),
                                  std::string(kKernelNames[it.first->k]) + "_batch"});
      } else {
        for (const LaunchRec* L : it.members)
          planned.push_back(Planned{L->fn, L->grid, L->block, L->shmem, L->params,
                                    kKernelNames[L->k]});
      }
    }
    for (auto* it : its) it->ff->graph_invalid = false;
    if (!capture_graph) return;
    check(hipStreamBeginCapture(master, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      issue_planned();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(master, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(master, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }

  // The merged plan's launches, in order, on the master stream.
  void issue_planned() {
    for (Planned& L : planned) {
      void* args[] = {L.args.data()};
      check(hipModuleLaunchKernel(L.fn, L.grid, 1, 1, L.block, 1, 1, (unsigned)L.shmem, master,
                                  args, nullptr),
            L.name.c_str());
    }
  }

  void step(uint32_t n) {
    if (mode == Mode::Batched && eager) {
      bool stale = planned.empty();
      for (auto* it : its) stale = stale || it->ff->graph_invalid;
      if (stale) capture_batched(false);
      for (uint32_t i = 0; i < n; ++i) issue_planned();
      check(hipStreamSynchronize(master), "sync");
    } else if (mode != Mode::PerReplica) {
      bool stale = !exec;
      for (auto* it : its) stale = stale || it->ff->graph_invalid;
      if (stale) {
        if (mode == Mode::Batched) capture_batched();
        else capture();
      }
      for (uint32_t i = 0; i < n; ++i) check(hipGraphLaunch(exec, master), "hipGraphLaunch");
      check(hipStreamSynchronize(master), "sync");
    } else {
      for (auto* it : its)
        if (!it->exec || it->ff->graph_invalid) it->capture();
      for (uint32_t i = 0; i < n; ++i)
        for (auto* it : its) check(hipGraphLaunch(it->exec, it->c.stream), "hipGraphLaunch");
      for (auto* it : its) check(hipStreamSynchronize(it->c.stream), "sync");
    }
    steps += n;
    for (auto* it : its) it->steps += n;
    // every replica's status is read (and grown on overflow) before the
    // first error is reported, so one re-run of the batch suffices
    std::string err;
    hmd_status code = HMD_OK;
    for (size_t r = 0; r < its.size(); ++r) {
      try {
        its[r]->ff->check_status();
      } catch (const Error& e) {
        if (code == HMD_OK) {
          code = e.code;
          err = "replica " + std::to_string(r) + ": " + e.what();
        }
      }
    }
    if (code != HMD_OK) throw Error(code, err);
  }
};

// Device copy of everything a batch of steps changes (positions, velocities,
// images, forces, the compensated low parts, the random-stream and thermostat
// counters), taken at the batch's start so that a batch that reported a list
// overflow is re-run from there: the caller-side protocol of
// HMD_ERR_NBL_OVERFLOW, kept inside entry points that run several batches.
struct BatchStart {
  DevBuf<float> posq, vel, force;
  DevBuf<int32_t> images;
  DevBuf<double> pos_lo, vel_lo, injection;
  DevBuf<unsigned long long> step_counter, csvr_counter, draw_counter, baro_counter;
  DevBuf<double> baro_state;
  HmdLattice lattice{};
  uint64_t steps = 0;

  template <class T>
  static void copy(DevBuf<T>& dst, const DevBuf<T>& src, hipStream_t s) {
    if (!src.count) return;
    if (dst.count != src.count) dst.alloc(src.count, false);
    check(hipMemcpyAsync(dst.ptr, src.ptr, src.count * sizeof(T), hipMemcpyDeviceToDevice, s),
          "D2D");
  }
  void save(Integrator& I) {
    Particles& p = I.p;
    hipStream_t s = I.c.stream;
    copy(posq, p.posq, s);
    copy(vel, p.vel, s);
    copy(force, p.force, s);
    copy(images, p.images, s);
    copy(pos_lo, p.pos_lo, s);
    copy(vel_lo, p.vel_lo, s);
    copy(draw_counter, p.draw_counter, s);
    copy(step_counter, I.step_counter, s);
    copy(csvr_counter, I.csvr_counter, s);
    copy(injection, I.csvr_injection, s);
    copy(baro_counter, I.baro_counter, s);
    copy(baro_state, I.baro_state, s);
    lattice = p.lattice;
    steps = I.steps;
  }
  void restore(Integrator& I) {
    Particles& p = I.p;
    hipStream_t s = I.c.stream;
    copy(p.posq, posq, s);
    copy(p.vel, vel, s);
    copy(p.force, force, s);
    copy(p.images, images, s);
    copy(p.pos_lo, pos_lo, s);
    copy(p.vel_lo, vel_lo, s);
    copy(p.draw_counter, draw_counter, s);
    copy(I.step_counter, step_counter, s);
    copy(I.csvr_counter, csvr_counter, s);
    copy(I.csvr_injection, injection, s);
    copy(I.baro_counter, baro_counter, s);
    copy(I.baro_state, baro_state, s);
    check(hipStreamSynchronize(s), "sync");
    I.steps = steps;
    const HmdLattice& L = p.lattice;
    if (L.lx != lattice.lx || L.ly != lattice.ly || L.lz != lattice.lz || L.xy != lattice.xy ||
        L.xz != lattice.xz || L.yz != lattice.yz) {
      p.lattice = lattice;   // a barostat moved the box within the batch
      for (ForceField* f : p.observers) f->on_box_change();
    }
    I.ff->restart_list();
  }
};

}  // namespace hmd

struct hmd_integrator {
  hmd::Integrator* it = nullptr;
  std::unique_ptr<hmd::BatchStart> mark;   // hmd_integrator_checkpoint
};

struct hmd_replicas {
  hmd::ReplicaSet* set = nullptr;
};

// A SETTLE constraint set on its own: the three hooks of the reference's
// Constraint interface (src/integrator/settle.rs:310-404) as operators.
struct hmd_settle {
  hmd::Particles* p = nullptr;
  hmd::DevBuf<uint32_t> group_atoms;
  hmd::DevBuf<float> snapshot, constraint_virial;
  uint32_t n_groups = 0;
  float ra = 0, rb = 0, rc = 0, m_o = 0, m_h = 0;
  SettleParams params(float dt) const {
    SettleParams s{};
    s.posq = p->posq.ptr;
    s.v = p->vel.ptr;
    s.snapshot = snapshot.ptr;
    s.group_atoms = group_atoms.ptr;
    s.n_groups = n_groups;
    s.n = p->n;
    s.ra = ra; s.rb = rb; s.rc = rc; s.m_o = m_o; s.m_h = m_h;
    s.L = p->lattice;
    s.dt = dt;
    s.constraint_virial = constraint_virial.ptr;
    return s;
  }
};

using namespace hmd;

// ------------------------------------------------------ velocity set-up ---
// ChaCha block function (Bernstein 2008), `rounds` rounds on a raw state.
static inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static void chacha_block(const uint32_t in[16], uint32_t rounds, uint32_t out[16]) {
  uint32_t x[16];
  std::memcpy(x, in, sizeof x);
#define QR(a, b, c, d)                   \
  a += b; d ^= a; d = rotl(d, 16);       \
  c += d; b ^= c; b = rotl(b, 12);       \
  a += b; d ^= a; d = rotl(d, 8);        \
  c += d; b ^= c; b = rotl(b, 7);
  for (uint32_t r = 0; r < rounds; r += 2) {
    QR(x[0], x[4], x[8], x[12]) QR(x[1], x[5], x[9], x[13])
    QR(x[2], x[6], x[10], x[14]) QR(x[3], x[7], x[11], x[15])
    QR(x[0], x[5], x[10], x[15]) QR(x[1], x[6], x[11], x[12])
    QR(x[2], x[7], x[8], x[13]) QR(x[3], x[4], x[9], x[14])
  }
#undef QR
  for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

// ChaCha8 stream seeded like rand_core 0.6 `seed_from_u64` (a PCG32 pass
// fills the 32-byte key) and read 64 bits at a time, low word first,
// through a 4-block buffer (rand_chacha 0.3), with rand 0.8's f64 mapping
// (u64 >> 11) * 2^-53.
struct ChaCha8 {
  uint32_t state[16];
  uint32_t buf[64];
  uint32_t idx = 64;
  explicit ChaCha8(uint64_t seed) {
    uint32_t key[8];
    uint64_t st = seed;
    for (int i = 0; i < 8; ++i) {
      st = st * 6364136223846793005ull + 11634580027462260723ull;
      uint32_t xorshifted = (uint32_t)(((st >> 18) ^ st) >> 27);
      uint32_t rot = (uint32_t)(st >> 59);
      key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
    state[0] = 0x61707865u; state[1] = 0x3320646eu;
    state[2] = 0x79622d32u; state[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i) state[4 + i] = key[i];
    state[12] = state[13] = state[14] = state[15] = 0u;
  }
  void refill() {
    for (int b = 0; b < 4; ++b) {
      chacha_block(state, 8, buf + 16 * b);
      uint64_t ctr = ((uint64_t)state[13] << 32 | state[12]) + 1;
      state[12] = (uint32_t)ctr;
      state[13] = (uint32_t)(ctr >> 32);
    }
    idx = 0;
  }
  uint64_t next_u64() {
    if (idx + 2 > 64) refill();
    uint64_t lo = buf[idx], hi = buf[idx + 1];
    idx += 2;
    return (hi << 32) | lo;
  }
  double next_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

extern "C" {

void hmd_chacha_block(const uint32_t state[16], uint32_t rounds, uint32_t out[16]) {
  chacha_block(state, rounds, out);
}

// src/runner.rs:3108-3200
hmd_status hmd_generate_velocities(uint32_t n, uint32_t n_constraints, double temperature,
                                   uint64_t seed, const double* masses, float* vel) {
  HMD_GUARD({
    std::vector<float> v(3 * (size_t)n, 0.0f);
    if (temperature != 0.0 && n != 0) {
      ChaCha8 rng(seed);
      for (uint32_t i = 0; i < n; ++i) {
        double sigma = std::sqrt(temperature / masses[i]);
        for (int a = 0; a < 3; ++a) {
          double u1 = 1.0 - rng.next_f64();
          double u2 = rng.next_f64();
          double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
          v[3 * (size_t)i + a] = (float)(z * sigma);
        }
      }
      double total = 0.0;
      for (uint32_t i = 0; i < n; ++i) total += masses[i];
      if (total > 0.0)
        for (int a = 0; a < 3; ++a) {
          double pm = 0.0;
          for (uint32_t i = 0; i < n; ++i) pm += masses[i] * (double)v[3 * (size_t)i + a];
          double vc = pm / total;
          for (uint32_t i = 0; i < n; ++i)
            v[3 * (size_t)i + a] = (float)((double)v[3 * (size_t)i + a] - vc);
        }
      int64_t dof = 3 * (int64_t)n - (int64_t)n_constraints - 3;
      if (dof <= 0) {
        std::fill(v.begin(), v.end(), 0.0f);
      } else {
        double ke = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
          double s = 0.0;
          for (int a = 0; a < 3; ++a) s += std::pow((double)v[3 * (size_t)i + a], 2);
          ke += 0.5 * masses[i] * s;
        }
        if (ke > 0.0) {
          double scale = std::sqrt(0.5 * (double)dof * temperature / ke);
          for (auto& x : v) x = (float)((double)x * scale);
        }
      }
    }
    std::copy(v.begin(), v.end(), vel);
  });
}

static hmd_status vv_half(hmd_particles* hp, float dt, int32_t compensated, bool drift) {
  HMD_GUARD({
    if (!hp) throw Error(HMD_ERR_INVALID, "null particles");
    Particles& p = hp->p;
    VvParams v{};
    v.posq = p.posq.ptr;
    v.v = p.vel.ptr;
    v.f = p.force.ptr;
    v.m = p.mass.ptr;
    v.images = p.images.ptr;
    v.pos_lo = p.pos_lo.ptr;
    v.vel_lo = p.vel_lo.ptr;
    v.L = p.lattice;
    v.dt = dt;
    v.n = p.n;
    v.compensated = compensated ? 1u : 0u;
    launch(*p.ctx, drift ? K_hmd_vv_kick_drift : K_hmd_vv_kick, ceil_div(p.n, 256), 256, 0, v);
    check(hipStreamSynchronize(p.ctx->stream), "sync");
  });
}

hmd_status hmd_settle_create(hmd_particles* hp, uint32_t n_groups, const uint32_t* group_atoms,
                             double d_oh, double d_hh, double m_o, double m_h, hmd_settle** out) {
  HMD_GUARD({
    if (!hp || !out || (n_groups && !group_atoms)) throw Error(HMD_ERR_INVALID, "null argument");
    if (!(d_oh > 0.0 && d_hh > 0.0 && d_hh < 2.0 * d_oh && m_o > 0.0 && m_h > 0.0))
      throw Error(HMD_ERR_INVALID, "SETTLE geometry: need 0 < d_HH < 2 d_OH and positive masses");
    for (uint64_t k = 0; k < 3ull * n_groups; ++k)
      if (group_atoms[k] >= hp->p.n) throw Error(HMD_ERR_INVALID, "SETTLE atom out of range");
    auto* h = new hmd_settle();
    h->p = &hp->p;
    h->n_groups = n_groups;
    // canonical geometry (src/integrator/settle.rs:83-93)
    const double total = m_o + 2.0 * m_h;
    const double ht = std::sqrt(d_oh * d_oh - 0.25 * d_hh * d_hh);
    h->ra = (float)(2.0 * m_h / total * ht);
    h->rb = (float)(m_o / total * ht);
    h->rc = (float)(0.5 * d_hh);
    h->m_o = (float)m_o;
    h->m_h = (float)m_h;
    if (n_groups) {
      hipStream_t s = hp->p.ctx->stream;
      h->group_atoms.alloc(3 * (size_t)n_groups, false);
      h->group_atoms.upload(group_atoms, 3 * (size_t)n_groups, s);
      h->snapshot.alloc(12 * (size_t)n_groups);
      h->constraint_virial.alloc(3 * (size_t)n_groups);
      check(hipStreamSynchronize(s), "sync");
    }
    *out = h;
  });
}

void hmd_settle_destroy(hmd_settle* h) { delete h; }

static hmd_status settle_op(hmd_settle* h, float dt, KernelId k) {
  HMD_GUARD({
    if (!h) throw Error(HMD_ERR_INVALID, "null settle");
    if (h->n_groups == 0) return HMD_OK;
    Context& c = *h->p->ctx;
    launch(c, k, ceil_div(h->n_groups, 256), 256, 0, h->params(dt));
    check(hipStreamSynchronize(c.stream), "sync");
  });
}

hmd_status hmd_settle_before_drift(hmd_settle* h) {
  return settle_op(h, 0.0f, K_hmd_settle_snapshot);
}
hmd_status hmd_settle_after_drift(hmd_settle* h, float dt) {
  return settle_op(h, dt, K_hmd_settle_positions);
}
hmd_status hmd_settle_after_kick(hmd_settle* h, float dt) {
  return settle_op(h, dt, K_hmd_settle_velocities);
}

hmd_status hmd_lan_drift_half(hmd_particles* hp, float dt) {
  HMD_GUARD({
    if (!hp) throw Error(HMD_ERR_INVALID, "null particles");
    Particles& p = hp->p;
    LangevinParams l{};
    l.posq = p.posq.ptr;
    l.v = p.vel.ptr;
    l.images = p.images.ptr;
    l.L = p.lattice;
    l.dt = dt;
    l.n = p.n;
    launch(*p.ctx, K_hmd_lan_drift_half, ceil_div(p.n, 256), 256, 0, l);
    check(hipStreamSynchronize(p.ctx->stream), "sync");
  });
}

hmd_status hmd_lan_ou_step(hmd_particles* hp, uint64_t seed, float alpha, float kt,
                           uint64_t draw_counter) {
  HMD_GUARD({
    if (!hp) throw Error(HMD_ERR_INVALID, "null particles");
    Particles& p = hp->p;
    if (p.n == 0) return HMD_OK;
    unsigned long long c = draw_counter;
    p.draw_counter.upload(&c, 1, p.ctx->stream);
    LangevinParams l{};
    l.v = p.vel.ptr;
    l.m = p.mass.ptr;
    l.a = alpha;
    l.kt = kt;
    l.seed_lo = (uint32_t)(seed & 0xFFFFFFFFull);
    l.seed_hi = (uint32_t)(seed >> 32);
    l.step = p.draw_counter.ptr;
    l.n = p.n;
    launch(*p.ctx, K_hmd_lan_ou, ceil_div(p.n, 256), 256, 0, l);
    check(hipStreamSynchronize(p.ctx->stream), "sync");
  });
}

hmd_status hmd_vv_kick_drift(hmd_particles* hp, float dt, int32_t compensated) {
  return vv_half(hp, dt, compensated, true);
}

hmd_status hmd_vv_kick(hmd_particles* hp, float dt, int32_t compensated) {
  return vv_half(hp, dt, compensated, false);
}

hmd_status hmd_integrator_create(hmd_ctx* h, hmd_particles* hp, hmd_forcefield* f,
                                 const hmd_integrator_desc* d, hmd_integrator** out) {
  HMD_GUARD({
    if (!d) throw Error(HMD_ERR_INVALID, "null descriptor");
    auto* w = new hmd_integrator();
    try {
      w->it = new Integrator(h->c, hp->p, *f->ff, *d);
    } catch (...) {
      delete w;
      throw;
    }
    *out = w;
  });
}

void hmd_integrator_destroy(hmd_integrator* w) {
  if (!w) return;
  delete w->it;
  delete w;
}

hmd_status hmd_replicas_create(hmd_integrator* const* its, uint32_t n, uint32_t flags,
                               hmd_replicas** out) {
  HMD_GUARD({
    if (!its || !out || n == 0) throw Error(HMD_ERR_INVALID, "null or empty replica list");
    std::vector<Integrator*> v;
    for (uint32_t r = 0; r < n; ++r) {
      if (!its[r]) throw Error(HMD_ERR_INVALID, "null integrator");
      if (its[r]->it->baro)
        throw Error(HMD_ERR_UNSUPPORTED,
                    "replica batches are captured: a per-step barostat's steps run eagerly");
      v.push_back(its[r]->it);
    }
    auto* h = new hmd_replicas();
    try {
      if ((flags & HMD_REPLICAS_PER_REPLICA_GRAPHS) && (flags & HMD_REPLICAS_BATCHED))
        throw Error(HMD_ERR_INVALID, "replica modes are exclusive");
      const ReplicaSet::Mode m = (flags & HMD_REPLICAS_BATCHED) ? ReplicaSet::Mode::Batched
                                 : (flags & HMD_REPLICAS_PER_REPLICA_GRAPHS)
                                     ? ReplicaSet::Mode::PerReplica
                                     : ReplicaSet::Mode::Fused;
      if ((flags & HMD_REPLICAS_EAGER) && m != ReplicaSet::Mode::Batched)
        throw Error(HMD_ERR_INVALID, "HMD_REPLICAS_EAGER needs HMD_REPLICAS_BATCHED");
      h->set = new ReplicaSet(std::move(v), m);
      h->set->eager = (flags & HMD_REPLICAS_EAGER) != 0;
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void hmd_replicas_destroy(hmd_replicas* h) {
  if (!h) return;
  delete h->set;
  delete h;
}

hmd_status hmd_replicas_step(hmd_replicas* h, uint32_t n) {
  HMD_GUARD({
    for (Integrator* I : h->set->its)
      if (I->baro)
        throw Error(HMD_ERR_UNSUPPORTED,
                    "replica batches are captured: a per-step barostat's steps run eagerly");
    h->set->step(n);
  });
}
uint64_t hmd_replicas_steps(hmd_replicas* h) { return h->set->steps; }

hmd_status hmd_replicas_plan(hmd_replicas* h, uint32_t* n_batched, uint32_t* n_single) {
  HMD_GUARD({
    if (!h || !n_batched || !n_single) throw Error(HMD_ERR_INVALID, "null argument");
    *n_batched = h->set->n_batched;
    *n_single = h->set->n_single;
  });
}

hmd_status hmd_integrator_step(hmd_integrator* w, uint32_t n) { HMD_GUARD(w->it->step(n)); }
hmd_status hmd_integrator_capture(hmd_integrator* w) { HMD_GUARD(w->it->capture()); }
hmd_status hmd_integrator_replay(hmd_integrator* w, uint32_t n) { HMD_GUARD(w->it->replay(n)); }

hmd_status hmd_integrator_thermostat_state(hmd_integrator* w, double out[2]) {
  HMD_GUARD({
    Integrator& I = *w->it;
    out[0] = 0.0;
    out[1] = 0.0;
    if (I.csvr) {
      I.csvr_injection.download(&out[0], 1, I.c.stream);
      unsigned long long k = 0;
      I.csvr_counter.download(&k, 1, I.c.stream);
      out[1] = (double)k;
    }
  });
}

hmd_status hmd_integrator_set_barostat(hmd_integrator* w, const hmd_barostat_desc* d) {
  HMD_GUARD({
    if (!w || !d) throw Error(HMD_ERR_INVALID, "null argument");
    w->it->set_barostat(*d);
  });
}

hmd_status hmd_integrator_barostat_state(hmd_integrator* w, double out[5]) {
  HMD_GUARD({
    if (!w || !out) throw Error(HMD_ERR_INVALID, "null argument");
    Integrator& I = *w->it;
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    if (I.baro_state.count) {
      double st[4];
      I.baro_state.download(st, 4, I.c.stream);
      unsigned long long k = 0;
      I.baro_counter.download(&k, 1, I.c.stream);
      out[0] = st[1];
      out[1] = st[2];
      out[2] = st[3];
      out[3] = (double)k;
      out[4] = st[0];
    }
  });
}

hmd_status hmd_integrator_set_forcefield(hmd_integrator* w, hmd_forcefield* f) {
  HMD_GUARD({
    if (!w || !f) throw Error(HMD_ERR_INVALID, "null argument");
    w->it->set_forcefield(*f->ff);
  });
}

// The learned leg's guarded run: batches of judge_every graph-replayed steps,
// the judge after each (hmd_mlp_judge: the integrator's corrected forces
// against fine_ff's at the same positions), and on the first judge above
// the threshold a switch of the integrator to fine_ff for the rest of the
// run.  The reference's nearest policy is the runner's fallback from graph
// replay to eager steps when a capture fails (src/runner.rs:1650-1656).
hmd_status hmd_integrator_run_judged(hmd_integrator* w, hmd_forcefield* fine, uint32_t n_steps,
                                     uint32_t judge_every, double threshold, double* judges,
                                     uint32_t max_judges, uint32_t* n_judges,
                                     uint64_t* reverted_at) {
  HMD_GUARD({
    if (!w || !fine || !reverted_at || judge_every == 0)
      throw Error(HMD_ERR_INVALID, "null argument or zero judge interval");
    Integrator& I = *w->it;
    if (&fine->ff->p != &I.p) throw Error(HMD_ERR_INVALID, "fine force field is on other particles");
    *reverted_at = UINT64_MAX;
    uint32_t k = 0;
    BatchStart start;
    for (uint32_t done = 0; done < n_steps;) {
      const uint32_t b = std::min(judge_every, n_steps - done);
      // a block that overflowed the list is re-run from its start with the
      // grown capacity (twice at most, as a batch runner would)
      start.save(I);
      for (int attempt = 0;; ++attempt) {
        try {
          I.replay(b);
          break;
        } catch (const Error& e) {
          if (e.code != HMD_ERR_NBL_OVERFLOW || attempt == 2) throw;
          start.restore(I);
        }
      }
      done += b;
      if (I.ff == fine->ff) continue;   // reverted: plain fine-mesh steps
      hmd_forcefield nn{I.ff};
      double e = 0.0;
      // (an overflowing list is grown and the judge evaluated again inside)
      const hmd_status st = hmd_mlp_judge(&nn, fine, &e);
      if (st != HMD_OK) throw Error(st, hmd_last_error());
      if (judges && k < max_judges) judges[k] = e;
      ++k;
      if (!(e <= threshold)) {
        *reverted_at = I.steps;
        I.set_forcefield(*fine->ff);
      }
    }
    if (n_judges) *n_judges = k;
  });
}
uint64_t hmd_integrator_steps(hmd_integrator* w) { return w->it->steps; }

hmd_status hmd_integrator_checkpoint(hmd_integrator* w) {
  HMD_GUARD({
    if (!w) throw Error(HMD_ERR_INVALID, "null integrator");
    if (!w->mark) w->mark.reset(new BatchStart());
    w->mark->save(*w->it);
    check(hipStreamSynchronize(w->it->c.stream), "sync");
  });
}

hmd_status hmd_integrator_rollback(hmd_integrator* w) {
  HMD_GUARD({
    if (!w) throw Error(HMD_ERR_INVALID, "null integrator");
    if (!w->mark) throw Error(HMD_ERR_INVALID, "no checkpoint to roll back to");
    w->mark->restore(*w->it);
  });
}

hmd_status hmd_integrator_profile(hmd_integrator* w, uint32_t n, double* us, int32_t max_stages,
                                  int32_t* n_stages, char* names, int32_t names_len) {
  HMD_GUARD({
    Integrator& I = *w->it;
    Context& c = I.c;
    I.prime();
    std::vector<std::string> order;
    std::map<std::string, double> tot;
    for (uint32_t s = 0; s < n; ++s) {
      for (auto& m : c.marks) (void)hipEventDestroy(m.second);
      c.marks.clear();
      c.profiling = true;
      try {
        c.mark("begin");
        I.enqueue_step();
      } catch (...) {
        c.profiling = false;
        throw;
      }
      c.profiling = false;
      I.publish_box();
      check(hipStreamSynchronize(c.stream), "sync");
      for (size_t k = 1; k < c.marks.size(); ++k) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, c.marks[k - 1].second, c.marks[k].second), "elapsed");
        const std::string& nm = c.marks[k].first;
        if (!tot.count(nm)) order.push_back(nm);
        tot[nm] += 1000.0 * ms;
      }
    }
    I.steps += n;
    I.ff->check_status();
    std::string all;
    int32_t k = 0;
    for (auto& nm : order) {
      if (k < max_stages) us[k] = tot[nm] / n;
      ++k;
      all += nm + "\n";
    }
    *n_stages = std::min(k, max_stages);
    if (names && names_len > 0) std::snprintf(names, (size_t)names_len, "%s", all.c_str());
  });
}

hmd_status hmd_integrator_timings(hmd_integrator* w, uint32_t n, int32_t max_stages,
                                  int32_t* n_stages, char* names, int32_t names_len,
                                  uint64_t* counts, uint64_t* total_ns, uint64_t* min_ns,
                                  uint64_t* max_ns) {
  HMD_GUARD({
    if (!w || !n_stages || max_stages < 0) throw Error(HMD_ERR_INVALID, "null argument");
    Integrator& I = *w->it;
    Context& c = I.c;
    I.prime();
    struct Acc { uint64_t count = 0, total = 0, lo = UINT64_MAX, hi = 0; };
    std::vector<std::string> order;
    std::map<std::string, Acc> acc;
    for (uint32_t s = 0; s < n; ++s) {
      for (auto& m : c.marks) (void)hipEventDestroy(m.second);
      c.marks.clear();
      c.profiling = true;
      try {
        c.mark("begin");
        I.enqueue_step();
      } catch (...) {
        c.profiling = false;
        throw;
      }
      c.profiling = false;
      I.publish_box();
      check(hipStreamSynchronize(c.stream), "sync");
      for (size_t k = 1; k < c.marks.size(); ++k) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, c.marks[k - 1].second, c.marks[k].second), "elapsed");
        const std::string& nm = c.marks[k].first;
        if (!acc.count(nm)) order.push_back(nm);
        Acc& a = acc[nm];
        const uint64_t ns = (uint64_t)std::llround((double)ms * 1.0e6);
        a.count += 1;
        a.total += ns;
        a.lo = std::min(a.lo, ns);
        a.hi = std::max(a.hi, ns);
      }
    }
    for (auto& m : c.marks) (void)hipEventDestroy(m.second);
    c.marks.clear();
    I.steps += n;
    I.ff->check_status();
    std::string all;
    int32_t k = 0;
    for (auto& nm : order) {
      if (k < max_stages) {
        const Acc& a = acc[nm];
        if (counts) counts[k] = a.count;
        if (total_ns) total_ns[k] = a.total;
        if (min_ns) min_ns[k] = a.lo;
        if (max_ns) max_ns[k] = a.hi;
        all += nm + "\n";
      }
      ++k;
    }
    *n_stages = std::min(k, max_stages);
    if (names && names_len > 0) std::snprintf(names, (size_t)names_len, "%s", all.c_str());
  });
}

hmd_status hmd_integrator_minimize(hmd_integrator* w, uint32_t n_iter, double max_step) {
  HMD_GUARD({
    // Steepest descent with an adaptive step: move every atom by h F (the
    // largest move is `step`), project with SETTLE, accept if the potential
    // energy went down (step * 1.2), otherwise restore and halve the step.
    Integrator& I = *w->it;
    Context& c = I.c;
    Particles& p = I.p;
    std::vector<float> f(3 * (size_t)p.n);
    DevBuf<float> backup;
    backup.alloc(p.posq.count, false);
    // A list overflow (the relaxing structure can outgrow the capacity probed
    // on the starting one) grows the capacity and marks a rebuild: the
    // evaluation is then repeated at the same positions.
    auto energy = [&]() {
      for (int attempt = 0;; ++attempt) {
        I.ff->compute(1);
        check(hipStreamSynchronize(c.stream), "sync");
        try {
          I.ff->check_status();
        } catch (const Error& e) {
          if (e.code != HMD_ERR_NBL_OVERFLOW || attempt >= 3) throw;
          continue;
        }
        return reduce_f64(c, p, p.energy.ptr, 0);
      }
    };
    auto fmax_now = [&]() {
      p.force.download(f.data(), f.size(), c.stream);
      double fm = 0.0;
      for (uint32_t i = 0; i < p.n; ++i) {
        double s2 = 0.0;
        for (int d = 0; d < 3; ++d) s2 += (double)f[(size_t)d * p.n + i] * f[(size_t)d * p.n + i];
        fm = std::max(fm, std::sqrt(s2));
      }
      return fm;
    };
    double e0 = energy();
    double step = max_step;
    for (uint32_t it = 0; it < n_iter; ++it) {
      const double fm = fmax_now();
      if (!(fm > 0.0) || !std::isfinite(fm)) break;
      check(hipMemcpyAsync(backup.ptr, p.posq.ptr, p.posq.count * sizeof(float),
                           hipMemcpyDeviceToDevice, c.stream), "D2D");
      SdParams sp{p.posq.ptr, p.force.ptr, p.n, p.lattice, (float)(step / fm), (float)step};
      I.settle_snapshot();
      launch(c, K_hmd_sd_step, ceil_div(p.n, 256), 256, 0, sp);
      I.settle_positions(0.0f);
      const double e1 = energy();
      if (std::isfinite(e1) && e1 < e0) {
        e0 = e1;
        step = std::min(max_step, step * 1.2);
      } else {
        check(hipMemcpyAsync(p.posq.ptr, backup.ptr, p.posq.count * sizeof(float),
                             hipMemcpyDeviceToDevice, c.stream), "D2D");
        step *= 0.5;
        e0 = energy();
        if (step < 1e-6 * max_step) break;
      }
    }
    check(hipStreamSynchronize(c.stream), "sync");
    I.primed = false;
  });
}

}  // extern "C"
