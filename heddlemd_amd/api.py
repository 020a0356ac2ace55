"""Thin Python handles over the C ABI, named after the reference's host types
(GpuContext, ParticleBuffers, ForceField, NeighborListState, Integrator).
All computation happens in libheddle_hip.so."""
import ctypes

import numpy as np

from ._lib import (BarostatDesc, BondedDesc, ForceFieldDesc, IntegratorDesc, MlpDesc, Tuning,
                   check, lib)

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class GpuContext:
    """init_device (src/gpu/device.rs:67): one device, one stream, the
    embedded gfx950 code object loaded."""

    def __init__(self, device=0, stream=None):
        """stream: a caller's HIP stream handle (an int, e.g. a torch
        stream's cuda_stream) to enqueue on instead of a stream of the
        context's own (hmd_ctx_create_on_stream)."""
        h = ctypes.c_void_p()
        if stream is None:
            check(lib().hmd_ctx_create(device, ctypes.byref(h)))
        else:
            check(lib().hmd_ctx_create_on_stream(device, ctypes.c_void_p(int(stream)),
                                                 ctypes.byref(h)))
        self.h = h

    def synchronize(self):
        check(lib().hmd_ctx_synchronize(self.h))

    def device_info(self):
        cus = ctypes.c_int()
        name = ctypes.create_string_buffer(64)
        check(lib().hmd_ctx_device_info(self.h, ctypes.byref(cus), name, 64))
        return cus.value, name.value.decode()

    def close(self):
        if self.h:
            lib().hmd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_names():
    L = lib()
    return [L.hmd_kernel_name(i).decode() for i in range(L.hmd_kernel_count())]


class ParticleBuffers:
    """ParticleBuffers (src/gpu/buffers.rs:11-178)."""

    def __init__(self, ctx, posq, mass, lattice, vel=None, types=None):
        self.ctx = ctx
        posq = np.ascontiguousarray(posq, np.float32).reshape(-1, 4)
        self.n = posq.shape[0]
        mass = np.ascontiguousarray(mass, np.float32)
        lat = np.ascontiguousarray(lattice, np.float32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        t = None if types is None else np.ascontiguousarray(types, np.uint32)
        h = ctypes.c_void_p()
        check(lib().hmd_particles_create(ctx.h, self.n, _ptr(posq, _fp), _ptr(v, _fp),
                                         _ptr(mass, _fp), _ptr(t, _u32p), _ptr(lat, _fp),
                                         ctypes.byref(h)))
        self.h = h

    @property
    def lattice(self):
        """The current box {lx, ly, lz, xy, xz, yz} (hmd_particles_get_box;
        a barostat moves it)."""
        lat = np.empty(6, np.float32)
        check(lib().hmd_particles_get_box(self.h, _ptr(lat, _fp)))
        return lat

    def download(self):
        n = self.n
        posq = np.empty((n, 4), np.float32)
        vel = np.empty((n, 3), np.float32)
        f = np.empty((n, 3), np.float32)
        im = np.empty((n, 3), np.int32)
        check(lib().hmd_particles_download(self.h, _ptr(posq, _fp), _ptr(vel, _fp),
                                           _ptr(f, _fp), _ptr(im, _i32p)))
        return {"posq": posq, "vel": vel, "force": f, "images": im}

    def forces(self):
        f = np.empty((self.n, 3), np.float32)
        check(lib().hmd_particles_download(self.h, None, None, _ptr(f, _fp), None))
        return f

    def upload(self, posq=None, vel=None):
        p = None if posq is None else np.ascontiguousarray(posq, np.float32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32)
        check(lib().hmd_particles_upload(self.h, _ptr(p, _fp), _ptr(v, _fp)))

    def energies(self):
        e = np.empty(self.n, np.float32)
        w = np.empty(self.n, np.float32)
        check(lib().hmd_particles_download_energies(self.h, _ptr(e, _fp), _ptr(w, _fp)))
        return e, w

    def lo_parts(self):
        a = np.empty((self.n, 3))
        b = np.empty((self.n, 3))
        check(lib().hmd_particles_download_lo(self.h, _ptr(a, _dp), _ptr(b, _dp)))
        return a, b

    def kinetic_energy(self):
        ke = ctypes.c_double()
        check(lib().hmd_particles_kinetic_energy(self.h, ctypes.byref(ke)))
        return ke.value

    def set_box(self, lattice):
        """New box {lx, ly, lz, xy, xz, yz} (hmd_particles_set_box)."""
        lat = np.ascontiguousarray(lattice, np.float32)
        check(lib().hmd_particles_set_box(self.h, _ptr(lat, _fp)))

    def upload_images(self, images):
        im = np.ascontiguousarray(images, np.int32).reshape(-1, 3)
        check(lib().hmd_particles_upload_images(self.h, _ptr(im, _i32p)))

    def upload_forces(self, force):
        f = np.ascontiguousarray(force, np.float32).reshape(-1, 3)
        check(lib().hmd_particles_upload_forces(self.h, _ptr(f, _fp)))

    def negate_velocities(self):
        check(lib().hmd_particles_negate_velocities(self.h))

    def kick_drift(self, dt, compensated=False):
        """v += F/m dt/2, x += v dt, wrap (src/gpu/kernels.rs:34, :2292)."""
        check(lib().hmd_vv_kick_drift(self.h, dt, 1 if compensated else 0))

    def lan_drift_half(self, dt):
        """BAOAB A step (src/gpu/kernels.rs:2339)."""
        check(lib().hmd_lan_drift_half(self.h, dt))

    def lan_ou_step(self, seed, alpha, kt, draw_counter):
        """BAOAB O step (src/gpu/kernels.rs:2374)."""
        check(lib().hmd_lan_ou_step(self.h, seed, alpha, kt, draw_counter))

    def kick(self, dt, compensated=False):
        """v += F/m dt/2 (src/gpu/kernels.rs:128, :2416)."""
        check(lib().hmd_vv_kick(self.h, dt, 1 if compensated else 0))

    def energy_sums(self):
        pe = ctypes.c_double()
        w = ctypes.c_double()
        check(lib().hmd_particles_energy_sums(self.h, ctypes.byref(pe), ctypes.byref(w)))
        return pe.value, w.value

    def close(self):
        if self.h:
            lib().hmd_particles_destroy(self.h)
            self.h = None


class ForceField:
    """ForceField (src/forces/mod.rs) built from the builtin potential registry;
    `system` is a dict from heddlemd_amd.systems."""

    def __init__(self, ctx, particles, ff):
        self.ctx = ctx
        self.particles = particles
        d = ForceFieldDesc()
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        nt = int(ff.get("n_types", 1))
        d.n_types = nt
        if ff.get("lj") is not None:
            lj = ff["lj"]
            d.lj_on = 1
            d.lj_sigma = _ptr(arr(lj["sigma"], np.float64), _dp)
            d.lj_epsilon = _ptr(arr(lj["epsilon"], np.float64), _dp)
            d.lj_cutoff = _ptr(arr(lj["cutoff"], np.float64), _dp)
            d.lj_switch = _ptr(arr(lj["switch"], np.float64), _dp)
        if ff.get("coulomb") is not None:
            co = ff["coulomb"]
            d.coulomb_on = 1
            d.alpha = co["alpha"]
            d.r_cut_real = co["r_cut"]
            d.k_coulomb = co.get("k_coulomb", 1.0)
        if ff.get("spme") is not None:
            sp = ff["spme"]
            d.spme_on = 1
            for k in range(3):
                d.grid[k] = int(sp["grid"][k])
                d.brick[k] = int(sp.get("brick", (0, 0, 0))[k])
            d.spline_order = int(sp.get("order", 4))
        d.r_skin = ff["r_skin"]
        if ff.get("exclusions") is not None:
            # (offsets, partners) or (offsets, partners, lj_scale, coul_scale)
            ex = ff["exclusions"]
            off, par = ex[0], ex[1]
            d.excl_offsets = _ptr(arr(off, np.uint32), _u32p)
            d.excl_partners = _ptr(arr(par, np.uint32), _u32p)
            if len(ex) == 4:
                d.excl_lj_scale = _ptr(arr(ex[2], np.float32), _fp)
                d.excl_coul_scale = _ptr(arr(ex[3], np.float32), _fp)
        self._keep = keep
        h = ctypes.c_void_p()
        check(lib().hmd_forcefield_create(ctx.h, particles.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def compute(self, level=0):
        check(lib().hmd_forcefield_compute(self.h, level))

    def slots(self):
        b = ctypes.create_string_buffer(256)
        check(lib().hmd_forcefield_slots(self.h, b, 256))
        return [s for s in b.value.decode().split("\n") if s]

    def spme_scalars(self):
        out = np.zeros(3)
        check(lib().hmd_spme_scalars(self.h, _ptr(out, _dp)))
        return {"w_recip": out[0], "u_recip": out[1], "u_self": out[2]}

    def spme_grids(self, shape):
        m = int(np.prod(shape))
        rho = np.empty(m, np.float32)
        v = np.empty(m, np.float32)
        check(lib().hmd_spme_grids(self.h, _ptr(rho, _fp), _ptr(v, _fp)))
        return rho.reshape(shape), v.reshape(shape)

    def spme_forces(self):
        f = np.empty((self.particles.n, 3), np.float32)
        check(lib().hmd_spme_forces(self.h, _ptr(f, _fp)))
        return f

    def set_pair_fragment(self, name, source, params=()):
        """Compose a user pair term into this force field's kernels
        (hmd_forcefield_set_pair_fragment)."""
        prm = np.ascontiguousarray(list(params) or [0.0], np.float32)
        check(lib().hmd_forcefield_set_pair_fragment(self.h, name.encode(), source.encode(),
                                                     _ptr(prm, _fp), len(params)))

    def set_fragment_exclusion_scales(self, scales):
        """The composed term's scale for each listed exclusion partner, in
        the exclusion lists' CSR order
        (hmd_forcefield_set_fragment_exclusion_scales)."""
        s = np.ascontiguousarray(scales, np.float32)
        check(lib().hmd_forcefield_set_fragment_exclusion_scales(self.h, _ptr(s, _fp), s.size))

    @staticmethod
    def _tuning_dict(t):
        return {"pair_blocks_per_cu": t.pair_blocks_per_cu,
                "fft_blocks_per_cu": t.fft_blocks_per_cu,
                "brick": tuple(t.brick), "spme_stream": t.spme_stream,
                "fft_slab": t.fft_slab, "n_candidates": t.n_candidates,
                "us_initial": t.us_initial, "us_tuned": t.us_tuned}

    def autotune(self, n_evals=20):
        """Time candidate launch shapes (pair grid, FFT grids, SPME brick) on
        the current positions and keep the fastest (hmd_forcefield_autotune)."""
        t = Tuning()
        check(lib().hmd_forcefield_autotune(self.h, n_evals, ctypes.byref(t)))
        return self._tuning_dict(t)

    def tuning(self):
        t = Tuning()
        check(lib().hmd_forcefield_get_tuning(self.h, ctypes.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, tuning):
        """Apply recorded launch shapes (a dict as autotune / tuning return)."""
        t = Tuning()
        t.pair_blocks_per_cu = int(tuning["pair_blocks_per_cu"])
        t.fft_blocks_per_cu = int(tuning.get("fft_blocks_per_cu", 0))
        br = tuning.get("brick", (0, 0, 0))
        for k in range(3):
            t.brick[k] = int(br[k])
        t.spme_stream = int(tuning.get("spme_stream", 1))
        t.fft_slab = int(tuning.get("fft_slab", 0))
        check(lib().hmd_forcefield_set_tuning(self.h, ctypes.byref(t)))

    def request_rebuild(self):
        check(lib().hmd_nbl_request_rebuild(self.h))

    def pair_snapshot(self):
        """Canonical (min, max) pairs within r_c + r_skin, sorted."""
        cap = 1 << 20
        while True:
            out = np.empty((cap, 2), np.uint32)
            k = ctypes.c_uint64()
            check(lib().hmd_nbl_pair_snapshot(self.h, _ptr(out, _u32p), cap, ctypes.byref(k)))
            if k.value <= cap:
                pairs = out[:k.value]
                key = np.sort((pairs[:, 0].astype(np.uint64) << np.uint64(32)) |
                              pairs[:, 1].astype(np.uint64))
                return np.stack([(key >> np.uint64(32)).astype(np.uint32),
                                 (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)], axis=1)
            cap = int(k.value)

    def count_pairs(self, radius):
        k = ctypes.c_uint64()
        check(lib().hmd_nbl_count_pairs(self.h, radius, ctypes.byref(k)))
        return int(k.value)

    def set_capacity(self, entries, lds_entries=0):
        """Seed the list capacity (entries per cluster) and, for tests of the
        build's direct-write path, the entries staged in LDS
        (hmd_nbl_set_capacity); the list is rebuilt at the next evaluation."""
        check(lib().hmd_nbl_set_capacity(self.h, int(entries), int(lds_entries)))

    def pre_step(self):
        """Skin check and device-gated rebuild, enqueued (hmd_nbl_pre_step)."""
        check(lib().hmd_nbl_pre_step(self.h))

    def pair_force(self, level, out):
        """The pair slots added into a SlotOutput's device arrays, enqueued."""
        check(lib().hmd_pair_force(self.h, level, ctypes.byref(out)))

    BONDED_KINDS = {"harmonic-bond": (1, 2), "morse": (2, 2), "harmonic-angle": (3, 3),
                    "periodic-dihedral": (4, 4)}

    def add_bonded(self, kind, terms, params):
        """One set of bonded terms (hmd_forcefield_add_bonded): kind one of
        BONDED_KINDS; terms (n, arity + 1) canonical atoms then the type;
        params (n_types, <= 3) per type, atomic units (harmonic-bond k, r0;
        morse De, a, r0; harmonic-angle k_theta, theta0; periodic-dihedral
        k_phi, phi0, n)."""
        code, arity = self.BONDED_KINDS[kind]
        t = np.ascontiguousarray(terms, np.uint32).reshape(-1, arity + 1)
        q = np.asarray(params, np.float64)
        q = q.reshape(q.shape[0] if q.ndim > 1 else 1, -1)
        pp = np.zeros((q.shape[0], 3))
        pp[:, :q.shape[1]] = q
        d = BondedDesc()
        d.kind, d.n_terms, d.terms = code, t.shape[0], _ptr(t, _u32p)
        d.n_types, d.params = pp.shape[0], _ptr(pp, _dp)
        check(lib().hmd_forcefield_add_bonded(self.h, ctypes.byref(d)))

    def bonded_compute(self, level, out):
        """Every bonded set as one slot, added into a SlotOutput, enqueued."""
        check(lib().hmd_bonded_compute(self.h, level, ctypes.byref(out)))

    def spme_compute(self, level, out):
        """The reciprocal slot added into a SlotOutput's device arrays, enqueued."""
        check(lib().hmd_spme_compute(self.h, level, ctypes.byref(out)))

    def check(self):
        """Synchronize and read the neighbour status (hmd_forcefield_check)."""
        check(lib().hmd_forcefield_check(self.h))

    def stats(self):
        out = np.zeros(8, np.uint32)
        check(lib().hmd_nbl_stats(self.h, _ptr(out, _u32p)))
        keys = ["n_work", "generation", "status", "high_water", "cap_entries",
                "n_clusters", "fine_cells", "subcluster_visits"]
        return dict(zip(keys, (int(x) for x in out)))

    def close(self):
        if self.h:
            lib().hmd_forcefield_destroy(self.h)
            self.h = None


class Integrator:
    """Integrator + constraint plan (src/integrator/mod.rs)."""

    def __init__(self, ctx, particles, forcefield, kind="velocity-verlet", dt=1.0,
                 compensated=False, temperature=0.0, gamma=0.0, seed=0, settle=None,
                 fused=True, scalars=False, thermostat=None, tau=0.0):
        """thermostat: None or "csvr" (temperature = k_B T, tau, seed)."""
        d = IntegratorDesc()
        d.thermostat = {None: 0, "csvr": 1}[thermostat]
        d.tau = tau
        # HMD_INTEGRATOR_UNFUSED = 1, HMD_INTEGRATOR_SCALARS = 2
        d.flags = (0 if fused else 1) | (2 if scalars else 0)
        self._kind = kind.encode()
        d.kind = self._kind
        d.dt = dt
        d.compensated = 1 if compensated else 0
        d.temperature = temperature
        d.gamma = gamma
        d.seed = seed
        self._groups = None
        if settle is not None:
            g = np.ascontiguousarray(settle["groups"], np.uint32)
            self._groups = g
            d.n_groups = g.shape[0]
            d.group_atoms = _ptr(g, _u32p)
            d.d_oh, d.d_hh, d.m_o, d.m_h = settle["d_oh"], settle["d_hh"], settle["m_o"], settle["m_h"]
        h = ctypes.c_void_p()
        check(lib().hmd_integrator_create(ctx.h, particles.h, forcefield.h, ctypes.byref(d),
                                          ctypes.byref(h)))
        self.h = h

    def step(self, n=1):
        check(lib().hmd_integrator_step(self.h, n))

    def capture(self):
        check(lib().hmd_integrator_capture(self.h))

    def replay(self, n):
        check(lib().hmd_integrator_replay(self.h, n))

    def steps(self):
        return int(lib().hmd_integrator_steps(self.h))

    def thermostat_state(self):
        """CSVR: (injected energy sum K' - K, draws so far)."""
        out = np.zeros(2)
        check(lib().hmd_integrator_thermostat_state(self.h, _ptr(out, _dp)))
        return float(out[0]), int(out[1])

    def set_barostat(self, kind="c-rescale", pressure=0.0, temperature=0.0, tau=0.0,
                     compressibility=0.0, seed=0):
        """Per-step c-rescale barostat (hmd_integrator_set_barostat): target
        pressure, k_B T of the noise term, coupling time, compressibility,
        Philox seed, all in atomic units; kind None removes it.  Barostat
        steps run eagerly (replay steps eagerly, capture is refused)."""
        d = BarostatDesc()
        d.kind = {None: 0, "c-rescale": 1}[kind]
        d.pressure, d.temperature, d.tau = pressure, temperature, tau
        d.compressibility, d.seed = compressibility, seed
        check(lib().hmd_integrator_set_barostat(self.h, ctypes.byref(d)))

    def barostat_state(self):
        """(pressure of the latest move, volume after it, accumulated
        P_target dV, draws so far, latest mu)."""
        out = np.zeros(5)
        check(lib().hmd_integrator_barostat_state(self.h, _ptr(out, _dp)))
        return float(out[0]), float(out[1]), float(out[2]), int(out[3]), float(out[4])

    def checkpoint(self):
        """Keep a device copy of the state a batch of steps changes
        (hmd_integrator_checkpoint)."""
        check(lib().hmd_integrator_checkpoint(self.h))

    def rollback(self):
        """Restore the last checkpoint, with a list rebuild at the next
        evaluation (hmd_integrator_rollback)."""
        check(lib().hmd_integrator_rollback(self.h))

    def set_forcefield(self, forcefield):
        """Step with another force field on the same particles from here on
        (hmd_integrator_set_forcefield)."""
        check(lib().hmd_integrator_set_forcefield(self.h, forcefield.h))

    def run_judged(self, fine, n_steps, judge_every=50, threshold=1e-3):
        """Graph-replayed steps with the fallback judge every judge_every
        steps; the first judge above threshold switches to `fine` for good
        (hmd_integrator_run_judged).  Returns (judge values, step of the
        switch or None)."""
        m = n_steps // judge_every + 1
        js = np.zeros(m)
        k = ctypes.c_uint32()
        at = ctypes.c_uint64()
        check(lib().hmd_integrator_run_judged(self.h, fine.h, n_steps, judge_every, threshold,
                                              _ptr(js, _dp), m, ctypes.byref(k),
                                              ctypes.byref(at)))
        return js[:k.value].tolist(), (None if at.value == 2 ** 64 - 1 else int(at.value))

    def minimize(self, n_iter, max_step):
        check(lib().hmd_integrator_minimize(self.h, n_iter, max_step))

    def profile(self, n_steps):
        us = np.zeros(64)
        k = ctypes.c_int32()
        names = ctypes.create_string_buffer(4096)
        check(lib().hmd_integrator_profile(self.h, n_steps, _ptr(us, _dp), 64,
                                           ctypes.byref(k), names, 4096))
        labels = [s for s in names.value.decode().split("\n") if s]
        return dict(zip(labels, us[:k.value].tolist()))

    def timings(self, n_steps):
        """Per-stage kernel timings of n eager steps: [{name, count,
        total_ns, min_ns, max_ns}] in plan order (src/timings.rs)."""
        m = 64
        arrs = [np.zeros(m, np.uint64) for _ in range(4)]
        k = ctypes.c_int32()
        names = ctypes.create_string_buffer(4096)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        check(lib().hmd_integrator_timings(self.h, n_steps, m, ctypes.byref(k), names, 4096,
                                           *[_ptr(a, u64p) for a in arrs]))
        labels = [s for s in names.value.decode().split("\n") if s]
        return [{"name": labels[i], "count": int(arrs[0][i]), "total_ns": int(arrs[1][i]),
                 "min_ns": int(arrs[2][i]), "max_ns": int(arrs[3][i])} for i in range(k.value)]

    def close(self):
        if self.h:
            lib().hmd_integrator_destroy(self.h)
            self.h = None


class Settle:
    """SETTLE constraint hooks as operators (src/integrator/settle.rs:310-404)."""

    def __init__(self, particles, groups, d_oh, d_hh, m_o, m_h):
        self._g = np.ascontiguousarray(groups, np.uint32).reshape(-1, 3)
        h = ctypes.c_void_p()
        check(lib().hmd_settle_create(particles.h, self._g.shape[0], _ptr(self._g, _u32p),
                                      d_oh, d_hh, m_o, m_h, ctypes.byref(h)))
        self.h = h

    def before_drift(self):
        check(lib().hmd_settle_before_drift(self.h))

    def after_drift(self, dt):
        check(lib().hmd_settle_after_drift(self.h, dt))

    def after_kick(self, dt):
        check(lib().hmd_settle_after_kick(self.h, dt))

    def close(self):
        if self.h:
            lib().hmd_settle_destroy(self.h)
            self.h = None


class ReplicaSet:
    """Independent systems stepped as one batch (hmd_replicas_*).  Each
    integrator must come from its own GpuContext.  mode "fused" captures every
    replica's step into one graph, "per_replica" launches one graph per
    replica on its own stream, "batched" merges the replicas' steps into one
    launch per stage (each kernel's batched twin over all replicas' blocks),
    "batched_eager" issues those launches directly every step (no graph; for
    kernel tracers).  fused=False is mode "per_replica"."""

    MODES = {"fused": 0, "per_replica": 1, "batched": 2, "batched_eager": 2 | 4}

    def __init__(self, integrators, fused=True, mode=None):
        if mode is None:
            mode = "fused" if fused else "per_replica"
        if mode not in self.MODES:
            raise ValueError(f"unknown replica mode {mode!r}")
        self._its = list(integrators)
        self.mode = mode
        arr = (ctypes.c_void_p * len(self._its))(*[it.h.value for it in self._its])
        h = ctypes.c_void_p()
        check(lib().hmd_replicas_create(arr, len(self._its), self.MODES[mode],
                                        ctypes.byref(h)))
        self.h = h

    def plan(self):
        """(batched launches, single launches) per step of the merged plan
        (batched mode, after the first step; zeros otherwise)."""
        nb, ns = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().hmd_replicas_plan(self.h, ctypes.byref(nb), ctypes.byref(ns)))
        return nb.value, ns.value

    def step(self, n=1):
        check(lib().hmd_replicas_step(self.h, n))

    def steps(self):
        return int(lib().hmd_replicas_steps(self.h))

    def close(self):
        if self.h:
            lib().hmd_replicas_destroy(self.h)
            self.h = None


class CorrectionNet:
    """Pair-scalar network correcting a coarse SPME mesh toward a fine one
    (the learned part of the hot path; no counterpart in the reference).
    Parameters are 7553 floats: W1[64][51], b1, W2[64][64], b2, w3, b3."""

    N_PARAMS = 7553

    def __init__(self, ctx, r_cut, force_scale, species_of_type, seed=31, lr=1e-3):
        sp = np.ascontiguousarray(species_of_type, np.uint32)
        d = MlpDesc()
        d.r_cut = r_cut
        d.force_scale = force_scale
        d.n_types = sp.size
        d.species_of_type = _ptr(sp, _u32p)
        d.seed = seed
        d.lr = lr
        h = ctypes.c_void_p()
        check(lib().hmd_mlp_create(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    @property
    def params(self):
        out = np.empty(self.N_PARAMS, np.float32)
        check(lib().hmd_mlp_get_params(self.h, _ptr(out, _fp)))
        return out

    @params.setter
    def params(self, values):
        v = np.ascontiguousarray(values, np.float32)
        assert v.size == self.N_PARAMS
        check(lib().hmd_mlp_set_params(self.h, _ptr(v, _fp)))

    def set_force_scale(self, s):
        check(lib().hmd_mlp_set_force_scale(self.h, s))

    def descriptors(self, ff):
        out = np.empty((ff.particles.n, 16), np.float32)
        check(lib().hmd_mlp_descriptors(self.h, ff.h, _ptr(out, _fp)))
        return out

    def forward(self, ff):
        out = np.empty((ff.particles.n, 3), np.float32)
        check(lib().hmd_mlp_forward(self.h, ff.h, _ptr(out, _fp)))
        return out

    def loss_grad(self, ff, target):
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = ctypes.c_double()
        g = np.empty(self.N_PARAMS, np.float32)
        check(lib().hmd_mlp_loss_grad(self.h, ff.h, _ptr(t, _fp), ctypes.byref(loss),
                                      _ptr(g, _fp)))
        return loss.value, g

    def adam_step(self, grad):
        g = np.ascontiguousarray(grad, np.float32)
        check(lib().hmd_mlp_adam_step(self.h, _ptr(g, _fp)))

    def attach(self, ff):
        check(lib().hmd_mlp_attach(ff.h, self.h))

    @staticmethod
    def detach(ff):
        check(lib().hmd_mlp_attach(ff.h, None))

    def reserve(self, n_atoms, n_frames):
        check(lib().hmd_mlp_dataset_reserve(self.h, n_atoms, n_frames))

    def record_frame(self, fine, coarse):
        check(lib().hmd_mlp_record_frame(self.h, fine.h, coarse.h))

    def dataset_stats(self):
        out = np.zeros(3)
        check(lib().hmd_mlp_dataset_stats(self.h, _ptr(out, _dp)))
        return {"frames": int(out[0]), "target_rms": out[1], "fine_recip_rms": out[2]}

    def train(self, ff, n_steps, batch_frames=8):
        losses = np.zeros(n_steps)
        check(lib().hmd_mlp_train(self.h, ff.h, n_steps, batch_frames, _ptr(losses, _dp)))
        return losses

    def close(self):
        if self.h:
            lib().hmd_mlp_destroy(self.h)
            self.h = None


def judge(nn_ff, fine_ff):
    """Fallback judge: ||F_nn - F_fine|| / ||F_fine|| at the current positions."""
    r = ctypes.c_double()
    check(lib().hmd_mlp_judge(nn_ff.h, fine_ff.h, ctypes.byref(r)))
    return r.value


def generate_velocities(masses, temperature, seed, n_constraints=0):
    m = np.ascontiguousarray(masses, np.float64)
    out = np.empty((m.size, 3), np.float32)
    check(lib().hmd_generate_velocities(m.size, n_constraints, temperature, seed,
                                        _ptr(m, _dp), _ptr(out, _fp)))
    return out


def chacha_block(state, rounds):
    s = np.ascontiguousarray(state, np.uint32)
    out = np.empty(16, np.uint32)
    lib().hmd_chacha_block(_ptr(s, _u32p), rounds, _ptr(out, _u32p))
    return out


class TrajectoryWriter:
    """Extended-XYZ trajectory writer (src/io/trajectory.rs:39-176), byte for
    byte: `units` "si" or "atomic"; frames take atomic-unit values."""

    def __init__(self, path, units="si", include_velocities=False, include_images=False,
                 type_names=("X",)):
        from ._lib import UNITS
        names = (ctypes.c_char_p * len(type_names))(*[n.encode() for n in type_names])
        h = ctypes.c_void_p()
        check(lib().hmd_traj_open(str(path).encode(), UNITS[units], int(include_velocities),
                                  int(include_images), names, len(type_names),
                                  ctypes.byref(h)))
        self.h = h
        self.velocities, self.images = include_velocities, include_images

    def write_frame(self, step, dt, lattice, types, positions, velocities=None, images=None):
        """positions, velocities: (n, 3) in atomic units (cast to f32 like the
        engine's Real); images: (n, 3) int; lattice {lx, ly, lz, xy, xz, yz}."""
        pos = np.asarray(positions, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        cols = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
        t = np.ascontiguousarray(types, np.uint32)
        lat = np.ascontiguousarray(lattice, np.float32)
        vel = [None] * 3
        img = [None] * 3
        if velocities is not None:
            v = np.asarray(velocities, np.float32).reshape(-1, 3)
            vel = [np.ascontiguousarray(v[:, k]) for k in range(3)]
        if images is not None:
            m = np.asarray(images, np.int32).reshape(-1, 3)
            img = [np.ascontiguousarray(m[:, k]) for k in range(3)]
        fp = lambda a: _ptr(a, _fp) if a is not None else None
        ip = lambda a: _ptr(a, _i32p) if a is not None else None
        check(lib().hmd_traj_write_frame(self.h, int(step), float(dt), _ptr(lat, _fp), n,
                                         _ptr(t, _u32p), *[fp(c) for c in cols],
                                         *[fp(c) for c in vel], *[ip(c) for c in img]))

    def flush(self):
        check(lib().hmd_traj_flush(self.h))

    def close(self):
        if self.h:
            lib().hmd_traj_close(self.h)
            self.h = None


def compile_pair_fragment(name, source):
    """Compile a pair fragment without a device: (ok, compiler log)."""
    buf = ctypes.create_string_buffer(1 << 16)
    st = lib().hmd_pair_fragment_compile(name.encode(), source.encode(), buf, len(buf))
    return st == 0, buf.value.decode(errors="replace")


def write_timings(path, stages):
    """The .timings report (src/timings.rs:513-553) of [{name, count,
    total_ns, min_ns, max_ns}]; stages with a zero count are left out."""
    n = len(stages)
    names = (ctypes.c_char_p * max(1, n))(*[st["name"].encode() for st in stages])
    cols = [np.ascontiguousarray([st[k] for st in stages] or [0], np.uint64)
            for k in ("count", "total_ns", "min_ns", "max_ns")]
    u64p = ctypes.POINTER(ctypes.c_uint64)
    check(lib().hmd_timings_write(str(path).encode(), n, names, *[_ptr(c, u64p) for c in cols]))


class LogWriter:
    """CSV log writer (src/io/log_output.rs:36-110), byte for byte; extra
    columns are (name, dimension) with dimension a name from
    heddlemd_amd._lib.DIMENSIONS."""

    def __init__(self, path, units="si", extra_columns=()):
        from ._lib import UNITS, DIMENSIONS
        names = (ctypes.c_char_p * max(1, len(extra_columns)))(
            *[c[0].encode() for c in extra_columns])
        dims = np.ascontiguousarray([DIMENSIONS.index(c[1]) for c in extra_columns] or [0],
                                    np.uint32)
        h = ctypes.c_void_p()
        check(lib().hmd_log_open(str(path).encode(), UNITS[units], names, _ptr(dims, _u32p),
                                 len(extra_columns), ctypes.byref(h)))
        self.h = h
        self.n_extra = len(extra_columns)

    def write_row(self, step, time, kinetic_energy, temperature, extras=()):
        if len(extras) != self.n_extra:
            raise ValueError("extras length does not match the count declared at open")
        ex = np.ascontiguousarray(list(extras) or [0.0], np.float64)
        check(lib().hmd_log_write_row(self.h, int(step), float(time), float(kinetic_energy),
                                      float(temperature), _ptr(ex, _dp)))

    def flush(self):
        check(lib().hmd_log_flush(self.h))

    def close(self):
        if self.h:
            lib().hmd_log_close(self.h)
            self.h = None
