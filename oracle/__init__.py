"""FP64 CPU oracle for the nonbonded hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it; the product library
(``heddlemd_amd``) never does.  Each wrapper calls one routine of
``heddle_oracle.c``, whose comments cite the reference file:line restated.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so from heddle_oracle.c (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint32)


class PairParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("pos", _dp), ("q", _dp), ("type", _ip), ("L", _dp),
        ("n_types", ctypes.c_int), ("lj_sigma", _dp), ("lj_epsilon", _dp),
        ("lj_cutoff", _dp), ("lj_switch", _dp), ("lj_on", ctypes.c_int),
        ("coul_on", ctypes.c_int), ("k_coulomb", ctypes.c_double),
        ("alpha", ctypes.c_double), ("r_cut_real", ctypes.c_double),
        ("erfc_mode", ctypes.c_int),
        ("excl_offsets", _ip), ("excl_partners", _ip),
        ("excl_lj_scale", _dp), ("excl_coul_scale", _dp),
        ("max_cutoff", ctypes.c_double),
    ]


def _declare(L):
    L.orc_r2_f32.restype = ctypes.c_float
    L.orc_r2_f32.argtypes = [_fp, _fp, _fp]
    L.orc_r2_f32_many.argtypes = [ctypes.c_int, _fp, _fp, _fp, _fp]
    L.orc_pair_set_f32.restype = ctypes.c_int64
    L.orc_pair_set_f32.argtypes = [ctypes.c_int, _fp, _fp, ctypes.c_float, _up, ctypes.c_int64]
    L.orc_pair_set_f32_cells.restype = ctypes.c_int64
    L.orc_pair_set_f32_cells.argtypes = [ctypes.c_int, _fp, _fp, ctypes.c_float, _up,
                                         ctypes.c_int64]
    L.orc_pair_forces.argtypes = [ctypes.POINTER(PairParams), _dp, _dp, _dp]
    L.orc_pair_forces_cells.argtypes = [ctypes.POINTER(PairParams), _dp, _dp, _dp]
    L.orc_ewald_recip_direct.restype = ctypes.c_double
    L.orc_ewald_recip_direct.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, _dp]
    L.orc_bspline.restype = ctypes.c_double
    L.orc_bspline.argtypes = [ctypes.c_int, ctypes.c_double]
    L.orc_b_factors.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_spme_spread.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_int, _dp]
    L.orc_fft3d.argtypes = [_ip, ctypes.c_int, _dp, _dp]
    L.orc_spme_recip.restype = ctypes.c_double
    L.orc_spme_recip.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]
    L.orc_settle_geometry.argtypes = [ctypes.c_double] * 4 + [_dp]
    L.orc_settle_positions_one.argtypes = [_dp, _dp, _dp] + [ctypes.c_double] * 6 + [_dp, _dp]
    L.orc_settle_velocities_one.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, _dp,
                                            ctypes.c_double, _dp]
    L.orc_philox4x32_10.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _up, _up]
    L.orc_philox_gaussian.restype = ctypes.c_double
    L.orc_philox_gaussian.argtypes = [ctypes.c_uint32] * 6
    L.orc_csvr_knew.restype = ctypes.c_double
    L.orc_csvr_knew.argtypes = [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double,
                                ctypes.c_double]
    L.orc_crescale_mu.argtypes = ([ctypes.c_double] * 3 + [ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint64] + [ctypes.c_double] * 6 + [_dp])
    L.orc_bonded.restype = ctypes.c_double
    L.orc_bonded.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _up, _dp, _dp, _dp, _dp]
    L.orc_kinetic_energy.restype = ctypes.c_double
    L.orc_kinetic_energy.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_vv_kick_drift_f32.argtypes = [ctypes.c_int, _fp, _fp, _dp, _dp, _fp, _fp, _ip, _fp,
                                        ctypes.c_float, ctypes.c_int]
    L.orc_lan_drift_half_f32.argtypes = [ctypes.c_int, _fp, _fp, _ip, _fp, ctypes.c_float]
    L.orc_lan_ou_step_f32.argtypes = [ctypes.c_int, _fp, _fp, ctypes.c_uint64, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_uint64]
    L.orc_vv_kick_f32.argtypes = [ctypes.c_int, _fp, _dp, _fp, _fp, ctypes.c_float, ctypes.c_int]


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def r2_f32(pi, pj, lattice):
    pi = np.ascontiguousarray(pi, np.float32)
    pj = np.ascontiguousarray(pj, np.float32)
    la = np.ascontiguousarray(lattice, np.float32)
    return lib().orc_r2_f32(_p(pi, _fp), _p(pj, _fp), _p(la, _fp))


def r2_f32_many(a, b, lattice):
    """Exact f32 min-image squared distance of each row pair (posq rows)."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1, 4)
    b = np.ascontiguousarray(b, np.float32).reshape(-1, 4)
    la = np.ascontiguousarray(lattice, np.float32)
    out = np.empty(a.shape[0], np.float32)
    lib().orc_r2_f32_many(a.shape[0], _p(a, _fp), _p(b, _fp), _p(la, _fp), _p(out, _fp))
    return out


def pair_set_f32(posq, lattice, r_search_sq, cells=False):
    """Canonical (i<j) pairs within the exact f32 search radius: uint32[k,2]
    in ascending order (cells=True: the cell-list form, same output)."""
    posq = np.ascontiguousarray(posq, np.float32).reshape(-1, 4)
    la = np.ascontiguousarray(lattice, np.float32)
    n = posq.shape[0]
    fn = lib().orc_pair_set_f32_cells if cells else lib().orc_pair_set_f32
    cap = 1 << 16
    while True:
        out = np.empty((cap, 2), np.uint32)
        k = fn(n, _p(posq, _fp), _p(la, _fp), ctypes.c_float(r_search_sq), _p(out, _up), cap)
        if k <= cap:
            return out[:k].copy()
        cap = int(k)


class PairSystem:
    """Holds the arrays a PairParams struct points to (keeps them alive)."""

    def __init__(self, pos, q, types, lattice, lj=None, coulomb=None,
                 exclusions=None, erfc_mode=0):
        self.pos = _f64(pos).reshape(-1, 3)
        n = self.pos.shape[0]
        self.q = _f64(q)
        self.types = np.ascontiguousarray(types, np.int32)
        self.L = _f64(lattice)
        p = PairParams()
        p.n = n
        p.pos = _p(self.pos, _dp)
        p.q = _p(self.q, _dp)
        p.type = _p(self.types, _ip)
        p.L = _p(self.L, _dp)
        cut = 0.0
        if lj is not None:
            self.lj = [_f64(lj[k]) for k in ("sigma", "epsilon", "cutoff", "switch")]
            p.n_types = int(round(np.sqrt(self.lj[0].size)))
            p.lj_sigma, p.lj_epsilon, p.lj_cutoff, p.lj_switch = [_p(a, _dp) for a in self.lj]
            p.lj_on = 1
            cut = max(cut, float(self.lj[2].max()))
        if coulomb is not None:
            p.coul_on = 1
            p.k_coulomb = coulomb.get("k_coulomb", 1.0)
            p.alpha = coulomb["alpha"]
            p.r_cut_real = coulomb["r_cut"]
            cut = max(cut, p.r_cut_real)
        p.erfc_mode = erfc_mode
        if exclusions is not None:
            off, par, slj, scl = exclusions
            self.ex = [np.ascontiguousarray(off, np.int32), np.ascontiguousarray(par, np.int32),
                       _f64(slj), _f64(scl)]
            p.excl_offsets = _p(self.ex[0], _ip)
            p.excl_partners = _p(self.ex[1], _ip)
            p.excl_lj_scale = _p(self.ex[2], _dp)
            p.excl_coul_scale = _p(self.ex[3], _dp)
        p.max_cutoff = cut
        self.params = p
        self.n = n

    def forces(self, cells=False):
        n = self.n
        F = np.zeros((n, 3)); E = np.zeros(n); W = np.zeros(n)
        fn = lib().orc_pair_forces_cells if cells else lib().orc_pair_forces
        fn(ctypes.byref(self.params), _p(F, _dp), _p(E, _dp), _p(W, _dp))
        return F, E, W


def ewald_recip_direct(pos, q, lattice, alpha, m_max, k_coulomb=1.0):
    pos = _f64(pos).reshape(-1, 3); q = _f64(q); la = _f64(lattice)
    n = pos.shape[0]
    F = np.zeros((n, 3)); E = np.zeros(n)
    u = lib().orc_ewald_recip_direct(n, _p(pos, _dp), _p(q, _dp), _p(la, _dp), alpha,
                                     k_coulomb, m_max, _p(F, _dp), _p(E, _dp))
    return u, F, E


def bspline(p, x):
    return lib().orc_bspline(p, float(x))


def b_factors(n, p):
    out = np.zeros(n)
    lib().orc_b_factors(n, p, _p(out, _dp))
    return out


def spme_spread(pos, q, lattice, grid, order):
    pos = _f64(pos).reshape(-1, 3); q = _f64(q); la = _f64(lattice)
    nk = np.ascontiguousarray(grid, np.int32)
    rho = np.zeros(int(np.prod(grid)))
    lib().orc_spme_spread(pos.shape[0], _p(pos, _dp), _p(q, _dp), _p(la, _dp),
                          _p(nk, _ip), order, _p(rho, _dp))
    return rho.reshape(tuple(grid))


def fft3d(re, im, sign):
    re = _f64(re).copy(); im = _f64(im).copy()
    nk = np.ascontiguousarray(re.shape, np.int32)
    lib().orc_fft3d(_p(nk, _ip), sign, _p(re, _dp), _p(im, _dp))
    return re, im


def spme_recip(pos, q, lattice, grid, order, alpha, k_coulomb=1.0, want_grids=False):
    """Returns (U_recip, F[n,3], E[n] (incl. -u_self), W_recip[, rho, V])."""
    pos = _f64(pos).reshape(-1, 3); q = _f64(q); la = _f64(lattice)
    nk = np.ascontiguousarray(grid, np.int32)
    n = pos.shape[0]
    F = np.zeros((n, 3)); E = np.zeros(n); w = np.zeros(1)
    M = int(np.prod(grid))
    rho = np.zeros(M) if want_grids else None
    V = np.zeros(M) if want_grids else None
    u = lib().orc_spme_recip(n, _p(pos, _dp), _p(q, _dp), _p(la, _dp), _p(nk, _ip), order,
                             alpha, k_coulomb, _p(F, _dp), _p(E, _dp), _p(w, _dp),
                             _p(rho, _dp) if want_grids else None,
                             _p(V, _dp) if want_grids else None)
    if want_grids:
        return u, F, E, float(w[0]), rho.reshape(tuple(grid)), V.reshape(tuple(grid))
    return u, F, E, float(w[0])


def settle_geometry(d_oh, d_hh, m_o, m_h):
    out = np.zeros(3)
    lib().orc_settle_geometry(d_oh, d_hh, m_o, m_h, _p(out, _dp))
    return out


def settle_positions_one(snap, cur, v, m_o, m_h, geom, dt, lattice):
    snap = _f64(snap).reshape(9); cur = _f64(cur).reshape(9).copy()
    v = _f64(v).reshape(9).copy(); la = _f64(lattice); vir = np.zeros(3)
    lib().orc_settle_positions_one(_p(snap, _dp), _p(cur, _dp), _p(v, _dp), m_o, m_h,
                                   geom[0], geom[1], geom[2], dt, _p(la, _dp), _p(vir, _dp))
    return cur.reshape(3, 3), v.reshape(3, 3), vir


def settle_velocities_one(pos, v, m_o, m_h, lattice, dt=None):
    """Projected velocities (3, 3); with dt, also the velocity-level
    constraint virial per atom (3,)."""
    pos = _f64(pos).reshape(9); v = _f64(v).reshape(9).copy(); la = _f64(lattice)
    vir = np.zeros(3)
    lib().orc_settle_velocities_one(_p(pos, _dp), _p(v, _dp), m_o, m_h, _p(la, _dp),
                                    0.0 if dt is None else float(dt), _p(vir, _dp))
    if dt is None:
        return v.reshape(3, 3)
    return v.reshape(3, 3), vir


def philox4x32_10(key, ctr):
    c = np.ascontiguousarray(ctr, np.uint32); o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(int(key[0]), int(key[1]), _p(c, _up), _p(o, _up))
    return o


def philox_gaussian(seed_lo, seed_hi, step_lo, step_hi, pid, axis):
    return lib().orc_philox_gaussian(seed_lo, seed_hi, step_lo, step_hi, pid, axis)


class VvState:
    """Particle state of the f32 velocity-Verlet oracle (orc_vv_kick_drift_f32 /
    orc_vv_kick_f32): posq float4[n], SoA [3][n] velocities, forces, f64 low
    parts and image counters; accessors return [n][3] copies."""

    def __init__(self, posq, vel, force, mass, lattice, images=None):
        self.posq = np.ascontiguousarray(posq, np.float32).reshape(-1, 4).copy()
        n = self.posq.shape[0]
        self.n = n
        soa = lambda a, t: np.ascontiguousarray(np.asarray(a, t).reshape(n, 3).T).copy()
        self._v = soa(vel, np.float32)
        self._f = soa(force, np.float32)
        self._m = np.ascontiguousarray(mass, np.float32).reshape(n).copy()
        self._L = np.ascontiguousarray(lattice, np.float32).reshape(6).copy()
        self._im = soa(np.zeros((n, 3)) if images is None else images, np.int32)
        self._xlo = np.zeros((3, n))
        self._vlo = np.zeros((3, n))

    def kick_drift(self, dt, compensated=False):
        lib().orc_vv_kick_drift_f32(self.n, _p(self.posq, _fp), _p(self._v, _fp),
                                    _p(self._xlo, _dp), _p(self._vlo, _dp), _p(self._f, _fp),
                                    _p(self._m, _fp), _p(self._im, _ip), _p(self._L, _fp),
                                    float(np.float32(dt)), 1 if compensated else 0)

    def kick(self, dt, compensated=False):
        lib().orc_vv_kick_f32(self.n, _p(self._v, _fp), _p(self._vlo, _dp), _p(self._f, _fp),
                              _p(self._m, _fp), float(np.float32(dt)), 1 if compensated else 0)

    def lan_drift_half(self, dt):
        lib().orc_lan_drift_half_f32(self.n, _p(self.posq, _fp), _p(self._v, _fp),
                                     _p(self._im, _ip), _p(self._L, _fp), float(np.float32(dt)))

    def lan_ou_step(self, seed, alpha, kt, draw_counter):
        lib().orc_lan_ou_step_f32(self.n, _p(self._v, _fp), _p(self._m, _fp), seed,
                                  float(np.float32(alpha)), float(np.float32(kt)), draw_counter)

    def negate_velocities(self):
        self._v = -self._v
        self._vlo = -self._vlo

    @property
    def vel(self):
        return self._v.T.copy()

    @property
    def images(self):
        return self._im.T.copy()

    @property
    def pos_lo(self):
        return self._xlo.T.copy()

    @property
    def vel_lo(self):
        return self._vlo.T.copy()


def csvr_knew(k_old, seed, counter, g_dof, c, kt_half):
    """CSVR's new kinetic energy (orc_csvr_knew)."""
    return lib().orc_csvr_knew(float(k_old), seed & 0xFFFFFFFF, seed >> 32, counter, g_dof,
                               float(c), float(kt_half))


def crescale_mu(k, w, v, seed, counter, p_target, tau, beta, kt, dt, mu3_min=1e-18):
    """One c-rescale move (orc_crescale_mu): (mu, P, V after, P_target dV)."""
    out = np.zeros(4)
    lib().orc_crescale_mu(float(k), float(w), float(v), seed & 0xFFFFFFFF, seed >> 32, counter,
                          float(p_target), float(tau), float(beta), float(kt), float(dt),
                          float(mu3_min), _p(out, _dp))
    return tuple(float(x) for x in out)


BONDED_KINDS = {"harmonic-bond": (1, 2), "morse": (2, 2), "harmonic-angle": (3, 3),
                "periodic-dihedral": (4, 4)}


def bonded(kind, pos, lattice, terms, params):
    """Bonded terms in FP64 (orc_bonded): (forces (n, 3), per-atom energy,
    per-atom virial, total energy)."""
    code, arity = BONDED_KINDS[kind]
    x = _f64(pos).reshape(-1, 3)
    L = _f64(lattice)
    t = np.ascontiguousarray(terms, np.uint32).reshape(-1, arity + 1)
    q = np.asarray(params, np.float64)
    q = q.reshape(q.shape[0] if q.ndim > 1 else 1, -1)
    pp = np.zeros((q.shape[0], 3))
    pp[:, :q.shape[1]] = q
    n = x.shape[0]
    f = np.zeros((n, 3)); e = np.zeros(n); w = np.zeros(n)
    u = lib().orc_bonded(code, _p(x, _dp), _p(L, _dp), t.shape[0], _p(t, _up), _p(pp, _dp),
                         _p(f, _dp), _p(e, _dp), _p(w, _dp))
    return f, e, w, u


def kinetic_energy(v, m):
    v = _f64(v).reshape(-1, 3); m = _f64(m)
    return lib().orc_kinetic_energy(v.shape[0], _p(v, _dp), _p(m, _dp))


def num_threads():
    return lib().orc_num_threads()


_STEP_PATH = os.path.join(_HERE, "libcpustep.so")
_step_lib = None


def step_lib():
    """The f32 CPU baseline of the MD step (cpu_step.c)."""
    global _step_lib
    if _step_lib is None:
        if not os.path.exists(_STEP_PATH):
            build()
        L = ctypes.CDLL(_STEP_PATH)
        L.cs_create.restype = ctypes.c_void_p
        L.cs_create.argtypes = ([ctypes.c_int, _fp, _fp, _fp, _ip, ctypes.c_int, _fp, _fp, _fp,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_double,
                                 ctypes.c_double, _ip, _ip, _ip, ctypes.c_int, _ip]
                                + [ctypes.c_double] * 5 + [ctypes.c_float, ctypes.c_int])
        L.cs_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cs_rebuilds.restype = ctypes.c_int
        L.cs_rebuilds.argtypes = [ctypes.c_void_p]
        L.cs_forces.argtypes = [ctypes.c_void_p, _fp]
        L.cs_destroy.argtypes = [ctypes.c_void_p]
        L.cs_set_threads.argtypes = [ctypes.c_int]
        L.cs_max_threads.restype = ctypes.c_int
        _step_lib = L
    return _step_lib


class CpuStep:
    """The same MD step on the host cores in f32 (cell + Verlet list, LJ +
    tabulated real-space Ewald, SPME order 4, compensated or f32 velocity
    Verlet, SETTLE): the CPU baseline bench.py times."""

    def __init__(self, posq, vel, mass, types, lattice, lj, coulomb, r_skin, mesh,
                 exclusions, settle, dt, compensated=True):
        L = step_lib()
        self.n = n = len(mass)
        nt = int(round(np.sqrt(np.asarray(lj["sigma"]).size)))
        sg = np.asarray(lj["sigma"], np.float64)
        ep = np.asarray(lj["epsilon"], np.float64)
        self._a = [np.ascontiguousarray(posq, np.float32).reshape(-1, 4),
                   np.ascontiguousarray(vel, np.float32).reshape(-1, 3),
                   np.ascontiguousarray(mass, np.float32),
                   np.ascontiguousarray(types, np.int32),
                   (4 * ep * sg ** 6).astype(np.float32), (4 * ep * sg ** 12).astype(np.float32),
                   np.ascontiguousarray(np.asarray(lattice)[:3], np.float32),
                   np.ascontiguousarray(exclusions[0], np.int32),
                   np.ascontiguousarray(exclusions[1], np.int32),
                   np.ascontiguousarray(mesh, np.int32),
                   np.ascontiguousarray(settle["groups"], np.int32)]
        a = self._a
        geom = settle_geometry(settle["d_oh"], settle["d_hh"], settle["m_o"], settle["m_h"])
        self.h = L.cs_create(n, _p(a[0], _fp), _p(a[1], _fp), _p(a[2], _fp), _p(a[3], _ip), nt,
                             _p(a[4], _fp), _p(a[5], _fp), _p(a[6], _fp),
                             float(np.asarray(lj["cutoff"]).flat[0]),
                             float(np.asarray(lj["switch"]).flat[0]), float(r_skin),
                             coulomb["alpha"], coulomb.get("k_coulomb", 1.0), _p(a[7], _ip),
                             _p(a[8], _ip), _p(a[9], _ip), len(a[10]), _p(a[10], _ip),
                             settle["m_o"], settle["m_h"], geom[0], geom[1], geom[2],
                             float(dt), 1 if compensated else 0)

    def step(self, n):
        step_lib().cs_step(self.h, n)

    def rebuilds(self):
        return step_lib().cs_rebuilds(self.h)

    def forces(self):
        f = np.empty((self.n, 3), np.float32)
        step_lib().cs_forces(self.h, _p(f, _fp))
        return f

    @staticmethod
    def set_threads(k):
        step_lib().cs_set_threads(int(k))

    @staticmethod
    def max_threads():
        return step_lib().cs_max_threads()

    def close(self):
        if self.h:
            step_lib().cs_destroy(self.h)
            self.h = None
