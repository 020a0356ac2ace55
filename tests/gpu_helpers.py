"""Shared set-up for the GPU parity tests: build a system on the device and
the matching FP64 oracle inputs."""
import numpy as np

import oracle
from heddlemd_amd import api, systems
from heddlemd_amd import units as U


def water_case(ctx, grid=(12, 12, 9), seed=7, mesh=(32, 32, 32), order=4, relax=0,
               r_skin_nm=0.1, spme=True):
    sysw = systems.spc_water(grid=grid, seed=seed)
    ff = systems.water_forcefield(grid=mesh if spme else None, order=order,
                                  r_skin_nm=r_skin_nm, exclusions=sysw["exclusions"])
    part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                               types=sysw["types"])
    field = api.ForceField(ctx, part, ff)
    integ = api.Integrator(ctx, part, field, dt=2.0 * U.FS, settle=sysw["settle"])
    if relax:
        integ.minimize(relax, 0.05 * U.ANGSTROM)
    return sysw, ff, part, field, integ


def replica_case(ctx, spec, relax=0, kind="velocity-verlet", temperature=0.0):
    """One replica box of systems.replica_specs on the device."""
    sysw = systems.spc_water_n(spec["n_mol"], spec["seed"])
    ff = systems.water_forcefield(r_cut_nm=spec["r_cut_nm"], grid=spec["mesh"],
                                  exclusions=sysw["exclusions"])
    part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                               types=sysw["types"])
    field = api.ForceField(ctx, part, ff)
    integ = api.Integrator(ctx, part, field, kind=kind, dt=2.0 * U.FS, settle=sysw["settle"],
                           temperature=temperature, gamma=1.0 / (1000.0 * U.FS),
                           seed=spec["seed"])
    if relax:
        integ.minimize(relax, 0.05 * U.ANGSTROM)
    return sysw, ff, part, field, integ


def oracle_pair(sysw, ff, posq, erfc_mode=1):
    lj = ff["lj"]
    exl = ff["exclusions"]
    off, par = exl[0], exl[1]
    slj = np.asarray(exl[2], np.float64) if len(exl) == 4 else np.zeros(par.size)
    sco = np.asarray(exl[3], np.float64) if len(exl) == 4 else np.zeros(par.size)
    ex = (off.astype(np.int32), par.astype(np.int32), slj, sco)
    ps = oracle.PairSystem(posq[:, :3].astype(np.float64), posq[:, 3].astype(np.float64),
                           sysw["types"].astype(np.int32), sysw["lattice"],
                           lj=lj, coulomb=ff["coulomb"], exclusions=ex, erfc_mode=erfc_mode)
    return ps.forces(cells=True)


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b ** 2)))


class HipRuntime:
    """Device arrays, a stream and stream capture through ctypes on the HIP
    runtime that the library itself links (the ROCm install's
    libamdhip64.so.7), as a C caller would use them.  torch's wheel bundles
    a separate copy of the runtime; two runtimes in one process do not see
    the same devices, so tests that hand device memory or a stream to the
    library take them from this one."""

    H2D, D2H, D2D = 1, 2, 3
    CAPTURE_GLOBAL = 0

    def __init__(self):
        import ctypes
        from heddlemd_amd._lib import lib
        lib()   # the library's own runtime is loaded first; the soname resolves to it
        self.ct = ctypes
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.allocs = []

    def _ok(self, code, what):
        if code != 0:
            raise RuntimeError(f"{what} failed with HIP error {code}")

    def malloc(self, nbytes):
        p = self.ct.c_void_p()
        self._ok(self.rt.hipMalloc(self.ct.byref(p), self.ct.c_size_t(nbytes)), "hipMalloc")
        self.allocs.append(p.value)
        return p.value

    def zeros(self, nbytes):
        p = self.malloc(nbytes)
        self._ok(self.rt.hipMemset(self.ct.c_void_p(p), 0, self.ct.c_size_t(nbytes)), "hipMemset")
        self.device_sync()
        return p

    def upload(self, dst, host):
        host = np.ascontiguousarray(host)
        self._ok(self.rt.hipMemcpy(self.ct.c_void_p(dst), self.ct.c_void_p(host.ctypes.data),
                                   self.ct.c_size_t(host.nbytes), self.H2D), "hipMemcpy H2D")
        self.device_sync()   # the copy has landed before any stream reads it

    def download(self, src, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        self._ok(self.rt.hipMemcpy(self.ct.c_void_p(out.ctypes.data), self.ct.c_void_p(src),
                                   self.ct.c_size_t(out.nbytes), self.D2H), "hipMemcpy D2H")
        return out

    def copy_async(self, dst, src, nbytes, stream):
        self._ok(self.rt.hipMemcpyAsync(self.ct.c_void_p(dst), self.ct.c_void_p(src),
                                        self.ct.c_size_t(nbytes), self.D2D,
                                        self.ct.c_void_p(stream)), "hipMemcpyAsync")

    def device_sync(self):
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def stream_create(self):
        s = self.ct.c_void_p()
        self._ok(self.rt.hipStreamCreate(self.ct.byref(s)), "hipStreamCreate")
        return s.value

    def stream_sync(self, s):
        self._ok(self.rt.hipStreamSynchronize(self.ct.c_void_p(s)), "hipStreamSynchronize")

    def stream_destroy(self, s):
        self._ok(self.rt.hipStreamDestroy(self.ct.c_void_p(s)), "hipStreamDestroy")

    def begin_capture(self, s):
        self._ok(self.rt.hipStreamBeginCapture(self.ct.c_void_p(s), self.CAPTURE_GLOBAL),
                 "hipStreamBeginCapture")

    def end_capture(self, s):
        """Ends the capture and instantiates it: (graph, executable)."""
        g = self.ct.c_void_p()
        self._ok(self.rt.hipStreamEndCapture(self.ct.c_void_p(s), self.ct.byref(g)),
                 "hipStreamEndCapture")
        ex = self.ct.c_void_p()
        self._ok(self.rt.hipGraphInstantiate(self.ct.byref(ex), g, None, None,
                                             self.ct.c_size_t(0)), "hipGraphInstantiate")
        return g.value, ex.value

    def launch(self, ex, s):
        self._ok(self.rt.hipGraphLaunch(self.ct.c_void_p(ex), self.ct.c_void_p(s)),
                 "hipGraphLaunch")

    def graph_destroy(self, g, ex):
        self._ok(self.rt.hipGraphExecDestroy(self.ct.c_void_p(ex)), "hipGraphExecDestroy")
        self._ok(self.rt.hipGraphDestroy(self.ct.c_void_p(g)), "hipGraphDestroy")

    def free_all(self):
        for p in self.allocs:
            self.rt.hipFree(self.ct.c_void_p(p))
        self.allocs = []
