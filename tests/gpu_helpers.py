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
