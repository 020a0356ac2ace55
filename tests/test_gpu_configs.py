"""GPU parity at the full sizes of the two water workloads, against the FP64
oracle on the same f32 positions:

* 8,192 SPC molecules (24,576 atoms, 6.26 nm box, seed 7), 64^3 SPME,
  Langevin BAOAB at 300 K + SETTLE;
* 65,536 SPC molecules (196,608 atoms, 12.53 nm box, seed 11), 128^3 SPME,
  compensated velocity Verlet + SETTLE, NVE from 300 K.

Each case checks the neighbour pair set bit for bit (cell-list form of the
oracle's exact f32 enumeration), forces (relative RMS <= 1e-5), potential
energy and W_recip (relative <= 1e-6), the per-atom virial (relative RMS
<= 1e-5) and the FP64 virial sum (within 1e-6 of sum |W_i|), at the start
and again after graph-replayed MD.  The 196,608-atom NVE run is 1,000 steps:
the total energy must stay within 5% of |E0| over the run (the reference's
bound, tests/e2e.rs:118-128) and, tighter, its drift within 2% of the mean
kinetic energy; a 50-step forward / velocity-reversed / 50-step backward
run returns every atom to its start within 1e-3 of the largest |x| (the
bound of the reference's approximate-reversibility test,
tests/integrators_mtk_npt.rs:512-625): SETTLE's projection is not
bit-reversible and the liquid's chaotic dynamics amplify the f32 force
roundings over 100 steps, so this bar is a tolerance, not bit identity.
"""
import time

import numpy as np
import pytest

import oracle
from heddlemd_amd import api, systems
from heddlemd_amd import units as U
from tests.gpu_helpers import oracle_pair, rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6


def _log(msg):
    print(msg, flush=True)


def _check_against_oracle(tag, sysw, ff, part, field):
    """Energy-level evaluation at the current positions vs the oracle.  The
    list is complete within r_c on its own (the count of listed pairs within
    r_c equals the oracle's); a rebuild at these positions then makes it the
    exact r_c + r_skin set, compared pair for pair."""
    field.compute(1)
    posq0 = part.download()["posq"]
    rc = np.float32(ff["coulomb"]["r_cut"])
    n_rc = len(oracle.pair_set_f32(posq0, sysw["lattice"].astype(np.float32), rc * rc,
                                   cells=True))
    n_rc_gpu = field.count_pairs(float(ff["coulomb"]["r_cut"]))
    assert n_rc_gpu == n_rc, (n_rc_gpu, n_rc)
    field.request_rebuild()
    field.compute(1)
    d = part.download()
    posq = d["posq"]
    e_gpu, w_gpu = part.energies()
    pe_gpu, vir_gpu = part.energy_sums()
    t0 = time.perf_counter()
    got = field.pair_snapshot()
    rs = np.float32(np.float32(ff["coulomb"]["r_cut"]) + np.float32(ff["r_skin"]))
    want = oracle.pair_set_f32(posq, sysw["lattice"].astype(np.float32), rs * rs, cells=True)
    same_pairs = got.shape == want.shape and np.array_equal(got, want)
    f_pair, e_pair, w_pair = oracle_pair(sysw, ff, posq)
    u_rec, f_rec, e_rec, w_rec = oracle.spme_recip(
        posq[:, :3], posq[:, 3], sysw["lattice"], ff["spme"]["grid"], ff["spme"]["order"],
        ff["coulomb"]["alpha"])
    err_f = rel_rms(d["force"], f_pair + f_rec)
    pe_ref = e_pair.sum() + e_rec.sum()
    err_e = abs(pe_gpu - pe_ref) / abs(pe_ref)
    sc = field.spme_scalars()
    err_u = abs(sc["u_recip"] - u_rec) / abs(u_rec)
    err_w = abs(sc["w_recip"] - w_rec) / abs(w_rec)
    w_ref = w_pair + w_rec / len(w_pair)
    err_wi = rel_rms(w_gpu, w_ref)
    err_ws = abs(vir_gpu - w_ref.sum()) / np.abs(w_ref).sum()
    _log(f"{tag}: pairs gpu {len(got)} oracle {len(want)} equal={same_pairs}; force rel RMS "
         f"{err_f:.2e}; PE {err_e:.2e}; U_recip {err_u:.2e}; W_recip {err_w:.2e}; virial "
         f"per atom {err_wi:.2e}, sum {err_ws:.2e} (oracle {time.perf_counter() - t0:.1f} s)")
    assert same_pairs
    assert err_f <= FORCE_TOL
    assert err_e <= ENERGY_TOL and err_u <= ENERGY_TOL and err_w <= ENERGY_TOL
    assert err_wi <= FORCE_TOL and err_ws <= ENERGY_TOL


def _case(ctx, grid, layout_seed, mesh, **integ_kw):
    sysw = systems.spc_water(grid=grid, seed=layout_seed)
    ff = systems.water_forcefield(grid=mesh, exclusions=sysw["exclusions"])
    part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                               types=sysw["types"])
    field = api.ForceField(ctx, part, ff)
    integ = api.Integrator(ctx, part, field, dt=2.0 * U.FS, settle=sysw["settle"], **integ_kw)
    return sysw, ff, part, field, integ


def test_water_24k_langevin_64_mesh(gpu_ctx):
    sysw, ff, part, field, integ = _case(gpu_ctx, (32, 32, 8), 7, (64, 64, 64),
                                         kind="langevin-baoab",
                                         temperature=U.kelvin(300.0),
                                         gamma=1.0 / (1000.0 * U.FS), seed=7)
    try:
        integ.minimize(100, 0.1 * U.ANGSTROM)
        part.upload(vel=api.generate_velocities(sysw["mass"], U.kelvin(300.0), 7,
                                                n_constraints=3 * sysw["n_mol"]))
        _check_against_oracle("24,576 atoms, start", sysw, ff, part, field)
        integ.capture()
        integ.replay(200)
        _check_against_oracle("24,576 atoms, after 200 Langevin steps", sysw, ff, part, field)
    finally:
        for h in (integ, field, part):
            h.close()


def test_water_196k_compensated_nve_128_mesh(gpu_ctx):
    sysw, ff, part, field, integ = _case(gpu_ctx, (64, 64, 16), 11, (128, 128, 128),
                                         compensated=True, scalars=True)
    try:
        integ.minimize(200, 0.1 * U.ANGSTROM)
        part.upload(vel=api.generate_velocities(sysw["mass"], U.kelvin(300.0), 11,
                                                n_constraints=3 * sysw["n_mol"]))
        _check_against_oracle("196,608 atoms, start", sysw, ff, part, field)
        integ.capture()
        energies = []
        for block in range(10):
            integ.replay(100)
            pe, _ = part.energy_sums()
            ke = part.kinetic_energy()
            energies.append((ke, pe, ke + pe))
        e = np.array(energies)
        e0 = e[0, 2]
        drift = np.abs(e[:, 2] - e0).max()
        ke_mean = e[:, 0].mean()
        _log(f"196,608-atom NVE, 1,000 steps: E0 {e0:.6e} Ha, max |E - E0| {drift:.3e} Ha "
             f"({drift / abs(e0):.2e} of |E0|, {drift / ke_mean:.2e} of <KE>); "
             f"KE {e[0, 0]:.4e} -> {e[-1, 0]:.4e}")
        assert drift <= 0.05 * abs(e0)
        assert drift <= 0.02 * ke_mean
        _check_against_oracle("196,608 atoms, after 1,000 NVE steps", sysw, ff, part, field)
        # time reversal: 50 steps forward, reverse velocities, 50 steps back
        start = part.download()["posq"][:, :3].astype(np.float64)
        integ.replay(50)
        part.negate_velocities()
        integ.replay(50)
        back = part.download()["posq"][:, :3].astype(np.float64)
        L = sysw["lattice"][0]
        dx = back - start
        dx -= L * np.round(dx / L)
        err = np.abs(dx).max()
        tol = 1e-3 * np.abs(start).max()
        _log(f"time reversal over 2 x 50 steps: max position error {err:.2e} bohr, median "
             f"{np.median(np.abs(dx)):.2e} (bound {tol:.2e})")
        assert err < tol
    finally:
        for h in (integ, field, part):
            h.close()
