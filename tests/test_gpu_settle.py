"""GPU SETTLE operators (hmd_settle_before_drift / after_drift / after_kick)
against the FP64 oracle (orc_settle_positions_one / _velocities_one, pinned
against converged SHAKE and the RATTLE linear solve), molecule by molecule,
on the 3,888-atom water box after a free drift of 2 fs at thermal speeds x 5
(large corrections).  The GPU works in f32, the oracle in f64 on the same
f32 inputs.  Bars: positions within 2e-5 bohr (a few f32 ulps of the ~20
bohr coordinates), bond lengths within 2e-6 relative of d_OH / d_HH, the
velocity corrections within 1e-6 of the largest velocity; after the
velocity projection every bond's relative velocity is < 1e-5 of it."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api, systems
from heddlemd_amd import units as U

pytestmark = pytest.mark.gpu


def test_settle_operators_match_oracle(gpu_ctx):
    w = systems.spc_water(grid=(12, 12, 9), seed=7)
    st = w["settle"]
    rng = np.random.default_rng(3)
    vel = api.generate_velocities(w["mass"], U.kelvin(300.0), 5) * 5.0
    part = api.ParticleBuffers(gpu_ctx, w["posq"], w["mass"], w["lattice"], vel=vel,
                               types=w["types"])
    settle = api.Settle(part, st["groups"], st["d_oh"], st["d_hh"], st["m_o"], st["m_h"])
    try:
        dt = np.float32(2.0 * U.FS)
        part.upload_forces(np.zeros((w["n"], 3), np.float32))
        snap = part.download()["posq"][:, :3].astype(np.float64)
        settle.before_drift()
        part.kick_drift(dt)                    # zero force: x += v dt, wrapped
        mid = part.download()
        settle.after_drift(dt)
        got = part.download()
        v_in = rng.standard_normal((w["n"], 3)).astype(np.float32) * np.float32(1e-3)
        part.upload(vel=v_in)
        settle.after_kick(dt)
        got_v = part.download()["vel"]
    finally:
        settle.close()
        part.close()
    geom = oracle.settle_geometry(st["d_oh"], st["d_hh"], st["m_o"], st["m_h"])
    L = w["lattice"]
    err_x = err_v = err_b = err_p = 0.0
    vmax = np.abs(mid["vel"]).max()
    for g in st["groups"]:
        s = snap[g]
        c = mid["posq"][g, :3].astype(np.float64)
        # unwrap the drifted molecule next to its snapshot (the kernels use
        # minimum images inside a group)
        c = c - np.round((c - s) / L[0]) * L[0]
        new, v2, _ = oracle.settle_positions_one(s, c, mid["vel"][g].astype(np.float64),
                                                 st["m_o"], st["m_h"], geom, float(dt), L)
        gx = got["posq"][g, :3].astype(np.float64)
        gx = gx - np.round((gx - new) / L[0]) * L[0]
        err_x = max(err_x, np.abs(gx - new).max())
        err_v = max(err_v, np.abs(got["vel"][g] - v2).max() / vmax)
        d = [np.linalg.norm(gx[a] - gx[b]) for a, b in ((0, 1), (0, 2), (1, 2))]
        err_b = max(err_b, abs(d[0] / st["d_oh"] - 1), abs(d[1] / st["d_oh"] - 1),
                    abs(d[2] / st["d_hh"] - 1))
        pos = got["posq"][g, :3].astype(np.float64)
        pos = pos - np.round((pos - pos[0]) / L[0]) * L[0]
        vw = oracle.settle_velocities_one(pos, v_in[g].astype(np.float64), st["m_o"], st["m_h"], L)
        err_p = max(err_p, np.abs(got_v[g] - vw).max() / np.abs(v_in).max())
    print(f"SETTLE vs oracle: positions {err_x:.2e} bohr, bonds {err_b:.2e} rel, "
          f"velocity fix {err_v:.2e}, velocity projection {err_p:.2e}")
    assert err_x < 2e-5 and err_b < 2e-6 and err_v < 1e-6 and err_p < 1e-5


@pytest.mark.parametrize("fused", [True, False])
def test_step_virial_adds_the_settle_constraint_virial(gpu_ctx, fused):
    """An energy-level MD step (HMD_INTEGRATOR_SCALARS) leaves in each atom's
    virial its pair and reciprocal shares at the new positions plus the
    step's SETTLE constraint virial (position level after the drift,
    velocity level after the kick; settle_virial_scatter,
    kernels/settle.cu:498-510).  Replayed on the host: the f32 kick-drift
    (oracle VvState, bit-exact with the device), the FP64 SETTLE oracle per
    molecule, the kick with the device's new forces, the FP64 pair and
    reciprocal virials.  Bars: the constraint part (device total minus the
    oracle force virial) within 1e-4 relative RMS of the oracle's SETTLE
    virial, and the per-atom total within 1e-5 relative RMS (f32 SETTLE on
    f32 inputs against FP64)."""
    from tests.gpu_helpers import oracle_pair, rel_rms
    w = systems.spc_water(grid=(12, 12, 9), seed=7)
    ff = systems.water_forcefield(grid=(32, 32, 32), exclusions=w["exclusions"])
    st = w["settle"]
    part = api.ParticleBuffers(gpu_ctx, w["posq"], w["mass"], w["lattice"], types=w["types"])
    field = api.ForceField(gpu_ctx, part, ff)
    dt = 2.0 * U.FS
    integ = api.Integrator(gpu_ctx, part, field, dt=dt, settle=st, fused=fused, scalars=True)
    try:
        integ.minimize(40, 0.05 * U.ANGSTROM)
        part.upload(vel=api.generate_velocities(w["mass"], U.kelvin(300.0), 9,
                                                n_constraints=3 * w["n_mol"]))
        integ.step(1)
        s0 = part.download()
        integ.step(1)
        s1 = part.download()
        _, w_gpu = part.energies()
    finally:
        for h in (integ, field, part):
            h.close()
    L = w["lattice"]
    mass = w["mass"].astype(np.float32)
    vv = oracle.VvState(s0["posq"], s0["vel"], s0["force"], mass, L, images=s0["images"])
    vv.kick_drift(np.float32(dt))
    unc = vv.posq[:, :3].astype(np.float64)
    v_half = vv.vel.astype(np.float64)
    snap = s0["posq"][:, :3].astype(np.float64)
    x1 = s1["posq"][:, :3].astype(np.float64)
    inv_m = 1.0 / mass.astype(np.float64)
    geom = oracle.settle_geometry(st["d_oh"], st["d_hh"], st["m_o"], st["m_h"])
    w_settle = np.zeros(w["n"])
    err_x = 0.0
    for g in st["groups"]:
        s = snap[g]
        c = unc[g] - np.round((unc[g] - s) / L[0]) * L[0]
        new, v2, vir_p = oracle.settle_positions_one(s, c, v_half[g], st["m_o"], st["m_h"],
                                                     geom, float(np.float32(dt)), L)
        gx = x1[g] - np.round((x1[g] - new) / L[0]) * L[0]
        err_x = max(err_x, np.abs(gx - new).max())
        v3 = v2 + s1["force"][g].astype(np.float64) * inv_m[g, None] * (0.5 * dt)
        _, vir_v = oracle.settle_velocities_one(gx, v3, st["m_o"], st["m_h"], L,
                                                dt=float(np.float32(dt)))
        w_settle[g] = vir_p + vir_v
    _, _, w_pair = oracle_pair(w, ff, s1["posq"])
    *_, w_rec = oracle.spme_recip(s1["posq"][:, :3], s1["posq"][:, 3], L, ff["spme"]["grid"], 4,
                                  ff["coulomb"]["alpha"])
    w_force = w_pair + w_rec / w["n"]
    err_c = rel_rms(w_gpu.astype(np.float64) - w_force, w_settle)
    err_t = rel_rms(w_gpu, w_force + w_settle)
    print(f"fused={fused}: positions vs oracle SETTLE {err_x:.2e} bohr; constraint virial "
          f"rel RMS {err_c:.2e} (RMS {np.sqrt(np.mean(w_settle ** 2)):.3e} vs force virial "
          f"{np.sqrt(np.mean(w_force ** 2)):.3e}); total {err_t:.2e}")
    assert err_x < 2e-5
    assert err_c <= 1e-4
    assert err_t <= 1e-5
