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
