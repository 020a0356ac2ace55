"""The c-rescale barostat (src/integrator/c_rescale_barostat.rs:172-232,
kernels/barostat.cu:93-227) on the GPU through the C ABI.

* One move against the oracle (orc_crescale_mu, pinned by its Berendsen and
  equilibrium limits and its noise statistics): the kinetic energy and the
  virial sum are those of a twin box stepped with the operator entry points
  and stopped at the barostat point (after the last kick, before the SETTLE
  velocity projection); positions and all six lattice components are the
  twin's multiplied by the device's (float) mu, to the bit.
* The target pressure drives the volume: two runs from the same state and
  seed, at +2,000 and -2,000 bar, end with the volume ordered as the targets
  and apart by roughly beta (dt / tau) (P_high - P_low) per step.
* The accumulated P_target dV matches the volume change; a run after a
  rollback repeats the first byte for byte, box included; capture is refused
  and replay steps eagerly; the Langevin integrator carries the barostat too.
"""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api
from heddlemd_amd import units as U
from heddlemd_amd._lib import HmdError
from tests.gpu_helpers import water_case

pytestmark = pytest.mark.gpu

DT = 2.0 * U.FS
TAU_P = 1000.0 * U.FS
BETA = 4.5e-10 * U.PRESSURE_AU_TO_PA     # liquid water, 1/Pa in atomic units
KT = U.kelvin(300.0)
SEED = 0x5EED0C0FFEE


def _bar(b):
    return b * 1.0e5 / U.PRESSURE_AU_TO_PA


def _volume(lat):
    return float(np.float64(lat[0]) * np.float64(lat[1]) * np.float64(lat[2]))


def _integrator(ctx, case, pressure, seed=SEED, **kw):
    sysw, _, part, field, _ = case
    it = api.Integrator(ctx, part, field, dt=DT, settle=sysw["settle"], **kw)
    it.set_barostat("c-rescale", pressure=pressure, temperature=KT, tau=TAU_P,
                    compressibility=BETA, seed=seed)
    return it


def _thermal(case, seed=5):
    sysw, _, part, _, _ = case
    part.upload(vel=api.generate_velocities(sysw["mass"], KT, seed,
                                            n_constraints=3 * sysw["n_mol"]))


def _close(case, *its):
    for it in its:
        it.close()
    for h in case[2:][::-1]:
        h.close()


def test_crescale_move_matches_the_oracle(gpu_ctx):
    a = water_case(gpu_ctx, relax=40)
    b = water_case(gpu_ctx, relax=40)
    _thermal(a)
    _thermal(b)
    sysw = a[0]
    st = sysw["settle"]
    it = _integrator(gpu_ctx, a, _bar(1.0), fused=False)
    op = api.Settle(b[2], st["groups"], st["d_oh"], st["d_hh"], st["m_o"], st["m_h"])
    try:
        lat0 = a[2].lattice.copy()
        it.step(1)
        pb, fb = b[2], b[3]
        fb.compute(1)
        op.before_drift()
        pb.kick_drift(DT)
        op.after_drift(DT)
        fb.compute(1)
        pb.kick(DT)
        k = pb.kinetic_energy()
        _, w = pb.energy_sums()
        op.after_kick(DT)
        x_b = pb.download()["posq"]
        v0 = _volume(lat0)
        mu, pres, v_post, inj = oracle.crescale_mu(k, w, v0, SEED, 0, _bar(1.0), TAU_P, BETA, KT,
                                                   float(np.float32(DT)))
        p_d, v_d, inj_d, draws, mu_d = it.barostat_state()
        print(f"K {k:.6e}, W {w:.6e}, P {pres * U.PRESSURE_AU_TO_PA / 1e5:.1f} bar; "
              f"mu oracle {mu!r}, device {mu_d!r}")
        assert draws == 1
        assert p_d == pytest.approx(pres, rel=1e-13)
        assert mu_d == pytest.approx(mu, rel=1e-14)
        assert v_d == pytest.approx(v_post, rel=1e-13)
        assert inj_d == pytest.approx(inj, rel=1e-9, abs=1e-12 * abs(_bar(1.0)) * v0)
        f = np.float32(mu_d)
        x_a = a[2].download()["posq"]
        assert np.array_equal(x_a[:, :3], x_b[:, :3] * f)
        assert np.array_equal(x_a[:, 3], x_b[:, 3])
        assert np.array_equal(a[2].lattice, lat0 * f)
    finally:
        op.close()
        it.close()
        for case in (a, b):
            _close(case)


def test_crescale_target_pressure_drives_the_volume(gpu_ctx):
    n_steps = 150
    rel = {}
    for bar in (2000.0, -2000.0):
        case = water_case(gpu_ctx, relax=40)
        _thermal(case)
        it = _integrator(gpu_ctx, case, _bar(bar))
        try:
            v0 = _volume(case[2].lattice)
            it.replay(n_steps)
            lat = case[2].lattice
            p, v, inj, draws, mu = it.barostat_state()
            assert draws == n_steps
            assert v == pytest.approx(_volume(lat), rel=1e-6)
            # sum of P_target (V_post - V_pre) over the moves: the volume change
            # up to the f32 rounding of the box after each move
            assert inj == pytest.approx(_bar(bar) * (v - v0), rel=2e-3)
            x = case[2].download()["posq"]
            assert np.isfinite(x).all()
            rel[bar] = v / v0
            print(f"target {bar:+.0f} bar: V/V0 {v / v0:.5f}, last P "
                  f"{p * U.PRESSURE_AU_TO_PA / 1e5:.0f} bar")
        finally:
            _close(case, it)
    open_loop = BETA * (float(np.float32(DT)) / TAU_P) * (_bar(2000.0) - _bar(-2000.0)) * n_steps
    gap = rel[-2000.0] - rel[2000.0]
    print(f"volume gap {gap:.5f}, open-loop {open_loop:.5f}")
    assert 0.3 * open_loop < gap < 1.3 * open_loop


def test_crescale_rollback_repeats_and_capture_is_refused(gpu_ctx):
    case = water_case(gpu_ctx, relax=40)
    _thermal(case)
    it = _integrator(gpu_ctx, case, _bar(1.0), thermostat="csvr", temperature=KT,
                     tau=100.0 * U.FS, seed=3)
    try:
        part = case[2]
        it.step(2)
        lat0 = part.lattice.copy()
        it.checkpoint()
        it.step(30)
        d1 = part.download()
        lat1 = part.lattice.copy()
        s1 = it.barostat_state()
        t1 = it.thermostat_state()
        assert not np.array_equal(lat1, lat0)
        it.rollback()
        assert np.array_equal(part.lattice, lat0)
        it.replay(30)
        d2 = part.download()
        assert np.array_equal(part.lattice, lat1)
        for key in ("posq", "vel", "images"):
            assert np.array_equal(d1[key], d2[key])
        assert it.barostat_state() == s1
        assert it.thermostat_state() == t1
        with pytest.raises(HmdError, match="eagerly"):
            it.capture()
        with pytest.raises(HmdError, match="eagerly"):
            api.ReplicaSet([it])
        it.set_barostat(None)
        it.capture()
        it.replay(5)
        assert np.array_equal(part.lattice, lat1)
    finally:
        _close(case, it)


def test_crescale_with_langevin(gpu_ctx):
    case = water_case(gpu_ctx, relax=40)
    _thermal(case)
    it = _integrator(gpu_ctx, case, _bar(1.0), kind="langevin-baoab", temperature=KT,
                     gamma=1.0 / (1000.0 * U.FS), seed=9)
    try:
        lat0 = case[2].lattice.copy()
        it.step(50)
        _, v, _, draws, _ = it.barostat_state()
        assert draws == 50
        assert v == pytest.approx(_volume(case[2].lattice), rel=1e-6)
        assert not np.array_equal(case[2].lattice, lat0)
        assert np.isfinite(case[2].download()["posq"]).all()
    finally:
        _close(case, it)
