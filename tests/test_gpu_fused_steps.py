"""The fused integrator launches (hmd_step_vv_first / hmd_step_baoa /
hmd_step_kick_last: every sub-step between two force evaluations in one
launch, one thread per SETTLE group or free atom) against the reference's
one-launch-per-sub-step plan (HMD_INTEGRATOR_UNFUSED).  Bar: positions,
velocities, image counters and compensated low parts byte-identical after
40 graph-replayed steps, for compensated VV + SETTLE, Langevin BAOAB +
SETTLE, VV without constraints (argon) and a water box with half of its
molecules left unconstrained (groups and free atoms in one launch; there the
hydrogens get a small LJ core); and the same number of list rebuilds, the
fused plans running the skin check inside their first kernel."""
import numpy as np
import pytest

from heddlemd_amd import api, systems
from heddlemd_amd import units as U

pytestmark = pytest.mark.gpu

STEPS = 40


def run(kind, fused):
    if kind == "argon":
        a = systems.argon()
        rng = np.random.default_rng(5)
        posq = a["posq"].copy()
        posq[:, :3] += rng.normal(scale=0.05 * 3.405 * U.ANGSTROM, size=(a["n"], 3))
        ctx = api.GpuContext(0)
        part = api.ParticleBuffers(ctx, posq, a["mass"], a["lattice"], types=a["types"])
        field = api.ForceField(ctx, part, a["ff"])
        integ = api.Integrator(ctx, part, field, dt=5.0 * U.FS, fused=fused)
        part.upload(vel=api.generate_velocities(a["mass"], a["temperature"], 42))
    else:
        sysw = systems.spc_water(grid=(12, 12, 9), seed=7)
        ff = systems.water_forcefield(grid=(32, 32, 32), exclusions=sysw["exclusions"])
        if kind == "half_constrained":
            # the unconstrained hydrogens get a small LJ core, or they fall
            # onto neighbouring oxygens
            ff["lj"]["sigma"] = np.array([ff["lj"]["sigma"][0], 2.0 * U.ANGSTROM,
                                          2.0 * U.ANGSTROM, 1.5 * U.ANGSTROM])
            e = ff["lj"]["epsilon"][0]
            ff["lj"]["epsilon"] = np.array([e, 0.25 * e, 0.25 * e, 0.1 * e])
        ctx = api.GpuContext(0)
        part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                                   types=sysw["types"])
        field = api.ForceField(ctx, part, ff)
        settle = dict(sysw["settle"])
        if kind == "half_constrained":
            settle["groups"] = sysw["settle"]["groups"][::2].copy()
        lang = kind == "langevin"
        integ = api.Integrator(ctx, part, field, kind="langevin-baoab" if lang else "velocity-verlet",
                               dt=2.0 * U.FS if kind != "half_constrained" else 0.5 * U.FS,
                               compensated=(kind == "compensated"),
                               temperature=U.kelvin(300.0), gamma=1.0 / (1000.0 * U.FS), seed=9,
                               settle=settle, fused=fused)
        integ.minimize(40, 0.05 * U.ANGSTROM)
        part.upload(vel=api.generate_velocities(sysw["mass"], U.kelvin(300.0), 11,
                                                n_constraints=3 * sysw["n_mol"]))
    try:
        integ.replay(STEPS)
        d = part.download()
        d["pos_lo"], d["vel_lo"] = part.lo_parts()
        d["generation"] = field.stats()["generation"]
        return d
    finally:
        integ.close(); field.close(); part.close(); ctx.close()


@pytest.mark.parametrize("kind", ["compensated", "langevin", "argon", "half_constrained"])
def test_fused_plan_byte_identical(kind):
    a = run(kind, fused=True)
    b = run(kind, fused=False)
    for k in ("posq", "vel", "images", "pos_lo", "vel_lo"):
        same = a[k].tobytes() == b[k].tobytes()
        print(f"{kind}: {k} identical {same}", flush=True)
        assert same
    # the fused plans run the skin check inside their first kernel: the
    # same list rebuilds as the displacement kernel's
    print(f"{kind}: list generation fused {a['generation']}, unfused {b['generation']}")
    assert a["generation"] == b["generation"]
    if kind in ("compensated", "langevin"):
        assert a["generation"] > 2
