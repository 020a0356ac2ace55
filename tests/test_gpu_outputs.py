"""Run-to-run byte reproducibility of the MD outputs on the GPU: two
identical graph-replayed runs of the 3,888-atom water box write their
extended-XYZ trajectory (positions, velocities, images) and CSV log
(kinetic energy, temperature, potential energy) through the format writers;
the files must be byte-identical."""
import numpy as np
import pytest

from heddlemd_amd import api
from heddlemd_amd import units as U
from tests.gpu_helpers import water_case

pytestmark = pytest.mark.gpu


def run_and_write(gpu_ctx, tmp_path, tag):
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=20)
    v = api.generate_velocities(sysw["mass"], U.kelvin(300.0), 7,
                                n_constraints=3 * sysw["n_mol"])
    part.upload(vel=v)
    n_dof = 3 * sysw["n"] - 3 * sysw["n_mol"] - 3
    traj = api.TrajectoryWriter(tmp_path / f"{tag}.out.xyz", "si", True, True, ["O", "H"])
    log = api.LogWriter(tmp_path / f"{tag}.out.log", "si", [("potential_energy", "energy")])
    dt = 2.0 * U.FS
    for step in range(0, 41, 5):
        if step:
            integ.replay(5)
        field.compute(1)
        d = part.download()
        ke = part.kinetic_energy()
        pe, _ = part.energy_sums()
        log.write_row(step, step * dt, ke, 2.0 * ke / n_dof, [pe])
        if step % 10 == 0:
            traj.write_frame(step, dt, sysw["lattice"], sysw["types"], d["posq"][:, :3],
                             velocities=d["vel"], images=d["images"])
    traj.close()
    log.close()
    for h in (integ, field, part):
        h.close()
    return ((tmp_path / f"{tag}.out.xyz").read_bytes(), (tmp_path / f"{tag}.out.log").read_bytes())


def test_trajectory_and_log_are_byte_identical_run_to_run(gpu_ctx, tmp_path):
    a = run_and_write(gpu_ctx, tmp_path, "a")
    b = run_and_write(gpu_ctx, tmp_path, "b")
    print(f"trajectory {len(a[0])} bytes, log {len(a[1])} bytes; "
          f"log tail: {a[1].decode().splitlines()[-1]}")
    assert len(a[0].splitlines()) == 5 * (2 + 3888)
    assert a[0] == b[0]
    assert a[1] == b[1]


def test_timings_report_of_eager_steps(gpu_ctx, tmp_path):
    """hmd_integrator_timings over 5 eager steps: every stage of the plan is
    counted once per step with min <= mean <= max; the .timings report has
    the reference's header and one row per stage (tests/timings.rs:136-151,
    :583-600)."""
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=5)
    try:
        stages = integ.timings(5)
    finally:
        for h in (integ, field, part):
            h.close()
    names = [s["name"] for s in stages]
    print("stages:", ", ".join(f"{s['name']} {s['total_ns'] / s['count'] / 1e3:.1f} us"
                               for s in stages))
    assert "pair_force" in names and "kick_settle" in names
    for s in stages:
        assert s["count"] == 5
        assert s["min_ns"] * 5 <= s["total_ns"] <= s["max_ns"] * 5
    p = tmp_path / "run.out.timings"
    api.write_timings(p, stages)
    lines = p.read_text().splitlines()
    assert lines[0].split() == ["stage", "count", "total_ms", "mean_us", "min_us", "max_us"]
    assert [ln.split()[0] for ln in lines[1:]] == names
