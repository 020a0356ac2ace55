"""Batched replicas (hmd_replicas_*): stepping independent boxes as one batch
changes nothing in any of them.  Three small boxes of the replica workload
(Langevin BAOAB + SETTLE at 300 K, each its own seed and context) are run
for 30 steps four ways: each on its own through its step graph, all as one
fused graph, all with per-replica graphs on their own streams, and all
merged into one batched launch per stage.  Bar:
positions and velocities byte-identical; per-replica CSV logs written from
two fused runs byte-identical."""
import numpy as np
import pytest

from heddlemd_amd import api, systems
from heddlemd_amd import units as U
from tests.gpu_helpers import replica_case

pytestmark = pytest.mark.gpu

RIDS = [25, 3, 6]
SPECS = {s["replica"]: s for s in systems.replica_specs()}
STEPS = 30


def build(rids):
    out = []
    for rid in rids:
        ctx = api.GpuContext(0)
        case = replica_case(ctx, SPECS[rid], relax=20, kind="langevin-baoab",
                            temperature=U.kelvin(300.0))
        sysw, ff, part, field, integ = case
        v = api.generate_velocities(sysw["mass"], U.kelvin(300.0), SPECS[rid]["seed"],
                                    n_constraints=3 * sysw["n_mol"])
        part.upload(vel=v)
        out.append((ctx, part, field, integ, sysw))
    return out


def close(reps):
    for ctx, part, field, integ, _ in reps:
        integ.close(); field.close(); part.close(); ctx.close()


def state(reps):
    return [(r[1].download()["posq"].copy(), r[1].download()["vel"].copy()) for r in reps]


def run(mode, log_dir=None):
    reps = build(RIDS)
    try:
        if mode == "alone":
            for r in reps:
                r[3].replay(STEPS)
        else:
            rs = api.ReplicaSet([r[3] for r in reps], mode=mode)
            try:
                if log_dir is None:
                    rs.step(STEPS)
                    if mode.startswith("batched"):
                        nb, ns = rs.plan()
                        print(f"batched plan: {nb} batched launches, {ns} single", flush=True)
                        assert nb > 0
                else:
                    logs = [api.LogWriter(log_dir / f"replica{rid}.out.log", units="si",
                                          extra_columns=[("potential_energy", "energy")])
                            for rid in RIDS]
                    for k in range(0, STEPS + 1, 10):
                        if k:
                            rs.step(10)
                        for (ctx, part, field, integ, sysw), lw in zip(reps, logs):
                            field.compute(1)
                            ke = part.kinetic_energy()
                            pe, _ = part.energy_sums()
                            dof = 3 * sysw["n"] - 3 * sysw["n_mol"] - 3
                            lw.write_row(k, k * 2.0 * U.FS, ke, 2.0 * ke / dof, [pe])
                    for lw in logs:
                        lw.close()
                assert rs.steps() == STEPS
            finally:
                rs.close()
        return state(reps)
    finally:
        close(reps)


def test_batched_steps_equal_solo_steps():
    alone = run("alone")
    for mode in ("fused", "per_replica", "batched", "batched_eager"):
        got = run(mode)
        for rid, (a, b) in zip(RIDS, zip(alone, got)):
            print(f"{mode}: replica {rid} positions/velocities identical: "
                  f"{np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])}", flush=True)
            assert a[0].tobytes() == b[0].tobytes()
            assert a[1].tobytes() == b[1].tobytes()


def test_per_replica_logs_byte_reproducible(tmp_path):
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    run("fused", tmp_path / "a")
    run("fused", tmp_path / "b")
    for rid in RIDS:
        a = (tmp_path / "a" / f"replica{rid}.out.log").read_bytes()
        b = (tmp_path / "b" / f"replica{rid}.out.log").read_bytes()
        assert a == b and a.count(b"\n") == 1 + STEPS // 10 + 1
