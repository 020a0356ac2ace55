"""hmd_integrator_checkpoint / hmd_integrator_rollback through the C ABI:
ten graph-replayed steps from a checkpoint, a rollback, and the same ten
steps again give byte-identical positions, velocities, forces, images, step
count and (CSVR) injected energy and draw count, for compensated velocity
Verlet with the CSVR thermostat (f64 low parts, the thermostat's device
counter) and Langevin BAOAB (the Philox step counter).  The list is rebuilt
at the first step of both runs (a rebuild request before the checkpoint, the
rollback's own after it), so both runs evaluate the same lists."""
import numpy as np
import pytest

from heddlemd_amd import api
from heddlemd_amd import units as U
from tests.gpu_helpers import water_case

pytestmark = pytest.mark.gpu


def _state(part, it, csvr):
    d = part.download()
    out = {k: d[k].view(np.uint32).copy() for k in ("posq", "vel", "force", "images")}
    out["steps"] = it.steps()
    if csvr:
        out["thermostat"] = it.thermostat_state()
    return out


@pytest.mark.parametrize("kind", ["compensated_csvr", "langevin"])
def test_rollback_reruns_the_batch_byte_for_byte(gpu_ctx, kind):
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=20)
    csvr = kind == "compensated_csvr"
    if csvr:
        it = api.Integrator(gpu_ctx, part, field, dt=2.0 * U.FS, settle=sysw["settle"],
                            compensated=True, temperature=U.kelvin(300.0), seed=77,
                            thermostat="csvr", tau=100.0 * U.FS)
    else:
        it = api.Integrator(gpu_ctx, part, field, kind="langevin-baoab", dt=2.0 * U.FS,
                            settle=sysw["settle"], temperature=U.kelvin(300.0),
                            gamma=1.0 / (1000.0 * U.FS), seed=77)
    try:
        part.upload(vel=api.generate_velocities(sysw["mass"], U.kelvin(300.0), 9,
                                                n_constraints=3 * sysw["n_mol"]))
        it.capture()
        it.replay(5)
        field.request_rebuild()
        it.checkpoint()
        it.replay(10)
        first = _state(part, it, csvr)
        it.rollback()
        assert it.steps() == first["steps"] - 10
        it.replay(10)
        second = _state(part, it, csvr)
        for k in first:
            if k in ("steps", "thermostat"):
                assert first[k] == second[k], k
            else:
                assert np.array_equal(first[k], second[k]), k
    finally:
        it.close()
        for h in (integ, field, part):
            h.close()


def test_rollback_without_checkpoint_is_refused(gpu_ctx):
    from heddlemd_amd._lib import HmdError
    sysw, ff, part, field, integ = water_case(gpu_ctx)
    try:
        with pytest.raises(HmdError):
            integ.rollback()
    finally:
        for h in (integ, field, part):
            h.close()
