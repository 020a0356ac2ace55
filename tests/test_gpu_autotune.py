"""Launch-shape autotuner (hmd_forcefield_autotune / set_tuning / get_tuning)
on a 3,888-atom SPC water box with a 32^3 mesh.  Bars: SPME brick and FFT
grid changes leave the forces byte-identical (exact integer mesh sums; the
same per-line transforms); a pair grid change regroups the f32 per-wave sums
and moves the forces by at most 1e-6 relative RMS; the tuned shapes applied
to a fresh force field (its list built from the same positions) give
byte-identical forces; 5 graph-replayed steps
after a brick / FFT change give byte-identical positions and velocities; an
invalid brick is refused and the previous shapes kept."""
import numpy as np
import pytest

from heddlemd_amd import api
from heddlemd_amd._lib import HmdError
from tests.gpu_helpers import rel_rms, water_case

pytestmark = pytest.mark.gpu

PAIR = {8, 16, 24, 32, 48, 64}
FFT = {1, 2, 3, 4, 6, 8}


def forces(field, part):
    field.compute(1)
    return part.download()["force"].copy(), part.energy_sums()[0]


def test_autotune_choice_reproduces(gpu_ctx):
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=20)
    try:
        f0, _ = forces(field, part)
        res = field.autotune(5)
        print(f"autotune: {res}", flush=True)
        assert res["n_candidates"] >= 3
        assert res["us_tuned"] <= res["us_initial"]
        assert res["pair_blocks_per_cu"] in PAIR and res["fft_blocks_per_cu"] in FFT
        now = field.tuning()
        assert all(now[k] == res[k] for k in ("pair_blocks_per_cu", "fft_blocks_per_cu", "brick"))
        # a fresh force field builds its list from the current positions: so
        # does this one, before the comparison (cluster membership, and so
        # the f32 summation order, follows the positions at the build)
        field.request_rebuild()
        f1, _ = forces(field, part)
        err = rel_rms(f1, f0)
        print(f"tuned vs initial shapes: rel RMS force {err:.2e}", flush=True)
        assert err <= 1e-6
        fresh = api.ForceField(gpu_ctx, part, ff)
        try:
            fresh.set_tuning(res)
            f2, _ = forces(fresh, part)
        finally:
            fresh.close()
        assert f2.tobytes() == f1.tobytes()
    finally:
        integ.close(); field.close(); part.close()


def test_shape_invariance(gpu_ctx):
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=20)
    try:
        base = field.tuning()
        f0, e0 = forces(field, part)
        for brick in [(16, 8, 8), (16, 16, 8), (16, 16, 16)]:
            field.set_tuning(dict(base, brick=brick))
            f, e = forces(field, part)
            print(f"brick {brick}: forces identical {f.tobytes() == f0.tobytes()}", flush=True)
            assert f.tobytes() == f0.tobytes()
        for fb in (1, 8):
            field.set_tuning(dict(base, fft_blocks_per_cu=fb))
            f, e = forces(field, part)
            print(f"fft blocks/CU {fb}: forces identical {f.tobytes() == f0.tobytes()}, "
                  f"rel PE {abs(e - e0) / abs(e0):.1e}", flush=True)
            assert f.tobytes() == f0.tobytes()
            assert abs(e - e0) <= 1e-12 * abs(e0)
        # each work-list entry's sums go to fixed point on their own, so the
        # split of the list among waves does not change a bit
        for pb in (8, 64):
            field.set_tuning(dict(base, pair_blocks_per_cu=pb))
            f, e = forces(field, part)
            print(f"pair blocks/CU {pb}: forces identical {f.tobytes() == f0.tobytes()}", flush=True)
            assert f.tobytes() == f0.tobytes()
            assert e == e0
        field.set_tuning(base)
        with pytest.raises(HmdError):
            field.set_tuning(dict(base, brick=(5, 8, 8)))
        assert field.tuning()["brick"] == base["brick"]
        f, _ = forces(field, part)
        assert f.tobytes() == f0.tobytes()
    finally:
        integ.close(); field.close(); part.close()


def test_replayed_steps_after_retuning(gpu_ctx):
    states = []
    for shape in [dict(brick=(16, 16, 8), fft_blocks_per_cu=1),
                  dict(brick=(8, 8, 8), fft_blocks_per_cu=4)]:
        sysw, ff, part, field, integ = water_case(gpu_ctx, relax=20)
        try:
            integ.replay(3)                       # captured with the initial shapes
            field.set_tuning(dict(field.tuning(), **shape))
            integ.replay(5)                       # re-captured with the new ones
            d = part.download()
            states.append((d["posq"].copy(), d["vel"].copy()))
        finally:
            integ.close(); field.close(); part.close()
    assert states[0][0].tobytes() == states[1][0].tobytes()
    assert states[0][1].tobytes() == states[1][1].tobytes()
