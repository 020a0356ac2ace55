"""The slot-level boundary (hmd_nbl_pre_step, hmd_pair_force,
hmd_spme_compute, hmd_forcefield_check) through the C ABI: the reference's
Potential::compute writes each slot's per-atom results with += into a
caller-owned SlotOutputView (src/forces/mod.rs:86-133, 206-212).  Device
arrays here come from torch (plumbing only).  Checks: the two slots composed
into zeroed external arrays equal hmd_forcefield_compute byte for byte
(forces, energies, virials); += keeps what the caller had there; and on a
context created on the caller's stream (hmd_ctx_create_on_stream) the slots
captured into the caller's graph replay to the eager bytes every time, with
eager work and synchronisations between the replays."""
import numpy as np
import pytest

from heddlemd_amd import api
from heddlemd_amd._lib import SlotOutput
from tests.gpu_helpers import water_case

pytestmark = pytest.mark.gpu


def _outputs(torch, n):
    t = torch.zeros((5, n), dtype=torch.float32, device="cuda")
    return t, SlotOutput(*(t[k].data_ptr() for k in range(5)))


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _want(part, field):
    field.compute(1)
    f = part.forces()
    e, w = part.energies()
    return np.stack([f[:, 0], f[:, 1], f[:, 2], e, w])


def test_slots_compose_to_the_force_field(gpu_ctx):
    import torch
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=40)
    try:
        want = _want(part, field)
        t, out = _outputs(torch, part.n)
        torch.cuda.synchronize()
        field.pre_step()
        field.pair_force(1, out)
        field.spme_compute(1, out)
        field.check()
        got = t.cpu().numpy()
        assert np.array_equal(_bits(got), _bits(want))
        # += : the pair slot twice into the same arrays doubles it exactly
        t.zero_()
        torch.cuda.synchronize()
        field.pre_step()
        field.pair_force(0, out)
        field.pre_step()
        field.pair_force(0, out)
        field.check()
        pair2 = t[:3].cpu().numpy()
        t.zero_()
        torch.cuda.synchronize()
        field.pre_step()
        field.pair_force(0, out)
        field.check()
        pair1 = t[:3].cpu().numpy()
        assert np.array_equal(_bits(pair2), _bits(2.0 * pair1))
    finally:
        for h in (integ, field, part):
            h.close()


def test_slots_on_the_callers_stream_replay_in_its_graph():
    import torch
    stream = torch.cuda.Stream()
    ctx = api.GpuContext(0, stream=stream.cuda_stream)
    sysw, ff, part, field, integ = water_case(ctx, relax=40)
    try:
        want = _want(part, field)
        t, out = _outputs(torch, part.n)
        torch.cuda.synchronize()
        # eager on the caller's stream first (the slots' one-time set-up)
        field.pre_step()
        field.pair_force(1, out)
        field.spme_compute(1, out)
        field.check()
        assert np.array_equal(_bits(t.cpu().numpy()), _bits(want))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            t.zero_()
            field.pre_step()
            field.pair_force(1, out)
            field.spme_compute(1, out)
        for rep in range(3):
            t.fill_(7.0)                     # eager work between the replays
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            field.check()
            assert np.array_equal(_bits(t.cpu().numpy()), _bits(want)), f"replay {rep}"
            field.compute(1)                 # the library's own eager evaluation
        del g
    finally:
        for h in (integ, field, part, ctx):
            h.close()
