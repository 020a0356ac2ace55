"""GPU parity of the velocity-Verlet kernels through the C ABI
(hmd_vv_kick_drift / hmd_vv_kick, f32 and compensated) against the oracle.

Bar: bit-exact.  Positions, velocities and image counters (f32 / i32) and
the f64 low parts of the compensated variant must equal the oracle's on
every case of tests/vv_cases.py: the exact-value cases of the reference's
integrator tests (tests/integration.rs, tests/integration_lossless.rs),
their time-reversal round trips, and seeded cubic / triclinic states that
cross the periodic boundary."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api
from tests import vv_cases as C

pytestmark = pytest.mark.gpu

CASES = C.cases()


def run_gpu(ctx, s, ops, comp):
    part = api.ParticleBuffers(ctx, s["posq"], s["mass"], s["lattice"], vel=s["vel"])
    try:
        part.upload_forces(s["force"])
        for op, dt in ops:
            if op == "kd":
                part.kick_drift(dt, comp)
            elif op == "k":
                part.kick(dt, comp)
            else:
                part.negate_velocities()
        d = part.download()
        d["pos_lo"], d["vel_lo"] = part.lo_parts()
        return d
    finally:
        part.close()


def run_oracle(s, ops, comp):
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    for op, dt in ops:
        if op == "kd":
            st.kick_drift(dt, comp)
        elif op == "k":
            st.kick(dt, comp)
        else:
            st.negate_velocities()
    return st


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_vv_bit_exact_vs_oracle(gpu_ctx, case):
    name, s, ops, comp = case
    print(f"vv case {name}: n={s['posq'].shape[0]} ops={len(ops)} compensated={comp}", flush=True)
    got = run_gpu(gpu_ctx, s, ops, comp)
    want = run_oracle(s, ops, comp)
    assert same(got["posq"], want.posq)
    assert same(got["vel"], want.vel)
    assert np.array_equal(got["images"], want.images)
    assert same(got["force"], s["force"])                   # forces are read-only
    assert same(got["pos_lo"], want.pos_lo)
    assert same(got["vel_lo"], want.vel_lo)


@pytest.mark.parametrize("name,tol", [("lossless_single_round_trip", 1e-12),
                                      ("lossless_multi_round_trip", 1e-10)])
def test_compensated_time_reversal(gpu_ctx, name, tol):
    """tests/integration_lossless.rs:301-372: forward, reverse velocities,
    the same steps, reverse again -> the f32 state is bit-identical to the
    start and the low parts are within f64 rounding of zero."""
    _, s, ops, comp = next(c for c in CASES if c[0] == name)
    got = run_gpu(gpu_ctx, s, ops, comp)
    assert np.array_equal(got["posq"], s["posq"])
    assert np.array_equal(got["vel"], s["vel"])
    assert np.max(np.abs(got["pos_lo"])) <= tol
    assert np.max(np.abs(got["vel_lo"])) <= tol


def test_compensated_runs_byte_identical(gpu_ctx):
    """tests/integration_lossless.rs:374-397."""
    s = C.diverse_state(64)
    ops = [("kd", 0.001), ("k", 0.001)] * 10
    a = run_gpu(gpu_ctx, s, ops, True)
    b = run_gpu(gpu_ctx, s, ops, True)
    for k in ("posq", "vel", "images", "pos_lo", "vel_lo"):
        assert a[k].tobytes() == b[k].tobytes()
