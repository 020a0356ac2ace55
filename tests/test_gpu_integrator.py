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


LAN_CASES = C.langevin_cases()


@pytest.mark.parametrize("case", LAN_CASES, ids=[c[0] for c in LAN_CASES])
def test_langevin_steps_vs_oracle(gpu_ctx, case):
    """BAOAB A / O steps through hmd_lan_drift_half / hmd_lan_ou_step.  The A
    step and the image counters are bit-exact.  The O step's normal is the
    f32 rounding of an f64 Box-Muller value whose log / cos come from the
    device math library on the GPU and from the host libm in the oracle, so
    a last-bit difference there can move a velocity by an ulp or two: the
    bar is >= 99.9% of components bit-exact and every one within 4 f32 ulps."""
    name, s, ops = case
    print(f"langevin case {name}: n={s['posq'].shape[0]} ops={len(ops)}", flush=True)
    part = api.ParticleBuffers(gpu_ctx, s["posq"], s["mass"], s["lattice"], vel=s["vel"])
    try:
        for op in ops:
            if op[0] == "A":
                part.lan_drift_half(op[1])
            else:
                part.lan_ou_step(*op[1:])
        got = part.download()
    finally:
        part.close()
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    for op in ops:
        if op[0] == "A":
            st.lan_drift_half(op[1])
        else:
            st.lan_ou_step(*op[1:])
    assert np.array_equal(got["images"], st.images)
    has_o = any(op[0] == "O" for op in ops)
    has_a_after_o = any(op[0] == "A" for op in ops[1:]) and has_o
    for key, want in (("vel", st.vel), ("posq", st.posq)):
        g = got[key]
        exact = np.mean(g == want)
        ulp = np.abs(np.spacing(np.maximum(np.abs(g), np.abs(want)).astype(np.float32)))
        worst = float(np.max(np.abs(g.astype(np.float64) - want) / ulp)) if g.size else 0.0
        print(f"  {key}: bit-exact {exact:.6f}, worst {worst:.1f} ulp")
        if not has_o or (key == "posq" and not has_a_after_o):
            assert exact == 1.0
        else:
            assert exact >= 0.999 and worst <= 4.0


@pytest.mark.parametrize("n", [1, 255, 8193, 196608])
def test_kinetic_energy_reduction_fp64(gpu_ctx, n):
    """FP64 fixed-shape reduction of 0.5 m |v|^2 (the reference's
    kinetic_energy_reduce, kernels/nose_hoover.cu:22-104, restated in FP64):
    within 1e-13 of the exactly rounded FP64 sum (math.fsum) and bit-identical
    on repeat (tests/reductions_multiblock.rs:59-88)."""
    import math
    rng = np.random.default_rng(n)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    m = rng.uniform(1.0, 30.0, size=n).astype(np.float32)
    posq = np.zeros((n, 4), np.float32)
    part = api.ParticleBuffers(gpu_ctx, posq, m, C.BIG, vel=v)
    try:
        ke = part.kinetic_energy()
        again = part.kinetic_energy()
    finally:
        part.close()
    v64 = v.astype(np.float64)
    terms = 0.5 * m.astype(np.float64) * (v64[:, 0] ** 2 + v64[:, 1] ** 2 + v64[:, 2] ** 2)
    want = math.fsum(terms.tolist())
    assert abs(ke - want) <= 1e-13 * abs(want)
    assert ke == again
