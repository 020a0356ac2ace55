"""Pins the velocity-Verlet oracle (orc_vv_kick_drift_f32 / orc_vv_kick_f32).

- the exact values asserted by the reference's own tests
  (tests/integration.rs, tests/integration_lossless.rs);
- an exact-rational model (Python Fractions, every operation rounded once to
  f32 or f64 with ties to even) of the kernel's arithmetic on seeded states,
  including a case where fusing the drift changes the result;
- closed-form constant-force kinematics of the compensated variant;
- time reversal of the compensated variant;
- the Langevin A / O steps: the reference's tests/langevin.rs cases, the
  Box-Muller normal rebuilt from the (Random123-pinned) Philox words, the
  exact-rational model of the O-step roundings and the OU variance."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from tests import vv_cases as C

F32 = np.float32


def run(case):
    name, s, ops, comp = case
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    for op, dt in ops:
        if op == "kd":
            st.kick_drift(dt, comp)
        elif op == "k":
            st.kick(dt, comp)
        else:
            st.negate_velocities()
    return st


def case(name):
    return next(c for c in C.cases() if c[0] == name)


# ---- exact values of the reference's tests ---------------------------------

def test_zero_force_advance():
    st = run(case("zero_force_advance"))
    assert st.vel.tolist() == [[0.5, -0.25, 0.125]]
    want = [F32(1.0) + F32(0.5) * F32(0.1), F32(2.0) + F32(-0.25) * F32(0.1),
            F32(3.0) + F32(0.125) * F32(0.1)]
    assert st.posq[0, :3].tolist() == [float(w) for w in want]


def test_zero_force_kick_leaves_state():
    st = run(case("zero_force_kick"))
    assert st.vel.tolist() == [[0.5, -0.25, 0.125]]
    assert st.posq[0, :3].tolist() == [1.0, 2.0, 3.0]


def test_constant_force_half_step():
    st = run(case("constant_force_half_step"))
    h = F32(0.1) * F32(0.5)
    v = [F32(2.0) * h, F32(-4.0) * h, F32(1.0) * h]
    assert st.vel[0].tolist() == [float(x) for x in v]
    assert st.posq[0, :3].tolist() == [float(x * F32(0.1)) for x in v]


def test_full_step_kinematics():
    st = run(case("full_step_kinematics"))
    h = F32(0.1) * F32(0.5)
    a = F32(2.0) / F32(1.0)
    vh = a * h
    assert st.posq[0, 0] == vh * F32(0.1)
    assert st.vel[0, 0] == vh + a * h
    assert st.posq[0, 1:3].tolist() == [0.0, 0.0] and st.vel[0, 1:].tolist() == [0.0, 0.0]


def test_mass_scaling():
    st = run(case("mass_scaling"))
    h = F32(0.2) * F32(0.5)
    assert st.vel[:, 0].tolist() == [float(F32(1.0) * h), float(F32(0.25) * h)]


def test_independent_particles():
    _, s, _, _ = case("independent_particles")
    st = run(case("independent_particles"))
    h = F32(0.1) * F32(0.5)
    for i in range(3):
        a = s["force"][i, 0] / s["mass"][i]
        vh = a * h
        assert st.posq[i, 0] == s["posq"][i, 0] + vh * F32(0.1)
        assert st.vel[i, 0] == vh + a * h


@pytest.mark.parametrize("name", ["dt_zero_kick_drift", "dt_zero_kick"])
def test_dt_zero_leaves_state(name):
    _, s, _, _ = case(name)
    st = run(case(name))
    assert np.array_equal(st.posq, s["posq"]) and np.array_equal(st.vel, s["vel"])


def test_nan_force_propagates():
    st = run(case("nan_force"))
    assert np.isnan(st.vel[0, 0]) and np.isnan(st.posq[0, 0])
    assert st.vel[0, 1:].tolist() == [0.0, 0.0] and st.posq[0, 1:3].tolist() == [2.0, 3.0]
    assert st.images[0].tolist() == [0, 0, 0]   # NaN image shift converts to 0


def test_forward_then_negated_dt_returns():
    _, s, _, _ = case("forward_then_negated_dt")
    st = run(case("forward_then_negated_dt"))
    assert np.max(np.abs(st.posq - s["posq"])) <= 1e-6
    assert np.max(np.abs(st.vel - s["vel"])) <= 1e-6


def test_wraps_and_image_counts():
    st = run(case("wrap_positive_x"))
    assert abs(st.posq[0, 0] - (-4.9)) < 1e-5 and st.images[0].tolist() == [1, 0, 0]
    st = run(case("wrap_negative_x"))
    assert abs(st.posq[0, 0] - 4.9) < 1e-5 and st.images[0].tolist() == [-1, 0, 0]
    # tests/integration.rs:638-667: unwrapped 0 + 25 lands on -5.0 with +3 images
    st = run(case("multi_period_crossing"))
    assert st.posq[0, 0] == -5.0 and st.images[0, 0] == 3


def test_lossless_block_non_aligned_exact():
    st = run(case("lossless_block_non_aligned"))
    assert np.all(st.posq[:, 0] == F32(0.0) + F32(0.1)) and np.all(st.vel[:, 0] == 1.0)


@pytest.mark.parametrize("name,tol", [("lossless_single_round_trip", 1e-12),
                                      ("lossless_multi_round_trip", 1e-10)])
def test_lossless_round_trip(name, tol):
    _, s, _, _ = case(name)
    st = run(case(name))
    assert np.array_equal(st.posq, s["posq"])
    assert np.array_equal(st.vel, s["vel"])
    assert np.max(np.abs(st.pos_lo)) <= tol and np.max(np.abs(st.vel_lo)) <= tol


# ---- exact-rational model ---------------------------------------------------

def r32(x):
    """Fraction -> nearest f32, ties to even (normal range)."""
    if x == 0:
        return F32(0.0)
    c = F32(float(x))
    cands = [np.nextafter(c, F32(-np.inf)), c, np.nextafter(c, F32(np.inf))]
    return min(cands, key=lambda y: (abs(Fraction(float(y)) - x),
                                     int(y.view(np.uint32)) & 1))


def r64(x):
    return float(x)   # int / int true division rounds once, ties to even


def Q(x):
    """Exact rational value of a float / f32 (or a Fraction, unchanged)."""
    return x if isinstance(x, Fraction) else Fraction(float(x))


def model_kick_drift(x, v, xl, vl, f, m, L, dt, comp):
    """One particle; x, v f32[3], xl, vl float[3]; returns the new state."""
    dt = F32(dt)
    h = r32(Q(dt) * Q(F32(0.5)))
    x = list(x); v = list(v); xl = list(xl); vl = list(vl)
    for d in range(3):
        a = r32(Q(f[d]) / Q(m))
        if comp:
            p = r32(Q(a) * Q(h))
            ev = r64(Q(r64(Q(v[d]) + Q(vl[d]))) + Q(p))
            vh = r32(Q(ev))
            vlo = r64(Q(ev) - Q(vh))
            s = r64(Q(vh) + Q(vlo))
            ex = r64(Q(s) * Q(dt) + Q(r64(Q(x[d]) + Q(xl[d]))))
            xh = r32(Q(ex))
            v[d], vl[d], x[d], xl[d] = vh, vlo, xh, r64(Q(ex) - Q(xh))
        else:
            v[d] = r32(Q(a) * Q(h) + Q(v[d]))
            x[d] = r32(Q(v[d]) * Q(dt) + Q(x[d]))
    lx, ly, lz, xy, xz, yz = [Q(F32(t)) for t in L]
    sc = r32(Q(x[2]) / lz)
    sb = r32(Q(r32(Q(x[1]) - Q(r32(Q(sc) * yz)))) / ly)
    t = r32(Q(x[0]) - Q(r32(Q(sb) * xy)))
    sa = r32(Q(r32(Q(t) - Q(r32(Q(sc) * xz)))) / lx)
    k = [np.floor(r32(Q(s) + Q(F32(0.5)))) for s in (sa, sb, sc)]
    sx = r32(Q(r32(Q(F32(k[0])) * lx)) + Q(r32(Q(F32(k[1])) * xy)))
    sx = r32(Q(sx) + Q(r32(Q(F32(k[2])) * xz)))
    x[0] = r32(Q(x[0]) - Q(sx))
    x[1] = r32(Q(x[1]) - Q(r32(Q(r32(Q(F32(k[1])) * ly)) + Q(r32(Q(F32(k[2])) * yz)))))
    x[2] = r32(Q(x[2]) - Q(r32(Q(F32(k[2])) * lz)))
    return x, v, xl, vl, [int(t) for t in k]


def model_kick(v, vl, f, m, dt, comp):
    h = r32(Q(F32(dt)) * Q(F32(0.5)))
    v = list(v); vl = list(vl)
    for d in range(3):
        a = r32(Q(f[d]) / Q(m))
        if comp:
            p = r32(Q(a) * Q(h))
            ev = r64(Q(r64(Q(v[d]) + Q(vl[d]))) + Q(p))
            v[d] = r32(Q(ev))
            vl[d] = r64(Q(ev) - Q(v[d]))
        else:
            v[d] = r32(Q(a) * Q(h) + Q(v[d]))
    return v, vl


@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("make", [lambda: C.random_state(24, 5), lambda: C.triclinic_state(24, 6)])
def test_oracle_matches_exact_rational_model(make, comp):
    s = make()
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    n = s["posq"].shape[0]
    X = [list(s["posq"][i, :3]) for i in range(n)]
    V = [list(s["vel"][i]) for i in range(n)]
    XL = [[0.0] * 3 for _ in range(n)]
    VL = [[0.0] * 3 for _ in range(n)]
    K = [[0, 0, 0] for _ in range(n)]
    for step in range(3):
        st.kick_drift(0.05, comp)
        st.kick(0.05, comp)
        for i in range(n):
            X[i], V[i], XL[i], VL[i], k = model_kick_drift(
                X[i], V[i], XL[i], VL[i], s["force"][i], s["mass"][i], s["lattice"], 0.05, comp)
            K[i] = [K[i][d] + k[d] for d in range(3)]
            V[i], VL[i] = model_kick(V[i], VL[i], s["force"][i], s["mass"][i], 0.05, comp)
        assert np.array_equal(st.posq[:, :3], np.array(X, F32)), f"step {step}"
        assert np.array_equal(st.vel, np.array(V, F32))
        assert np.array_equal(st.pos_lo, np.array(XL)) and np.array_equal(st.vel_lo, np.array(VL))
        assert np.array_equal(st.images, np.array(K))
    assert np.any(np.array(K) != 0), "no boundary crossing exercised"


def test_fused_drift_rounding_is_pinned():
    """Find f32 (x, v, dt) where x + v dt rounded once (fused) differs from
    rounding v dt first; the oracle must give the fused value."""
    rng = np.random.default_rng(3)
    found = 0
    for _ in range(2000):
        x, v = F32(rng.uniform(-3, 3)), F32(rng.uniform(-3, 3))
        fused = r32(Q(v) * Q(F32(0.1)) + Q(x))
        split = r32(Q(x) + Q(r32(Q(v) * Q(F32(0.1)))))
        if fused == split:
            continue
        st = oracle.VvState([[x, 0, 0, 0]], [[v, 0, 0]], [[0, 0, 0]], [1.0], C.BIG)
        st.kick_drift(0.1)
        assert st.posq[0, 0] == fused and st.posq[0, 0] != split
        found += 1
        if found == 5:
            break
    assert found == 5


def test_compensated_constant_force_closed_form():
    """x(t) = x0 + v0 t + a_k t^2 / 2 with a_k = 2 fl32(a dt/2) / dt, the
    kick the kernel applies; the compensated (hi + lo) position tracks it to
    f64 rounding over 4000 steps, the f32 variant does not."""
    n = 64
    rng = np.random.default_rng(9)
    x0 = rng.uniform(-1, 1, (n, 3)).astype(F32)
    v0 = rng.uniform(-1, 1, (n, 3)).astype(F32)
    f = rng.uniform(-1, 1, (n, 3)).astype(F32)
    m = rng.uniform(1, 4, n).astype(F32)
    s = C.state(x0, v0, m, f)
    dt, steps = 1e-3, 4000
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    plain = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    for _ in range(steps):
        st.kick_drift(dt, True); st.kick(dt, True)
        plain.kick_drift(dt); plain.kick(dt)
    dt32 = F32(dt)
    p = (f / m[:, None]).astype(F32) * (dt32 * F32(0.5))
    t = float(dt32) * steps
    ak = 2.0 * p.astype(np.float64) / float(dt32)
    x_exact = x0.astype(np.float64) + v0.astype(np.float64) * t + 0.5 * ak * t * t
    v_exact = v0.astype(np.float64) + ak * t
    err_c = np.max(np.abs(st.posq[:, :3].astype(np.float64) + st.pos_lo - x_exact))
    err_v = np.max(np.abs(st.vel.astype(np.float64) + st.vel_lo - v_exact))
    err_p = np.max(np.abs(plain.posq[:, :3].astype(np.float64) - x_exact))
    assert err_c < 1e-11 and err_v < 1e-11
    assert err_p > 1e3 * err_c


# ---- Langevin BAOAB A / O steps ---------------------------------------------

def run_lan(name):
    _, s, ops = next(c for c in C.langevin_cases() if c[0] == name)
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    for op in ops:
        if op[0] == "A":
            st.lan_drift_half(op[1])
        else:
            st.lan_ou_step(*op[1:])
    return s, st


def test_lan_drift_half_reference_cases():
    s, st = run_lan("lan_drift_half_kat")   # tests/langevin.rs:120-134
    assert abs(st.posq[0, 0] - (1.0 + 0.5 * 0.05)) < 1e-6
    assert abs(st.posq[0, 1] - (2.0 - 0.25 * 0.05)) < 1e-6
    assert abs(st.posq[0, 2] - (3.0 + 0.125 * 0.05)) < 1e-6
    assert np.array_equal(st.vel, s["vel"])
    assert st.posq[0, 0] == r32(Q(F32(0.5)) * Q(F32(0.1) * F32(0.5)) + Q(F32(1.0)))
    s, st = run_lan("lan_drift_half_dt_zero")
    assert np.array_equal(st.posq, s["posq"])
    s, st = run_lan("lan_drift_half_wrap")  # :641-664
    assert abs(st.posq[0, 0] - (-4.95)) < 1e-5 and st.images[0].tolist() == [1, 0, 0]


def test_lan_ou_alpha_one_kt_zero_is_identity():
    s, st = run_lan("lan_ou_identity")      # :185-198
    assert np.array_equal(st.vel, s["vel"])


def test_lan_ou_normal_from_philox_words():
    """xi = (float)(sqrt(-2 ln u1) cos(2 pi u2)), u = (word + 1/2) 2^-32 from
    words 0 and 1 of Philox(key = seed, ctr = (draw lo, draw hi, id, axis));
    with alpha = 0, kT = m = 1 the O step writes xi itself."""
    n, seed, draw = 40, 0x1234_5678_9ABC_DEF0, 0x0000_0003_0000_0011
    s = C.state(np.zeros((n, 3)), np.ones((n, 3)), np.ones(n))
    st = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
    st.lan_ou_step(seed, 0.0, 1.0, draw)
    for i in range(n):
        for d in range(3):
            w = oracle.philox4x32_10((seed & 0xFFFFFFFF, seed >> 32),
                                     (draw & 0xFFFFFFFF, draw >> 32, i, d))
            u1, u2 = (w[0] + 0.5) / 2**32, (w[1] + 0.5) / 2**32
            xi = F32(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))
            assert abs(st.vel[i, d] - xi) <= abs(xi) * 2.0**-23, (i, d)


def test_lan_ou_roundings_match_exact_rational_model():
    s, st = run_lan("lan_ou_n64_kt1")
    st0 = oracle.VvState(s["posq"], np.zeros((64, 3)), s["force"], s["mass"], s["lattice"])
    st0.lan_ou_step(1, 0.0, 1.0, 1)          # xi itself (m = 1)
    xi = st0.vel
    a, kt = F32(0.5), F32(1.0)
    for i in range(64):
        one_minus = r32(Q(1) - Q(a) * Q(a))
        sig = F32(np.sqrt(np.float64(r32(Q(r32(Q(one_minus) * Q(kt))) / Q(s["mass"][i])))))
        for d in range(3):
            want = r32(Q(a) * Q(s["vel"][i, d]) + Q(r32(Q(sig) * Q(xi[i, d]))))
            assert st.vel[i, d] == want, (i, d)


def test_lan_ou_seed_and_counter_change_the_draws():
    _, a = run_lan("lan_ou_n64_kt1")
    for seed, draw in [(2, 1), (1, 2)]:           # tests/langevin.rs:238-286
        _, s, _ = next(c for c in C.langevin_cases() if c[0] == "lan_ou_n64_kt1")
        b = oracle.VvState(s["posq"], s["vel"], s["force"], s["mass"], s["lattice"])
        b.lan_ou_step(seed, 0.5, 1.0, draw)
        assert np.mean(a.vel[:, 0] != b.vel[:, 0]) >= 0.9


def test_lan_ou_variance():
    _, st = run_lan("lan_ou_variance")      # tests/langevin.rs:289-327
    kt = float(F32(1.380649e-23 * 300.0))
    want = (1.0 - 0.25) * kt / float(F32(6.6335e-26))
    for d in range(3):
        v = st.vel[:, d].astype(np.float64)
        assert abs(v.var() - want) / want < 0.05
