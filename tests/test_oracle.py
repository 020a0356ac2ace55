"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Each check is independent of the oracle's own formulas: published
known-answer vectors, closed forms, finite differences, invariants, numpy's
FFT, and the reference's own fixtures (tests/spme_pipeline.rs,
tests/forces_framework.rs).
"""
import math

import numpy as np
import pytest

import oracle
from heddlemd_amd import units as U

BOHR_PER_M = 1.0 / U.LENGTH_BOHR_TO_M


# --------------------------------------------------------------- Philox ---
@pytest.mark.parametrize("key,ctr,want", [
    ((0, 0), (0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff), (0xffffffff,) * 4,
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0xa4093822, 0x299f31d0), (0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_random123_known_answers(key, ctr, want):
    """Random123 kat_vectors for philox4x32 with 10 rounds."""
    assert tuple(int(x) for x in oracle.philox4x32_10(key, ctr)) == want


def test_philox_gaussian_moments():
    xs = np.array([oracle.philox_gaussian(7, 0, s, 0, i, a)
                   for s in range(4) for i in range(1500) for a in range(3)])
    assert abs(xs.mean()) < 0.02
    assert abs(xs.var() - 1.0) < 0.03
    # Box-Muller maps the two leading words through (x + 1/2) 2^-32
    o = oracle.philox4x32_10((7, 0), (3, 0, 11, 2))
    u1, u2 = (o[0] + 0.5) / 2 ** 32, (o[1] + 0.5) / 2 ** 32
    assert oracle.philox_gaussian(7, 0, 3, 0, 11, 2) == pytest.approx(
        math.sqrt(-2 * math.log(u1)) * math.cos(2 * math.pi * u2), rel=1e-14)


# ------------------------------------------------------------------ FFT ---
@pytest.mark.parametrize("shape", [(6, 8, 10), (16, 16, 16), (24, 20, 30), (9, 5, 7)])
def test_fft3d_matches_numpy(shape):
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(shape), rng.standard_normal(shape)
    re, im = oracle.fft3d(a, b, -1)
    ref = np.fft.fftn(a + 1j * b)
    assert np.abs(re + 1j * im - ref).max() < 1e-10 * np.abs(ref).max()
    re2, im2 = oracle.fft3d(re, im, +1)
    assert np.abs((re2 + 1j * im2) / a.size - (a + 1j * b)).max() < 1e-12


# ------------------------------------------------------------ B-splines ---
@pytest.mark.parametrize("p", [4, 5, 6, 8])
def test_bspline_partition_of_unity_and_symmetry(p):
    for t in np.linspace(0, 0.999, 17):
        assert sum(oracle.bspline(p, d + t) for d in range(p)) == pytest.approx(1.0, abs=1e-14)
        assert oracle.bspline(p, t + 1.3) == pytest.approx(oracle.bspline(p, p - t - 1.3), abs=1e-14)


def test_cubic_bspline_closed_form():
    assert oracle.bspline(4, 1.0) == pytest.approx(1 / 6)
    assert oracle.bspline(4, 2.0) == pytest.approx(2 / 3)
    x = 0.37   # M_4(x) = x^3/6 on [0, 1)
    assert oracle.bspline(4, x) == pytest.approx(x ** 3 / 6, abs=1e-15)


def test_b_factors_cubic_closed_form():
    """|sum_j M_4(j+1) e^{i theta j}|^2 = (2/3 + cos(theta)/3)^2 for p = 4."""
    n = 16
    got = oracle.b_factors(n, 4)
    th = 2 * np.pi * np.arange(n) / n
    assert np.allclose(got, 1.0 / (2 / 3 + np.cos(th) / 3) ** 2, rtol=1e-13)
    assert got[0] == pytest.approx(1.0)


def test_b_factors_odd_order_finite():
    """Odd orders have a near-zero modulus at k = n/2; the neighbour average
    keeps every factor finite (src/forces/spme.rs:545-554)."""
    f = oracle.b_factors(16, 5)
    assert np.all(np.isfinite(f)) and np.all(f > 0)


# ------------------------------------------------------------------ SPME ---
def _box(l_nm=1.0):
    L = l_nm * 1e-9 * BOHR_PER_M
    return L, np.array([L, L, L, 0, 0, 0])


@pytest.mark.parametrize("positions,charges", [
    ([[0.2e-9, 0, 0], [-0.2e-9, 0, 0]], [1.0, -1.0]),
    ([[0.1e-9, 0, 0], [-0.1e-9, 0.2e-9, 0], [0, -0.2e-9, 0.15e-9],
      [-0.15e-9, 0.1e-9, -0.1e-9]], [1.0, -1.0, 2.0, -2.0]),
])
@pytest.mark.parametrize("order", [4, 5, 6, 8])
def test_spme_matches_explicit_ewald(positions, charges, order):
    """tests/spme_pipeline.rs:101-189 and :875-913: L = 1 nm, alpha L = 4,
    16^3 mesh, |m| <= 8, relative energy error < 5e-3."""
    L, lat = _box()
    alpha = 4.0e9 / BOHR_PER_M
    pos = np.array(positions) * BOHR_PER_M
    q = np.array(charges)
    u, F, E, w = oracle.spme_recip(pos, q, lat, (16, 16, 16), order, alpha)
    ud, Fd, Ed = oracle.ewald_recip_direct(pos, q, lat, alpha, 8)
    assert abs(u - ud) / abs(ud) < 5e-3
    assert np.abs(F - Fd).max() < 1e-2 * np.abs(Fd).max()


def test_single_particle_spread_matches_bspline_closed_form():
    """tests/spme_pipeline.rs:320-382."""
    L, lat = _box()
    n, p, q = 16, 4, 0.5
    s = np.array([0.11, -0.18, 0.37])
    rho = oracle.spme_spread((s * L)[None, :], [q], lat, (n, n, n), p)
    u = (s + 0.5) * n
    b = np.floor(u).astype(int)
    t = u - b

    def w(axis, g):
        da = (b[axis] - g) % n
        return oracle.bspline(p, da + t[axis]) if da < p else 0.0
    ref = np.array([[[q * w(0, a) * w(1, bb) * w(2, c) for c in range(n)] for bb in range(n)]
                    for a in range(n)])
    assert np.abs(rho - ref).max() < 1e-14
    assert rho.sum() == pytest.approx(q)


def test_spread_zero_outside_support_and_conserves_charge():
    """tests/spme_pipeline.rs:192-231 and :385-409."""
    L, lat = _box()
    rho = oracle.spme_spread(np.array([[0.2 * L, 0.2 * L, 0.2 * L]]), [1.0], lat, (16,) * 3, 4)
    assert rho[0, 0, 0] == 0.0
    pos = np.array([[0.1, 0.05, -0.2], [-0.1, 0.2, 0.15], [0.25, -0.1, 0.0]]) * L
    q = [0.3, -0.5, 0.4]
    assert oracle.spme_spread(pos, q, lat, (16,) * 3, 4).sum() == pytest.approx(sum(q), rel=1e-12)


def test_direct_ewald_forces_are_energy_gradient():
    rng = np.random.default_rng(5)
    L, lat = _box()
    alpha = 4.0e9 / BOHR_PER_M
    pos = (rng.random((5, 3)) - 0.5) * L
    q = np.array([1.0, -1.0, 0.5, -0.25, -0.25])
    _, F, _ = oracle.ewald_recip_direct(pos, q, lat, alpha, 6)
    h = 1e-5
    for i, d in [(0, 0), (2, 1), (4, 2)]:
        pp, pm = pos.copy(), pos.copy()
        pp[i, d] += h
        pm[i, d] -= h
        ep = oracle.ewald_recip_direct(pp, q, lat, alpha, 6)[0]
        em = oracle.ewald_recip_direct(pm, q, lat, alpha, 6)[0]
        assert F[i, d] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-10)


def test_spme_forces_are_energy_gradient():
    rng = np.random.default_rng(6)
    L, lat = _box(1.2)
    alpha = 3.5 / (0.4e-9 * BOHR_PER_M)
    pos = (rng.random((6, 3)) - 0.5) * L
    q = np.array([1.0, -1.0, 0.5, -0.5, 0.3, -0.3])
    _, F, _, _ = oracle.spme_recip(pos, q, lat, (20, 20, 20), 6, alpha)
    h = 1e-5
    for i, d in [(1, 0), (3, 2)]:
        pp, pm = pos.copy(), pos.copy()
        pp[i, d] += h
        pm[i, d] -= h
        ep = oracle.spme_recip(pp, q, lat, (20,) * 3, 6, alpha)[0]
        em = oracle.spme_recip(pm, q, lat, (20,) * 3, 6, alpha)[0]
        assert F[i, d] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-9)


def test_spme_energy_is_half_sum_of_per_atom_shares():
    rng = np.random.default_rng(8)
    L, lat = _box()
    alpha = 4.0e9 / BOHR_PER_M
    pos = (rng.random((8, 3)) - 0.5) * L
    q = rng.standard_normal(8)
    q -= q.mean()
    u, F, E, w = oracle.spme_recip(pos, q, lat, (16,) * 3, 4, alpha)
    u_self = alpha / math.sqrt(math.pi) * np.sum(q ** 2)
    assert E.sum() == pytest.approx(u - u_self, rel=1e-10)


# ---------------------------------------------------------- pair forces ---
def test_lj_cross_term_known_answer():
    """tests/forces_framework.rs:590-627: r = 1.5, sigma = 1.3, eps = 2, no
    switch; F = 24 eps / r (2 (s/r)^12 - (s/r)^6), equal and opposite."""
    r, sig, eps = 1.5, 1.3, 2.0
    ps = oracle.PairSystem([[0, 0, 0], [r, 0, 0]], [0, 0], [0, 1], [40, 40, 40, 0, 0, 0],
                           lj={"sigma": [sig] * 4, "epsilon": [eps] * 4,
                               "cutoff": [10.0] * 4, "switch": [10.0] * 4})
    F, E, W = ps.forces()
    want = 24 * eps / r * (2 * (sig / r) ** 12 - (sig / r) ** 6)
    assert F[1, 0] == pytest.approx(want, rel=1e-12)
    assert F[0, 0] == pytest.approx(-want, rel=1e-12)
    assert E.sum() == pytest.approx(4 * eps * ((sig / r) ** 12 - (sig / r) ** 6), rel=1e-12)


def _random_fluid(n=64, L=30.0, seed=1):
    """Jittered simple-cubic lattice (no overlapping pairs, so the energy
    finite differences are well conditioned)."""
    rng = np.random.default_rng(seed)
    k = int(round(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = (g + 0.5) * (L / k) - 0.5 * L + rng.uniform(-0.8, 0.8, (n, 3))
    q = rng.choice([-0.8, 0.4], n)
    types = rng.integers(0, 2, n)
    lj = {"sigma": [3.0, 2.5, 2.5, 2.0], "epsilon": [1e-3, 5e-4, 5e-4, 0.0],
          "cutoff": [9.0] * 4, "switch": [8.0] * 4}
    return pos, q, types, [L, L, L, 0, 0, 0], lj


def test_pair_forces_are_energy_gradient_with_switch_and_erfc():
    pos, q, types, lat, lj = _random_fluid()
    co = {"alpha": 0.35, "r_cut": 9.0}
    h = 1e-6
    base = oracle.PairSystem(pos, q, types, lat, lj=lj, coulomb=co)
    F, _, _ = base.forces()
    for i, d in [(0, 0), (7, 1), (13, 2)]:
        pp, pm = pos.copy(), pos.copy()
        pp[i, d] += h
        pm[i, d] -= h
        ep = oracle.PairSystem(pp, q, types, lat, lj=lj, coulomb=co).forces()[1].sum()
        em = oracle.PairSystem(pm, q, types, lat, lj=lj, coulomb=co).forces()[1].sum()
        assert F[i, d] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-9)


def test_pair_set_cells_equals_brute_force():
    """The cell-list pair set (used for the large GPU parity cases) is the
    brute-force f32 set, pair for pair, in the same order."""
    rng = np.random.default_rng(3)
    for n, L, rs in [(3000, 40.0, 12.0), (1500, 30.0, 9.5)]:
        posq = np.zeros((n, 4), np.float32)
        posq[:, :3] = (rng.random((n, 3)) - 0.5) * L
        lat = np.array([L, L, L, 0, 0, 0], np.float32)
        rs2 = np.float32(rs * rs)
        a = oracle.pair_set_f32(posq, lat, rs2)
        b = oracle.pair_set_f32(posq, lat, rs2, cells=True)
        assert len(a) > 1000
        assert np.array_equal(a, b)


def test_pair_forces_newton3_and_cells_match_brute_force():
    pos, q, types, lat, lj = _random_fluid(n=343, L=32.0)
    co = {"alpha": 0.35, "r_cut": 9.0}
    ps = oracle.PairSystem(pos, q, types, lat, lj=lj, coulomb=co)
    F, E, W = ps.forces()
    Fc, Ec, Wc = ps.forces(cells=True)
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).max() * len(F)
    assert np.allclose(F, Fc, rtol=1e-12, atol=1e-15)
    assert np.allclose(E, Ec, rtol=1e-12, atol=1e-15)


def _scaled_energy(pos, q, types, lat, lj, co, eps, mesh, alpha):
    """Pair + reciprocal energy with the box and every position scaled by
    (1 + eps)."""
    f = 1.0 + eps
    lat_s = [x * f for x in lat]
    e_pair = oracle.PairSystem(pos * f, q, types, lat_s, lj=lj, coulomb=co).forces()[1].sum()
    u_rec = oracle.spme_recip(pos * f, q, lat_s, mesh, 4, alpha)[0]
    return e_pair, u_rec


def test_virials_are_minus_energy_derivative_under_uniform_scaling():
    """The scalar virial W = sum over pairs of F . r (pair slots, with the
    LJ switch and erfc) and W_recip = 1/2 sum vf |rho_hat|^2 both equal
    -dU/d eps when the box and all positions scale by (1 + eps): central
    differences of the oracle's energies, no pair within 1e-4 of a cutoff."""
    pos, q, types, lat, lj = _random_fluid(n=64, L=30.0, seed=5)
    q = q - q.mean()
    co = {"alpha": 0.35, "r_cut": 9.0}
    # no pair close to the cutoffs, where the truncated energy jumps
    d = pos[:, None, :] - pos[None, :, :]
    d -= 30.0 * np.round(d / 30.0)
    r = np.sqrt((d ** 2).sum(-1))[np.triu_indices(64, 1)]
    assert np.min(np.abs(r / 9.0 - 1.0)) > 1e-4
    mesh = (24, 24, 24)
    _, _, W = oracle.PairSystem(pos, q, types, lat, lj=lj, coulomb=co).forces()
    _, _, _, w_rec = oracle.spme_recip(pos, q, lat, mesh, 4, co["alpha"])
    h = 1e-6
    ep, up = _scaled_energy(pos, q, types, lat, lj, co, h, mesh, co["alpha"])
    em, um = _scaled_energy(pos, q, types, lat, lj, co, -h, mesh, co["alpha"])
    assert W.sum() == pytest.approx(-(ep - em) / (2 * h), rel=1e-6)
    assert w_rec == pytest.approx(-(up - um) / (2 * h), rel=1e-6)


def test_erfc_approximation_error_bound():
    """A&S 7.1.26: |erfc_AS - erfc| <= 1.5e-7; the oracle's two modes differ
    by at most that, scaled by the Coulomb factor."""
    pos = np.array([[0, 0, 0], [0, 0, 0]], float)
    for r in [0.5, 1.0, 2.0, 4.0, 8.0]:
        pos[1, 0] = r
        co = {"alpha": 0.4, "r_cut": 9.0}
        e_as = oracle.PairSystem(pos, [1, -1], [0, 0], [40] * 3 + [0] * 3, coulomb=co,
                                 erfc_mode=1).forces()[1].sum()
        e_ex = oracle.PairSystem(pos, [1, -1], [0, 0], [40] * 3 + [0] * 3, coulomb=co,
                                 erfc_mode=0).forces()[1].sum()
        assert e_ex == pytest.approx(-math.erfc(0.4 * r) / r, rel=1e-13)
        assert abs(e_as - e_ex) <= 1.5e-7 / r


def test_exclusions_zero_the_pair():
    """A listed partner contributes nothing (scale 0); every other pair is
    untouched: the three-atom energy equals the two non-excluded two-atom
    energies."""
    pos = np.array([[0, 0, 0], [1.8, 0, 0], [0, 2.5, 0]], float)
    q = [-0.8, 0.4, 0.4]
    box = [40] * 3 + [0] * 3
    co = {"alpha": 0.3, "r_cut": 9.0}
    ex = ([0, 1, 2, 2], [1, 0], [0.0, 0.0], [0.0, 0.0])
    F, E, W = oracle.PairSystem(pos, q, [0, 1, 1], box, coulomb=co, exclusions=ex).forces()

    def two(a, b):
        return oracle.PairSystem(pos[[a, b]], [q[a], q[b]], [0, 0], box, coulomb=co).forces()
    F02, E02, _ = two(0, 2)
    F12, E12, _ = two(1, 2)
    assert E.sum() == pytest.approx(E02.sum() + E12.sum(), rel=1e-14)
    assert np.allclose(F[0], F02[0], rtol=1e-14)
    assert np.allclose(F[1], F12[0], rtol=1e-14)
    assert np.allclose(F[2], F02[1] + F12[1], rtol=1e-14)


# ---------------------------------------------------------- pair set -------
def test_pair_set_agrees_with_f64_away_from_the_boundary():
    rng = np.random.default_rng(2)
    L = 30.0
    posq = np.zeros((400, 4), np.float32)
    posq[:, :3] = (rng.random((400, 3)) - 0.5) * L
    lat = np.array([L, L, L, 0, 0, 0], np.float32)
    rs = np.float32(9.5)
    pairs = oracle.pair_set_f32(posq, lat, rs * rs)
    got = set(map(tuple, pairs.tolist()))
    p = posq[:, :3].astype(np.float64)
    d = p[:, None, :] - p[None, :, :]
    d -= L * np.floor(d / L + 0.5)
    r2 = np.sum(d * d, axis=2)
    i, j = np.triu_indices(400, 1)
    margin = 1e-4 * float(rs) ** 2
    inside = r2[i, j] < float(rs) ** 2 - margin
    outside = r2[i, j] > float(rs) ** 2 + margin
    assert all((a, b) in got for a, b in zip(i[inside], j[inside]))
    assert not any((a, b) in got for a, b in zip(i[outside], j[outside]))
    assert all(a < b for a, b in got)


# ---------------------------------------------------------------- SETTLE --
def _water(rng):
    d_oh, ang = 1.8897261, math.radians(109.47)
    loc = np.array([[0, 0, 0], [d_oh * math.sin(ang / 2), d_oh * math.cos(ang / 2), 0],
                    [-d_oh * math.sin(ang / 2), d_oh * math.cos(ang / 2), 0]])
    A = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    return loc @ A.T + rng.standard_normal(3) * 5, d_oh, 2 * d_oh * math.sin(ang / 2)


def test_settle_position_reset_restores_rigidity_and_com():
    rng = np.random.default_rng(4)
    m_o, m_h = 15.9994 * U.AMU, 1.008 * U.AMU
    lat = [60, 60, 60, 0, 0, 0]
    for _ in range(5):
        snap, d_oh, d_hh = _water(rng)
        geom = oracle.settle_geometry(d_oh, d_hh, m_o, m_h)
        cur = snap + rng.standard_normal((3, 3)) * 0.05
        v = np.zeros((3, 3))
        new, v2, vir = oracle.settle_positions_one(snap, cur, v, m_o, m_h, geom, 80.0, lat)
        dist = [np.linalg.norm(new[1] - new[0]), np.linalg.norm(new[2] - new[0]),
                np.linalg.norm(new[2] - new[1])]
        assert dist == pytest.approx([d_oh, d_oh, d_hh], rel=1e-10)
        m = np.array([m_o, m_h, m_h])[:, None]
        assert np.allclose((m * new).sum(0), (m * cur).sum(0), rtol=1e-12)
        # corrections change momentum by zero in total
        assert np.abs((m * v2).sum(0)).max() < 1e-12


def test_settle_velocity_projection_zeroes_bond_velocities():
    rng = np.random.default_rng(9)
    m_o, m_h = 15.9994 * U.AMU, 1.008 * U.AMU
    pos, _, _ = _water(rng)
    v = rng.standard_normal((3, 3)) * 1e-3
    v2 = oracle.settle_velocities_one(pos, v, m_o, m_h, [60, 60, 60, 0, 0, 0])
    for a, b in [(0, 1), (0, 2), (1, 2)]:
        assert abs(np.dot(v2[a] - v2[b], pos[a] - pos[b])) < 1e-14
    m = np.array([m_o, m_h, m_h])[:, None]
    assert np.allclose((m * v2).sum(0), (m * v).sum(0), rtol=1e-12, atol=1e-18)


def _shake(snap, cur, m, d, tol=1e-13, max_iter=100000):
    """Textbook SHAKE (Ryckaert, Ciccotti & Berendsen 1977): displace along
    the reference bond vectors until every |r_a - r_b| = d_ab."""
    r = cur.copy()
    bonds = [(0, 1, d[0]), (0, 2, d[1]), (1, 2, d[2])]
    for _ in range(max_iter):
        worst = 0.0
        for a, b, dd in bonds:
            s = r[a] - r[b]
            diff = dd * dd - s @ s
            worst = max(worst, abs(diff) / (dd * dd))
            r0 = snap[a] - snap[b]
            g = diff / (2.0 * (s @ r0) * (1.0 / m[a] + 1.0 / m[b]))
            r[a] += g * r0 / m[a]
            r[b] -= g * r0 / m[b]
        if worst < tol:
            return r
    raise AssertionError("SHAKE did not converge")


def test_settle_positions_equal_converged_shake():
    """SETTLE solves the SHAKE equations in closed form: the reset positions
    equal the converged iterative solution (tests/settle.rs:930), and the
    velocity correction is (reset - unconstrained) / dt."""
    rng = np.random.default_rng(12)
    m_o, m_h = 15.9994 * U.AMU, 1.008 * U.AMU
    m = np.array([m_o, m_h, m_h])
    lat = [60, 60, 60, 0, 0, 0]
    dt = 80.0
    for _ in range(20):
        snap, d_oh, d_hh = _water(rng)
        geom = oracle.settle_geometry(d_oh, d_hh, m_o, m_h)
        cur = snap + rng.standard_normal((3, 3)) * 0.03
        v = rng.standard_normal((3, 3)) * 1e-3
        new, v2, _ = oracle.settle_positions_one(snap, cur, v, m_o, m_h, geom, dt, lat)
        ref = _shake(snap, cur, m, (d_oh, d_oh, d_hh))
        assert np.max(np.abs(new - ref)) < 1e-9 * d_oh
        assert np.allclose(v2, v + (new - cur) / dt, rtol=0, atol=1e-12)


def test_settle_velocities_equal_linear_constraint_solve():
    """The velocity projection is the RATTLE step: impulses along the three
    bonds from the 3x3 linear system that zeroes every relative velocity
    along its bond (solved here with numpy)."""
    rng = np.random.default_rng(13)
    m_o, m_h = 15.9994 * U.AMU, 1.008 * U.AMU
    m = np.array([m_o, m_h, m_h])
    bonds = [(0, 1), (0, 2), (1, 2)]
    for _ in range(20):
        pos, _, _ = _water(rng)
        v = rng.standard_normal((3, 3)) * 1e-3
        got = oracle.settle_velocities_one(pos, v, m_o, m_h, [60, 60, 60, 0, 0, 0])
        r = [pos[a] - pos[b] for a, b in bonds]
        A = np.zeros((3, 3))
        rhs = np.zeros(3)
        for k, (a, b) in enumerate(bonds):
            rhs[k] = -(v[a] - v[b]) @ r[k]
            for l, (c, d) in enumerate(bonds):
                # d/dg_l of (v_a - v_b) . r_k with v_c += g_l r_l / m_c, v_d -= g_l r_l / m_d
                coef = 0.0
                for atom, sgn in ((c, 1.0), (d, -1.0)):
                    if atom == a:
                        coef += sgn / m[atom]
                    if atom == b:
                        coef -= sgn / m[atom]
                A[k, l] = coef * (r[l] @ r[k])
        g = np.linalg.solve(A, rhs)
        want = v.copy()
        for l, (c, d) in enumerate(bonds):
            want[c] += g[l] * r[l] / m[c]
            want[d] -= g[l] * r[l] / m[d]
        assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(v)) + 1e-18


def _bond_form(G, ref_pos, x, bonds):
    """Write per-atom constraint forces G as sum_b g_b (+/-) r_b(ref) (one
    multiplier per bond, least squares: exact when the forces are bond
    forces) and return (residual, sum_b g_b r_b(ref) . r_b(x)), the virial
    sum_a G_a . x_a in bond form, independent of the origin."""
    B = np.zeros((9, len(bonds)))
    for k, (a, b) in enumerate(bonds):
        rb = ref_pos[a] - ref_pos[b]
        B[3 * a:3 * a + 3, k] += rb
        B[3 * b:3 * b + 3, k] -= rb
    g, *_ = np.linalg.lstsq(B, G.reshape(9), rcond=None)
    resid = np.abs(B @ g - G.reshape(9)).max() / np.abs(G).max()
    w = sum(g[k] * ((ref_pos[a] - ref_pos[b]) @ (x[a] - x[b])) for k, (a, b) in enumerate(bonds))
    return resid, w


def test_settle_constraint_virials_equal_bond_form():
    """The SETTLE virials (position level: forces m corr / dt^2; velocity
    level: m dv / dt, each dotted with the atom's position relative to the
    COM, kernels/settle.cu) equal the SHAKE / RATTLE bond form
    sum_b g_b r_b(ref) . r_b: the forces lie along the reference bonds and
    the per-atom virials sum to an origin-free total."""
    rng = np.random.default_rng(21)
    m_o, m_h = 15.9994 * U.AMU, 1.008 * U.AMU
    m = np.array([m_o, m_h, m_h])
    bonds = [(0, 1), (0, 2), (1, 2)]
    lat = [60, 60, 60, 0, 0, 0]
    dt = 80.0
    for _ in range(10):
        snap, d_oh, d_hh = _water(rng)
        geom = oracle.settle_geometry(d_oh, d_hh, m_o, m_h)
        cur = snap + rng.standard_normal((3, 3)) * 0.03
        new, _, vir = oracle.settle_positions_one(snap, cur, np.zeros((3, 3)), m_o, m_h, geom,
                                                  dt, lat)
        G = m[:, None] * (new - cur) / dt ** 2
        resid, w = _bond_form(G, snap, new, bonds)
        assert resid < 1e-9
        assert vir.sum() == pytest.approx(w, rel=1e-9)
        # velocity level at the reset positions
        v = rng.standard_normal((3, 3)) * 1e-3
        v2, vir_v = oracle.settle_velocities_one(new, v, m_o, m_h, lat, dt=dt)
        G = m[:, None] * (v2 - v) / dt
        resid, w = _bond_form(G, new, new, bonds)
        assert resid < 1e-9
        assert vir_v.sum() == pytest.approx(w, rel=1e-9)
        assert oracle.settle_velocities_one(new, v, m_o, m_h, lat) == pytest.approx(v2)


def test_csvr_limits_and_stationarity():
    """CSVR's K' (Bussi, Donadio & Parrinello 2007): c = 1 keeps K; c = 0
    gives kT/2 times the sum of g_dof squared Philox normals (summed here
    from the Random123-pinned normals); and the canonical distribution of K
    (Gamma(N_f/2, kT)) is stationary: K drawn from it maps to K' with the
    same mean N_f kT/2 and variance N_f (kT)^2/2."""
    kt_half = 0.5e-3
    assert oracle.csvr_knew(0.37, 7, 3, 40, 1.0, kt_half) == pytest.approx(0.37, rel=1e-15)
    xs = [oracle.philox_gaussian(7, 0, 3, 0, i, 0) for i in range(40)]
    assert oracle.csvr_knew(0.37, 7, 3, 40, 0.0, kt_half) == pytest.approx(
        kt_half * sum(x * x for x in xs), rel=1e-13)
    g = 30
    kt = 2.0 * kt_half
    rng = np.random.default_rng(1)
    k0 = rng.gamma(g / 2.0, kt, size=20000)
    k1 = np.array([oracle.csvr_knew(k, 11, n, g, 0.8, kt_half) for n, k in enumerate(k0)])
    assert k1.mean() == pytest.approx(g * kt / 2, rel=0.02)
    assert k1.var() == pytest.approx(g * kt ** 2 / 2, rel=0.05)


def test_kinetic_energy_closed_form():
    v = np.array([[1.0, 2.0, 3.0], [0.5, 0, 0]])
    m = np.array([2.0, 4.0])
    assert oracle.kinetic_energy(v, m) == pytest.approx(0.5 * 2 * 14 + 0.5 * 4 * 0.25)


def _pressure_system(p_target):
    """The reference's 8-particle construction (tests/barostats_c_rescale.rs:
    132-158): K from +-500 m/s velocities of 1.66e-27 kg particles in a 1 nm
    cubic box, W chosen so that (2K + W)/(3V) is the given pressure."""
    mass = 1.66e-27 / U.MASS_ME_TO_KG
    v_mag = 500.0 / 2187691.2636411153
    k = 0.5 * mass * 8 * v_mag ** 2
    v = (1.0e-9 / U.LENGTH_BOHR_TO_M) ** 3
    return k, 3.0 * v * p_target - 2.0 * k, v


def test_crescale_limits_and_noise():
    """The c-rescale move (orc_crescale_mu).  At kT -> 0 it is the Berendsen
    move mu^3 = 1 - beta (dt/tau)(P_target - P) with P recovered from the
    constructed K and W (the reference's
    temperature_zero_limit_matches_berendsen_barostat, P = P_target / 2);
    at P = P_target and kT -> 0 mu is 1; the noise over draw counters has mean
    0 and variance 2 beta kT dt / (tau V); a collapse is floored at mu3_min;
    the injected energy is P_target (V_post - V)."""
    p_si = 1.0e6
    p_au = p_si / U.PRESSURE_AU_TO_PA
    tau = 1.0e-12 / U.TIME_AU_TO_S
    beta = 4.5e-10 * U.PRESSURE_AU_TO_PA
    dt = float(np.float32(1.0e-13 / U.TIME_AU_TO_S))
    k, w, v = _pressure_system(p_au / 2.0)
    mu, pres, v_post, inj = oracle.crescale_mu(k, w, v, 1, 0, p_au, tau, beta, 1e-30, dt)
    assert pres == pytest.approx(p_au / 2.0, rel=1e-12)
    assert mu ** 3 == pytest.approx(1.0 - beta * (dt / tau) * (p_au / 2.0), rel=1e-12)
    assert v_post == pytest.approx(v * mu ** 3, rel=1e-14)
    assert inj == pytest.approx(p_au * (v_post - v), rel=1e-12)
    k, w, v = _pressure_system(p_au)
    mu, pres, _, _ = oracle.crescale_mu(k, w, v, 1, 0, p_au, tau, beta, 1e-30, dt)
    assert abs(mu - 1.0) < 1e-14
    kt = 300.0 / U.TEMPERATURE_AU_TO_K
    x = np.array([oracle.crescale_mu(k, w, v, 5, n, p_au, tau, beta, kt, dt)[0] ** 3 - 1.0
                  for n in range(20000)])
    var = 2.0 * beta * kt * dt / (tau * v)
    assert abs(x.mean()) < 4.0 * np.sqrt(var / x.size)
    assert x.var() == pytest.approx(var, rel=0.05)
    mu, _, v_post, _ = oracle.crescale_mu(k, -1e6 * abs(w) - 1.0, v, 1, 0, p_au, tau,
                                          beta * 1e6, 1e-30, dt)
    assert mu == pytest.approx(1e-6, rel=1e-12) and v_post == pytest.approx(1e-18 * v)
