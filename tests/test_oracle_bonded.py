"""The bonded terms of the oracle (orc_bonded: harmonic and Morse bonds,
harmonic angles, periodic dihedrals) pinned by what the mathematics fixes:
forces equal to minus the central-difference gradient of the energy, zero
net force and torque, the virial equal to the sum of r . F, closed-form
energies at known geometries (bond stretched by delta, the Morse minimum
and plateau, a 120-degree angle, cis and trans dihedrals) and a bond
across the periodic boundary equal to the same bond unwrapped."""
import numpy as np
import pytest

import oracle

BIG = np.array([1000.0, 1000.0, 1000.0, 0.0, 0.0, 0.0])
CASES = {
    "harmonic-bond": ([[0, 1, 0], [1, 2, 1], [3, 4, 0]], [[0.35, 1.9], [0.5, 2.2]]),
    "morse": ([[0, 1, 0], [2, 5, 1]], [[0.17, 1.1, 1.9], [0.1, 0.8, 2.4]]),
    "harmonic-angle": ([[0, 1, 2, 0], [1, 2, 3, 1], [4, 5, 6, 0]],
                       [[0.12, 1.91], [0.08, 2.3]]),
    "periodic-dihedral": ([[0, 1, 2, 3, 0], [1, 2, 3, 4, 1], [4, 5, 6, 7, 2]],
                          [[0.01, 0.0, 3], [0.02, 3.14159, 1], [0.005, 0.4, 2]]),
}


def _chain(seed=3, n=8):
    rng = np.random.default_rng(seed)
    steps = rng.normal(size=(n, 3))
    steps *= 2.0 / np.linalg.norm(steps, axis=1, keepdims=True)
    return 500.0 + np.cumsum(steps, axis=0)


@pytest.mark.parametrize("kind", sorted(CASES))
def test_bonded_forces_are_minus_the_gradient(kind):
    terms, params = CASES[kind]
    x = _chain()
    f, e, w, u = oracle.bonded(kind, x, BIG, terms, params)
    h = 1e-5
    g = np.zeros_like(x)
    for a in range(x.shape[0]):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            g[a, k] = (oracle.bonded(kind, xp, BIG, terms, params)[3]
                       - oracle.bonded(kind, xm, BIG, terms, params)[3]) / (2 * h)
    assert np.abs(f + g).max() < 1e-7 * max(1.0, np.abs(f).max()) + 1e-9
    assert np.abs(f).max() > 1e-4                       # the case is not trivial
    assert np.abs(f.sum(axis=0)).max() < 1e-12
    assert np.abs(np.cross(x - 500.0, f).sum(axis=0)).max() < 1e-9
    assert e.sum() == pytest.approx(u, rel=1e-12)
    assert w.sum() == pytest.approx(np.sum((x - 500.0) * f), rel=1e-9, abs=1e-12)


def test_bonded_closed_forms():
    k, r0, d = 0.4, 2.0, 0.1
    x = np.array([[1.0, 1.0, 1.0], [1.0 + r0 + d, 1.0, 1.0]])
    f, _, _, u = oracle.bonded("harmonic-bond", x, BIG, [[0, 1, 0]], [[k, r0]])
    assert u == pytest.approx(0.5 * k * d * d, rel=1e-12)
    assert f[0, 0] == pytest.approx(k * d, rel=1e-12) and f[1, 0] == pytest.approx(-k * d)
    de, a = 0.2, 1.3
    x = np.array([[1.0, 1.0, 1.0], [1.0 + r0, 1.0, 1.0]])
    f, _, _, u = oracle.bonded("morse", x, BIG, [[0, 1, 0]], [[de, a, r0]])
    assert abs(u) < 1e-15 and np.abs(f).max() < 1e-15
    x[1, 0] = 1.0 + r0 + 40.0
    assert oracle.bonded("morse", x, BIG, [[0, 1, 0]], [[de, a, r0]])[3] == pytest.approx(de)
    # 120 degrees at the middle atom against theta0 = 90 degrees
    x = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0],
                  [np.cos(2 * np.pi / 3), np.sin(2 * np.pi / 3), 0.0]]) + 5.0
    u = oracle.bonded("harmonic-angle", x, BIG, [[0, 1, 2, 0]], [[0.3, np.pi / 2]])[3]
    assert u == pytest.approx(0.5 * 0.3 * (np.pi / 6) ** 2, rel=1e-12)
    # cis (phi = 0) and trans (phi = pi): k (1 + cos(phi)) with n = 1, phi0 = 0
    cis = np.array([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]) + 5.0
    trans = cis.copy(); trans[3, 0] = 4.0
    kp = 0.05
    assert oracle.bonded("periodic-dihedral", cis, BIG, [[0, 1, 2, 3, 0]],
                         [[kp, 0.0, 1]])[3] == pytest.approx(2 * kp, rel=1e-12)
    assert abs(oracle.bonded("periodic-dihedral", trans, BIG, [[0, 1, 2, 3, 0]],
                             [[kp, 0.0, 1]])[3]) < 1e-15
    # a right-angle dihedral with n = 2: cos(2 * 90 degrees) = -1
    right = cis.copy(); right[3] = [5.0, 5.0, 6.0]
    assert abs(oracle.bonded("periodic-dihedral", right, BIG, [[0, 1, 2, 3, 0]],
                             [[kp, 0.0, 2]])[3]) < 1e-15


def test_bonded_across_the_boundary():
    L = np.array([20.0, 20.0, 20.0, 0.0, 0.0, 0.0])
    x = np.array([[19.5, 3.0, 3.0], [1.2, 3.4, 2.8], [2.9, 4.4, 3.1], [3.5, 6.0, 4.0]])
    xu = x.copy(); xu[0, 0] -= 20.0
    for kind, terms, params in [("harmonic-bond", [[0, 1, 0]], [[0.3, 2.0]]),
                                ("harmonic-angle", [[0, 1, 2, 0]], [[0.1, 1.8]]),
                                ("periodic-dihedral", [[0, 1, 2, 3, 0]], [[0.02, 0.3, 3]])]:
        a = oracle.bonded(kind, x, L, terms, params)
        b = oracle.bonded(kind, xu, BIG, terms, params)
        assert np.allclose(a[0], b[0], rtol=1e-12, atol=1e-14)
        assert a[3] == pytest.approx(b[3], rel=1e-12)
