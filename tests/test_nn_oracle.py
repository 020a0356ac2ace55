"""CPU pins of the learned-correction oracle (oracle/nn.py): closed forms of
the basis, cutoff and descriptors, the exact network output on hand-set
weights, symmetry of the pair forces, finite-difference gradients, and the
first Adam step."""
import numpy as np
import pytest

from oracle import nn as N


def cubic(L):
    return np.array([L, L, L, 0.0, 0.0, 0.0])


def random_box(n, L, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, (n, 3))
    q = rng.choice([-0.82, 0.41], n)
    sp = (q > 0).astype(np.int64)
    return pos, q, sp


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    got = N.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                       # tie -> even mantissa
    assert got[2] == 1 + 2 ** -6               # tie -> even (rounds up)
    assert got[3] == -2.5
    assert abs(got[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3       # half an ulp


def test_rbf_and_cutoff_closed_forms():
    s = np.array([0.0, 1 / 15, 0.5, 1.0])
    phi = N.rbf(s)
    assert phi[0, 0] == 1.0 and phi[1, 1] == 1.0 and phi[3, 15] == 1.0
    assert phi[1, 0] == pytest.approx(np.exp(-0.5))
    assert phi[2, 7] == pytest.approx(np.exp(-0.125))       # u = 7.5
    fc = N.cutoff(np.array([0.0, 0.5, 1.0, 1.2]))
    assert fc.tolist() == pytest.approx([1.0, 0.5, 0.0, 0.0])


def test_descriptors_two_and_three_atoms():
    L, r_nn = 40.0, 18.0
    pos = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0 + 7.0], [1.0, 2.0 + 12.0, 3.0]])
    q = np.array([-0.8, 0.4, 0.3])
    D = N.descriptors(pos, q, cubic(L), r_nn)
    def w(r):
        return N.rbf(np.array([r / r_nn]))[0] * N.cutoff(np.array([r / r_nn]))[0]
    r13 = 12.0
    r23 = np.hypot(7.0, 12.0)
    assert np.allclose(D[0], q[1] * w(7.0) + q[2] * w(r13))
    assert np.allclose(D[1], q[0] * w(7.0) + q[2] * w(r23))
    assert np.allclose(D[2], q[0] * w(r13) + q[1] * w(r23))


def test_descriptors_periodic_image():
    L, r_nn = 20.0, 9.0
    pos = np.array([[0.5, 0.0, 0.0], [19.0, 0.0, 0.0]])   # 1.5 apart through x = 0
    q = np.array([1.0, 2.0])
    D = N.descriptors(pos, q, cubic(L), r_nn)
    s = np.array([1.5 / r_nn])
    assert np.allclose(D[0], 2.0 * N.rbf(s)[0] * N.cutoff(s)[0])


def test_forward_constant_network_closed_form():
    """W1 = W2 = 0: the network is the constant c = w3.SiLU(b2) + b3, so
    dF_1 = s_out c fc(r/r_nn) r^_12 and dF_2 = -dF_1."""
    P = {k: np.zeros(s) for k, s in N.SIZES}
    rng = np.random.default_rng(1)
    P["b1"] = rng.normal(size=64)
    P["b2"] = rng.normal(size=64)
    P["w3"] = rng.normal(size=64)
    P["b3"] = np.array([0.3])
    c = P["w3"] @ N.silu(P["b2"]) + 0.3
    pos = np.array([[1.0, 1.0, 1.0], [4.0, 5.0, 1.0]])      # r = 5
    r_nn, s_out = 10.0, 2e-3
    dF = N.forward(N.pack(P), pos, np.array([-0.8, 0.4]), np.array([0, 1]), cubic(30.0),
                   r_nn, s_out)
    rhat = (pos[0] - pos[1]) / 5.0
    want = s_out * c * N.cutoff(np.array([0.5]))[0] * rhat
    assert np.allclose(dF[0], want) and np.allclose(dF[1], -want)


def test_forward_species_one_hot_selects_column():
    """Only the one-hot columns of W1 set: N(x) depends on the species pair
    alone, y_k = w3.SiLU(W2 SiLU(W1[:, 16 + k]) + b2) + b3 for k = s_i + s_j,
    and dF_0 = s_out fc y_k r^_01."""
    rng = np.random.default_rng(2)
    P = {k: np.zeros(s) for k, s in N.SIZES}
    P["W1"][:, 16:19] = rng.normal(size=(64, 3))
    P["W2"] = rng.normal(scale=0.2, size=(64, 64))
    P["b2"] = rng.normal(size=64)
    P["w3"] = rng.normal(size=64)
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    fc = N.cutoff(np.array([0.3]))[0]
    for sp in ([0, 0], [0, 1], [1, 1]):
        k = sp[0] + sp[1]
        y = P["w3"] @ N.silu(P["W2"] @ N.silu(P["W1"][:, 16 + k]) + P["b2"])
        dF = N.forward(N.pack(P), pos, np.zeros(2), np.array(sp), cubic(30.0), 10.0, 1.0)
        assert dF[0] == pytest.approx([-fc * y, 0.0, 0.0], abs=1e-14)


def test_forward_is_antisymmetric_and_permutation_equivariant():
    pos, q, sp = random_box(40, 15.0, 3)
    params = np.random.default_rng(4).normal(scale=0.3, size=N.N_PARAMS)
    dF = N.forward(params, pos, q, sp, cubic(15.0), 7.0, 1e-3)
    assert np.abs(dF.sum(0)).max() < 1e-12 * np.abs(dF).sum()
    perm = np.random.default_rng(5).permutation(40)
    dFp = N.forward(params, pos[perm], q[perm], sp[perm], cubic(15.0), 7.0, 1e-3)
    assert np.allclose(dFp, dF[perm], rtol=1e-12, atol=1e-18)


def test_loss_grad_matches_finite_differences():
    pos, q, sp = random_box(24, 12.0, 6)
    rng = np.random.default_rng(7)
    params = rng.normal(scale=0.3, size=N.N_PARAMS)
    target = rng.normal(scale=1e-3, size=(24, 3))
    lat, r_nn, s_out = cubic(12.0), 6.0, 1e-3
    loss, g, _ = N.loss_grad(params, pos, q, sp, lat, r_nn, s_out, target)
    idx = rng.choice(N.N_PARAMS, 40, replace=False)
    idx = np.concatenate([idx, [N.N_PARAMS - 1, 64 * 51 + 3, 64 * 51 + 64 + 5]])
    h = 1e-5
    for k in idx:
        p1 = params.copy(); p1[k] += h
        p2 = params.copy(); p2[k] -= h
        l1 = N.loss_grad(p1, pos, q, sp, lat, r_nn, s_out, target)[0]
        l2 = N.loss_grad(p2, pos, q, sp, lat, r_nn, s_out, target)[0]
        fd = (l1 - l2) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9 * np.abs(g).max())


def test_bf16_mode_close_to_exact():
    pos, q, sp = random_box(30, 12.0, 8)
    params = np.random.default_rng(9).normal(scale=0.3, size=N.N_PARAMS)
    a = N.forward(params, pos, q, sp, cubic(12.0), 6.0, 1e-3)
    b = N.forward(params, pos, q, sp, cubic(12.0), 6.0, 1e-3, bf16_mode=True)
    rel = np.sqrt(((a - b) ** 2).sum() / (a ** 2).sum())
    assert 0 < rel < 2e-2


def test_adam_first_step_is_lr_times_sign():
    g = np.array([3.0, -1e-3, 0.5])
    p, m, v = N.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=1e-3)
    assert np.allclose(p, -1e-3 * np.sign(g), rtol=1e-4)
    p2, _, _ = N.adam_step(p, g, m, v, 2, lr=1e-3)
    assert np.allclose(p2, 2 * p, rtol=1e-4)        # constant gradient: same step
