"""FP64 CPU oracle of the learned coarse-mesh SPME correction -- TEST
INFRASTRUCTURE ONLY (imported by tests/ alone; the product never imports it).

The reference has no learned component: this restates the model defined for
the 196,608-atom / 64^3-mesh workload (a pair-scalar MLP correcting a coarse
SPME mesh toward a fine one), in plain numpy, so that the HIP kernels can be
checked against it:

* s = r / r_nn, u = 15 s; radial basis phi_k(s) = exp(-(u - k)^2 / 2),
  k = 0..15 (16 Gaussians with centres evenly spaced over 0..r_nn, width one
  spacing); smooth cutoff fc(s) = (1 + cos(pi s)) / 2 for s < 1, else 0.
* per-atom descriptor D_i[k] = sum over j != i with r_ij < r_nn of
  q_j phi_k(s_ij) fc(s_ij) (every pair, intramolecular ones included);
* pair input x_ij = [phi(16), one-hot(species_i + species_j) (3), D_i (16),
  D_j (16)] = 51 values;
* N(x) = w3 . SiLU(W2 SiLU(W1 x + b1) + b2) + b3 (51 -> 64 -> 64 -> 1);
* g_ij = s_out fc(s_ij) (N(x_ij) + N(x_ji)) / 2 (symmetric in i, j), and the
  force correction dF_i = sum_j g_ij (r_i - r_j) / r_ij -- antisymmetric pair
  forces, so the correction carries no net force;
* loss per frame L = sum_i |dF_i - T_i|^2 / (3 N s_out^2); Adam with the
  textbook bias-corrected update (Kingma & Ba 2015, Algorithm 1).

``bf16=True`` rounds (round to nearest even, from f32) at exactly the points
where the GPU kernels hold bf16 operands: the weights W1, W2, w3, the inputs
x, the hidden activations h1 and, in the backward pass, the error signals
dz2 and dz1 that enter the matrix-core products.  ``bf16=False`` is the exact
FP64 model, pinned by finite differences in the tests.

Parameter vector layout (7,553 values): W1[64][51], b1[64], W2[64][64],
b2[64], w3[64], b3[1].
"""
import numpy as np

N_RBF = 16
N_HID = 64
N_IN = 2 * N_RBF + N_RBF + 3   # 51
SIZES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)),
         ("b2", (N_HID,)), ("w3", (N_HID,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(params):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[n], np.float64).ravel() for n, _ in SIZES])


def bf16(a):
    """Round to the nearest bf16 (ties to even) from the f32 value of a."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def rbf(s):
    u = 15.0 * np.asarray(s, np.float64)[:, None]
    k = np.arange(N_RBF, dtype=np.float64)[None, :]
    return np.exp(-0.5 * (u - k) ** 2)


def cutoff(s):
    s = np.asarray(s, np.float64)
    return np.where(s < 1.0, 0.5 * (1.0 + np.cos(np.pi * s)), 0.0)


def min_image(d, lattice):
    """Minimum image for the lower-triangular lattice {lx, ly, lz, xy, xz, yz}:
    fractional coordinates by back-substitution, nearest integer by
    floor(s + 1/2) (the device sequence, in FP64)."""
    lx, ly, lz, xy, xz, yz = (float(v) for v in lattice)
    d = np.array(d, dtype=np.float64, copy=True)
    sc = d[:, 2] / lz
    sb = (d[:, 1] - sc * yz) / ly
    sa = (d[:, 0] - sb * xy - sc * xz) / lx
    ka, kb, kc = np.floor(sa + 0.5), np.floor(sb + 0.5), np.floor(sc + 0.5)
    d[:, 0] -= ka * lx + kb * xy + kc * xz
    d[:, 1] -= kb * ly + kc * yz
    d[:, 2] -= kc * lz
    return d


def neighbor_pairs(pos, lattice, r_nn):
    """All unordered pairs i < j with minimum-image distance < r_nn
    (brute force in chunks): returns i, j, d = r_i - r_j, r."""
    pos = np.asarray(pos, np.float64)
    n = len(pos)
    I, J, D = [], [], []
    chunk = max(1, 2_000_000 // max(n, 1))
    for a in range(0, n, chunk):
        ii = np.arange(a, min(a + chunk, n))
        jj = np.arange(n)
        gi, gj = np.meshgrid(ii, jj, indexing="ij")
        m = gj > gi
        gi, gj = gi[m], gj[m]
        d = min_image(pos[gi] - pos[gj], lattice)
        r2 = np.einsum("ij,ij->i", d, d)
        keep = r2 < r_nn * r_nn
        I.append(gi[keep]); J.append(gj[keep]); D.append(d[keep])
    i = np.concatenate(I); j = np.concatenate(J); d = np.concatenate(D)
    return i, j, d, np.sqrt(np.einsum("ij,ij->i", d, d))


def descriptors(pos, q, lattice, r_nn, pairs=None):
    i, j, d, r = pairs if pairs is not None else neighbor_pairs(pos, lattice, r_nn)
    s = r / r_nn
    w = rbf(s) * cutoff(s)[:, None]
    q = np.asarray(q, np.float64)
    D = np.zeros((len(pos), N_RBF))
    np.add.at(D, i, q[j][:, None] * w)
    np.add.at(D, j, q[i][:, None] * w)
    return D


def _inputs(phi, species, i, j, D, a, b):
    """Rows x for orientation (a, b) of every pair."""
    oh = np.zeros((len(i), 3))
    oh[np.arange(len(i)), species[i] + species[j]] = 1.0
    return np.concatenate([phi, oh, D[a], D[b]], axis=1)


def _mlp(P, x, bf):
    W1, W2, w3 = P["W1"], P["W2"], P["w3"]
    if bf:
        W1, W2, w3, x = bf16(W1), bf16(W2), bf16(w3), bf16(x)
    z1 = x @ W1.T + P["b1"]
    h1 = silu(z1)
    if bf:
        h1 = bf16(h1)
    z2 = h1 @ W2.T + P["b2"]
    h2 = silu(z2)
    y = h2 @ w3 + P["b3"][0]
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, W2=W2, w3=w3)


def forward(params, pos, q, species, lattice, r_nn, s_out, D=None, bf16_mode=False,
            pairs=None):
    """Force correction dF (N, 3).  D (the descriptors) may be passed in so
    that the network can be checked on identical inputs."""
    P = unpack(params)
    pr = pairs if pairs is not None else neighbor_pairs(pos, lattice, r_nn)
    i, j, d, r = pr
    if D is None:
        D = descriptors(pos, q, lattice, r_nn, pr)
    species = np.asarray(species)
    s = r / r_nn
    phi = rbf(s)
    y_ij, _ = _mlp(P, _inputs(phi, species, i, j, D, i, j), bf16_mode)
    y_ji, _ = _mlp(P, _inputs(phi, species, i, j, D, j, i), bf16_mode)
    g = s_out * cutoff(s) * 0.5 * (y_ij + y_ji)
    f = g[:, None] * d / r[:, None]
    dF = np.zeros((len(pos), 3))
    np.add.at(dF, i, f)
    np.add.at(dF, j, -f)
    return dF


def _mlp_backward(P, c, dy, bf):
    """Gradients of sum(dy * y) for rows with cache c."""
    g = {}
    g["b3"] = np.array([dy.sum()])
    g["w3"] = c["h2"].T @ dy
    dz2 = (dy[:, None] * c["w3"][None, :]) * silu_grad(c["z2"])
    g["b2"] = dz2.sum(0)
    dz2m = bf16(dz2) if bf else dz2
    g["W2"] = dz2m.T @ c["h1"]
    dh1 = dz2m @ c["W2"]
    dz1 = dh1 * silu_grad(c["z1"])
    g["b1"] = dz1.sum(0)
    dz1m = bf16(dz1) if bf else dz1
    g["W1"] = dz1m.T @ c["x"]
    return g


def loss_grad(params, pos, q, species, lattice, r_nn, s_out, target, D=None,
              bf16_mode=False, pairs=None):
    """Loss of one frame and its gradient with respect to the parameters."""
    P = unpack(params)
    pr = pairs if pairs is not None else neighbor_pairs(pos, lattice, r_nn)
    i, j, d, r = pr
    if D is None:
        D = descriptors(pos, q, lattice, r_nn, pr)
    species = np.asarray(species)
    n = len(pos)
    s = r / r_nn
    phi = rbf(s)
    fc = cutoff(s)
    y_ij, c_ij = _mlp(P, _inputs(phi, species, i, j, D, i, j), bf16_mode)
    y_ji, c_ji = _mlp(P, _inputs(phi, species, i, j, D, j, i), bf16_mode)
    g = s_out * fc * 0.5 * (y_ij + y_ji)
    rhat = d / r[:, None]
    dF = np.zeros((n, 3))
    np.add.at(dF, i, g[:, None] * rhat)
    np.add.at(dF, j, -g[:, None] * rhat)
    res = dF - np.asarray(target, np.float64)
    den = 3.0 * n * s_out * s_out
    loss = float(np.sum(res * res) / den)
    R = 2.0 * res / den                       # dL/d dF
    dg = np.einsum("pk,pk->p", R[i] - R[j], rhat)
    dy = dg * s_out * fc * 0.5
    ga = _mlp_backward(P, c_ij, dy, bf16_mode)
    gb = _mlp_backward(P, c_ji, dy, bf16_mode)
    return loss, pack({k: ga[k] + gb[k] for k in ga}), dF


def adam_step(params, grad, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (t counts from 1); returns new params, m, v."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
