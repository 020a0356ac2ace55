"""Synthetic inputs for the benchmark workloads (SI in, atomic units out).

  argon(...)       -- LJ argon on a simple-cubic lattice (10,000 atoms).
  spc_water(...)   -- rigid SPC water on a grid of molecule centres scaled
                      to a cubic box at 997 kg/m^3, random orientations.
  spc_water_n(...) -- the same for any molecule count.
  replica_specs    -- the 64-box replica workload (sizes, cutoffs, meshes).
  water_forcefield -- LJ O-O + real-space Ewald + SPME descriptor.

Each returns plain numpy arrays and dicts accepted by heddlemd_amd.api.
"""
import math

import numpy as np

from . import units as U

# SPC (Berendsen 1981) geometry and parameters.
SPC_D_OH = 1.0 * U.ANGSTROM
SPC_ANGLE = 109.47
SPC_Q_O, SPC_Q_H = -0.82, 0.41
SPC_SIGMA_OO = 3.166 * U.ANGSTROM
SPC_EPS_OO = 0.650 * U.KJ_PER_MOL
M_O, M_H = 15.9994, 1.008
WATER_DENSITY = 997.0  # kg/m^3


def spc_d_hh():
    return 2.0 * SPC_D_OH * math.sin(U.deg(SPC_ANGLE) / 2.0)


def _random_rotations(n, rng):
    """Uniformly random rotations from unit quaternions (Shoemake 1992)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
                  np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3),
                  np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _wrap(x, L):
    return x - L * np.floor(x / L + 0.5)


def water_edge(n_mol, density=WATER_DENSITY):
    """Edge (bohr) of the cubic box holding n_mol SPC molecules at `density`."""
    m_mol = (M_O + 2 * M_H) * U.DALTON_KG
    return (n_mol * m_mol / density) ** (1.0 / 3.0) / U.LENGTH_BOHR_TO_M


def spc_water(grid=(32, 32, 8), seed=7, density=WATER_DENSITY):
    """Molecule centres (centre of mass) on grid[0] x grid[1] x grid[2] cells of
    a cubic box at `density`; atoms ordered O, H1, H2 per molecule."""
    nx, ny, nz = grid
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    frac = np.stack([(ix.ravel() + 0.5) / nx, (iy.ravel() + 0.5) / ny,
                     (iz.ravel() + 0.5) / nz], axis=1)
    return _water_from_centres(frac, seed, density)


def spc_water_n(n_mol, seed, density=WATER_DENSITY):
    """n_mol molecules at `density` on a seeded random choice of the sites of
    the smallest k^3 grid with at least n_mol sites (any molecule count)."""
    k = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
    ix, iy, iz = np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij")
    sites = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    pick = np.sort(np.random.default_rng(seed + 1_000_003).permutation(k ** 3)[:n_mol])
    return _water_from_centres((sites[pick] + 0.5) / k, seed, density)


def _water_from_centres(frac, seed, density):
    n_mol = frac.shape[0]
    L = water_edge(n_mol, density)
    centres = frac * L - 0.5 * L
    half = U.deg(SPC_ANGLE) / 2.0
    local = np.array([[0.0, 0.0, 0.0],
                      [SPC_D_OH * math.sin(half), SPC_D_OH * math.cos(half), 0.0],
                      [-SPC_D_OH * math.sin(half), SPC_D_OH * math.cos(half), 0.0]])
    com = (M_O * local[0] + M_H * local[1] + M_H * local[2]) / (M_O + 2 * M_H)
    local = local - com
    rng = np.random.default_rng(seed)
    R = _random_rotations(n_mol, rng)
    pos = centres[:, None, :] + np.einsum("mij,aj->mai", R, local)
    pos = _wrap(pos.reshape(-1, 3), L)
    n = 3 * n_mol
    q = np.tile([SPC_Q_O, SPC_Q_H, SPC_Q_H], n_mol)
    posq = np.concatenate([pos, q[:, None]], axis=1).astype(np.float32)
    mass = np.tile([M_O, M_H, M_H], n_mol) * U.AMU
    types = np.tile([0, 1, 1], n_mol).astype(np.uint32)
    groups = np.arange(n, dtype=np.uint32).reshape(n_mol, 3)
    # intramolecular exclusions (implicit constraint-group exclusions,
    # src/forces/topology.rs:1171-1177), symmetric CSR
    off = np.arange(0, 2 * n + 1, 2, dtype=np.uint32)
    par = np.empty(2 * n, np.uint32)
    base = 3 * np.repeat(np.arange(n_mol), 3)
    loc = np.tile([0, 1, 2], n_mol)
    par[0::2] = base + (loc + 1) % 3
    par[1::2] = base + (loc + 2) % 3
    par = par.reshape(n, 2)
    par.sort(axis=1)
    return {
        "n": n, "n_mol": n_mol, "L": L, "lattice": np.array([L, L, L, 0, 0, 0], np.float64),
        "posq": posq, "mass": mass, "types": types, "groups": groups,
        "exclusions": (off, par.ravel()),
        "settle": {"groups": groups, "d_oh": SPC_D_OH, "d_hh": spc_d_hh(),
                   "m_o": M_O * U.AMU, "m_h": M_H * U.AMU},
    }


def water_forcefield(r_cut_nm=1.0, r_skin_nm=0.1, grid=(64, 64, 64), order=4,
                     ewald_tol=1e-5, r_switch_nm=None, exclusions=None):
    """LJ O-O (H has no LJ) with the reference's default r_switch = 0.9 r_c
    (src/forces/lj.rs:67), real-space Ewald with alpha from
    erfc(alpha r_c) = ewald_tol, SPME of the given grid and order."""
    rc = r_cut_nm * U.NM
    rs = (0.9 * r_cut_nm if r_switch_nm is None else r_switch_nm) * U.NM
    sig = np.full(4, SPC_SIGMA_OO)
    eps = np.array([SPC_EPS_OO, 0.0, 0.0, 0.0])
    ff = {
        "n_types": 2,
        "lj": {"sigma": sig, "epsilon": eps, "cutoff": np.full(4, rc), "switch": np.full(4, rs)},
        "coulomb": {"alpha": U.erfc_alpha(rc, ewald_tol), "r_cut": rc, "k_coulomb": 1.0},
        "spme": {"grid": tuple(grid), "order": order} if grid is not None else None,
        "r_skin": r_skin_nm * U.NM,
        "exclusions": exclusions,
    }
    return ff


def argon(n_atoms=10000, n_side=22, edge_nm=7.84, seed=42, temperature_k=94.4):
    """First n_atoms sites of an n_side^3 simple-cubic lattice filling a cubic
    box of edge edge_nm, shifted into [-L/2, L/2)."""
    L = edge_nm * U.NM
    a = L / n_side
    ix, iy, iz = np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side),
                             indexing="ij")
    sites = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)[:n_atoms]
    pos = (sites + 0.5) * a - 0.5 * L
    posq = np.concatenate([pos, np.zeros((n_atoms, 1))], axis=1).astype(np.float32)
    mass = np.full(n_atoms, 6.6335e-26 / U.MASS_ME_TO_KG)
    sigma = 3.405 * U.ANGSTROM
    eps = 119.8 / U.TEMPERATURE_AU_TO_K
    rc = 2.5 * sigma
    ff = {
        "n_types": 1,
        "lj": {"sigma": [sigma], "epsilon": [eps], "cutoff": [rc], "switch": [rc]},
        "coulomb": None, "spme": None,
        "r_skin": 0.3 * rc,   # reference default skin (src/io/config.rs:1569-1577)
        "exclusions": None,
    }
    return {"n": n_atoms, "L": L, "lattice": np.array([L, L, L, 0, 0, 0]), "posq": posq,
            "mass": mass, "types": np.zeros(n_atoms, np.uint32), "ff": ff,
            "temperature": U.kelvin(temperature_k), "seed": seed}


def _smooth_even(n):
    """Smallest even 2-3-5-smooth integer >= n whose half is also 2-3-5
    smooth (the FFT's supported mesh lengths)."""
    def smooth(m):
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        return m == 1
    while not (n % 2 == 0 and smooth(n) and smooth(n // 2)):
        n += 1
    return n


def replica_specs(n_replicas=64, seed=23, lo=216, hi=4096, spacing_nm=0.1, mesh_lo=24,
                  mesh_hi=48, small_edge_nm=2.0):
    """Per-replica SPC water boxes: molecule counts log-uniform in [lo, hi]
    from `seed`; a cubic SPME mesh at about `spacing_nm` (the next supported
    length, clamped to [mesh_lo, mesh_hi]); r_c = 1.0 nm, or 0.8 nm for
    edges under small_edge_nm; Langevin seed = seed + replica id."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_replicas))).astype(int)
    out = []
    for r, n_mol in enumerate(counts.tolist()):
        L = water_edge(n_mol)
        edge_nm = L / U.NM
        mesh = min(max(_smooth_even(int(math.ceil(edge_nm / spacing_nm))), mesh_lo), mesh_hi)
        out.append({"replica": r, "n_mol": n_mol, "edge_nm": edge_nm,
                    "r_cut_nm": 0.8 if edge_nm < small_edge_nm else 1.0,
                    "mesh": (mesh, mesh, mesh), "seed": seed + r})
    return out
