"""Small `.in.toml` runs written into a directory: an argon-like LJ box and
a rigid SPC water box with SPME and SETTLE (the structure of the
reference's examples/lj-10000-argon and examples/spc-water-* inputs, at
test sizes)."""
import numpy as np

from heddlemd_amd import systems
from heddlemd_amd import units as U


def _fmt(x):
    return f"{x:.9e}"


def write_xyz(path, lattice_m, species, pos_m, vel=None):
    lx, ly, lz = lattice_m
    props = "species:S:1:pos:R:3" + (":velo:R:3" if vel is not None else "")
    lines = [str(len(species)),
             f'Lattice="{_fmt(lx)} 0 0 0 {_fmt(ly)} 0 0 0 {_fmt(lz)}" Properties={props}']
    for k, (s, p) in enumerate(zip(species, pos_m)):
        row = f"{s} {_fmt(p[0])} {_fmt(p[1])} {_fmt(p[2])}"
        if vel is not None:
            row += f" {_fmt(vel[k][0])} {_fmt(vel[k][1])} {_fmt(vel[k][2])}"
        lines.append(row)
    path.write_text("\n".join(lines) + "\n")


def argon_run(d, n_side=8, n_steps=20, lossless=False, extra=""):
    L = 2.7 * n_side * 1e-10 * 1.3
    a = L / n_side
    idx = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (idx + 0.5) * a - 0.5 * L
    write_xyz(d / "argon.in.xyz", (L, L, L), ["Ar"] * len(pos), pos)
    (d / "argon.in.toml").write_text(f"""schema_version = 1
init = "argon.in.xyz"

[simulation]
seed = 1
temperature = 100.0

[[phase]]
name = "run"
n_steps = {n_steps}
dt = 2.0e-15

[phase.integrator]
kind = "velocity-verlet"
lossless = {str(lossless).lower()}

[phase.output]
trajectory_every = 10
include_velocities = true
log_every = 5

[[particle_types]]
name = "Ar"
mass = 6.6335e-26

[[pair_interactions]]
between = ["Ar", "Ar"]
kind = "lennard-jones"
sigma = 3.40e-10
epsilon = 1.65e-21
cutoff = 8.5e-10
r_switch = 8.5e-10
{extra}""")
    return d / "argon.in.toml"


NPT_TABLES = """
[phase.thermostat]
kind = "csvr"
temperature = 298.15
tau = 1.0e-13
seed = 12

[phase.barostat]
kind = "c-rescale"
pressure = 1.01325e5
temperature = 298.15
tau = 1.0e-12
compressibility = 4.5e-10
seed = 13
"""


def water_run(d, grid=(8, 8, 8), n_steps=20, kind="langevin-baoab"):
    w = systems.spc_water(grid=grid, seed=3)
    L_m = w["L"] * U.LENGTH_BOHR_TO_M
    pos = w["posq"][:, :3].astype(np.float64) * U.LENGTH_BOHR_TO_M
    write_xyz(d / "water.in.xyz", (L_m,) * 3, ["O", "H", "H"] * w["n_mol"], pos)
    rows = "\n".join(f"{3 * m} {3 * m + 1} {3 * m + 2} SPC" for m in range(w["n_mol"]))
    (d / "water.in.topology").write_text("# rigid SPC\n[constraints]\n" + rows + "\n")
    integ = ('kind = "langevin-baoab"\nfriction = 1.0e12\ntemperature = 300.0\nseed = 7'
             if kind == "langevin-baoab" else 'kind = "velocity-verlet"\nlossless = true'
             if kind == "velocity-verlet" else 'kind = "velocity-verlet"\nlossless = false')
    # kind "npt": velocity Verlet with the CSVR thermostat and the c-rescale
    # barostat of the reference's examples/spc-water-8192 production phase
    coupling = NPT_TABLES if kind == "npt" else ""
    (d / "water.in.toml").write_text(f"""schema_version = 1
init = "water.in.xyz"
topology = "water.in.topology"

[simulation]
seed = 5
temperature = 300.0

[[minimization]]
name = "min"
[minimization.algorithm]
kind = "steepest-descent"
max_iterations = 30

[[phase]]
name = "md"
n_steps = {n_steps}
dt = 2.0e-15
[phase.integrator]
{integ}
{coupling}
[phase.output]
trajectory_every = 10
log_every = 10

[[particle_types]]
name = "O"
mass = {15.9994 * U.DALTON_KG!r}
charge = {-0.82 * U.CHARGE_E_TO_C!r}

[[particle_types]]
name = "H"
mass = {1.008 * U.DALTON_KG!r}
charge = {0.41 * U.CHARGE_E_TO_C!r}

[[pair_interactions]]
between = ["O", "O"]
kind = "lennard-jones"
sigma = 3.166e-10
epsilon = {0.650e3 / 6.02214076e23!r}
cutoff = 9.0e-10

[[pair_interactions]]
between = ["H", "H"]
kind = "lennard-jones"
sigma = 1.0e-10
epsilon = 0.0
cutoff = 9.0e-10

[[pair_interactions]]
between = ["H", "O"]
kind = "lennard-jones"
sigma = 1.0e-10
epsilon = 0.0
cutoff = 9.0e-10

[spme]
alpha = 3.47e9
r_cut_real = 9.0e-10
grid = [24, 24, 24]

[[constraint_types]]
name = "SPC"
kind = "settle"
d_OH = 1.0e-10
d_HH = 1.633e-10

[neighbor_list]
mode = "cell-list"
r_skin = 1.0e-10
""")
    return d / "water.in.toml"


def butane_run(d, n_side=4, n_steps=20):
    """United-atom butane (4 LJ sites per chain) with harmonic bonds and
    angles and a periodic dihedral from `[[bond_types]]` / `[[angle_types]]`
    / `[[dihedral_types]]` and the topology's `[bonds]` / `[angles]` /
    `[dihedrals]` rows; no constraints, no charges."""
    L = 25.0e-10
    a = L / n_side
    th = np.deg2rad(112.0)
    b = 1.54e-10
    chain = np.array([[0, 0, 0], [b, 0, 0], [b - b * np.cos(th), b * np.sin(th), 0],
                      [2 * b - b * np.cos(th), b * np.sin(th), 0]])
    idx = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = np.concatenate([(c + 0.2) * a + chain - 0.5 * L for c in idx])
    n_mol = idx.shape[0]
    write_xyz(d / "butane.in.xyz", (L, L, L), ["C"] * (4 * n_mol), pos)
    rows = ["[bonds]"] + [f"{4 * m + k} {4 * m + k + 1} cc" for m in range(n_mol) for k in range(3)]
    rows += ["[angles]"] + [f"{4 * m + k} {4 * m + k + 1} {4 * m + k + 2} ccc"
                            for m in range(n_mol) for k in range(2)]
    rows += ["[dihedrals]"] + [f"{4 * m} {4 * m + 1} {4 * m + 2} {4 * m + 3} cccc"
                               for m in range(n_mol)]
    (d / "butane.in.topology").write_text("\n".join(rows) + "\n")
    na = 6.02214076e23
    (d / "butane.in.toml").write_text(f"""schema_version = 1
init = "butane.in.xyz"
topology = "butane.in.topology"

[simulation]
seed = 2
temperature = 300.0

[[phase]]
name = "md"
n_steps = {n_steps}
dt = 1.0e-15
[phase.integrator]
kind = "velocity-verlet"
lossless = false
[phase.output]
trajectory_every = 10
log_every = 10

[[particle_types]]
name = "C"
mass = {14.027 * U.DALTON_KG!r}

[[pair_interactions]]
between = ["C", "C"]
kind = "lennard-jones"
sigma = 3.95e-10
epsilon = {0.38e3 / na!r}
cutoff = 9.0e-10

[[bond_types]]
name = "cc"
kind = "harmonic"
k = {2.5e5 * 1e3 / na / 1e-18!r}
r0 = 1.54e-10

[[angle_types]]
name = "ccc"
kind = "harmonic"
k_theta = {520e3 / na!r}
theta_0 = 1.9548

[[dihedral_types]]
name = "cccc"
kind = "periodic"
k_phi = {2.9e3 / na!r}
n = 3
phi_0 = 0.0
""")
    return d / "butane.in.toml"
