"""The `.in.toml` loader subset and the extended-XYZ init reader (CPU):
unit conversion (value / SI factor, src/units/mod.rs:110-125), the
reference's defaults (r_switch = 0.9 cutoff, src/forces/lj.rs:67;
r_skin = 0.3 max cutoff, src/io/config.rs:1569-1577; spline order 4;
trajectory / log cadence 100), Lorentz-Berthelot combining, implicit
exclusions of constraint groups (src/forces/topology.rs:1171-1177), the
init reader's checks (src/io/init_state.rs) and the errors for what the
subset does not cover."""
import numpy as np
import pytest

from heddlemd_amd import config as C
from heddlemd_amd import units as U
from tests import config_cases as K


def test_argon_config_values(tmp_path):
    cfg = C.load(K.argon_run(tmp_path))
    assert cfg["units"] == "si" and cfg["seed"] == 1
    assert cfg["temperature"] == 100.0 / U.TEMPERATURE_AU_TO_K
    lj = cfg["lj"]
    assert lj["sigma"][0] == 3.40e-10 / U.LENGTH_BOHR_TO_M
    assert lj["epsilon"][0] == 1.65e-21 / U.ENERGY_HARTREE_TO_J
    assert lj["cutoff"][0] == lj["switch"][0] == 8.5e-10 / U.LENGTH_BOHR_TO_M
    assert cfg["r_skin"] == 0.3 * lj["cutoff"][0]
    ph = cfg["phases"][0]
    assert ph["dt"] == 2.0e-15 / U.TIME_AU_TO_S and ph["n_steps"] == 20
    assert (ph["trajectory_every"], ph["log_every"], ph["include_images"]) == (10, 5, True)
    assert ph["integrator"] == {"kind": "velocity-verlet", "lossless": False}


def test_defaults_and_combining(tmp_path):
    p = K.argon_run(tmp_path)
    text = p.read_text().replace("r_switch = 8.5e-10\n", "")
    text = text.replace("""[phase.output]
trajectory_every = 10
include_velocities = true
log_every = 5
""", "")
    p.write_text(text)
    cfg = C.load(p)
    c = cfg["lj"]["cutoff"][0]
    assert cfg["lj"]["switch"][0] == 0.9 * c
    assert (cfg["phases"][0]["trajectory_every"], cfg["phases"][0]["log_every"]) == (100, 100)
    # Lorentz-Berthelot for pairs without an explicit entry
    p2 = tmp_path / "lb.in.toml"
    p2.write_text("""schema_version = 1
init = "x.xyz"
[simulation]
seed = 1
temperature = 1.0
[[phase]]
name = "a"
n_steps = 1
dt = 1.0e-15
[phase.integrator]
kind = "velocity-verlet"
[[particle_types]]
name = "A"
mass = 1.0e-26
sigma = 3.0e-10
epsilon = 1.0e-21
[[particle_types]]
name = "B"
mass = 2.0e-26
sigma = 4.0e-10
epsilon = 4.0e-21
[lennard_jones]
combining_rule = "lorentz-berthelot"
cutoff = 1.0e-9
""")
    cfg = C.load(p2)
    s = cfg["lj"]["sigma"].reshape(2, 2)
    e = cfg["lj"]["epsilon"].reshape(2, 2)
    sa, sb = 3.0e-10 / U.LENGTH_BOHR_TO_M, 4.0e-10 / U.LENGTH_BOHR_TO_M
    ea, eb = 1.0e-21 / U.ENERGY_HARTREE_TO_J, 4.0e-21 / U.ENERGY_HARTREE_TO_J
    assert s[0, 1] == s[1, 0] == 0.5 * (sa + sb)
    assert e[0, 1] == e[1, 0] == np.sqrt(ea * eb)
    assert s[0, 0] == sa and e[1, 1] == eb


def test_water_build_settle_and_exclusions(tmp_path):
    cfg = C.load(K.water_run(tmp_path))
    assert cfg["spme"]["order"] == 4 and cfg["spme"]["grid"] == (24, 24, 24)
    assert cfg["spme"]["alpha"] == 3.47e9 * U.LENGTH_BOHR_TO_M
    assert [p["kind"] for p in cfg["phases"]] == ["minimization", "md"]
    sysd = C.build(cfg)
    n = sysd["n"]
    assert n == 3 * 512 and sysd["n_constraints"] == 3 * 512
    g = sysd["settle"]["groups"]
    assert g.shape == (512, 3) and g[5].tolist() == [15, 16, 17]
    off, par, slj, sco = sysd["ff"]["exclusions"]
    assert off[-1] == 2 * n and par[off[16]:off[17]].tolist() == [15, 17]
    assert not slj.any() and not sco.any()
    assert sysd["posq"][1, 3] == np.float32(0.41 * U.CHARGE_E_TO_C / U.CHARGE_E_TO_C)


def test_init_reader_units_and_checks(tmp_path):
    p = tmp_path / "a.xyz"
    K.write_xyz(p, (1e-9, 2e-9, 3e-9), ["A", "B"], [[1e-10, -2e-10, 3e-10], [0, 0, -1.4e-9]],
                vel=[[100.0, 0, 0], [0, -50.0, 0]])
    st = C.read_init(p, ["A", "B"])
    assert st["lattice"].tolist() == [1e-9 / U.LENGTH_BOHR_TO_M, 2e-9 / U.LENGTH_BOHR_TO_M,
                                      3e-9 / U.LENGTH_BOHR_TO_M, 0.0, 0.0, 0.0]
    assert st["types"].tolist() == [0, 1]
    assert st["pos"][0, 1] == np.float32(-2e-10 / U.LENGTH_BOHR_TO_M)
    assert st["vel"][1, 1] == np.float32(-50.0 / 2187691.2636411153)
    assert st["images"] is None
    bad = [("x\n", "invalid particle count"), ("3\n", "missing the comment line"),
           ("1\nProperties=species:S:1:pos:R:3\nA 0 0 0\n", "Lattice"),
           ('1\nLattice="1 0 0 0 1 0 0 0 1" Properties=species:S:1:pos:R:3:foo:R:1\nA 0 0 0\n',
            "Properties"),
           ('1\nLattice="1 1 0 0 1 0 0 0 1" Properties=species:S:1:pos:R:3\nA 0 0 0\n',
            "upper-triangular"),
           ('2\nLattice="1 0 0 0 1 0 0 0 1" Properties=species:S:1:pos:R:3\nA 0 0 0\n',
            "expected 2 particle rows"),
           ('1\nLattice="1 0 0 0 1 0 0 0 1" Properties=species:S:1:pos:R:3\nC 0 0 0\n',
            "unknown particle type"),
           ('1\nLattice="1 0 0 0 1 0 0 0 1" Properties=species:S:1:pos:R:3\nA 0.5 0 0\n',
            "outside the primary image")]
    for text, msg in bad:
        p.write_text(text)
        with pytest.raises(C.InitStateError, match=msg):
            C.read_init(p, ["A", "B"], units="atomic")


@pytest.mark.parametrize("patch,msg", [
    ("schema_version = 1", "schema_version = 2"),
    ('kind = "velocity-verlet"', 'kind = "nose-hoover-chain"'),
    ("[[particle_types]]", "[phase.thermostat]\nkind = \"csvr\"\n\n[[particle_types]]"),
    ("[[particle_types]]", "[[bond_types]]\nname = \"x\"\n\n[[particle_types]]"),
    ('kind = "lennard-jones"', 'kind = "morse"'),
])
def test_unsupported_or_invalid_is_an_error(tmp_path, patch, msg):
    p = K.argon_run(tmp_path)
    p.write_text(p.read_text().replace(patch, msg, 1))
    with pytest.raises(C.ConfigError):
        C.load(p)


def test_thermostat_and_barostat_tables(tmp_path):
    """[phase.thermostat] csvr and [phase.barostat] c-rescale in SI, converted
    once as value / factor (the values of the reference's
    examples/spc-water-8192 production phase); CSVR on a Langevin integrator
    and a non-positive compressibility are errors."""
    p = K.water_run(tmp_path, kind="npt")
    cfg = C.load(p)
    ph = [x for x in cfg["phases"] if x["kind"] == "md"][0]
    assert ph["thermostat"] == {"kind": "csvr", "temperature": 298.15 / U.TEMPERATURE_AU_TO_K,
                                "tau": 1.0e-13 / U.TIME_AU_TO_S, "seed": 12}
    assert ph["barostat"] == {"kind": "c-rescale", "pressure": 1.01325e5 / U.PRESSURE_AU_TO_PA,
                              "temperature": 298.15 / U.TEMPERATURE_AU_TO_K,
                              "tau": 1.0e-12 / U.TIME_AU_TO_S,
                              "compressibility": 4.5e-10 / (1.0 / U.PRESSURE_AU_TO_PA),
                              "seed": 13}
    text = p.read_text()
    p.write_text(text.replace('kind = "velocity-verlet"\nlossless = false',
                              'kind = "langevin-baoab"\nfriction = 1.0e12\n'
                              'temperature = 300.0\nseed = 7', 1))
    with pytest.raises(C.ConfigError, match="owns its own thermostat"):
        C.load(p)
    p.write_text(text.replace("compressibility = 4.5e-10", "compressibility = 0.0", 1))
    with pytest.raises(C.ConfigError, match="compressibility"):
        C.load(p)


def test_host_kinetic_energy_order():
    """Sequential f64 sum of m |v|^2 in index order, halved
    (src/io/log_output.rs:121-138)."""
    rng = np.random.default_rng(1)
    v = rng.normal(size=(1000, 3)).astype(np.float32)
    m = rng.uniform(1, 5, 1000)
    s = 0.0
    for i in range(1000):
        a, b, c = (float(x) for x in v[i])
        s += float(np.float32(m[i])) * (a * a + b * b + c * c)
    assert C.host_kinetic_energy(m, v) == 0.5 * s


def test_bonded_topology_and_effective_exclusions(tmp_path):
    """Bonded types and topology rows (src/io/config.rs:505-600,
    src/forces/topology.rs): one set per kind with its types in declaration
    order, parameters converted once; the effective exclusions of
    topology.rs:1166-1260: bonds and the angles' 1-3 pairs at scale 0, the
    dihedral's 1-4 pair at its type's default AMBER scales (0.5, 1/1.2)."""
    cfg = C.load(K.butane_run(tmp_path))
    sysd = C.build(cfg)
    kinds = {k: (t, q) for k, t, q in sysd["bonded"]}
    assert sorted(kinds) == ["harmonic-angle", "harmonic-bond", "periodic-dihedral"]
    t, q = kinds["harmonic-bond"]
    assert t.shape == (3 * 64, 3) and t[1].tolist() == [1, 2, 0]
    k_au = 2.5e5 * 1e3 / 6.02214076e23 / 1e-18 / (U.ENERGY_HARTREE_TO_J / U.LENGTH_BOHR_TO_M ** 2)
    assert q[0][0] == pytest.approx(k_au, rel=1e-14) and q[0][1] == pytest.approx(1.54e-10 / U.LENGTH_BOHR_TO_M)
    t, q = kinds["periodic-dihedral"]
    assert t[0].tolist() == [0, 1, 2, 3, 0] and q[0][1:] == [0.0, 3.0]
    off, par, slj, sco = sysd["ff"]["exclusions"]
    assert par[off[0]:off[1]].tolist() == [1, 2, 3]
    assert slj[off[0]:off[1]].tolist() == [0.0, 0.0, 0.5]
    assert sco[off[0]:off[1]] == pytest.approx([0.0, 0.0, 1.0 / 1.2])
    assert par[off[5]:off[6]].tolist() == [4, 6, 7]
