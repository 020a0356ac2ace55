"""The `.in.toml` run configuration (the subset on the nonbonded path), the
extended-XYZ init reader and a phase runner over the C ABI.

Follows the reference's loader (src/io/config.rs), init-state reader
(src/io/init_state.rs:119-494), topology `[constraints]` / `[charges]`
sections (src/forces/topology.rs) and runner output cadence
(src/runner.rs:1300-1530).  Values are read in the file's unit system
(`units = "si"` by default, or "atomic") and converted once, as
`value / factor` in f64 (src/units/mod.rs:110-125).

Supported: `[simulation]` seed / temperature; `[[phase]]` with a
velocity-verlet (optionally `lossless`) or langevin-baoab integrator and
its `[phase.output]` cadence; `[[minimization]]` steepest descent;
`[[particle_types]]`; `[[pair_interactions]]` of kind lennard-jones and a
`[lennard_jones]` Lorentz-Berthelot table; `[spme]`; `[[constraint_types]]`
of kind settle with topology `[constraints]` rows; `[neighbor_list]`;
`[[bond_types]]` / `[[angle_types]]` / `[[dihedral_types]]` with topology
`[bonds]` / `[angles]` / `[dihedrals]` / `[exclusions]` rows and the derived
exclusion list (src/forces/topology.rs:1166-1260);
`[phase.thermostat]` of kind csvr (src/integrator/csvr.rs) and
`[phase.barostat]` of kind c-rescale (src/integrator/c_rescale_barostat.rs),
with their log columns (csvr_conserved; pressure, box_volume,
c_rescale_conserved).  Everything else the reference accepts (other
thermostats and barostats, SHAKE, analysis) raises
ConfigError naming the unsupported key.
"""
import math
import os
import time

import numpy as np

try:
    import tomllib as _toml          # Python >= 3.11
except ImportError:                  # pragma: no cover - 3.10 image
    import tomli as _toml

from . import units as U

SUPPORTED_SCHEMA_VERSION = 1
# SI value of one atomic unit per dimension (src/units/mod.rs:85-108)
_SI = {"length": U.LENGTH_BOHR_TO_M, "inverse_length": 1.0 / U.LENGTH_BOHR_TO_M,
       "mass": U.MASS_ME_TO_KG, "charge": U.CHARGE_E_TO_C,
       "energy": U.ENERGY_HARTREE_TO_J, "time": U.TIME_AU_TO_S,
       "inverse_time": 1.0 / U.TIME_AU_TO_S, "temperature": U.TEMPERATURE_AU_TO_K,
       "velocity": 2187691.2636411153, "pressure": U.PRESSURE_AU_TO_PA,
       "stiffness": U.ENERGY_HARTREE_TO_J / U.LENGTH_BOHR_TO_M ** 2,
       "inverse_pressure": 1.0 / U.PRESSURE_AU_TO_PA}


class ConfigError(ValueError):
    pass


class InitStateError(ValueError):
    pass


def _from_user(units, dim, value):
    return float(value) if units == "atomic" else float(value) / _SI[dim]


def _keys(table, where, allowed):
    extra = set(table) - set(allowed)
    if extra:
        raise ConfigError(f"{where}: unsupported key(s) {sorted(extra)}")


def _require(table, key, where):
    if key not in table:
        raise ConfigError(f"{where}: missing `{key}`")
    return table[key]


def _positive(name, v):
    if not (math.isfinite(v) and v > 0.0):
        raise ConfigError(f"{name} must be finite and > 0, got {v}")
    return v


# --------------------------------------------------------------- init file --

def _attributes(line):
    """key=value and key="quoted value" pairs (init_state.rs:119-176)."""
    out, i, n = {}, 0, len(line)
    while i < n:
        while i < n and line[i].isspace():
            i += 1
        if i >= n:
            break
        k0 = i
        while i < n and line[i] != "=" and not line[i].isspace():
            i += 1
        if i >= n or line[i] != "=":
            while i < n and not line[i].isspace():
                i += 1
            continue
        key = line[k0:i]
        i += 1
        if i < n and line[i] == '"':
            j = line.find('"', i + 1)
            j = n if j < 0 else j
            out[key] = line[i + 1:j]
            i = j + 1
        else:
            v0 = i
            while i < n and not line[i].isspace():
                i += 1
            out[key] = line[v0:i]
    return out


_PROPERTIES = {"species:S:1:pos:R:3": (False, False),
               "species:S:1:pos:R:3:image:I:3": (False, True),
               "species:S:1:pos:R:3:velo:R:3": (True, False),
               "species:S:1:pos:R:3:velo:R:3:image:I:3": (True, True)}


def read_init(path, type_names, units="si"):
    """Extended-XYZ init state (init_state.rs:250-420): lattice
    {lx, ly, lz, xy, xz, yz}, f32 positions (and velocities, images when
    the Properties carry them), type indices by species name."""
    with open(path) as f:
        lines = f.read().splitlines()
    if not lines:
        raise InitStateError("init-state file is empty")
    try:
        n = int(lines[0].strip())
        if n < 0:
            raise ValueError
    except ValueError:
        raise InitStateError(f"invalid particle count on line 1: `{lines[0].strip()}`")
    if len(lines) < 2:
        raise InitStateError("init-state file is missing the comment line")
    attrs = _attributes(lines[1])
    for name in ("Lattice", "Properties"):
        if name not in attrs:
            raise InitStateError(f"comment line is missing the required `{name}` attribute")
    parts = attrs["Lattice"].split()
    if len(parts) != 9:
        raise InitStateError(f"invalid `Lattice` attribute: expected 9 components, got {len(parts)}")
    try:
        vals = [_from_user(units, "length", p) for p in parts]
    except ValueError:
        raise InitStateError("invalid `Lattice` attribute: a component is not a number")
    if not all(math.isfinite(v) for v in vals):
        raise InitStateError("invalid `Lattice` attribute: a component is not finite")
    for i in (1, 2, 5):
        if vals[i] != 0.0:
            raise InitStateError(f"invalid `Lattice` attribute: upper-triangular component {i} "
                                 f"must be 0.0, got {vals[i]}")
    lattice = np.array([vals[0], vals[4], vals[8], vals[3], vals[6], vals[7]])
    if attrs["Properties"] not in _PROPERTIES:
        raise InitStateError(f"invalid `Properties` attribute: unsupported Properties value "
                             f"{attrs['Properties']!r}")
    has_v, has_im = _PROPERTIES[attrs["Properties"]]
    ncol = 4 + 3 * has_v + 3 * has_im
    rows = [(k + 3, ln.split()) for k, ln in enumerate(lines[2:]) if ln.strip()]
    if len(rows) != n:
        raise InitStateError(f"expected {n} particle rows, found {len(rows)}")
    types = np.empty(n, np.uint32)
    pos = np.empty((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32) if has_v else None
    img = np.zeros((n, 3), np.int32) if has_im else None
    L32 = lattice.astype(np.float32)
    for r, (line_no, cols) in enumerate(rows):
        if len(cols) != ncol:
            raise InitStateError(f"line {line_no} has {len(cols)} columns, expected {ncol}")
        if cols[0] not in type_names:
            raise InitStateError(f"line {line_no}: unknown particle type `{cols[0]}`")
        types[r] = type_names.index(cols[0])
        p = [_from_user(units, "length", c) for c in cols[1:4]]
        if not all(math.isfinite(x) for x in p):
            raise InitStateError(f"line {line_no}: position is not finite")
        pos[r] = p
        # fractional coordinates of the f32 position in [-1/2, 1/2)
        sc = np.float32(pos[r, 2]) / L32[2]
        sb = (np.float32(pos[r, 1]) - sc * L32[5]) / L32[1]
        sa = (np.float32(pos[r, 0]) - sb * L32[3] - sc * L32[4]) / L32[0]
        for d, s in zip("abc", (sa, sb, sc)):
            if not (-0.5 <= float(s) < 0.5):
                raise InitStateError(f"line {line_no}: position outside the primary image "
                                     f"along {d} (fractional {float(s)})")
        if has_v:
            vel[r] = [_from_user(units, "velocity", c) for c in cols[4:7]]
        if has_im:
            img[r] = [int(c) for c in cols[ncol - 3:ncol]]
    return {"n": n, "lattice": lattice, "types": types, "pos": pos, "vel": vel, "images": img}


# ---------------------------------------------------------------- topology --

def read_topology(path, n, constraint_type_names, bonded_types=None):
    """`[constraints]` rows (`i j k TYPE`, one group each), an optional
    `[charges]` section (`atom charge`, every atom or none), `[bonds]`
    (`i j TYPE`), `[angles]` (`i j k TYPE`), `[dihedrals]` (`i j k l TYPE`)
    with types declared in the config, and `[exclusions]` (`i j s_lj
    s_coul`) (src/forces/topology.rs).  The bonded rows land in
    `bonded_types`-keyed lists returned as a fourth value when
    `bonded_types` is given."""
    groups, group_types, charges, section = [], [], {}, None
    bt = bonded_types or {"bond": {}, "angle": {}, "dihedral": {}}
    bonded = {"bonds": [], "angles": [], "dihedrals": [], "exclusions": {}}
    arity = {"bonds": (2, "bond"), "angles": (3, "angle"), "dihedrals": (4, "dihedral")}
    with open(path) as f:
        for ln_no, raw in enumerate(f, 1):
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("[") and line.endswith("]"):
                section = line[1:-1].strip()
                if section not in ("constraints", "charges", "bonds", "angles", "dihedrals",
                                   "exclusions"):
                    raise ConfigError(f"topology line {ln_no}: section `[{section}]` is not "
                                      "supported by this loader")
                continue
            if section is None:
                raise ConfigError(f"topology line {ln_no}: content appears outside any section")
            cols = line.split()
            if section == "constraints":
                if len(cols) != 4 or cols[3] not in constraint_type_names:
                    raise ConfigError(f"topology line {ln_no}: expected `i j k TYPE` with a "
                                      "declared settle constraint type")
                atoms = [int(c) for c in cols[:3]]
                if any(a < 0 or a >= n for a in atoms):
                    raise ConfigError(f"topology line {ln_no}: atom index out of range")
                groups.append(atoms)
                group_types.append(cols[3])
            elif section in arity:
                k, cls = arity[section]
                if len(cols) != k + 1 or cols[k] not in bt[cls]:
                    raise ConfigError(f"topology line {ln_no}: expected {k} atoms and a "
                                      f"declared {cls} type")
                atoms = [int(c) for c in cols[:k]]
                if any(a < 0 or a >= n for a in atoms) or len(set(atoms)) != k:
                    raise ConfigError(f"topology line {ln_no}: atom index out of range or "
                                      "repeated")
                bonded[section].append(atoms + [cols[k]])
            elif section == "exclusions":
                if len(cols) != 4:
                    raise ConfigError(f"topology line {ln_no}: expected `i j s_lj s_coul`")
                i, j = int(cols[0]), int(cols[1])
                if not (0 <= i < n and 0 <= j < n) or i == j:
                    raise ConfigError(f"topology line {ln_no}: atom index out of range")
                bonded["exclusions"][(min(i, j), max(i, j))] = (float(cols[2]), float(cols[3]))
            else:
                charges[int(cols[0])] = float(cols[1])
    flat = [a for g in groups for a in g]
    if len(set(flat)) != len(flat):
        raise ConfigError("an atom appears in more than one [constraints] row")
    if charges and len(charges) != n:
        raise ConfigError(f"[charges] section assigns charges to {len(charges)} of {n} atoms")
    if bonded_types is not None:
        return groups, group_types, charges, bonded
    return groups, group_types, charges


# ------------------------------------------------------------------ config --

_TOP = {"schema_version", "units", "init", "topology", "simulation", "phase", "minimization",
        "particle_types", "pair_interactions", "lennard_jones", "spme", "constraint_types",
        "neighbor_list", "bond_types", "angle_types", "dihedral_types"}


def load(path):
    """Parse and validate a `.in.toml` file; every value in atomic units."""
    with open(path, "rb") as f:
        raw = _toml.load(f)
    unsupported = set(raw) - _TOP
    if unsupported:
        raise ConfigError(f"unsupported top-level key(s) {sorted(unsupported)}")
    sv = _require(raw, "schema_version", "config")
    if sv != SUPPORTED_SCHEMA_VERSION:
        raise ConfigError(f"unsupported schema_version {sv} (this loader reads "
                          f"{SUPPORTED_SCHEMA_VERSION})")
    units = raw.get("units", "si")
    if units not in ("si", "atomic"):
        raise ConfigError(f"unknown `units` value `{units}`: expected one of `si`, `atomic`")
    base = os.path.dirname(os.path.abspath(path))
    cfg = {"path": os.path.abspath(path), "units": units,
           "init": os.path.join(base, _require(raw, "init", "config")),
           "topology": os.path.join(base, raw["topology"]) if "topology" in raw else None}
    sim = _require(raw, "simulation", "config")
    _keys(sim, "[simulation]", {"seed", "temperature", "graph_batch_size", "cuda_graphs_disable",
                                "fast_math"})
    cfg["seed"] = int(_require(sim, "seed", "[simulation]"))
    cfg["temperature"] = _from_user(units, "temperature", _require(sim, "temperature",
                                                                   "[simulation]"))
    cfg["graph_batch_size"] = int(sim.get("graph_batch_size", 50))
    if cfg["graph_batch_size"] < 1:
        raise ConfigError("graph_batch_size must be >= 1")
    # particle types
    types = _require(raw, "particle_types", "config")
    cfg["types"] = []
    for t in types:
        _keys(t, "[[particle_types]]", {"name", "mass", "charge", "sigma", "epsilon"})
        if ("sigma" in t) != ("epsilon" in t):
            raise ConfigError(f"particle type {t.get('name')}: sigma and epsilon go together")
        cfg["types"].append({
            "name": _require(t, "name", "[[particle_types]]"),
            "mass": _positive("mass", _from_user(units, "mass", _require(t, "mass", "type"))),
            "charge": _from_user(units, "charge", t.get("charge", 0.0)),
            "sigma": _from_user(units, "length", t["sigma"]) if "sigma" in t else None,
            "epsilon": _from_user(units, "energy", t["epsilon"]) if "epsilon" in t else None})
    names = [t["name"] for t in cfg["types"]]
    if len(set(names)) != len(names):
        raise ConfigError("duplicate particle type name")
    cfg["lj"] = _lj_table(raw, units, names, cfg["types"])
    # SPME
    if "spme" in raw:
        s = raw["spme"]
        _keys(s, "[spme]", {"alpha", "r_cut_real", "grid", "spline_order"})
        cfg["spme"] = {"alpha": _positive("spme.alpha",
                                          _from_user(units, "inverse_length", s["alpha"])),
                       "r_cut_real": _positive("spme.r_cut_real",
                                               _from_user(units, "length", s["r_cut_real"])),
                       "grid": tuple(int(g) for g in s["grid"]),
                       "order": int(s.get("spline_order", 4))}
    else:
        cfg["spme"] = None
    max_cut = max([c for c in cfg["lj"]["cutoff"]] +
                  ([cfg["spme"]["r_cut_real"]] if cfg["spme"] else []))
    # neighbour list (config.rs:1569-1577: r_skin defaults to 0.3 max_cutoff)
    nl = raw.get("neighbor_list", {})
    _keys(nl, "[neighbor_list]", {"mode", "r_skin"})
    mode = nl.get("mode", "cell-list")
    if mode not in ("cell-list", "all-pairs"):
        raise ConfigError(f"unknown neighbor_list mode `{mode}`")
    if mode == "all-pairs" and cfg["spme"]:
        raise ConfigError("[neighbor_list] mode = \"all-pairs\" is incompatible with a [spme] "
                          "table")
    cfg["r_skin"] = (_positive("neighbor_list.r_skin", _from_user(units, "length", nl["r_skin"]))
                     if "r_skin" in nl else 0.3 * max_cut)
    cfg["neighbor_mode"] = mode
    # constraint types
    cfg["constraint_types"] = {}
    for ct in raw.get("constraint_types", []):
        kind = ct.get("kind")
        if kind != "settle":
            raise ConfigError(f"constraint type `{ct.get('name')}`: kind `{kind}` is not "
                              "supported (settle only)")
        _keys(ct, "[[constraint_types]]", {"name", "kind", "d_OH", "d_HH"})
        d_oh = _positive("d_OH", _from_user(units, "length", _require(ct, "d_OH", "settle")))
        d_hh = _positive("d_HH", _from_user(units, "length", _require(ct, "d_HH", "settle")))
        if d_hh >= 2.0 * d_oh:
            raise ConfigError("settle: d_HH must be below 2 d_OH")
        cfg["constraint_types"][ct["name"]] = {"d_oh": d_oh, "d_hh": d_hh}
    cfg["bonded_types"] = _bonded_types(raw, units)
    # phases, in document order ([[minimization]] tables keep their position
    # only relative to each other here: they run before the MD phases)
    cfg["phases"] = []
    for m in raw.get("minimization", []):
        _keys(m, "[[minimization]]", {"name", "algorithm", "output"})
        alg = m.get("algorithm", {})
        if alg.get("kind", "steepest-descent") != "steepest-descent":
            raise ConfigError("minimization: only steepest-descent is supported")
        cfg["phases"].append({"kind": "minimization", "name": m["name"],
                              "max_iterations": int(alg.get("max_iterations", 1000))})
    for p in raw.get("phase", []):
        _keys(p, "[[phase]]", {"name", "n_steps", "dt", "integrator", "output", "thermostat",
                               "barostat"})
        integ = dict(_require(p, "integrator", "[[phase]]"))
        kind = integ.pop("kind", None)
        if kind == "velocity-verlet":
            _keys(integ, "[phase.integrator]", {"lossless"})
            it = {"kind": kind, "lossless": bool(integ.get("lossless", False))}
        elif kind == "langevin-baoab":
            _keys(integ, "[phase.integrator]", {"friction", "temperature", "seed"})
            it = {"kind": kind, "lossless": False,
                  "friction": _from_user(units, "inverse_time", _require(integ, "friction", kind)),
                  "temperature": _from_user(units, "temperature",
                                            _require(integ, "temperature", kind)),
                  "seed": int(_require(integ, "seed", kind))}
        else:
            raise ConfigError(f"integrator kind `{kind}` is not supported")
        th = _thermostat(p.get("thermostat"), units, kind)
        baro = _barostat(p.get("barostat"), units)
        out = p.get("output", {})
        _keys(out, "[phase.output]", {"trajectory_every", "include_velocities", "include_images",
                                      "log_every"})
        cfg["phases"].append({
            "kind": "md", "name": _require(p, "name", "[[phase]]"),
            "n_steps": int(_require(p, "n_steps", "[[phase]]")),
            "dt": _positive("dt", _from_user(units, "time", _require(p, "dt", "[[phase]]"))),
            "integrator": it, "thermostat": th, "barostat": baro,
            "trajectory_every": int(out.get("trajectory_every", 100)),
            "include_velocities": bool(out.get("include_velocities", True)),
            "include_images": bool(out.get("include_images", True)),
            "log_every": int(out.get("log_every", 100))})
    if not cfg["phases"]:
        raise ConfigError("no [[phase]] or [[minimization]] entries")
    return cfg


def _bonded_types(raw, units):
    """[[bond_types]] (harmonic k, r0; morse de, a, re), [[angle_types]]
    (harmonic k_theta, theta_0) and [[dihedral_types]] (periodic k_phi, n,
    phi_0, scale_lj_14 = 0.5, scale_coul_14 = 1/1.2) by name
    (src/io/config.rs:505-600; harmonic_bond.rs, morse.rs, angle.rs,
    dihedral.rs parameter tables).  Values: (force-field kind, parameters
    in atomic units[, 1-4 scales])."""
    out = {"bond": {}, "angle": {}, "dihedral": {}}
    spec = {("bond_types", "harmonic"): ("harmonic-bond", [("k", "stiffness"), ("r0", "length")]),
            ("bond_types", "morse"): ("morse", [("de", "energy"), ("a", "inverse_length"),
                                                ("re", "length")]),
            ("angle_types", "harmonic"): ("harmonic-angle", [("k_theta", "energy"),
                                                             ("theta_0", None)]),
            ("dihedral_types", "periodic"): ("periodic-dihedral", [("k_phi", "energy"),
                                                                   ("phi_0", None), ("n", None)])}
    for table, cls in (("bond_types", "bond"), ("angle_types", "angle"),
                       ("dihedral_types", "dihedral")):
        for t in raw.get(table, []):
            t = dict(t)
            name = _require(t, "name", f"[[{table}]]")
            key = (table, t.get("kind"))
            if key not in spec:
                raise ConfigError(f"[[{table}]] `{name}`: kind `{t.get('kind')}` is not supported")
            kind, fields = spec[key]
            allowed = {"name", "kind"} | {f for f, _ in fields}
            if cls == "dihedral":
                allowed |= {"scale_lj_14", "scale_coul_14"}
            _keys(t, f"[[{table}]] `{name}`", allowed)
            vals = []
            for f, dim in fields:
                v = _require(t, f, f"[[{table}]] `{name}`")
                vals.append(float(v) if dim is None else _from_user(units, dim, v))
            if cls == "dihedral":
                if float(vals[2]) != int(vals[2]) or vals[2] < 0:
                    raise ConfigError(f"[[dihedral_types]] `{name}`: n must be a non-negative "
                                      "integer")
                # the force field's parameter order: k_phi, phi_0, n
                out[cls][name] = (kind, vals, (float(t.get("scale_lj_14", 0.5)),
                                               float(t.get("scale_coul_14", 1.0 / 1.2))))
            else:
                out[cls][name] = (kind, vals, None)
    return out


def _effective_exclusions(n, explicit, bonds, angles, groups, dihedrals):
    """The effective exclusion list of src/forces/topology.rs:1166-1260:
    explicit rows as given; bonds and angles' 1-3 pairs at scale 0; every
    pair inside a constraint group at scale 0; then each dihedral's 1-4 pair
    (in sorted order, first wins) at its type's 1-4 scales, where no entry
    covers it yet.  Returns a symmetric CSR (offsets, partners, lj scales,
    coulomb scales)."""
    eff = {}
    for (i, j), sc in explicit.items():
        eff[(min(i, j), max(i, j))] = sc
    def add(i, j, sc):
        key = (min(i, j), max(i, j))
        if key not in eff:
            eff[key] = sc
    for i, j, _ in bonds:
        add(i, j, (0.0, 0.0))
    for i, _, k, _ in angles:
        add(i, k, (0.0, 0.0))
    for g in groups:
        for a in range(len(g)):
            for b in range(a + 1, len(g)):
                add(g[a], g[b], (0.0, 0.0))
    for i, _, _, l, sc in sorted(dihedrals, key=lambda d: (min(d[0], d[3]), max(d[0], d[3]))):
        add(i, l, sc)
    rows = [[] for _ in range(n)]
    for (i, j), (slj, sco) in sorted(eff.items()):
        rows[i].append((j, slj, sco))
        rows[j].append((i, slj, sco))
    off = np.zeros(n + 1, np.uint32)
    off[1:] = np.cumsum([len(r) for r in rows])
    par = np.array([p for r in rows for p, _, _ in sorted(r)], np.uint32)
    slj = np.array([s for r in rows for _, s, _ in sorted(r)], np.float32)
    sco = np.array([s for r in rows for _, _, s in sorted(r)], np.float32)
    return off, par, slj, sco


def _thermostat(t, units, integrator_kind):
    """`[phase.thermostat]`: CSVR (csvr.rs:21-30 params, :282-300 checks),
    coupling every step."""
    if t is None:
        return None
    t = dict(t)
    kind = t.get("kind")
    if kind != "csvr":
        raise ConfigError(f"thermostat kind `{kind}` is not supported (csvr only)")
    if integrator_kind != "velocity-verlet":
        raise ConfigError(f"integrator `{integrator_kind}` owns its own thermostat and is "
                          "incompatible with `[phase.thermostat]`")
    _keys(t, "[phase.thermostat]", {"kind", "temperature", "tau", "seed", "coupling_interval"})
    if int(t.get("coupling_interval", 1)) != 1:
        raise ConfigError("thermostat.coupling_interval: only 1 (every step) is supported")
    return {"kind": kind,
            "temperature": _positive("thermostat.temperature", _from_user(
                units, "temperature", _require(t, "temperature", "[phase.thermostat]"))),
            "tau": _positive("thermostat.tau", _from_user(
                units, "time", _require(t, "tau", "[phase.thermostat]"))),
            "seed": int(_require(t, "seed", "[phase.thermostat]"))}


def _barostat(b, units):
    """`[phase.barostat]`: c-rescale (CRescaleBarostatParams,
    c_rescale_barostat.rs:21-28; validate_params :282-290)."""
    if b is None:
        return None
    b = dict(b)
    kind = b.get("kind")
    if kind != "c-rescale":
        raise ConfigError(f"barostat kind `{kind}` is not supported (c-rescale only)")
    _keys(b, "[phase.barostat]", {"kind", "pressure", "temperature", "tau", "compressibility",
                                  "seed"})
    where = "[phase.barostat]"
    pressure = _from_user(units, "pressure", _require(b, "pressure", where))
    if not math.isfinite(pressure):
        raise ConfigError(f"barostat.pressure must be finite, got {pressure}")
    return {"kind": kind, "pressure": pressure,
            "temperature": _positive("barostat.temperature", _from_user(
                units, "temperature", _require(b, "temperature", where))),
            "tau": _positive("barostat.tau", _from_user(units, "time",
                                                        _require(b, "tau", where))),
            "compressibility": _positive("barostat.compressibility", _from_user(
                units, "inverse_pressure", _require(b, "compressibility", where))),
            "seed": int(_require(b, "seed", where))}


def _lj_table(raw, units, names, types):
    """Per type-pair LJ parameters: explicit [[pair_interactions]] entries,
    then Lorentz-Berthelot combining from [lennard_jones] for the rest
    (r_switch defaults to 0.9 cutoff, src/forces/lj.rs:67)."""
    nt = len(names)
    sig = np.zeros(nt * nt)
    eps = np.zeros(nt * nt)
    cut = np.zeros(nt * nt)
    sw = np.zeros(nt * nt)
    have = np.zeros(nt * nt, bool)
    for e in raw.get("pair_interactions", []):
        if e.get("kind") != "lennard-jones":
            raise ConfigError(f"pair interaction kind `{e.get('kind')}` is not supported")
        _keys(e, "[[pair_interactions]]", {"between", "kind", "sigma", "epsilon", "cutoff",
                                           "r_switch"})
        a, b = e["between"]
        if a not in names or b not in names:
            raise ConfigError(f"pair interaction between unknown types {a!r}, {b!r}")
        i, j = names.index(a), names.index(b)
        c = _positive("cutoff", _from_user(units, "length", e["cutoff"]))
        rs = _from_user(units, "length", e["r_switch"]) if "r_switch" in e else 0.9 * c
        if have[i * nt + j]:
            raise ConfigError(f"duplicate pair interaction {a}-{b}")
        for k in (i * nt + j, j * nt + i):
            sig[k] = _from_user(units, "length", e["sigma"])
            eps[k] = _from_user(units, "energy", e["epsilon"])
            cut[k], sw[k], have[k] = c, rs, True
    lj = raw.get("lennard_jones")
    if lj is not None:
        _keys(lj, "[lennard_jones]", {"combining_rule", "cutoff", "r_switch"})
        if lj.get("combining_rule", "lorentz-berthelot") != "lorentz-berthelot":
            raise ConfigError("only the lorentz-berthelot combining rule is supported")
        c = _positive("lennard_jones.cutoff", _from_user(units, "length", lj["cutoff"]))
        rs = _from_user(units, "length", lj["r_switch"]) if "r_switch" in lj else 0.9 * c
        for i in range(nt):
            for j in range(nt):
                k = i * nt + j
                if have[k]:
                    continue
                ti, tj = types[i], types[j]
                if ti["sigma"] is None or tj["sigma"] is None:
                    raise ConfigError(f"pair {names[i]}-{names[j]}: no interaction declared and "
                                      "a type has no sigma/epsilon to combine")
                sig[k] = 0.5 * (ti["sigma"] + tj["sigma"])
                eps[k] = math.sqrt(ti["epsilon"] * tj["epsilon"])
                cut[k], sw[k], have[k] = c, rs, True
    if not have.all():
        missing = [f"{names[k // nt]}-{names[k % nt]}" for k in range(nt * nt) if not have[k]]
        raise ConfigError(f"every unordered type pair needs an interaction; missing {missing}")
    return {"sigma": sig, "epsilon": eps, "cutoff": cut, "switch": sw}


def build(cfg):
    """Arrays and descriptors for heddlemd_amd.api from a loaded config."""
    names = [t["name"] for t in cfg["types"]]
    init = read_init(cfg["init"], names, cfg["units"])
    n = init["n"]
    q = np.array([cfg["types"][t]["charge"] for t in init["types"]], np.float64)
    groups, gtypes = [], []
    bonded = {"bonds": [], "angles": [], "dihedrals": [], "exclusions": {}}
    if cfg["topology"]:
        groups, gtypes, charges, bonded = read_topology(cfg["topology"], n,
                                                        list(cfg["constraint_types"]),
                                                        cfg["bonded_types"])
        if charges:
            q = np.array([charges[i] for i in range(n)], np.float64)
    mass = np.array([cfg["types"][t]["mass"] for t in init["types"]], np.float64)
    posq = np.zeros((n, 4), np.float32)
    posq[:, :3] = init["pos"]
    posq[:, 3] = q.astype(np.float32)
    ff = {"n_types": len(names), "lj": cfg["lj"],
          "coulomb": None, "spme": None, "r_skin": cfg["r_skin"], "exclusions": None}
    if cfg["spme"]:
        sp = cfg["spme"]
        ff["coulomb"] = {"alpha": sp["alpha"], "r_cut": sp["r_cut_real"], "k_coulomb": 1.0}
        ff["spme"] = {"grid": sp["grid"], "order": sp["order"]}
    settle = None
    if groups:
        if len(set(gtypes)) != 1:
            raise ConfigError("one settle constraint type per run in this loader")
        geo = cfg["constraint_types"][gtypes[0]]
        g = np.array(groups, np.uint32)
        settle = {"groups": g, "d_oh": geo["d_oh"], "d_hh": geo["d_hh"],
                  "m_o": float(mass[g[0, 0]]), "m_h": float(mass[g[0, 1]])}
    # bonded sets (one per force-field kind, types numbered in declaration
    # order) and the effective exclusion list
    bt = cfg["bonded_types"]
    sets = []
    for section, cls in (("bonds", "bond"), ("angles", "angle"), ("dihedrals", "dihedral")):
        by_kind = {}
        for row in bonded[section]:
            kind = bt[cls][row[-1]][0]
            by_kind.setdefault(kind, []).append(row)
        for kind, rows in by_kind.items():
            names_k = [nm for nm, v in bt[cls].items() if v[0] == kind]
            params = [bt[cls][nm][1] for nm in names_k]
            terms = np.array([r[:-1] + [names_k.index(r[-1])] for r in rows], np.uint32)
            sets.append((kind, terms, params))
    dihed = [r[:4] + [bt["dihedral"][r[4]][2]] for r in bonded["dihedrals"]]
    if groups or bonded["exclusions"] or bonded["bonds"] or bonded["angles"] or dihed:
        ff["exclusions"] = _effective_exclusions(
            n, bonded["exclusions"], bonded["bonds"], bonded["angles"], groups, dihed)
    n_constraints = 3 * len(groups)
    return {"n": n, "lattice": init["lattice"], "posq": posq, "mass": mass,
            "types": init["types"], "vel": init["vel"], "images": init["images"], "ff": ff,
            "settle": settle, "n_constraints": n_constraints, "type_names": names,
            "bonded": sets}


def host_kinetic_energy(mass, vel):
    """0.5 sum m |v|^2 summed in index order in f64 (log_output.rs:121-138)."""
    m = mass.astype(np.float32).astype(np.float64)
    v = vel.astype(np.float64)
    terms = m * (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    return 0.5 * float(np.add.accumulate(terms)[-1]) if len(terms) else 0.0


def run(path, device=0):
    """Run every phase of a config: outputs `<stem>.out.<phase>.xyz`, `.log`
    and `.timings` beside the config (create-new, as the reference)."""
    from . import api
    t_all = time.perf_counter()
    t0 = time.perf_counter()
    cfg = load(path)
    t_cfg = time.perf_counter() - t0
    t0 = time.perf_counter()
    sysd = build(cfg)
    t_init = time.perf_counter() - t0
    stem = os.path.basename(path)
    stem = stem[:-len(".in.toml")] if stem.endswith(".in.toml") else os.path.splitext(stem)[0]
    out_base = os.path.join(os.path.dirname(os.path.abspath(path)), stem + ".out.")
    t0 = time.perf_counter()
    ctx = api.GpuContext(device)
    t_gpu = time.perf_counter() - t0
    n = sysd["n"]
    vel = sysd["vel"]
    t_vel = 0.0
    if vel is None:
        t0 = time.perf_counter()
        vel = api.generate_velocities(sysd["mass"], cfg["temperature"], cfg["seed"],
                                      n_constraints=sysd["n_constraints"])
        t_vel = time.perf_counter() - t0
    part = api.ParticleBuffers(ctx, sysd["posq"], sysd["mass"], sysd["lattice"], vel=vel,
                               types=sysd["types"])
    if sysd["images"] is not None:
        part.upload_images(sysd["images"])
    field = api.ForceField(ctx, part, sysd["ff"])
    for kind, terms, params in sysd["bonded"]:
        field.add_bonded(kind, terms, params)
    dof = max(0, 3 * n - sysd["n_constraints"] - 3)
    summary = []
    try:
        for k, ph in enumerate(cfg["phases"]):
            pre = {"config_load": t_cfg, "init_load": t_init, "gpu_init": t_gpu,
                   "velocity_generation": t_vel} if k == 0 else {}
            summary.append(_run_phase(api, ph, cfg, sysd, ctx, part, field, dof, out_base, pre))
    finally:
        field.close(); part.close(); ctx.close()
    return {"phases": summary, "seconds": time.perf_counter() - t_all}


def _run_phase(api, ph, cfg, sysd, ctx, part, field, dof, out_base, pre):
    host = {"device_to_host_download": [], "trajectory_write": [], "log_write": []}
    t_phase = time.perf_counter()
    if ph["kind"] == "minimization":
        integ = api.Integrator(ctx, part, field, dt=1.0, settle=sysd["settle"])
        integ.minimize(ph["max_iterations"], 0.1 * U.ANGSTROM)
        integ.close()
        return {"name": ph["name"], "kind": "minimization"}
    it = ph["integrator"]
    th, baro = ph.get("thermostat"), ph.get("barostat")
    if th is not None:
        integ = api.Integrator(ctx, part, field, kind=it["kind"], dt=ph["dt"],
                               compensated=it["lossless"], temperature=th["temperature"],
                               seed=th["seed"], thermostat="csvr", tau=th["tau"],
                               settle=sysd["settle"], scalars=True)
    else:
        integ = api.Integrator(ctx, part, field, kind=it["kind"], dt=ph["dt"],
                               compensated=it["lossless"],
                               temperature=it.get("temperature", 0.0),
                               gamma=it.get("friction", 0.0), seed=it.get("seed", 0),
                               settle=sysd["settle"], scalars=baro is not None)
    if baro is not None:
        integ.set_barostat("c-rescale", pressure=baro["pressure"],
                           temperature=baro["temperature"], tau=baro["tau"],
                           compressibility=baro["compressibility"], seed=baro["seed"])
    # extra log columns in the reference's order: integrator (none),
    # thermostat, barostat (runner.rs:1426-1433)
    extra = ([("csvr_conserved", "energy")] if th is not None else []) + (
        [("pressure", "pressure"), ("box_volume", "dimensionless"),
         ("c_rescale_conserved", "energy")] if baro is not None else [])
    traj = log = None
    if ph["trajectory_every"] > 0:
        traj = api.TrajectoryWriter(out_base + ph["name"] + ".xyz", cfg["units"],
                                    ph["include_velocities"], ph["include_images"],
                                    sysd["type_names"])
    if ph["log_every"] > 0:
        log = api.LogWriter(out_base + ph["name"] + ".log", cfg["units"], extra_columns=extra)

    def outputs(step):
        want_t = traj is not None and step % ph["trajectory_every"] == 0
        want_l = log is not None and step % ph["log_every"] == 0
        if not (want_t or want_l):
            return
        t0 = time.perf_counter()
        d = part.download()
        host["device_to_host_download"].append(time.perf_counter() - t0)
        if want_t:
            t0 = time.perf_counter()
            traj.write_frame(step, ph["dt"], part.lattice, sysd["types"], d["posq"][:, :3],
                             velocities=d["vel"] if ph["include_velocities"] else None,
                             images=d["images"] if ph["include_images"] else None)
            host["trajectory_write"].append(time.perf_counter() - t0)
        if want_l:
            t0 = time.perf_counter()
            ke = host_kinetic_energy(sysd["mass"], d["vel"])
            vals = []
            if extra:
                pe, _ = part.energy_sums()
                if th is not None:      # csvr.rs:242-248
                    vals.append(ke + pe - integ.thermostat_state()[0])
                if baro is not None:    # c_rescale_barostat.rs:240-252
                    p_last, v_last, inj, _, _ = integ.barostat_state()
                    vals += [p_last, v_last, ke + pe + baro["pressure"] * v_last - inj]
            log.write_row(step, step * ph["dt"], ke, 2.0 * ke / dof if dof else 0.0, vals)
            host["log_write"].append(time.perf_counter() - t0)

    try:
        field.compute(1)                       # warm-up: forces at the current positions
        outputs(0)
        if baro is None:               # barostat steps run eagerly
            integ.capture()
        step = 0
        cad = [c for c in (ph["trajectory_every"], ph["log_every"]) if c > 0]
        while step < ph["n_steps"]:
            nxt = min([ph["n_steps"], step + cfg["graph_batch_size"]] +
                      [(step // c + 1) * c for c in cad])
            integ.replay(nxt - step)
            step = nxt
            outputs(step)
    finally:
        for w in (traj, log):
            if w is not None:
                w.close()
        integ.close()
    stages = []
    for name in ("host_to_device_upload", "device_to_host_download", "neighbor_list_rebuild",
                 "trajectory_write", "log_write", "velocity_generation", "config_load",
                 "init_load", "gpu_init", "total_runtime"):
        if name in host:
            ts = host[name]
        elif name in pre and pre[name] > 0.0:
            ts = [pre[name]]
        elif name == "total_runtime":
            ts = [time.perf_counter() - t_phase]
        else:
            ts = []
        if ts:
            ns = [int(round(t * 1e9)) for t in ts]
            stages.append({"name": name, "count": len(ns), "total_ns": sum(ns),
                           "min_ns": min(ns), "max_ns": max(ns)})
    api.write_timings(out_base + ph["name"] + ".timings", stages)
    return {"name": ph["name"], "kind": "md", "steps": ph["n_steps"]}
