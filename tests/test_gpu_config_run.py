"""End-to-end runs of `.in.toml` configurations through heddlemd_amd.config:
an argon-like LJ box (f32 and lossless VV) and a rigid SPC water box
(minimization, then Langevin BAOAB + SETTLE + SPME, or velocity Verlet with
the CSVR thermostat and the c-rescale barostat).  Each writes
`<stem>.out.<phase>.xyz`, `.log` and `.timings`; two runs in separate
directories must give byte-identical trajectories and logs, with the frame
and row counts the cadence implies (step 0 included)."""
import pytest

from heddlemd_amd import config as C
from tests import config_cases as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["argon", "argon_lossless", "water", "water_npt", "butane"])
def test_config_run_outputs_reproducible(tmp_path, case):
    outs = []
    for tag in ("a", "b"):
        d = tmp_path / tag
        d.mkdir()
        if case == "butane":
            p, stem, phase, n_atoms, rows, frames = K.butane_run(d), "butane", "md", 256, 3, 3
        elif case.startswith("water"):
            p = K.water_run(d, kind="npt" if case == "water_npt" else "langevin-baoab")
            stem, phase, n_atoms, rows, frames = "water", "md", 1536, 3, 3
        else:
            p = K.argon_run(d, lossless=(case == "argon_lossless"))
            stem, phase, n_atoms, rows, frames = "argon", "run", 512, 5, 3
        print(f"config run {case}/{tag}", flush=True)
        C.run(p)
        xyz = (d / f"{stem}.out.{phase}.xyz").read_bytes()
        log = (d / f"{stem}.out.{phase}.log").read_bytes()
        tim = (d / f"{stem}.out.{phase}.timings").read_text()
        assert len(xyz.splitlines()) == frames * (n_atoms + 2)
        assert len(log.splitlines()) == 1 + rows
        assert tim.splitlines()[0].split()[:2] == ["stage", "count"]
        if case == "water_npt":
            head = log.decode().splitlines()[0]
            for col in ("csvr_conserved", "pressure", "box_volume", "c_rescale_conserved"):
                assert col in head
            lat = [x.split('Lattice="')[1].split('"')[0] for x in xyz.decode().splitlines()
                   if x.startswith("Lattice=") or 'Lattice="' in x]
            assert lat[0] != lat[-1]          # the barostat moved the box
        outs.append((xyz, log))
    print(f"{case}: last log row {outs[0][1].decode().splitlines()[-1]}")
    assert outs[0] == outs[1]


def test_bonded_config_forces_match_the_oracle(tmp_path, gpu_ctx):
    """The butane configuration's force field (LJ with the derived scaled
    1-4 exclusions, bonds, angles, dihedrals) against the FP64 oracle at the
    init positions: forces 1e-5 relative RMS, energy 1e-5 relative (the
    bonded terms' energies are f32 arithmetic, as the reference's functors).
    Measured: forces 3.0e-6, energy 1.4e-6."""
    import numpy as np
    import oracle
    from heddlemd_amd import api
    from tests.gpu_helpers import rel_rms
    sysd = C.build(C.load(K.butane_run(tmp_path)))
    part = api.ParticleBuffers(gpu_ctx, sysd["posq"], sysd["mass"], sysd["lattice"],
                               types=sysd["types"])
    field = api.ForceField(gpu_ctx, part, sysd["ff"])
    try:
        for kind, t, q in sysd["bonded"]:
            field.add_bonded(kind, t, q)
        field.compute(1)
        f = part.forces()
        pe, _ = part.energy_sums()
        posq = part.download()["posq"]
    finally:
        field.close()
        part.close()
    ex = sysd["ff"]["exclusions"]
    x = posq[:, :3].astype(np.float64)
    ps = oracle.PairSystem(x, posq[:, 3].astype(np.float64), sysd["types"].astype(np.int32),
                           sysd["lattice"], lj=sysd["ff"]["lj"],
                           exclusions=(ex[0].astype(np.int32), ex[1].astype(np.int32),
                                       ex[2].astype(np.float64), ex[3].astype(np.float64)))
    f_o, e_o, _ = ps.forces(cells=True)
    f_o = np.array(f_o, np.float64)
    u_o = float(np.sum(e_o))
    for kind, t, q in sysd["bonded"]:
        fb, _, _, ub = oracle.bonded(kind, x, np.asarray(sysd["lattice"], np.float64), t, q)
        f_o += fb
        u_o += ub
    err = rel_rms(f, f_o)
    print(f"butane: forces rel RMS {err:.2e}, PE {pe:.9e} vs {u_o:.9e}")
    assert err < 1e-5
    assert pe == pytest.approx(u_o, rel=1e-5)
