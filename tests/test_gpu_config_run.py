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


@pytest.mark.parametrize("case", ["argon", "argon_lossless", "water", "water_npt"])
def test_config_run_outputs_reproducible(tmp_path, case):
    outs = []
    for tag in ("a", "b"):
        d = tmp_path / tag
        d.mkdir()
        if case.startswith("water"):
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
