"""End-to-end runs of `.in.toml` configurations through heddlemd_amd.config:
an argon-like LJ box (f32 and lossless VV) and a rigid SPC water box
(minimization, then Langevin BAOAB + SETTLE + SPME).  Each writes
`<stem>.out.<phase>.xyz`, `.log` and `.timings`; two runs in separate
directories must give byte-identical trajectories and logs, with the frame
and row counts the cadence implies (step 0 included)."""
import pytest

from heddlemd_amd import config as C
from tests import config_cases as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["argon", "argon_lossless", "water"])
def test_config_run_outputs_reproducible(tmp_path, case):
    outs = []
    for tag in ("a", "b"):
        d = tmp_path / tag
        d.mkdir()
        if case == "water":
            p, stem, phase, n_atoms, rows, frames = K.water_run(d), "water", "md", 1536, 3, 3
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
        outs.append((xyz, log))
    print(f"{case}: last log row {outs[0][1].decode().splitlines()[-1]}")
    assert outs[0] == outs[1]
