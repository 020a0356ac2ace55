"""Box changes (hmd_particles_set_box, a barostat's rescale): positions and
box scaled together, then the force field's pair set, forces and energy at
the new box against the oracle (pair set bit-exact, relative RMS force
<= 1e-5, relative energy <= 1e-6), and a graph-replayed integrator across a
box change (re-captured, then forces still match).  A box the force field
cannot use (r_c not below half a width) is refused and the old box kept."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api
from heddlemd_amd._lib import HmdError
from tests.gpu_helpers import oracle_pair, rel_rms, water_case

pytestmark = pytest.mark.gpu


def check_against_oracle(sysw, ff, part, field, lattice):
    field.compute(1)
    d = part.download()
    posq = d["posq"]
    got = field.pair_snapshot()
    rs = np.float32(np.float32(ff["coulomb"]["r_cut"]) + np.float32(ff["r_skin"]))
    want = oracle.pair_set_f32(posq, np.asarray(lattice, np.float32), rs * rs)
    sysl = dict(sysw, lattice=np.asarray(lattice, np.float64))
    f_pair, e_pair, _ = oracle_pair(sysl, ff, posq)
    _, f_rec, e_rec, _ = oracle.spme_recip(posq[:, :3], posq[:, 3], sysl["lattice"],
                                           ff["spme"]["grid"], ff["spme"]["order"],
                                           ff["coulomb"]["alpha"])
    pe, _ = part.energy_sums()
    pe_ref = e_pair.sum() + e_rec.sum()
    return (got.shape == want.shape and np.array_equal(got, want),
            rel_rms(d["force"], f_pair + f_rec), abs(pe - pe_ref) / abs(pe_ref))


def test_box_rescale_eager_and_graph(gpu_ctx):
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=30)
    try:
        lat0 = np.asarray(sysw["lattice"], np.float64)
        for scale in (1.02, 0.985):
            d = part.download()
            posq = d["posq"].copy()
            posq[:, :3] *= np.float32(scale)
            lat = (lat0 * scale).astype(np.float32)
            part.upload(posq=posq)
            part.set_box(lat)
            same, err_f, err_e = check_against_oracle(sysw, ff, part, field, lat)
            print(f"box x{scale}: pair set equal {same}, rel RMS force {err_f:.2e}, "
                  f"rel PE {err_e:.2e}", flush=True)
            assert same and err_f <= 1e-5 and err_e <= 1e-6
            integ.replay(5)                       # re-captured for the new box
            same, err_f, err_e = check_against_oracle(sysw, ff, part, field, lat)
            print(f"  after 5 replayed steps: rel RMS force {err_f:.2e}", flush=True)
            assert err_f <= 1e-5
            lat0 = lat.astype(np.float64)
        with pytest.raises(HmdError):
            part.set_box(np.array([1.0, 1.0, 1.0, 0, 0, 0], np.float32))
        same, err_f, _ = check_against_oracle(sysw, ff, part, field, part.lattice)
        assert err_f <= 1e-5
    finally:
        for h in (integ, field, part):
            h.close()
