"""GPU parity on the small boxes of the 64-replica workload
(systems.replica_specs: SPC water, 24^3 meshes): the smallest box with the
0.8 nm cutoff (edge 1.887 nm), a 1.0 nm box whose edge (2.085 nm) is below
twice the search radius, and a 2.27 nm box.  Same bars as the large box:
pair set bit-exact against the oracle's brute-force f32 enumeration,
relative RMS force <= 1e-5 and relative potential energy <= 1e-6 against
FP64."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import systems
from tests.gpu_helpers import oracle_pair, rel_rms, replica_case

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6
SPECS = {s["replica"]: s for s in systems.replica_specs()}


@pytest.mark.parametrize("rid", [25, 3, 6])
def test_small_replica_pairs_forces_energy(gpu_ctx, rid):
    spec = SPECS[rid]
    print(f"replica {rid}: {spec['n_mol']} molecules, edge {spec['edge_nm']:.3f} nm, "
          f"r_c {spec['r_cut_nm']} nm, mesh {spec['mesh']}", flush=True)
    case = replica_case(gpu_ctx, spec, relax=40)
    sysw, ff, part, field, integ = case
    try:
        field.request_rebuild()
        field.compute(1)
        got = field.pair_snapshot()
        d = part.download()
        posq = d["posq"]
        rs = np.float32(np.float32(ff["coulomb"]["r_cut"]) + np.float32(ff["r_skin"]))
        want = oracle.pair_set_f32(posq, sysw["lattice"].astype(np.float32), rs * rs)
        print(f"  pairs gpu={len(got)} oracle={len(want)}")
        assert got.shape == want.shape and np.array_equal(got, want)
        f_pair, e_pair, _ = oracle_pair(sysw, ff, posq, erfc_mode=1)
        u_rec, f_rec, e_rec, _ = oracle.spme_recip(
            posq[:, :3], posq[:, 3], sysw["lattice"], ff["spme"]["grid"], ff["spme"]["order"],
            ff["coulomb"]["alpha"])
        err_f = rel_rms(d["force"], f_pair + f_rec)
        pe_gpu, _ = part.energy_sums()
        pe_ref = e_pair.sum() + e_rec.sum()
        err_e = abs(pe_gpu - pe_ref) / abs(pe_ref)
        print(f"  rel RMS force {err_f:.3e}, rel PE {err_e:.3e}")
        assert err_f <= FORCE_TOL and err_e <= ENERGY_TOL
    finally:
        for h in case[2:][::-1]:
            h.close()
