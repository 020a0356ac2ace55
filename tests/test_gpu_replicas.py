"""GPU parity on the boxes of the 64-replica workload (systems.replica_specs:
SPC water): the smallest box with the 0.8 nm cutoff (edge 1.887 nm), a
1.0 nm box whose edge (2.085 nm) is below twice the search radius, a
2.27 nm box (24^3 meshes) and the largest box (4,006 molecules, 12,018
atoms, 4.94 nm, 48^3 mesh).  Same bars as the large box: pair set bit-exact
against the oracle's brute-force f32 enumeration, relative RMS force <= 1e-5
and relative potential energy <= 1e-6 against FP64.  Then the whole
64-box batch (257,244 atoms) steps as one batched launch per stage, and
its smallest and largest boxes end byte-identical to the same boxes stepped
alone, with forces still at the oracle's bars."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import systems
from tests.gpu_helpers import oracle_pair, rel_rms, replica_case

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6
SPECS = {s["replica"]: s for s in systems.replica_specs()}


@pytest.mark.parametrize("rid", [25, 3, 6, 52])
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


def _check_forces(sysw, ff, part, field, tag):
    field.compute(1)
    d = part.download()
    posq = d["posq"]
    f_pair, e_pair, _ = oracle_pair(sysw, ff, posq, erfc_mode=1)
    u_rec, f_rec, e_rec, _ = oracle.spme_recip(
        posq[:, :3], posq[:, 3], sysw["lattice"], ff["spme"]["grid"], ff["spme"]["order"],
        ff["coulomb"]["alpha"])
    err_f = rel_rms(d["force"], f_pair + f_rec)
    pe_gpu, _ = part.energy_sums()
    pe_ref = e_pair.sum() + e_rec.sum()
    err_e = abs(pe_gpu - pe_ref) / abs(pe_ref)
    print(f"  {tag}: rel RMS force {err_f:.3e}, rel PE {err_e:.3e}", flush=True)
    assert err_f <= FORCE_TOL and err_e <= ENERGY_TOL


def test_full_batch_of_64_boxes(gpu_ctx):
    from heddlemd_amd import api
    from heddlemd_amd import units as U
    steps = 20
    specs = systems.replica_specs()

    def build(spec):
        ctx = api.GpuContext(0)
        case = replica_case(ctx, spec, relax=20, kind="langevin-baoab",
                            temperature=U.kelvin(300.0))
        sysw, ff, part, field, integ = case
        part.upload(vel=api.generate_velocities(sysw["mass"], U.kelvin(300.0), spec["seed"],
                                                n_constraints=3 * sysw["n_mol"]))
        return ctx, sysw, ff, part, field, integ

    def close(r):
        ctx, sysw, ff, part, field, integ = r
        integ.close(); field.close(); part.close(); ctx.close()

    reps = [build(s) for s in specs]
    print(f"{len(reps)} boxes, {sum(r[1]['n'] for r in reps)} atoms", flush=True)
    try:
        rs = api.ReplicaSet([r[5] for r in reps], mode="batched")
        try:
            rs.step(steps)
            nb, ns = rs.plan()
            print(f"batched plan: {nb} batched launches, {ns} single per step", flush=True)
            assert rs.steps() == steps
        finally:
            rs.close()
        ends = {}
        for k, spec in enumerate(specs):
            if spec["replica"] in (25, 52):
                d = reps[k][3].download()
                ends[spec["replica"]] = (d["posq"].copy(), d["vel"].copy())
                _check_forces(*reps[k][1:5], f"replica {spec['replica']} after {steps} batched steps")
    finally:
        for r in reps:
            close(r)
    for rid in (25, 52):
        r = build(SPECS[rid])
        try:
            r[5].replay(steps)
            d = r[3].download()
            same = (d["posq"].tobytes() == ends[rid][0].tobytes() and
                    d["vel"].tobytes() == ends[rid][1].tobytes())
            print(f"  replica {rid} alone: byte-identical to the batch {same}", flush=True)
            assert same
        finally:
            close(r)
