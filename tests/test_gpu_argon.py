"""GPU parity of the plain LJ argon case (10,000 atoms on a simple-cubic
lattice, r_c = 2.5 sigma with no switch, no charges, skin 0.3 r_c): the
LJ-only pair kernel variant through the C ABI against the FP64 oracle.

Tolerances (the hot-path contract): relative RMS force error <= 1e-5,
relative potential energy error <= 1e-6; neighbour pair set bit-exact."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api, systems
from heddlemd_amd import units as U
from tests.gpu_helpers import rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6


@pytest.fixture(scope="module")
def argon_case(gpu_ctx):
    a = systems.argon()
    # displace the lattice sites (0.05 sigma RMS per axis, seeded) so that
    # the forces are not fixed by symmetry
    rng = np.random.default_rng(42)
    posq = a["posq"].copy()
    posq[:, :3] += rng.normal(scale=0.05 * 3.405 * U.ANGSTROM, size=(a["n"], 3)).astype(np.float32)
    part = api.ParticleBuffers(gpu_ctx, posq, a["mass"], a["lattice"], types=a["types"])
    field = api.ForceField(gpu_ctx, part, a["ff"])
    yield a, part, field
    field.close()
    part.close()


def test_argon_pair_set_bit_exact(argon_case):
    a, part, field = argon_case
    field.request_rebuild()
    field.compute(0)
    got = field.pair_snapshot()
    posq = part.download()["posq"]
    rs = np.float32(np.float32(a["ff"]["lj"]["cutoff"][0]) + np.float32(a["ff"]["r_skin"]))
    want = oracle.pair_set_f32(posq, a["lattice"].astype(np.float32), rs * rs)
    print(f"argon pairs gpu={len(got)} oracle={len(want)}")
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_argon_lj_forces_and_energy_match_oracle(argon_case):
    a, part, field = argon_case
    field.compute(1)
    d = part.download()
    posq = d["posq"].astype(np.float64)
    ps = oracle.PairSystem(posq[:, :3], posq[:, 3], a["types"].astype(np.int32), a["lattice"],
                           lj=a["ff"]["lj"])
    f_ref, e_ref, _ = ps.forces(cells=True)
    pe_gpu, _ = part.energy_sums()
    err_f = rel_rms(d["force"], f_ref)
    err_e = abs(pe_gpu - e_ref.sum()) / abs(e_ref.sum())
    print(f"argon LJ: rel RMS force {err_f:.3e}, rel PE {err_e:.3e} (PE {e_ref.sum():.6e} Ha)")
    assert err_f <= FORCE_TOL
    assert err_e <= ENERGY_TOL
