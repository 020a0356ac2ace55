"""GPU parity tests: the HIP path through the C ABI against the FP64 oracle.

Tolerances (stated by the hot-path contract): relative RMS force error
<= 1e-5 and relative energy error <= 1e-6 against the FP64 oracle run on the
same f32 input positions; neighbour pair sets bit-exact.
"""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api
from heddlemd_amd import units as U
from tests.gpu_helpers import oracle_pair, rel_rms, water_case

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6


@pytest.fixture(scope="module")
def water(gpu_ctx):
    case = water_case(gpu_ctx, relax=60)
    yield case
    for h in case[2:][::-1]:
        h.close()


def test_pair_set_bit_exact(water):
    sysw, ff, part, field, integ = water
    field.request_rebuild()
    field.compute(0)
    got = field.pair_snapshot()
    posq = part.download()["posq"]
    rs = np.float32(np.float32(ff["coulomb"]["r_cut"]) + np.float32(ff["r_skin"]))
    want = oracle.pair_set_f32(posq, sysw["lattice"].astype(np.float32), rs * rs)
    print(f"pairs gpu={len(got)} oracle={len(want)}")
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_pair_plus_spme_forces_match_oracle(water):
    sysw, ff, part, field, integ = water
    field.compute(1)
    d = part.download()
    posq = d["posq"]
    e_gpu, w_gpu = part.energies()
    f_pair, e_pair, _ = oracle_pair(sysw, ff, posq, erfc_mode=1)
    u_rec, f_rec, e_rec, w_rec = oracle.spme_recip(
        posq[:, :3], posq[:, 3], sysw["lattice"], ff["spme"]["grid"], ff["spme"]["order"],
        ff["coulomb"]["alpha"])
    f_ref = f_pair + f_rec
    err_f = rel_rms(d["force"], f_ref)
    f_spme = field.spme_forces()
    err_rec = rel_rms(f_spme, f_rec)
    err_pair = rel_rms(d["force"] - f_spme, f_pair)
    pe_ref = e_pair.sum() + e_rec.sum()
    pe_gpu, vir_gpu = part.energy_sums()
    err_e = abs(pe_gpu - pe_ref) / abs(pe_ref)
    sc = field.spme_scalars()
    err_u = abs(sc["u_recip"] - u_rec) / abs(u_rec)
    err_w = abs(sc["w_recip"] - w_rec) / abs(w_rec)
    # per-atom virial: the pair slots' half shares plus W_recip / N
    _, _, w_pair = oracle_pair(sysw, ff, posq, erfc_mode=1)
    w_ref = w_pair + w_rec / len(w_pair)
    err_wi = rel_rms(w_gpu, w_ref)
    err_wsum = abs(vir_gpu - w_ref.sum()) / np.abs(w_ref).sum()
    print(f"rel RMS force: total {err_f:.3e} pair {err_pair:.3e} recip {err_rec:.3e}; "
          f"rel PE {err_e:.3e}; U_recip {err_u:.3e}; W_recip {err_w:.3e}; per-atom virial "
          f"{err_wi:.3e}; virial sum {err_wsum:.3e} of sum |W_i| (W = {vir_gpu:.6e}, "
          f"oracle {w_ref.sum():.6e})")
    assert err_pair <= FORCE_TOL
    assert err_rec <= FORCE_TOL
    assert err_f <= FORCE_TOL
    assert err_u <= ENERGY_TOL
    assert err_e <= ENERGY_TOL
    assert err_w <= ENERGY_TOL
    assert err_wi <= FORCE_TOL
    assert err_wsum <= ENERGY_TOL


def test_pair_forces_vs_exact_erfc(water):
    """Against libm erfc: the A&S 7.1.26 approximation (max abs error 1.5e-7)
    stays far below the force tolerance."""
    sysw, ff, part, field, integ = water
    field.compute(0)
    d = part.download()
    f_spme = field.spme_forces()
    f_pair, _, _ = oracle_pair(sysw, ff, d["posq"], erfc_mode=0)
    err = rel_rms(d["force"] - f_spme, f_pair)
    print(f"pair vs exact erfc: rel RMS {err:.3e}")
    assert err <= FORCE_TOL


def test_forces_are_reproducible_and_newton3(water):
    sysw, ff, part, field, integ = water
    field.compute(0)
    f1 = part.forces().copy()
    field.request_rebuild()
    field.compute(0)
    f2 = part.forces()
    assert np.array_equal(f1.view(np.uint32), f2.view(np.uint32))
    # Newton's third law: the pair forces sum to ~0 (f32 partial sums are
    # converted to fixed point separately on the i and j sides).
    f_pair = f1.astype(np.float64) - field.spme_forces()
    scale = np.sqrt(np.mean(np.sum(f_pair ** 2, axis=1)))
    net = np.abs(f_pair.sum(axis=0)).max()
    print(f"net pair force {net:.3e} vs RMS {scale:.3e}")
    assert net < 1e-5 * scale * np.sqrt(f_pair.shape[0])


def test_md_graph_matches_eager_and_is_deterministic(gpu_ctx):
    runs = []
    for mode in ("eager", "graph", "graph"):
        sysw, ff, part, field, integ = water_case(gpu_ctx, relax=30)
        if mode == "eager":
            integ.step(10)
        else:
            integ.replay(10)
        d = part.download()
        runs.append((d["posq"].copy(), d["vel"].copy()))
        for h in (integ, field, part):
            h.close()
    for k in (1, 2):
        assert np.array_equal(runs[0][0].view(np.uint32), runs[k][0].view(np.uint32))
        assert np.array_equal(runs[0][1].view(np.uint32), runs[k][1].view(np.uint32))


def test_settle_keeps_water_rigid(gpu_ctx):
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=30)
    integ.replay(20)
    pos = part.download()["posq"][:, :3].astype(np.float64).reshape(-1, 3, 3)
    L = sysw["L"]

    def mi(d):
        return d - L * np.floor(d / L + 0.5)
    d_oh1 = np.linalg.norm(mi(pos[:, 1] - pos[:, 0]), axis=1)
    d_oh2 = np.linalg.norm(mi(pos[:, 2] - pos[:, 0]), axis=1)
    d_hh = np.linalg.norm(mi(pos[:, 2] - pos[:, 1]), axis=1)
    s = sysw["settle"]
    err = max(np.abs(d_oh1 / s["d_oh"] - 1).max(), np.abs(d_oh2 / s["d_oh"] - 1).max(),
              np.abs(d_hh / s["d_hh"] - 1).max())
    print(f"max relative bond error {err:.2e}")
    assert err < 2e-5
    for h in (integ, field, part):
        h.close()


def test_ctx_reports_gfx950(gpu_ctx):
    cus, name = gpu_ctx.device_info()
    assert name.startswith("gfx950")
    assert cus >= 200


@pytest.mark.parametrize("mesh", [(64, 64, 64), (128, 128, 128), (256, 32, 32), (48, 40, 24)])
def test_spme_mesh_shapes_match_oracle(gpu_ctx, mesh):
    """Reciprocal forces and energy on meshes that exercise each FFT kernel:
    compile-time power-of-two lengths 16-256 (both stages of the two-stage
    transform, radix 4, 8 and 16) and the generic mixed-radix path (48, 40,
    24 -> c-axis length 12)."""
    sysw, ff, part, field, integ = water_case(gpu_ctx, mesh=mesh)
    field.compute(0)
    posq = part.download()["posq"]
    u_rec, f_rec, _, _ = oracle.spme_recip(posq[:, :3], posq[:, 3], sysw["lattice"], mesh, 4,
                                           ff["coulomb"]["alpha"])
    err_f = rel_rms(field.spme_forces(), f_rec)
    err_u = abs(field.spme_scalars()["u_recip"] - u_rec) / abs(u_rec)
    print(f"mesh {mesh}: recip force rel RMS {err_f:.2e}, U_recip rel {err_u:.2e}")
    assert err_f <= FORCE_TOL
    assert err_u <= ENERGY_TOL
    for h in (integ, field, part):
        h.close()


@pytest.mark.parametrize("mesh", [(64, 64, 64), (128, 128, 128)])
def test_slab_fft_matches_oracle_and_line_passes(gpu_ctx, mesh):
    """The three-kernel slab FFT (tuning fft_slab = 1) on the meshes it
    serves: reciprocal forces, U_recip and W_recip against the oracle to the
    same bars as the five line passes, and against the line passes
    themselves (the same per-line arithmetic, lines grouped differently)."""
    sysw, ff, part, field, integ = water_case(gpu_ctx, mesh=mesh)
    try:
        out = {}
        for slab in (0, 1):
            field.set_tuning(dict(field.tuning(), fft_slab=slab))
            assert field.tuning()["fft_slab"] == slab
            field.compute(1)
            out[slab] = (field.spme_forces(), field.spme_scalars())
        posq = part.download()["posq"]
        u_rec, f_rec, _, w_rec = oracle.spme_recip(posq[:, :3], posq[:, 3], sysw["lattice"], mesh,
                                                   4, ff["coulomb"]["alpha"])
        f1, s1 = out[1]
        f0, s0 = out[0]
        err_f = rel_rms(f1, f_rec)
        err_u = abs(s1["u_recip"] - u_rec) / abs(u_rec)
        err_w = abs(s1["w_recip"] - w_rec) / abs(w_rec)
        d_f = rel_rms(f1, f0)
        print(f"mesh {mesh} slab: recip force {err_f:.2e}, U {err_u:.2e}, W {err_w:.2e}; "
              f"vs line passes {d_f:.2e}")
        assert err_f <= FORCE_TOL and err_u <= ENERGY_TOL and err_w <= ENERGY_TOL
        assert d_f <= 1e-6
        assert abs(s1["u_recip"] - s0["u_recip"]) <= 1e-9 * abs(s0["u_recip"])
    finally:
        for h in (integ, field, part):
            h.close()
