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
    themselves to f32 rounding (the same transform with its radix stages
    grouped differently)."""
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
        # the two transforms round differently: f32 rounding (2^-24) over the
        # 21 butterfly stages of a 128^3 transform and the solve is ~1e-6 RMS
        assert d_f <= 3e-6
        # the energy sums |F(k)|^2 G(k) over the f32 spectra of the two transforms
        assert abs(s1["u_recip"] - s0["u_recip"]) <= 1e-7 * abs(s0["u_recip"])
    finally:
        for h in (integ, field, part):
            h.close()


def _f32_wrap_index(x, lx):
    """Image index of the device's orthorhombic wrap (f32 divide, floor of
    s + 1/2), as the kernels evaluate it."""
    f = np.float32
    return float(np.floor(f(f(x / lx) + f(0.5))))


def _f32_primary_point(x, lx, n):
    """Mesh point g = floor((s + 1/2) n) mod n of the device's SPME set-up
    for a coordinate x on an orthorhombic axis of length lx (f32)."""
    f = np.float32
    xw = f(x - f(f(_f32_wrap_index(x, lx)) * lx))
    u = f(f(f(xw / lx) + f(0.5)) * f(n))
    return int(np.floor(u)) % n


def _face_case(lx, n):
    """A coordinate a within a few f32 steps of a box face and a coordinate b
    of the same point at the last list build, 0.25 bohr away and one box
    length over (as after a wrap), such that the list's image-shifted copy
    of a, a - k_b L, has its primary mesh point on the opposite face from
    a's own."""
    f = np.float32
    for face in (1.0, -1.0):
        base = f(f(0.5 * face) * lx)
        for k in range(-64, 65):
            a = f(base + f(k) * np.spacing(base))
            for db in (0.25, -0.25):
                for jump in (1.0, -1.0):
                    b = f(f(a + f(db)) + f(jump) * lx)
                    kb = _f32_wrap_index(b, lx)
                    if kb == _f32_wrap_index(a, lx):
                        continue
                    shifted = f(a - f(f(kb) * lx))
                    g_a, g_s = _f32_primary_point(a, lx, n), _f32_primary_point(shifted, lx, n)
                    if {g_a, g_s} == {0, n - 1}:
                        return a, b, g_a, g_s
    return None


def test_spme_atom_within_a_rounding_of_a_box_face(gpu_ctx):
    """An atom one box length and 0.25 bohr from where it stood at the last
    list build (a wrap) that now lies within a few f32 steps of a box face,
    where the list's image-shifted copy of its position wraps to the
    opposite face (mesh point 0 against n - 1 from the position itself).
    Binning, spreading and gathering use one primary point, so the forces
    match the oracle."""
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=20)
    try:
        f = np.float32
        lx = f(sysw["lattice"][0])
        n = ff["spme"]["grid"][0]
        case = _face_case(lx, n)
        assert case is not None, "no face case for this box length"
        a, b, g_a, g_s = case
        print(f"x now {a!r}, at the build {b!r}: mesh point {g_a} from the position, "
              f"{g_s} from the shifted copy")
        posq0 = part.download()["posq"].copy()
        i = int(np.argmax(posq0[:, 0]))
        p_build = posq0.copy()
        p_build[:, 0] = (p_build[:, 0] + f(b - posq0[i, 0])).astype(f)
        p_build[i, 0] = b
        part.upload(posq=p_build)
        field.request_rebuild()
        field.compute(0)
        gen = field.stats()["generation"]
        p_now = p_build.copy()
        p_now[:, 0] = (p_now[:, 0] + f(a - b)).astype(f)   # 0.25 bohr by minimum image
        p_now[i, 0] = a
        part.upload(posq=p_now)
        field.compute(0)
        assert field.stats()["generation"] == gen      # the build's images are kept
        got = part.forces().astype(np.float64)
        f_pair, _, _ = oracle_pair(sysw, ff, p_now, erfc_mode=1)
        _, f_rec, _, _ = oracle.spme_recip(p_now[:, :3], p_now[:, 3], sysw["lattice"],
                                           ff["spme"]["grid"], ff["spme"]["order"],
                                           ff["coulomb"]["alpha"])
        ref = f_pair + f_rec
        err_i = float(np.linalg.norm(got[i] - ref[i]))
        rms = float(np.sqrt(np.mean(np.sum(ref ** 2, axis=1))))
        print(f"atom {i}: force error {err_i:.3e} (its reciprocal force {np.linalg.norm(f_rec[i]):.3e}, "
              f"RMS force {rms:.3e}); all atoms rel {rel_rms(got, ref):.3e}")
        assert err_i <= 1e-4 * rms
        assert rel_rms(got, ref) <= FORCE_TOL
    finally:
        for h in (integ, field, part):
            h.close()
