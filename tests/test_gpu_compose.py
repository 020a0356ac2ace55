"""Run-time composed pair terms on the GPU (hmd_forcefield_set_pair_fragment),
on the argon case (10,000 atoms, LJ, r_c = 2.5 sigma, displaced lattice):

* the zero fragment: the composed kernels give byte-identical forces and
  energies to the built-in ones;
* a Morse term added to LJ: forces and potential energy against the FP64
  oracle (LJ) plus the Morse term summed in FP64 over the oracle's
  brute-force pair set within r_c.  Bars: relative RMS force <= 1e-5,
  relative energy <= 1e-6."""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api, systems
from heddlemd_amd import units as U
from tests import pair_fragments as F
from tests.gpu_helpers import rel_rms

pytestmark = pytest.mark.gpu


def argon_setup(ctx):
    a = systems.argon()
    rng = np.random.default_rng(42)
    posq = a["posq"].copy()
    posq[:, :3] += rng.normal(scale=0.05 * 3.405 * U.ANGSTROM, size=(a["n"], 3)).astype(np.float32)
    part = api.ParticleBuffers(ctx, posq, a["mass"], a["lattice"], types=a["types"])
    field = api.ForceField(ctx, part, a["ff"])
    return a, part, field


def evaluate(field, part):
    field.request_rebuild()
    field.compute(1)
    d = part.download()
    e, w = part.energies()
    return d, e, w


def test_zero_fragment_reproduces_builtin_kernels(gpu_ctx):
    a, part, field = argon_setup(gpu_ctx)
    try:
        d0, e0, w0 = evaluate(field, part)
        field.set_pair_fragment("ZeroPair", F.ZERO)
        d1, e1, w1 = evaluate(field, part)
    finally:
        field.close(); part.close()
    assert d0["force"].tobytes() == d1["force"].tobytes()
    assert e0.tobytes() == e1.tobytes() and w0.tobytes() == w1.tobytes()


def test_morse_fragment_matches_oracle(gpu_ctx):
    a, part, field = argon_setup(gpu_ctx)
    D, alpha, r0 = 2.0e-4, 1.1, 7.2          # Hartree, 1/bohr, bohr
    try:
        field.set_pair_fragment("MorsePair", F.MORSE, [D, alpha, r0])
        d, e, _ = evaluate(field, part)
        pe_gpu, _ = part.energy_sums()
    finally:
        field.close(); part.close()
    posq = d["posq"]
    ps = oracle.PairSystem(posq[:, :3].astype(np.float64), posq[:, 3].astype(np.float64),
                           a["types"].astype(np.int32), a["lattice"], lj=a["ff"]["lj"])
    f_lj, e_lj, _ = ps.forces(cells=True)
    rc = float(a["ff"]["lj"]["cutoff"][0])
    pairs = oracle.pair_set_f32(posq, a["lattice"].astype(np.float32),
                                np.float32(np.float32(rc) * np.float32(rc) * 1.01))
    x = posq[:, :3].astype(np.float64)
    L = a["lattice"][0]
    dvec = x[pairs[:, 0]] - x[pairs[:, 1]]
    dvec -= L * np.round(dvec / L)
    r = np.linalg.norm(dvec, axis=1)
    keep = r <= rc
    dvec, r, pi, pj = dvec[keep], r[keep], pairs[keep, 0], pairs[keep, 1]
    ex = np.exp(-alpha * (r - r0))
    e_m = D * (1 - ex) ** 2 - D
    fr = -2.0 * D * alpha * (1 - ex) * ex / r
    f_m = np.zeros_like(x)
    np.add.at(f_m, pi, fr[:, None] * dvec)
    np.add.at(f_m, pj, -fr[:, None] * dvec)
    f_ref = f_lj + f_m
    pe_ref = e_lj.sum() + e_m.sum()
    err_f = rel_rms(d["force"], f_ref)
    err_e = abs(pe_gpu - pe_ref) / abs(pe_ref)
    print(f"LJ + Morse fragment: rel RMS force {err_f:.3e}, rel PE {err_e:.3e} "
          f"({keep.sum()} pairs within r_c)")
    assert err_f <= 1e-5 and err_e <= 1e-6


def test_fragment_exclusion_scale_on_listed_pairs(gpu_ctx):
    """The composed term's own exclusion_scale (lj.rs:451): on the 3,888-atom
    water box with a Morse fragment, all-zero scales give the bytes of the
    fragment without scales (listed partners skipped), and scale 0.5 adds
    0.5 x the Morse term of every listed pair (i < j from i's list), summed
    in FP64 at the same positions: forces to 1e-5 relative RMS, energy to
    1e-5 relative.  Scales need a composed term and one entry per listed
    partner."""
    from heddlemd_amd._lib import HmdError
    from tests.gpu_helpers import water_case
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=40)
    D, alpha, r0 = 0.01, 1.0, 2.5
    off = np.asarray(ff["exclusions"][0], np.int64)
    par = np.asarray(ff["exclusions"][1], np.int64)
    n_entries = int(off[-1])
    try:
        with pytest.raises(HmdError, match="no composed"):
            field.set_fragment_exclusion_scales(np.zeros(n_entries))
        field.set_pair_fragment("MorsePair", F.MORSE, [D, alpha, r0])
        d0, _, _ = evaluate(field, part)
        pe0, _ = part.energy_sums()
        with pytest.raises(HmdError, match="one scale per listed"):
            field.set_fragment_exclusion_scales(np.zeros(n_entries + 1))
        field.set_fragment_exclusion_scales(np.zeros(n_entries))
        dz, _, _ = evaluate(field, part)
        assert dz["force"].tobytes() == d0["force"].tobytes()
        s = 0.5
        field.set_fragment_exclusion_scales(np.full(n_entries, s))
        d1, _, _ = evaluate(field, part)
        pe1, _ = part.energy_sums()
    finally:
        for h in (integ, field, part):
            h.close()
    x = d0["posq"][:, :3].astype(np.float64)
    L = float(sysw["lattice"][0])
    owner = np.repeat(np.arange(off.size - 1), np.diff(off))
    keep = par > owner
    pi, pj = owner[keep], par[keep]
    dvec = x[pi] - x[pj]
    dvec -= L * np.round(dvec / L)
    r = np.linalg.norm(dvec, axis=1)
    ex = np.exp(-alpha * (r - r0))
    fr = -2.0 * D * alpha * (1 - ex) * ex / r
    f_m = np.zeros_like(x)
    np.add.at(f_m, pi, fr[:, None] * dvec)
    np.add.at(f_m, pj, -fr[:, None] * dvec)
    e_m = np.sum(D * (1 - ex) ** 2 - D)
    got = d1["force"].astype(np.float64) - d0["force"].astype(np.float64)
    err = rel_rms(got, s * f_m)
    print(f"{pi.size} listed pairs; scaled Morse forces rel RMS {err:.2e}; "
          f"energy {pe1 - pe0:.9e} vs {s * e_m:.9e}")
    assert err < 1e-5
    assert pe1 - pe0 == pytest.approx(s * e_m, rel=1e-5)
