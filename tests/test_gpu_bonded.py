"""Bonded terms on the GPU (hmd_forcefield_add_bonded, hmd_bonded_compute)
against the FP64 oracle (orc_bonded, pinned by gradients, invariances and
closed forms in tests/test_oracle_bonded.py), on the 3,888-atom water box
with bonds, Morse bonds and angles inside every molecule and dihedrals
across neighbouring molecules (many of them across the periodic boundary):
the bonded slot alone to f32 accuracy (forces 1e-5 relative RMS, energy and
virial sums 1e-5 relative), and the force field's total with the bonded
sets equal to its total without them plus the bonded slot (1e-6 of the
largest force), eager and in a replayed step graph."""
import numpy as np
import pytest

import oracle
from heddlemd_amd._lib import SlotOutput
from tests.gpu_helpers import HipRuntime, water_case

pytestmark = pytest.mark.gpu


def _terms(n_mol):
    m = np.arange(n_mol, dtype=np.uint32)
    o, h1, h2 = 3 * m, 3 * m + 1, 3 * m + 2
    bonds = np.concatenate([np.stack([o, h1, m % 2], 1), np.stack([o, h2, m % 2], 1)])
    morse = np.stack([h1, h2, np.zeros_like(m)], 1)
    angles = np.stack([h1, o, h2, m % 2], 1)
    nxt = (m + 1) % n_mol
    dihed = np.stack([h1, o, 3 * nxt, 3 * nxt + 1, m % 3], 1)
    return {"harmonic-bond": (bonds, [[0.45, 1.7], [0.6, 2.0]]),
            "morse": (morse, [[0.15, 1.2, 3.3]]),
            "harmonic-angle": (angles, [[0.1, 1.75], [0.12, 2.0]]),
            "periodic-dihedral": (dihed, [[0.004, 0.0, 3], [0.006, 1.0, 1], [0.003, 2.5, 2]])}


def test_bonded_slot_matches_the_oracle(gpu_ctx):
    hip = HipRuntime()
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=40)
    try:
        n = part.n
        sets = _terms(sysw["n_mol"])
        for kind, (t, q) in sets.items():
            field.add_bonded(kind, t, q)
        buf = hip.zeros(5 * 4 * n)
        out = SlotOutput(*(buf + 4 * n * k for k in range(5)))
        field.pre_step()
        field.bonded_compute(1, out)
        field.check()
        got = hip.download(buf, (5, n))
        x = part.download()["posq"][:, :3].astype(np.float64)
        lat = part.lattice.astype(np.float64)
        f = np.zeros((n, 3)); e = np.zeros(n); w = np.zeros(n); u = 0.0
        for kind, (t, q) in sets.items():
            r = oracle.bonded(kind, x, lat, t, q)
            f += r[0]; e += r[1]; w += r[2]; u += r[3]
        err = np.sqrt(np.mean((got[:3].T - f) ** 2)) / np.sqrt(np.mean(f ** 2))
        print(f"bonded forces rel RMS {err:.2e}; energy {got[3].sum():.9e} vs {u:.9e}; "
              f"virial {got[4].sum():.9e} vs {w.sum():.9e}")
        assert err < 1e-5
        assert got[3].sum() == pytest.approx(u, rel=1e-5)
        assert got[4].sum() == pytest.approx(w.sum(), rel=1e-5)
        assert np.abs(got[3] - e).max() < 1e-5 * np.abs(e).max()
        # the whole force field: pair + reciprocal + bonded in one evaluation
        field.compute(1)
        tot = part.forces()
        field2 = type(field)(gpu_ctx, part, ff)
        try:
            field2.compute(1)
            base = part.forces()
        finally:
            field2.close()
        gap = np.abs(tot - (base + got[:3].T)).max() / np.abs(tot).max()
        print(f"total vs pair + reciprocal + bonded slot: {gap:.2e} of the largest force")
        assert gap < 1e-6
        # a replayed step graph carries the bonded launches: it steps as eager
        integ.capture()
        integ.replay(3)
        assert np.isfinite(part.download()["posq"]).all()
    finally:
        for h in (integ, field, part):
            h.close()
        hip.free_all()


def test_bonded_input_checks(gpu_ctx):
    from heddlemd_amd._lib import HmdError
    sysw, ff, part, field, integ = water_case(gpu_ctx)
    try:
        n = part.n
        with pytest.raises(HmdError, match="out of range"):
            field.add_bonded("harmonic-bond", [[0, n, 0]], [[0.1, 1.0]])
        with pytest.raises(HmdError, match="repeats"):
            field.add_bonded("harmonic-angle", [[0, 1, 0, 0]], [[0.1, 1.0]])
        with pytest.raises(HmdError, match="invalid parameters"):
            field.add_bonded("harmonic-bond", [[0, 1, 0]], [[-0.1, 1.0]])
        with pytest.raises(HmdError, match="invalid parameters"):
            field.add_bonded("periodic-dihedral", [[0, 1, 2, 3, 0]], [[0.1, 0.0, 1.5]])
    finally:
        for h in (integ, field, part):
            h.close()
