"""GPU tests of the neighbour list's skin trigger and capacity handling,
through the C ABI.

The skin trigger follows the reference's displacement tests
(tests/displacement_check.rs:102-205; kernels/neighbor.cu:50-76): an atom
that moved more than r_skin/2 (exact f32 minimum-image distance) from its
position at the last rebuild starts a rebuild in the same evaluation, one
that moved less does not, a move across the periodic boundary counts by its
minimum image, and a rebuild takes the current positions as the new
reference.  Here the flag is consumed by the device-gated rebuild of the same
step, so the observable is the rebuild generation (hmd_nbl_stats).

Capacity (src/forces/neighbor_list.rs:235-244, 415-445): an overflow grows
the per-cluster capacity by 1.5x and is reported as HMD_ERR_NBL_OVERFLOW; a
build past 0.8 of the capacity grows it ahead of any dropped entry; list
positions past the build's LDS staging are written to memory directly.
"""
import numpy as np
import pytest

import oracle
from heddlemd_amd import api
from heddlemd_amd import units as U
from heddlemd_amd._lib import HMD_ERR_NBL_OVERFLOW, HmdError
from tests.gpu_helpers import oracle_pair, rel_rms, water_case

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def _threshold_sq(ff):
    """(r_skin / 2)^2 as the library forms it in f32: 0.25 r_skin r_skin."""
    rs = np.float32(ff["r_skin"])
    return np.float32(np.float32(0.25) * rs) * rs


def _gen(field):
    return field.stats()["generation"]


def _pair_set_matches_oracle(sysw, ff, part, field):
    got = field.pair_snapshot()
    posq = part.download()["posq"]
    rs = np.float32(np.float32(ff["coulomb"]["r_cut"]) + np.float32(ff["r_skin"]))
    want = oracle.pair_set_f32(posq, sysw["lattice"].astype(np.float32), rs * rs)
    return got.shape == want.shape and np.array_equal(got, want)


@pytest.fixture()
def water(gpu_ctx):
    case = water_case(gpu_ctx, relax=40)
    yield case
    for h in case[2:][::-1]:
        h.close()


def test_no_rebuild_at_the_reference(water):
    sysw, ff, part, field, integ = water
    field.request_rebuild()
    field.compute(0)
    g0 = _gen(field)
    assert field.stats()["status"] & 1 == 0
    field.compute(0)
    field.compute(0)
    assert _gen(field) == g0


def test_threshold_is_half_the_skin(water):
    """One atom moved by 0.9 r_skin/2 keeps the list; by 1.1 r_skin/2 it
    starts a rebuild, after which the pair set is the oracle's, bit for bit,
    and the flag is clear."""
    sysw, ff, part, field, integ = water
    field.request_rebuild()
    field.compute(0)
    ref = part.download()["posq"].copy()
    g0 = _gen(field)
    half = 0.5 * ff["r_skin"]
    thr = _threshold_sq(ff)
    k = 123
    for frac, rebuilt in ((0.9, 0), (1.1, 1)):
        posq = ref.copy()
        posq[k, 1] += np.float32(frac * half)
        d2 = oracle.r2_f32_many(posq, ref, sysw["lattice"])
        assert bool((d2 > thr).any()) == bool(rebuilt)
        part.upload(posq=posq)
        field.compute(0)
        assert _gen(field) == g0 + rebuilt, f"move {frac} x r_skin/2"
    assert field.stats()["status"] & 1 == 0
    assert _pair_set_matches_oracle(sysw, ff, part, field)
    # the reference is now the moved configuration: putting the atom back
    # (1.1 r_skin/2 from it) rebuilds again
    part.upload(posq=ref)
    field.compute(0)
    assert _gen(field) == g0 + 2


def test_displacement_uses_the_minimum_image(water):
    """An atom carried across the periodic boundary (raw difference about one
    box length, minimum-image difference 0.4 r_skin/2) keeps the list, and
    the forces on the unchanged list still match the oracle."""
    sysw, ff, part, field, integ = water
    field.request_rebuild()
    field.compute(0)
    ref = part.download()["posq"].copy()
    g0 = _gen(field)
    L = float(sysw["lattice"][0])
    k = int(np.argmax(np.abs(ref[:, 0])))          # the atom nearest an x face
    posq = ref.copy()
    shift = -np.sign(ref[k, 0]) * L
    posq[k, 0] = np.float32(ref[k, 0] + shift + np.sign(ref[k, 0]) * 0.2 * ff["r_skin"])
    assert abs(posq[k, 0] - ref[k, 0]) > 0.9 * L
    d2 = oracle.r2_f32_many(posq, ref, sysw["lattice"])
    assert not (d2 > _threshold_sq(ff)).any()
    part.upload(posq=posq)
    field.compute(0)
    assert _gen(field) == g0
    d = part.download()
    f_pair, _, _ = oracle_pair(sysw, ff, d["posq"])
    u, f_rec, _, _ = oracle.spme_recip(d["posq"][:, :3], d["posq"][:, 3], sysw["lattice"],
                                       ff["spme"]["grid"], 4, ff["coulomb"]["alpha"])
    err = rel_rms(d["force"], f_pair + f_rec)
    print(f"forces after a move across the boundary: rel RMS {err:.2e}")
    assert err <= FORCE_TOL


def test_rebuilds_follow_the_skin_rule_across_graph_replays(gpu_ctx):
    """Hot water (900 K) stepped one graph replay at a time: each step
    rebuilds exactly when some atom is more than r_skin/2 (exact f32
    minimum image) from its position at the previous rebuild, as predicted
    on the host; the device flag never lingers; after the last step every
    pair within r_c is in the list (the count of listed pairs within r_c
    equals the oracle's count over all pairs)."""
    sysw, ff, part, field, integ = water_case(gpu_ctx, relax=40)
    try:
        v = api.generate_velocities(sysw["mass"], U.kelvin(900.0), 5,
                                    n_constraints=3 * sysw["n_mol"])
        part.upload(vel=v)
        field.request_rebuild()
        field.compute(0)
        integ.capture()
        ref = part.download()["posq"].copy()
        g = _gen(field)
        thr = _threshold_sq(ff)
        n_rebuilt = n_kept = 0
        for t in range(30):
            integ.replay(1)
            x = part.download()["posq"]
            want = bool((oracle.r2_f32_many(x, ref, sysw["lattice"]) > thr).any())
            g1 = _gen(field)
            assert g1 - g == int(want), f"step {t}: rebuilt {g1 - g}, predicted {want}"
            assert field.stats()["status"] & 1 == 0
            if want:
                ref = x.copy()
                n_rebuilt += 1
            else:
                n_kept += 1
            g = g1
        print(f"30 replays: {n_rebuilt} rebuilds, {n_kept} steps on the old list")
        assert n_rebuilt >= 2 and n_kept >= 2
        rc = np.float32(ff["coulomb"]["r_cut"])
        posq = part.download()["posq"]
        want = oracle.pair_set_f32(posq, sysw["lattice"].astype(np.float32), rc * rc)
        assert field.count_pairs(float(ff["coulomb"]["r_cut"])) == len(want)
    finally:
        for h in (integ, field, part):
            h.close()


def test_overflow_grows_the_capacity_and_reports_the_need(water):
    sysw, ff, part, field, integ = water
    field.compute(0)
    need = field.stats()["high_water"]
    field.set_capacity(2)
    with pytest.raises(HmdError) as e:
        field.compute(0)
    assert e.value.code == HMD_ERR_NBL_OVERFLOW
    assert f"needed {need} entries" in str(e.value)
    st = field.stats()
    print(f"need {need}: capacity 2 -> {st['cap_entries']} ({e.value})")
    assert st["cap_entries"] >= int(np.ceil(1.5 * need))
    field.compute(0)
    assert _pair_set_matches_oracle(sysw, ff, part, field)


def test_fill_threshold_grows_ahead_and_direct_writes_match_lds_staging(water):
    """A capacity just above the need (past the 0.8 fill threshold) is grown
    without an error and without changing a bit of the forces; staging 1 or
    3 entries in LDS and writing the rest of each list directly gives the
    same list and the same forces."""
    sysw, ff, part, field, integ = water
    field.request_rebuild()
    field.compute(0)
    f_ref = part.forces().copy()
    need = field.stats()["high_water"]
    field.set_capacity(need + 1)
    field.compute(0)
    assert field.stats()["cap_entries"] > need + 1
    assert np.array_equal(part.forces().view(np.uint32), f_ref.view(np.uint32))
    cap = field.stats()["cap_entries"]
    for lds in (1, 3):
        field.set_capacity(cap, lds_entries=lds)
        field.compute(0)
        assert np.array_equal(part.forces().view(np.uint32), f_ref.view(np.uint32)), lds
        assert _pair_set_matches_oracle(sysw, ff, part, field)
    field.set_capacity(cap, lds_entries=0)
