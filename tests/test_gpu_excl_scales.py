"""Exclusion scales other than 0 (the reference's per-pair exclusion scale,
src/forces/jit_composed.rs:939-972, :1260-1272): the fused pair kernel skips
every listed partner and a correction pass adds s_lj x LJ + s_coul x
real-space Coulomb back for each scaled pair.  A 3,888-atom SPC water box
with O-H pairs scaled (LJ 0.5, Coulomb 0.8333) and H-H pairs (LJ 0,
Coulomb 0.5), real space only: forces and energy against the FP64 oracle
with the same scales (relative RMS force <= 1e-5, relative energy <= 1e-6),
eager and after graph-replayed steps."""
import numpy as np
import pytest

from heddlemd_amd import api, systems
from heddlemd_amd import units as U
from tests.gpu_helpers import oracle_pair, rel_rms

pytestmark = pytest.mark.gpu


def scaled_case(ctx):
    sysw = systems.spc_water(grid=(12, 12, 9), seed=7)
    off, par = sysw["exclusions"]
    is_o = sysw["types"] == 0
    slj = np.zeros(par.size, np.float32)
    sco = np.zeros(par.size, np.float32)
    for i in range(sysw["n"]):
        for m in range(off[i], off[i + 1]):
            j = par[m]
            if is_o[i] != is_o[j]:          # O-H
                slj[m], sco[m] = 0.5, 0.8333
            else:                           # H-H
                slj[m], sco[m] = 0.0, 0.5
    ff = systems.water_forcefield(grid=None, exclusions=(off, par, slj, sco))
    part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                               types=sysw["types"])
    field = api.ForceField(ctx, part, ff)
    integ = api.Integrator(ctx, part, field, dt=2.0 * U.FS, settle=sysw["settle"])
    return sysw, ff, part, field, integ


def check(sysw, ff, part, field):
    field.compute(1)
    d = part.download()
    f_ref, e_ref, _ = oracle_pair(sysw, ff, d["posq"])
    pe, _ = part.energy_sums()
    return rel_rms(d["force"], f_ref), abs(pe - e_ref.sum()) / abs(e_ref.sum())


def test_scaled_exclusions_match_oracle(gpu_ctx):
    sysw, ff, part, field, integ = scaled_case(gpu_ctx)
    try:
        err_f, err_e = check(sysw, ff, part, field)
        print(f"scaled exclusions, eager: rel RMS force {err_f:.2e}, rel PE {err_e:.2e}",
              flush=True)
        assert err_f <= 1e-5 and err_e <= 1e-6
        integ.replay(5)
        err_f, err_e = check(sysw, ff, part, field)
        print(f"after 5 replayed steps: rel RMS force {err_f:.2e}, rel PE {err_e:.2e}",
              flush=True)
        assert err_f <= 1e-5 and err_e <= 1e-6
    finally:
        integ.close(); field.close(); part.close()
