"""The f32 CPU baseline of the MD step (oracle/cpu_step.c), which bench.py
times beside the GPU: its forces match the FP64 oracle as the GPU's must
(relative RMS <= 1e-5), and it steps (list rebuilds by the skin rule, rigid
water kept) -- so that the baseline times the same arithmetic."""
import numpy as np

import oracle
from heddlemd_amd import systems
from heddlemd_amd import units as U


def test_cpu_step_forces_match_the_oracle_and_it_steps():
    s = systems.spc_water(grid=(12, 12, 9), seed=7)
    mesh = (32, 32, 32)
    ff = systems.water_forcefield(grid=mesh, exclusions=s["exclusions"])
    rng = np.random.default_rng(2)
    v = (rng.standard_normal((s["n"], 3)) * 2e-4).astype(np.float32)
    cs = oracle.CpuStep(s["posq"], v, s["mass"], s["types"], s["lattice"], ff["lj"],
                        ff["coulomb"], ff["r_skin"], mesh, ff["exclusions"], s["settle"],
                        1.0 * U.FS)
    try:
        f = cs.forces().astype(np.float64)
        off, par = ff["exclusions"]
        ex = (off.astype(np.int32), par.astype(np.int32), np.zeros(par.size), np.zeros(par.size))
        ps = oracle.PairSystem(s["posq"][:, :3], s["posq"][:, 3], s["types"], s["lattice"],
                               lj=ff["lj"], coulomb=ff["coulomb"], exclusions=ex, erfc_mode=1)
        fp, _, _ = ps.forces(cells=True)
        _, fr, _, _ = oracle.spme_recip(s["posq"][:, :3], s["posq"][:, 3], s["lattice"], mesh, 4,
                                        ff["coulomb"]["alpha"])
        ref = fp + fr
        err = np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum())
        assert err <= 1e-5
        cs.step(20)
        assert cs.rebuilds() >= 1
        assert np.isfinite(cs.forces()).all()
    finally:
        cs.close()
