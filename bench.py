"""Benchmark: ns/day of 196,608-atom rigid SPC water on one MI355X.

Workload (the single-GPU configuration the metric is quoted on):
65,536 SPC molecules (196,608 atoms) with centres on a 64 x 64 x 16 grid
scaled to a cubic box at 997 kg/m^3 (edge 12.528 nm, seed 11); LJ O-O
(sigma 3.166 A, eps 0.650 kJ/mol, switch 0.9-1.0 nm) + real-space Ewald
(r_c 1.0 nm, erfc(alpha r_c) = 1e-5) + SPME 128^3 order 4; cell list +
packed neighbour list (skin 0.1 nm, displacement-triggered rebuilds);
SETTLE; compensated (f32, f64) velocity Verlet, NVE from 300 K; dt = 2 fs.

One "step" is one MD step (neighbour check/rebuild, pair forces, SPME,
integration, constraints), replayed from a captured hipGraph.  The synthetic
lattice is relaxed by steepest descent and thermalised before timing
(setup, untimed).

Prints one JSON line (see the keys below).  --gpus N > 1 exits: this build
runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DT_FS = 2.0
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0
# Algorithmic FLOP per in-cutoff unordered pair of the fused LJ + real-space
# Ewald force evaluation (forces only): the counting rule +,-,* = 1,
# FMA = 2, rsqrt/exp/rcp = 1 each, tallied from the reference functors.
FLOP_PER_PAIR = 59


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(api, systems, U, mesh=128, seed=11, minimize_iter=200):
    sysw = systems.spc_water(grid=(64, 64, 16), seed=seed)
    ff = systems.water_forcefield(grid=(mesh,) * 3, exclusions=sysw["exclusions"])
    ctx = api.GpuContext(0)
    part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                               types=sysw["types"])
    field = api.ForceField(ctx, part, ff)
    integ = api.Integrator(ctx, part, field, dt=DT_FS * U.FS, compensated=True,
                           settle=sysw["settle"])
    integ.minimize(minimize_iter, 0.1 * U.ANGSTROM)
    v = api.generate_velocities(sysw["mass"], U.kelvin(300.0), seed,
                                n_constraints=3 * sysw["n_mol"])
    part.upload(vel=v)
    return ctx, sysw, ff, part, field, integ


def cpu_baseline(sysw, ff, posq, budget_s=20.0):
    """The FP64 oracle (the CPU restatement of the same force arithmetic) on
    the host cores: full pair forces (cell list) + SPME 128^3 for the same
    196,608-atom configuration, repeated while within budget_s."""
    import oracle
    lj = ff["lj"]
    off, par = ff["exclusions"]
    ex = (off.astype(np.int32), par.astype(np.int32), np.zeros(par.size), np.zeros(par.size))
    ps = oracle.PairSystem(posq[:, :3].astype(np.float64), posq[:, 3].astype(np.float64),
                           sysw["types"].astype(np.int32), sysw["lattice"], lj=lj,
                           coulomb=ff["coulomb"], exclusions=ex, erfc_mode=1)
    evals = 0
    t0 = time.perf_counter()
    while True:
        ps.forces(cells=True)
        oracle.spme_recip(posq[:, :3], posq[:, 3], sysw["lattice"], ff["spme"]["grid"],
                          ff["spme"]["order"], ff["coulomb"]["alpha"])
        evals += 1
        el = time.perf_counter() - t0
        if el > budget_s or evals >= 3:
            break
    steps_per_s = evals / el
    from heddlemd_amd import units as U
    return {"value": round(U.ns_per_day(steps_per_s, DT_FS), 5), "unit": "ns/day",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{evals} force evaluation(s) (pair cell list + SPME 128^3, FP64) of "
                      f"the 196,608-atom box in {el:.1f} s; integration excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of graph replays (for kernel "
                         "tracers that do not follow graph launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from heddlemd_amd import api, systems
    from heddlemd_amd import units as U

    t_setup = time.perf_counter()
    ctx, sysw, ff, part, field, integ = build_system(api, systems, U)
    log(f"setup {time.perf_counter() - t_setup:.1f} s")
    run = integ.step if args.eager else integ.replay
    if not args.eager:
        integ.capture()
    run(args.warmup)
    ctx.synchronize()
    gen0 = field.stats()["generation"]
    t0 = time.perf_counter()
    run(args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    gen1 = field.stats()["generation"]
    ms_per_step = el / args.steps * 1e3
    ns_day = U.ns_per_day(args.steps / el, DT_FS)
    log(f"{args.steps} steps in {el:.3f} s: {ms_per_step:.3f} ms/step, {ns_day:.2f} ns/day, "
        f"{gen1 - gen0} rebuilds")

    # Kernel timing with HIP events on the context stream (eager steps, same
    # kernels): average duration of the pair kernel.
    prof = integ.profile(20)
    pair_us = prof["pair_force"]
    n_in_cut = field.count_pairs(ff["coulomb"]["r_cut"])
    achieved = n_in_cut * FLOP_PER_PAIR / (pair_us * 1e-6) / 1e12
    log("stage us: " + ", ".join(f"{k} {v:.1f}" for k, v in prof.items()))

    posq = part.download()["posq"]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sysw, ff, posq)

    out = {
        "metric": "ns/day, 196k-atom SPC water: plain 128^3 SPME vs NN-corrected 64^3, dF<=1e-3",
        "value": round(ns_day, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain SPME 128^3, 196,608-atom rigid SPC water, compensated "
                               "VV NVE 300 K, dt 2 fs (the NN-corrected 64^3 leg is not built yet)",
                   "launch": "eager" if args.eager else "hipGraph replay",
                   "atoms": int(sysw["n"]), "mesh": 128, "r_cut_nm": 1.0, "r_skin_nm": 0.1,
                   "atom_steps_per_s": round(sysw["n"] * args.steps / el, 1),
                   "rebuilds": int(gen1 - gen0)},
        "roofline": {"bound": "valu", "kernel": "hmd_pair_force_f",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": None, "pairs_in_cutoff": n_in_cut,
                     "kernel_us": round(pair_us, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
