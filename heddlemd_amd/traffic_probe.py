"""HBM traffic probe for hardware-counter runs (rocprofv3 --pmc).

Runs, in one process on one GPU:
  1. a device-to-device copy of a known size (CALIB_BYTES read + written),
     whose FETCH_SIZE / WRITE_SIZE per dispatch calibrate the counters;
  2. the 196,608-atom water workload's force evaluation (pair kernel, SPME
     128^3), eagerly, a few times, so that each kernel's counters can be read
     per dispatch.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m heddlemd_amd.traffic_probe
    rocprofv3 --pmc WRITE_SIZE -- python3 -m heddlemd_amd.traffic_probe
"""
import sys

CALIB_BYTES = 256 << 20


def main():
    from heddlemd_amd import api, systems
    from heddlemd_amd import units as U
    import torch   # device buffers for the calibration copy only

    a = torch.empty(CALIB_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration copies: 3 x {CALIB_BYTES} bytes", flush=True)

    sysw = systems.spc_water(grid=(64, 64, 16), seed=11)
    ff = systems.water_forcefield(grid=(128,) * 3, exclusions=sysw["exclusions"])
    ctx = api.GpuContext(0)
    part = api.ParticleBuffers(ctx, sysw["posq"], sysw["mass"], sysw["lattice"],
                               types=sysw["types"])
    field = api.ForceField(ctx, part, ff)
    # SPME on the main stream: every kernel's counters without a concurrent one
    field.set_tuning(dict(field.tuning(), spme_stream=0))
    integ = api.Integrator(ctx, part, field, dt=2.0 * U.FS, compensated=True,
                           settle=sysw["settle"])
    integ.minimize(20, 0.1 * U.ANGSTROM)
    for _ in range(5):
        field.compute(0)
    ctx.synchronize()
    print("5 force evaluations of the 196,608-atom box", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
