"""Velocity-Verlet cases shared by the oracle tests and the GPU parity tests.

Each case is a state (posq, vel, force, mass, lattice) and a list of
operations ("kd" kick-drift, "k" kick, "neg" velocity reversal) with a time
step; the states are those of the reference's own integrator tests
(tests/integration.rs, tests/integration_lossless.rs) plus seeded random
states that cross periodic boundaries."""
import numpy as np

F32 = np.float32
BIG = [1.0e6, 1.0e6, 1.0e6, 0.0, 0.0, 0.0]
BOX10 = [10.0, 10.0, 10.0, 0.0, 0.0, 0.0]


def state(pos, vel, mass, force=None, lattice=BIG):
    pos = np.asarray(pos, F32).reshape(-1, 3)
    n = pos.shape[0]
    posq = np.zeros((n, 4), F32)
    posq[:, :3] = pos
    return {
        "posq": posq,
        "vel": np.asarray(vel, F32).reshape(n, 3),
        "force": np.zeros((n, 3), F32) if force is None else np.asarray(force, F32).reshape(n, 3),
        "mass": np.asarray(mass, F32).reshape(n),
        "lattice": np.asarray(lattice, F32),
    }


def columns(x, y, z):
    return np.stack([np.asarray(x, F32), np.asarray(y, F32), np.asarray(z, F32)], axis=1)


def diverse_state(n):
    """tests/integration.rs:31-61 (distinct values in every field)."""
    i = np.arange(n, dtype=F32)
    pos = columns(F32(1.0) + i, F32(2.0) + i, F32(3.0) + i)
    vel = columns(F32(0.1) * (i + F32(1.0)), F32(-0.2) * (i + F32(1.0)),
                  F32(0.05) * (i + F32(1.0)))
    mass = F32(1.0) + F32(0.5) * i
    force = columns(F32(0.5) + i * F32(0.1), F32(-0.3) + i * F32(0.2), F32(0.7) - i * F32(0.05))
    return state(pos, vel, mass, force)


def random_state(n, seed, box=40.0, compensated_scale=1.0):
    """Seeded state in a cubic box with speeds that carry some atoms across
    the boundary within a few steps, and masses over two decades."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5 * box, 0.5 * box, size=(n, 3)).astype(F32)
    vel = (rng.normal(size=(n, 3)) * 3.0 * compensated_scale).astype(F32)
    force = (rng.normal(size=(n, 3)) * 50.0).astype(F32)
    mass = np.exp(rng.uniform(0.0, np.log(100.0), size=n)).astype(F32)
    return state(pos, vel, mass, force, lattice=[box, box, box, 0.0, 0.0, 0.0])


def triclinic_state(n, seed):
    rng = np.random.default_rng(seed)
    lat = [30.0, 28.0, 26.0, 5.0, -4.0, 3.0]
    pos = rng.uniform(-13.0, 13.0, size=(n, 3)).astype(F32)
    vel = (rng.normal(size=(n, 3)) * 8.0).astype(F32)
    force = (rng.normal(size=(n, 3)) * 20.0).astype(F32)
    mass = rng.uniform(1.0, 20.0, size=n).astype(F32)
    return state(pos, vel, mass, force, lattice=lat)


def cases():
    """(name, state, [(op, dt)], compensated)."""
    out = []
    # tests/integration.rs:76-107, :109-146, :148-177
    s = state([[1.0, 2.0, 3.0]], [[0.5, -0.25, 0.125]], [1.0])
    out.append(("zero_force_advance", s, [("kd", 0.1)], False))
    out.append(("zero_force_kick", s, [("k", 0.1)], False))
    s = state([[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0]], [1.0], force=[[2.0, -4.0, 1.0]])
    out.append(("constant_force_half_step", s, [("kd", 0.1)], False))
    # :179-219, :221-249, :251-290
    s = state([[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0]], [1.0], force=[[2.0, 0.0, 0.0]])
    out.append(("full_step_kinematics", s, [("kd", 0.1), ("k", 0.1)], False))
    s = state(np.zeros((2, 3)), np.zeros((2, 3)), [1.0, 4.0], force=[[1.0, 0, 0], [1.0, 0, 0]])
    out.append(("mass_scaling", s, [("kd", 0.2)], False))
    s = state([[0.0, 0, 0], [1.0, 0, 0], [-2.0, 0, 0]], np.zeros((3, 3)), [1.0, 1.0, 1.0],
              force=[[1.0, 0, 0], [-2.0, 0, 0], [0.5, 0, 0]])
    out.append(("independent_particles", s, [("kd", 0.1), ("k", 0.1)], False))
    # :292-339 dt = 0; :390-434 block non-aligned; :474-502 repeat
    out.append(("dt_zero_kick_drift", diverse_state(8), [("kd", 0.0)], False))
    out.append(("dt_zero_kick", diverse_state(8), [("k", 0.0)], False))
    out.append(("block_non_aligned", diverse_state(1000), [("kd", 0.01), ("k", 0.01)], False))
    # :504-537 NaN force
    s = state([[1.0, 2.0, 3.0]], [[0.0, 0.0, 0.0]], [1.0], force=[[np.nan, 0.0, 0.0]])
    out.append(("nan_force", s, [("kd", 0.1)], False))
    # :539-582 forward then negated dt
    s = state(columns([1.0, -2.0, 3.5, 0.25], [0.5, 1.25, -0.75, 2.0], [-1.5, 0.0, 0.5, -0.25]),
              columns([0.1, -0.2, 0.05, 0.3], [0.4, 0.05, -0.15, 0.0], [-0.1, 0.2, 0.1, -0.05]),
              [1.0] * 4)
    out.append(("forward_then_negated_dt", s, [("kd", 0.1), ("kd", -0.1)], False))
    # :584-667 wraps
    out.append(("wrap_positive_x", state([[4.9, 0, 0]], [[2.0, 0, 0]], [1.0], lattice=BOX10),
                 [("kd", 0.1)], False))
    out.append(("wrap_negative_x", state([[-4.9, 0, 0]], [[-2.0, 0, 0]], [1.0], lattice=BOX10),
                 [("kd", 0.1)], False))
    out.append(("multi_period_crossing", state([[0.0, 0, 0]], [[250.0, 0, 0]], [1.0],
                                               lattice=BOX10), [("kd", 0.1)], False))
    # tests/integration_lossless.rs:198-223 (x += 0.1 exactly), :301-337,
    # :339-372 round trips
    s = state(np.zeros((1000, 3)), columns(np.ones(1000), np.zeros(1000), np.zeros(1000)),
              np.ones(1000))
    out.append(("lossless_block_non_aligned", s, [("kd", 0.1)], True))
    i = np.arange(8, dtype=F32)
    s = state(columns(F32(0.5) + i * F32(0.3), F32(-1.0) + i * F32(0.2), F32(0.7) + i * F32(0.15)),
              columns(F32(0.1) * (i + 1), F32(-0.05) * (i + 1), F32(0.07) * (i + 1)), np.ones(8))
    out.append(("lossless_single_round_trip", s,
                [("kd", 0.1), ("neg", 0), ("kd", 0.1), ("neg", 0)], True))
    ops = [("kd", 0.001), ("k", 0.001)] * 50
    out.append(("lossless_multi_round_trip", diverse_state(16),
                ops + [("neg", 0)] + ops + [("neg", 0)], True))
    # seeded states: boundary crossings, cubic and triclinic, both variants
    for comp in (False, True):
        tag = "lossless" if comp else "f32"
        out.append((f"random_cubic_{tag}", random_state(4099, 11), [("kd", 0.05), ("k", 0.05)] * 20,
                    comp))
        out.append((f"random_triclinic_{tag}", triclinic_state(2051, 12),
                    [("kd", 0.05), ("k", 0.05)] * 10, comp))
    return out


def langevin_cases():
    """(name, state, ops) for the BAOAB A / O steps; ops are ("A", dt) and
    ("O", seed, alpha, kT, draw counter).  States follow tests/langevin.rs
    (one_particle_state :40-55, n_particle_state :57-75, the OU variance case
    :289-327, the +L/2 wrap :641-664) plus a seeded box."""
    out = []
    one = state([[1.0, 2.0, 3.0]], [[0.5, -0.25, 0.125]], [1.0])
    out.append(("lan_drift_half_kat", one, [("A", 0.1)]))
    out.append(("lan_drift_half_dt_zero", one, [("A", 0.0)]))
    out.append(("lan_ou_identity", one, [("O", 42, 1.0, 0.0, 7)]))
    i = np.arange(64, dtype=F32)
    nps = state(columns(i * F32(0.1), i * F32(0.1), i * F32(0.1)),
                columns(i * F32(0.05), i * F32(0.05), i * F32(0.05)), np.ones(64))
    out.append(("lan_ou_n64_seed42", nps, [("O", 42, 0.5, 1.0e-21, 7)]))
    out.append(("lan_ou_n64_kt1", nps, [("O", 1, 0.5, 1.0, 1)]))
    heavy = state(np.zeros((10000, 3)), np.zeros((10000, 3)), np.full(10000, 6.6335e-26))
    out.append(("lan_ou_variance", heavy, [("O", 42, 0.5, float(F32(1.380649e-23 * 300.0)), 1)]))
    out.append(("lan_drift_half_wrap", state([[4.95, 0, 0]], [[2.0, 0, 0]], [1.0], lattice=BOX10),
                [("A", 0.1)]))
    rs = random_state(4099, 21)
    steps = []
    for k in range(5):
        steps += [("A", 0.05), ("O", 2**40 + 7, 0.9, 0.3, 2**33 + k), ("A", 0.05)]
    out.append(("lan_baoa_random", rs, steps))
    return out
