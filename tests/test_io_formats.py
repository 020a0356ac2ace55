"""Output formats (CPU): the extended-XYZ trajectory writer and the CSV log
writer through the C ABI, against the strings the reference's own tests
expect (tests/io_log_output.rs, tests/io_trajectory.rs) and the format rule
they encode: Rust `{:.9e}` -- the exact binary value rounded once to ten
significant digits, exponent without '+' or leading zeros."""
import numpy as np
import pytest

from heddlemd_amd import api
from heddlemd_amd._lib import HMD_ERR_OUTPUT_EXISTS, HmdError

SI_TIME = 2.4188843265857195e-17
SI_ENERGY = 4.359744722207101e-18
SI_TEMPERATURE = 315775.0248040668
SI_LENGTH = 5.29177210903e-11
SI_VELOCITY = 2187691.2636411153


def read(path):
    with open(path) as f:
        return f.read()


def test_log_header_and_step_zero_row(tmp_path):
    # tests/io_log_output.rs:74-89
    p = tmp_path / "run.log"
    w = api.LogWriter(p, "atomic")
    w.write_row(0, 0.0, 0.0, 300.0)
    w.flush()
    assert read(p) == ("step,time,kinetic_energy,temperature\n"
                       "0,0.000000000e0,0.000000000e0,3.000000000e2\n")
    w.close()


def test_log_non_trivial_row(tmp_path):
    # tests/io_log_output.rs:131-149
    p = tmp_path / "run.log"
    w = api.LogWriter(p, "atomic")
    w.write_row(100, 1.0e-13, 4.123456789e-21, 298.7654321)
    w.close()
    assert read(p).splitlines()[-1] == "100,1.000000000e-13,4.123456789e-21,2.987654321e2"


def test_log_extra_column(tmp_path):
    # tests/io_log_output.rs:270-302
    p = tmp_path / "run.log"
    w = api.LogWriter(p, "atomic", [("nhc_conserved", "energy")])
    w.write_row(100, 1.0e-13, 4.0e-21, 298.0, [5.5e-20])
    w.close()
    lines = read(p).splitlines()
    assert lines[0] == "step,time,kinetic_energy,temperature,nhc_conserved"
    assert lines[1] == "100,1.000000000e-13,4.000000000e-21,2.980000000e2,5.500000000e-20"
    with pytest.raises(ValueError):
        api.LogWriter(tmp_path / "b.log", "atomic", [("x", "energy")]).write_row(0, 0, 0, 0)


def test_log_si_factors(tmp_path):
    # tests/io_log_output.rs:92-128
    p = tmp_path / "run.log"
    w = api.LogWriter(p, "si")
    w.write_row(100, 41.34, 0.0128, 9.5e-4)
    w.close()
    cols = read(p).splitlines()[-1].split(",")
    assert cols[0] == "100"
    for got, want in zip(cols[1:], (41.34 * SI_TIME, 0.0128 * SI_ENERGY, 9.5e-4 * SI_TEMPERATURE)):
        assert float(got) == pytest.approx(want, rel=1e-9)


def test_writers_never_overwrite(tmp_path):
    # tests/io_log_output.rs:45-57, tests/io_trajectory.rs:45-57
    for cls in (api.LogWriter, api.TrajectoryWriter):
        p = tmp_path / f"{cls.__name__}.out"
        p.write_text("existing")
        with pytest.raises(HmdError) as e:
            cls(p, "atomic")
        assert e.value.code == HMD_ERR_OUTPUT_EXISTS
        assert read(p) == "existing"


def test_trajectory_frame_layout(tmp_path):
    # tests/io_trajectory.rs:70-100 and :102-135
    p = tmp_path / "traj.xyz"
    w = api.TrajectoryWriter(p, "atomic", type_names=["Ar"])
    lat = [18.9, 18.9, 18.9, 0, 0, 0]
    w.write_frame(0, 1.0e-15, lat, [0, 0], [[0.1, 0.0, 0.0], [-1.5, 2.0, 300.0]])
    w.close()
    lines = read(p).splitlines()
    assert len(lines) == 4 and lines[0] == "2"
    # f32(18.9) = 18.8999996185302734375, f32(0.1) = 0.100000001490116119384765625
    assert lines[1] == ('Lattice="1.889999962e1 0.000000000e0 0.000000000e0 0.000000000e0 '
                        '1.889999962e1 0.000000000e0 0.000000000e0 0.000000000e0 1.889999962e1" '
                        'Properties=species:S:1:pos:R:3 Step=0 Time=0.000000000e0')
    assert lines[2] == "Ar 1.000000015e-1 0.000000000e0 0.000000000e0"
    assert lines[3] == "Ar -1.500000000e0 2.000000000e0 3.000000000e2"

    p2 = tmp_path / "traj2.xyz"
    w = api.TrajectoryWriter(p2, "atomic", include_velocities=True, type_names=["Ar"])
    w.write_frame(10, 1.0e-15, lat, [0], [[0.0, 0.0, 0.0]], velocities=[[100.0, 0.0, 0.0]])
    w.close()
    lines = read(p2).splitlines()
    assert "Properties=species:S:1:pos:R:3:velo:R:3" in lines[1]
    assert "Step=10 Time=1.000000000e-14" in lines[1]      # 10 * 1e-15 in f64
    assert lines[2].split() == ["Ar", "0.000000000e0", "0.000000000e0", "0.000000000e0",
                                "1.000000000e2", "0.000000000e0", "0.000000000e0"]


def test_trajectory_images_and_species(tmp_path):
    # tests/io_trajectory.rs:452-476 and :180-212
    p = tmp_path / "traj.xyz"
    w = api.TrajectoryWriter(p, "atomic", True, True, ["Ar", "Kr"])
    w.write_frame(0, 1.0e-15, [10, 10, 10, 0, 0, 0], [0, 1, 1, 0],
                  [[0.1, 0.2, 0.3]] * 4, velocities=[[1.0, 2.0, 3.0]] * 4,
                  images=[[4, -5, 6], [0, 0, 0], [1, 1, 1], [-1, 0, 2]])
    w.close()
    lines = read(p).splitlines()
    assert "Properties=species:S:1:pos:R:3:velo:R:3:image:I:3" in lines[1]
    cols = lines[2].split()
    assert len(cols) == 10 and cols[7:] == ["4", "-5", "6"]
    assert [l.split()[0] for l in lines[2:]] == ["Ar", "Kr", "Kr", "Ar"]


def test_trajectory_si_units_and_triclinic_lattice(tmp_path):
    # tests/io_trajectory.rs:215-302: values times the f32 factor; lattice in
    # lower-triangular row-major order with the upper slots zero
    p = tmp_path / "traj.xyz"
    w = api.TrajectoryWriter(p, "si", include_velocities=True, type_names=["Ar"])
    lat = [20.0, 21.0, 22.0, 1.5, -2.5, 3.5]
    w.write_frame(3, 41.0, lat, [0], [[1.25, -2.5, 3.0]], velocities=[[1e-4, 2e-4, -3e-4]])
    w.close()
    lines = read(p).splitlines()
    vals = lines[1].split('"')[1].split()
    f32 = np.float32
    lf = f32(SI_LENGTH)
    want = [f32(20.0) * lf, 0.0, 0.0, f32(1.5) * lf, f32(21.0) * lf,
            0.0, f32(-2.5) * lf, f32(3.5) * lf, f32(22.0) * lf]
    assert vals == [rust_exp(float(v)) for v in want]
    assert f"Step=3 Time={rust_exp(3 * 41.0 * SI_TIME)}" in lines[1]
    cols = lines[2].split()
    assert cols[1] == rust_exp(float(f32(1.25) * lf))
    assert cols[4] == rust_exp(float(f32(1e-4) * f32(SI_VELOCITY)))


def rust_exp(v, digits=9):
    """Reference formatting rule restated independently of the writer: the
    correctly rounded decimal of the exact value (Python's %e) with the
    exponent rewritten."""
    s = f"{v:.{digits}e}"
    m, e = s.split("e")
    return f"{m}e{int(e)}"


def test_format_rule_on_random_f32(tmp_path):
    """Every value of a random f32 sample (magnitudes 1e-30 .. 1e30, signed)
    against the rule: correctly rounded to ten significant digits."""
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(300) * 10.0 ** rng.uniform(-30, 30, 300)).astype(np.float32)
    x[:3] = [0.0, -0.0, 1.0]
    p = tmp_path / "traj.xyz"
    w = api.TrajectoryWriter(p, "atomic", type_names=["X"])
    pos = np.stack([x, x, x], axis=1)
    w.write_frame(0, 0.0, [1, 1, 1, 0, 0, 0], np.zeros(len(x), np.uint32), pos)
    w.close()
    got = [l.split()[1] for l in read(p).splitlines()[2:]]
    assert got[1] == "-0.000000000e0"
    assert got == [rust_exp(float(v)) for v in x]


# ---- .timings report (src/timings.rs:513-553) -------------------------------

def test_timings_header_only_when_empty(tmp_path):
    """tests/timings.rs:851-862: an empty report is the header line alone."""
    p = tmp_path / "run.timings"
    api.write_timings(p, [])
    lines = read(p).splitlines()
    assert len(lines) == 1
    assert lines[0].split() == ["stage", "count", "total_ms", "mean_us", "min_us", "max_us"]
    assert lines[0] == ("stage" + " " * 23 + " " + "count".rjust(10) + " " + "total_ms".rjust(14)
                        + " " + "mean_us".rjust(13) + " " + "min_us".rjust(11) + " "
                        + "max_us".rjust(11))


def test_timings_rows_widths_and_precision(tmp_path):
    """Row layout {:<28} {:>10} {:>14.3} {:>13.1} {:>11.1} {:>11.1}; the
    reference's accumulator example (tests/timings.rs:925-939: 3 calls of
    50, 100 and 200 us) and a zero-count stage, which is left out."""
    p = tmp_path / "run.timings"
    api.write_timings(p, [
        {"name": "jit_composed_pair_force", "count": 3, "total_ns": 350_000,
         "min_ns": 50_000, "max_ns": 200_000},
        {"name": "log_write", "count": 0, "total_ns": 0, "min_ns": 0, "max_ns": 0},
        {"name": "vv_kick", "count": 10, "total_ns": 1_234_567, "min_ns": 1_250,
         "max_ns": 987_654_321},
    ])
    lines = read(p).splitlines()
    assert len(lines) == 3
    assert lines[1] == ("jit_composed_pair_force".ljust(28) + " " + "3".rjust(10) + " "
                        + "0.350".rjust(14) + " " + "116.7".rjust(13) + " " + "50.0".rjust(11)
                        + " " + "200.0".rjust(11))
    # 1.25 us is a tie at one decimal: round half to even gives 1.2
    assert lines[2].split() == ["vv_kick", "10", "1.235", "123.5", "1.2", "987654.3"]
    for line in lines[1:]:
        cols = line.split()
        assert [len(c.split(".")[1]) for c in cols[2:]] == [3, 1, 1, 1]


def test_timings_create_new_and_missing_parent(tmp_path):
    """tests/timings.rs:824-849: an existing file is kept and reported; a
    missing parent directory is an I/O error."""
    p = tmp_path / "run.timings"
    p.write_text("existing")
    with pytest.raises(HmdError) as e:
        api.write_timings(p, [])
    assert e.value.code == HMD_ERR_OUTPUT_EXISTS
    assert p.read_text() == "existing"
    with pytest.raises(HmdError):
        api.write_timings(tmp_path / "missing" / "run.timings", [])
