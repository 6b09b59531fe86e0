"""Run-record formats (io.jl:38-180): the conservation CSV layout and its
migration, truncation and time lookup, checkpoint naming and the portable
checkpoint.  The cases follow the reference's own tests of these helpers
(test/integration/io_helpers.jl:12-161); rclone mirroring is not part of the
library.
"""
import os

import numpy as np
import pytest

import collop.records as R


def legacy_csv(path, steps_dt, with_r0=True, repeat_steps=()):
    """A conservation CSV in the layout before the ``dt`` column (time = step * DT)."""
    cols = ["step", "time", "entropy", "energy", "momentum_1", "momentum_2", "iter", "residual", "fp_minus_fs",
            "neg_part"]
    if with_r0:
        cols.append("r0")
    with open(path, "w") as f:
        f.write(",".join(cols) + "\n")
        for step, dt in steps_dt:
            row = [str(step), repr(step * dt), "1.0", "2.0", "0.1", "0.2", "7", "1e-12", "0.5", "0.25"]
            if with_r0:
                row.append("0.75")
            f.write(",".join(row) + "\n")
            if step in repeat_steps:
                f.write(",".join(row) + "\n")
    return path


def read_col(path, name):
    with open(path) as f:
        lines = f.read().splitlines()
    idx = lines[0].split(",").index(name)
    return np.array([float(ln.split(",")[idx]) for ln in lines[1:]])


def test_cons_cols_layout():
    assert R.CONS_COLS[0] == "step" and R.CONS_COLS[1] == "time" and R.CONS_COLS[-1] == "dt"
    assert len(set(R.CONS_COLS)) == len(R.CONS_COLS)
    assert R.CONS_COLS[:10] == ["step", "time", "entropy", "energy", "momentum_1", "momentum_2", "iter", "residual",
                                "fp_minus_fs", "neg_part"]
    assert R.STATS_COLS == ["step", "iter", "inner", "t_solve", "t_step"]


def test_migrate_constant_dt_and_idempotent(tmp_path):
    f = legacy_csv(tmp_path / "const.csv", [(s, 0.001) for s in range(11)])
    R.migrate_cons_csv(str(f))
    lines = f.read_text().splitlines()
    assert lines[0].split(",") == R.CONS_COLS
    assert all(len(ln.split(",")) == len(R.CONS_COLS) for ln in lines[1:])
    np.testing.assert_allclose(read_col(f, "time"), [s * 0.001 for s in range(11)], rtol=1e-12)
    np.testing.assert_allclose(read_col(f, "dt"), [0.0] + [0.001] * 10, rtol=1e-12)
    np.testing.assert_allclose(np.cumsum(read_col(f, "dt")), read_col(f, "time"), rtol=1e-12)
    before = f.read_text()
    R.migrate_cons_csv(str(f))
    assert f.read_text() == before


def test_migrate_recovers_dt_change(tmp_path):
    steps = [(s, 0.001) for s in range(1001)] + [(s, 0.002) for s in range(1001, 1501)]
    f = legacy_csv(tmp_path / "dtchange.csv", steps)
    R.migrate_cons_csv(str(f))
    t, dt, st = read_col(f, "time"), read_col(f, "dt"), read_col(f, "step")
    assert abs(t[np.flatnonzero(st == 1000.0)[0]] - 1.0) < 1e-9
    assert abs(t[-1] - 2.0) < 1e-9                     # not the 3.0 the old layout stored
    assert abs(dt[-1] - 0.002) < 1e-15
    np.testing.assert_allclose(np.cumsum(dt), t, rtol=1e-9)


def test_migrate_without_r0(tmp_path):
    f = legacy_csv(tmp_path / "nor0.csv", [(s, 0.001) for s in range(6)], with_r0=False)
    R.migrate_cons_csv(str(f))
    lines = f.read_text().splitlines()
    assert lines[0].split(",") == R.CONS_COLS
    assert all(len(ln.split(",")) == len(R.CONS_COLS) for ln in lines[1:])
    assert np.all(read_col(f, "r0") == 0.0)


def test_migrate_repeated_steps(tmp_path):
    f = legacy_csv(tmp_path / "dup.csv", [(s, 0.001) for s in range(11)], repeat_steps=(3, 4, 5))
    R.migrate_cons_csv(str(f))
    st, t = read_col(f, "step"), read_col(f, "time")
    assert len(st) == 14
    np.testing.assert_allclose(t, st * 0.001, rtol=1e-12)


def test_migrate_edge_cases(tmp_path):
    assert R.migrate_cons_csv(str(tmp_path / "absent.csv")) is None
    empty = tmp_path / "empty.csv"
    empty.write_text("")
    assert R.migrate_cons_csv(str(empty)) is None
    hdr = tmp_path / "hdr.csv"
    hdr.write_text(",".join(R.CONS_COLS[:11]) + "\n")
    assert R.migrate_cons_csv(str(hdr)) is None
    assert hdr.read_text() == ",".join(R.CONS_COLS[:11]) + "\n"


def test_cons_time_at(tmp_path):
    f = legacy_csv(tmp_path / "t.csv", [(s, 0.004) for s in range(21)])
    R.migrate_cons_csv(str(f))
    assert abs(R.cons_time_at(str(f), 0)) < 1e-15
    assert abs(R.cons_time_at(str(f), 20) - 0.08) < 1e-12
    assert R.cons_time_at(str(f), 999) is None
    assert R.cons_time_at(str(tmp_path / "absent.csv"), 1) is None


def test_truncate_csv_after(tmp_path):
    f = legacy_csv(tmp_path / "trunc.csv", [(s, 0.001) for s in range(21)])
    R.truncate_csv_after(str(f), 12)
    assert list(read_col(f, "step")) == list(np.arange(13.0))
    assert f.read_text().splitlines()[0] == ",".join(R.CONS_COLS[:11])
    R.truncate_csv_after(str(f), 100)
    assert list(read_col(f, "step")) == list(np.arange(13.0))
    assert R.truncate_csv_after(str(tmp_path / "absent.csv"), 3) is None


def test_checkpoint_path():
    assert R.checkpoint_path("run", 0) == "checkpoint_run_step0000.npz"
    assert R.checkpoint_path("run", 25) == "checkpoint_run_step0025.npz"
    assert R.checkpoint_path("run", 20233) == "checkpoint_run_step20233.npz"
    names = [R.checkpoint_path("r", s) for s in (0, 25, 1500)]
    assert sorted(names) == names


def test_checkpoint_round_trip_and_auto(tmp_path):
    rng = np.random.default_rng(3)
    v = np.asfortranarray(rng.standard_normal((17, 2)))
    w = np.full(17, 1 / 17)
    fc = rng.standard_normal(36)
    h = {"entropy_history": [1.0, 1.1], "energy_history": [2.0, 2.0], "momentum_history": [(0.1, 0.2), (0.1, 0.2)],
         "iter_history": [5], "res_history": [1e-12], "fp_l2_history": [0.3, 0.31], "neg_history": [0.0, 1e-4]}
    for step in (0, 25, 10500):            # 10500 sorts before 25 by name: the choice parses the step
        R.save_checkpoint("bim", step, 0.02 * step, v, w, fc, h, {"seed": 42}, str(tmp_path))
    R.save_checkpoint("other", 99999, 0.0, v, w, fc, None, None, str(tmp_path))
    c = R.load_checkpoint("bim", "auto", str(tmp_path))
    assert c["step"] == 10500 and abs(c["t"] - 210.0) < 1e-12
    assert R.load_checkpoint("bim", 0, str(tmp_path))["step"] == 10500
    c = R.load_checkpoint("bim", 25, str(tmp_path))
    assert c["step"] == 25
    np.testing.assert_array_equal(c["v_particles"], v)
    np.testing.assert_array_equal(c["w_particles"], w)
    np.testing.assert_array_equal(c["f_coeffs"], fc)
    assert c["momentum_history"].shape == (2, 2) and c["iter_history"].dtype == np.int64
    assert c["rng_state"] == {"seed": 42}
    assert not any(p.endswith(".tmp.npz") for p in os.listdir(tmp_path))
    with pytest.raises(FileNotFoundError):
        R.load_checkpoint("bim", 26, str(tmp_path))
    with pytest.raises(FileNotFoundError):
        R.load_checkpoint("none", "auto", str(tmp_path))


def test_run_record_rows_and_resume(tmp_path):
    rec = R.RunRecord("bim", str(tmp_path))
    rec.write_initial(1.5, 2.0, (0.0, 0.0))
    st = {"entropy": 1.6, "energy": 2.0, "momentum": (1e-17, -2e-17), "iters": 9, "residual": 3e-13, "r0": 0.5,
          "n_restart": 0, "inner_maps": 44}
    for k in (1, 2, 3):
        rec.write_step(k, 0.02 * k, st, {"fp_minus_fs": 0.31, "neg_part": 1e-5}, 0.02, 0.1, 0.12)
    cons = tmp_path / "conservation_history_bim.csv"
    lines = cons.read_text().splitlines()
    assert lines[0] == ",".join(R.CONS_COLS) and len(lines) == 5
    # the step-0 row as main.jl:191-194 writes it, byte for byte
    assert lines[1] == "0,0.0,1.5,2.0,0.0,0.0,0,0.0,0.0,0.0,0.0,0.0"
    assert lines[2].split(",") == ["1", "0.02", "1.6", "2.0", "1.0e-17", "-2.0e-17", "9", "3.0e-13", "0.31",
                                   "1.0e-5", "0.5", "0.02"]
    stats = (tmp_path / "solver_stats_bim.csv").read_text().splitlines()
    assert stats[0] == ",".join(R.STATS_COLS) and stats[1].split(",")[:3] == ["1", "9", "44"]
    # resume from step 2: rows past it go, the next row appends
    rec2 = R.RunRecord("bim", str(tmp_path), start_step=2)
    rec2.write_step(3, 0.06, st, {"fp_minus_fs": 0.32, "neg_part": 0.0}, 0.02, 0.1, 0.12)
    assert list(read_col(cons, "step")) == [0.0, 1.0, 2.0, 3.0]
    assert read_col(cons, "fp_minus_fs")[-1] == 0.32
    assert list(read_col(tmp_path / "solver_stats_bim.csv", "step")) == [1.0, 2.0, 3.0]


def test_number_text_follows_julia():
    """Julia's string interpolation of Float64 / Int: shortest round-trip digits,
    plain decimal in [1e-4, 1e6) with a fractional digit, else d.ddde+-x without
    '+' or padding."""
    cases = {0.0: "0.0", 1.0: "1.0", -2.5: "-2.5", 0.02: "0.02", 0.0001: "0.0001", 1e-5: "1.0e-5",
             123456.0: "123456.0", 1e6: "1.0e6", 1.5e-7: "1.5e-7", 1e16: "1.0e16", -3.2e-12: "-3.2e-12",
             0.1 + 0.2: "0.30000000000000004", 2.9999999999999995e-10: "2.9999999999999995e-10"}
    for x, t in cases.items():
        assert R._num(x) == t, (x, R._num(x), t)
    assert R._num(7) == "7" and R._num(np.int64(-3)) == "-3"
    for x in (1e-5, 3.14159e8, 0.30000000000000004, 6.02e23):
        assert float(R._num(x)) == x
