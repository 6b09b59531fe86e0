"""Run records: the conservation-history and solver-statistics CSVs and a
portable checkpoint (io.jl:38-180, the per-step rows of main.jl:296-346).

The CSV columns are the reference's (io.jl:104-107), so its dashboards and
readers take these files unchanged.  The checkpoint replaces the reference's
Julia ``Serialization`` file (io.jl:66-86), which only Julia can read, by an
``.npz`` archive: plain ``.npy`` arrays in a zip, loadable without pickles from
numpy, Julia (NPZ.jl) or any zip + npy reader.  The state fields are the same:
step, t, markers, weights, projection coefficients and the seven histories.
"""
from __future__ import annotations

import glob
import json
import os
import re
import time

import numpy as np

# io.jl:104 (solver_stats_<suffix>.csv)
STATS_COLS = ["step", "iter", "inner", "t_solve", "t_step"]

# io.jl:106-107 (conservation_history_<suffix>.csv)
CONS_COLS = ["step", "time", "entropy", "energy", "momentum_1", "momentum_2", "iter", "residual", "fp_minus_fs",
             "neg_part", "r0", "dt"]

_HISTORIES = ("entropy_history", "energy_history", "momentum_history", "iter_history", "res_history",
              "fp_l2_history", "neg_history")


def _num(x) -> str:
    """Text of a number as Julia's string interpolation writes it: integers
    plainly; a Float64 as its shortest round-trip digits, in plain decimal with
    at least one fractional digit (``1.0``, ``0.0001``, ``123456.0``) when
    1e-4 <= |x| < 1e6 or x == 0, and otherwise as ``d.ddde-x`` / ``d.ddde+x``
    with Julia's exponent style (``1.0e-5``, ``1.0e16``: no ``+``, no padding)."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if not np.isfinite(x):
        return "NaN" if np.isnan(x) else ("Inf" if x > 0 else "-Inf")
    r = repr(x)
    if x == 0.0 or 1e-4 <= abs(x) < 1e6:
        if "e" in r:  # repr switches to exponents outside [1e-4, 1e16)
            r = np.format_float_positional(x, unique=True, trim="0")
        return r if ("." in r or "e" in r) else r + ".0"
    m, e = np.format_float_scientific(x, unique=True, trim="0").split("e")
    if "." not in m:
        m += ".0"
    elif m.endswith("."):
        m += "0"
    return f"{m}e{int(e)}"


def _read_lines(fname: str) -> list[str]:
    with open(fname) as f:
        return f.read().splitlines()


def _step_of(line: str):
    try:
        return int(line.split(",", 1)[0])
    except ValueError:
        return None


def truncate_csv_after(fname: str, step: int) -> None:
    """Drop the rows past ``step`` (column 1), keeping the header (io.jl:88-97)."""
    if not os.path.isfile(fname):
        return None
    lines = _read_lines(fname)
    keep = [ln for ln in lines if (_step_of(ln) if _step_of(ln) is not None else -1) <= step]
    if len(keep) != len(lines):
        with open(fname, "w") as f:
            f.write("\n".join(keep) + "\n")
    return None


def migrate_cons_csv(fname: str) -> None:
    """Bring a conservation CSV without the ``dt`` column up to CONS_COLS
    (io.jl:109-150): missing columns become 0.0, each step's DT is recovered as
    time / step from the legacy ``time = step * DT`` (the last row of a repeated
    step wins), and ``time`` becomes the accumulated physical time.  A no-op on
    a file that has ``dt``."""
    if not os.path.isfile(fname):
        return None
    lines = _read_lines(fname)
    if not lines:
        return None
    old = lines[0].split(",")
    if "dt" in old:
        return None
    rows = [dict(zip(old, ln.split(","))) for ln in lines[1:]]
    step_dt: dict[int, float] = {}
    for r in rows:
        try:
            k = int(r.get("step", ""))
            tk = float(r.get("time", ""))
        except ValueError:
            continue
        if k >= 0:
            step_dt[k] = tk / k if k else 0.0
    if not step_dt:
        return None
    step_time = dict(zip(range(max(step_dt) + 1),
                         np.cumsum([step_dt.get(k, 0.0) for k in range(max(step_dt) + 1)]).tolist()))
    upgraded = fname + ".upgrade.tmp"
    with open(upgraded, "w") as f:
        f.write(",".join(CONS_COLS) + "\n")
        for r in rows:
            try:
                k = int(r.get("step", ""))
            except ValueError:
                continue
            if k in step_time:
                r["time"], r["dt"] = _num(step_time[k]), _num(step_dt[k])
                f.write(",".join(r.get(c, "0.0") for c in CONS_COLS) + "\n")
    os.replace(upgraded, fname)
    print(f"{fname}: conservation CSV upgraded from {len(old)} to {len(CONS_COLS)} columns "
          f"(accumulated time and per-step dt)")
    return None


def cons_time_at(fname: str, step: int):
    """Accumulated time recorded for ``step`` (last row wins), or None (io.jl:152-163)."""
    if not os.path.isfile(fname):
        return None
    t = None
    for ln in _read_lines(fname)[1:]:
        f = ln.split(",")
        if len(f) >= 2 and _step_of(ln) == step:
            try:
                t = float(f[1])
            except ValueError:
                pass
    return t


# ---- checkpoints ---------------------------------------------------------------

def checkpoint_path(suffix: str, step: int) -> str:
    """``checkpoint_<suffix>_step<NNNN>.npz`` (io.jl:66-68, with the portable suffix)."""
    return f"checkpoint_{suffix}_step{step:04d}.npz"


def save_checkpoint(suffix: str, step: int, t: float, v, w, f_coeffs, histories: dict | None = None,
                    rng_state: dict | None = None, directory: str = ".") -> str:
    """Write the state of io.jl:70-86 (save_checkpoint) as an ``.npz``.

    ``histories`` maps the reference's names (entropy_history, energy_history,
    momentum_history, iter_history, res_history, fp_l2_history, neg_history) to
    sequences; ``rng_state`` is any JSON-serialisable description of the marker
    generator (kept, like the reference's, for forward compatibility)."""
    v = np.asarray(v, dtype=np.float64)
    if v.ndim != 2 or v.shape[1] != 2:
        raise ValueError(f"v must be (N, 2), got {v.shape}")
    arrays = {
        "step": np.int64(step), "t": np.float64(t),
        "v_particles": v, "w_particles": np.asarray(w, dtype=np.float64),
        "f_coeffs": np.asarray(f_coeffs, dtype=np.float64),
        "rng_state": np.frombuffer(json.dumps(rng_state).encode(), dtype=np.uint8),
    }
    h = histories or {}
    for name in _HISTORIES:
        dt = np.int64 if name == "iter_history" else np.float64
        a = np.asarray(h.get(name, []), dtype=dt)
        if name == "momentum_history":
            a = a.reshape(-1, 2)
        arrays[name] = a
    fname = os.path.join(directory, checkpoint_path(suffix, step))
    tmp = fname + ".tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, fname)
    print(f"Saved {fname}")
    return fname


def load_checkpoint(suffix: str, step="auto", directory: str = ".") -> dict:
    """Read a checkpoint (io.jl:165-180).  ``step="auto"`` or a step <= 0 picks
    the highest-step file for the suffix; the step is parsed from the name, so
    the choice does not depend on the width of the padding."""
    if step == "auto" or (isinstance(step, (int, np.integer)) and step <= 0):
        pat = re.compile(rf"^checkpoint_{re.escape(suffix)}_step(\d+)\.npz$")
        found = [(int(mt.group(1)), f) for f in glob.glob(os.path.join(directory, "checkpoint_*.npz"))
                 if (mt := pat.match(os.path.basename(f)))]
        if not found:
            raise FileNotFoundError(f"No checkpoint files for suffix={suffix}")
        fname = max(found)[1]
    else:
        fname = os.path.join(directory, checkpoint_path(suffix, int(step)))
    if not os.path.isfile(fname):
        raise FileNotFoundError(f"Checkpoint not found: {fname}")
    print(f"Loading checkpoint: {fname}")
    with np.load(fname, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    out["step"] = int(out["step"])
    out["t"] = float(out["t"])
    out["rng_state"] = json.loads(out["rng_state"].tobytes().decode())
    return out


# ---- the per-step record of a run ------------------------------------------------

class RunRecord:
    """The conservation-history and solver-statistics CSVs of one run
    (main.jl:44-48,186-222,296-346).  A fresh record writes the headers; a
    resumed one (``start_step`` given) migrates the conservation CSV, drops rows
    past ``start_step`` and appends."""

    def __init__(self, suffix: str, directory: str = ".", start_step: int | None = None):
        self.suffix = suffix
        self.cons_csv = os.path.join(directory, f"conservation_history_{suffix}.csv")
        self.stats_csv = os.path.join(directory, f"solver_stats_{suffix}.csv")
        if start_step is None:
            for fn, cols in ((self.cons_csv, CONS_COLS), (self.stats_csv, STATS_COLS)):
                with open(fn, "w") as f:
                    f.write(",".join(cols) + "\n")
        else:
            migrate_cons_csv(self.cons_csv)
            for fn, cols in ((self.cons_csv, CONS_COLS), (self.stats_csv, STATS_COLS)):
                if os.path.isfile(fn):
                    truncate_csv_after(fn, start_step)
                else:
                    with open(fn, "w") as f:
                        f.write(",".join(cols) + "\n")

    def write_initial(self, entropy: float, energy: float, momentum):
        """Row of step 0: the initial state, with zeros in the diagnostic, residual
        and dt columns as the reference writes it (main.jl:189-194)."""
        self._cons_row(0, 0.0, entropy, energy, momentum, 0, 0.0, 0.0, 0.0, 0.0, 0.0)

    def write_step(self, step: int, t: float, stats: dict, diag: dict, dt: float, t_solve: float, t_step: float):
        """Rows of one step from the device step's statistics and diagnostics."""
        self._cons_row(step, t, stats["entropy"], stats["energy"], stats["momentum"], stats["iters"],
                       stats["residual"], diag["fp_minus_fs"], diag["neg_part"], stats["r0"], dt)
        with open(self.stats_csv, "a") as f:
            # the third column is step_defect!'s inner-map count, or
            # step_anderson!'s restart count (main.jl:256-294)
            inner = stats.get("inner_maps", 0) or stats.get("n_restart", 0)
            f.write(",".join(_num(x) for x in (step, stats["iters"], inner, t_solve, t_step)) + "\n")

    def _cons_row(self, step, t, S, E, P, it, res, fp, neg, r0, dt):
        with open(self.cons_csv, "a") as f:   # appended and closed per step: crash-safe
            f.write(",".join(_num(x) for x in (step, t, S, E, P[0], P[1], int(it), res, fp, neg, r0, dt)) + "\n")


def run_recorded(sim, params, n_steps: int, record: RunRecord, step0: int = 0, t0: float = 0.0,
                 checkpoint_every: int = 0, checkpoint_dir: str = ".", histories: dict | None = None) -> list[dict]:
    """Advance ``sim`` (a resident Simulation) by n_steps, one device step at a
    time, writing each step's conservation and statistics rows (main.jl:296-346)
    and, every ``checkpoint_every`` steps and at the end, a checkpoint."""
    h = histories if histories is not None else {k: [] for k in _HISTORIES}
    out = []
    t = t0
    for k in range(1, n_steps + 1):
        step = step0 + k
        t_step0 = time.perf_counter()
        st = sim.run(params, 1)[0]
        t_solve = time.perf_counter() - t_step0
        diag = sim.diagnostics()
        t += params.dt
        record.write_step(step, t, st, diag, params.dt, t_solve, time.perf_counter() - t_step0)
        h["entropy_history"].append(st["entropy"])
        h["energy_history"].append(st["energy"])
        h["momentum_history"].append(st["momentum"])
        h["iter_history"].append(st["iters"])
        h["res_history"].append(st["residual"])
        h["fp_l2_history"].append(diag["fp_minus_fs"])
        h["neg_history"].append(diag["neg_part"])
        if checkpoint_every and (k % checkpoint_every == 0 or k == n_steps):
            save_checkpoint(record.suffix, step, t, sim.markers(), sim.w, sim.coeffs(), h, None, checkpoint_dir)
        st = dict(st, **diag, t_solve=t_solve)
        out.append(st)
    return out
