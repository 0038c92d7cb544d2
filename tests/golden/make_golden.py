"""Writes the reference's own outputs for seeded workloads as golden fixtures.

Runs the reference CPU build (oracle/_ref/ref_driver, built from the reference
sources by oracle/ref.mk) and stores its results as JSON data under
tests/golden/.  Run from the repository root after
`make -f oracle/ref.mk REF=/root/reference`:

    python tests/golden/make_golden.py [solves] [prepare] [steps] [controls] [stalls]
    (default: all but stalls, which runs the reference for about 15 minutes)
"""
import hashlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
DRIVER = ROOT / "oracle" / "_ref" / "ref_driver"
OUT = ROOT / "tests" / "golden"

SOLVES = [
    # (config, seed, scale, index, keep_xy)
    (0, 42, 1, 0, False),
    (0, 7, 4, 0, True),
    (0, 3, 20, 0, True),
    (1, 42, 400, 0, False),
    (1, 42, 200, 0, False),
    (2, 3, 1, 0, True),
    (2, 3, 1, 1, True),
    (2, 3, 1, 2, False),
    (2, 42, 1, 0, False),
    (2, 42, 1, 7, False),
    (4, 1, 200, 0, True),
    (4, 42, 200, 0, False),
]


def run(*args):
    out = subprocess.run([str(DRIVER), *map(str, args)], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def case_args(config, seed, scale, index):
    return ["--config", config, "--seed", seed, "--scale", scale, "--index", index, "--threads", 1]


def array_sha256(values, dtype):
    """sha256 of the little-endian bytes of an array printed by the driver
    (infinities arrive as "inf" / "-inf" strings)."""
    a = np.asarray([float(v) if isinstance(v, str) else v for v in values], dtype=dtype)
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<")).tobytes()).hexdigest()


PREPARES = [(0, 3, 20, 0, "ruiz"), (4, 1, 200, 0, "ruiz"), (2, 3, 1, 1, "combined"), (1, 42, 50, 0, "ruiz")]


def prepares():
    out = []
    for config, seed, scale, index, scaling in PREPARES:
        r = run("prepare", *case_args(config, seed, scale, index), "--scaling", scaling)
        keep = {k: r[k] for k in ("failure", "objective_scale", "offset", "rows", "cols")}
        if r["cols"] <= 5000:
            keep.update({k: r[k] for k in ("col_map", "row_map", "column_scale", "row_scale")})
        # the whole prepared model, as hashes of its arrays
        keep["model_sha256"] = {k: array_sha256(r[k], np.int64 if k in ("ptr", "index", "col_map", "row_map")
                                                else np.float64)
                                for k in ("ptr", "index", "value", "c", "lb", "ub", "rl", "ru", "column_scale",
                                          "row_scale", "col_map", "row_map")}
        keep.update(config=config, seed=seed, scale=scale, index=index, scaling=scaling)
        out.append(keep)
        print("prepare", config, seed, scale, index, scaling, r["rows"], r["cols"])
    (OUT / "reference_prepare.json").write_text(json.dumps(out) + "\n")


def main():
    parts = set(sys.argv[1:]) or {"solves", "prepare", "steps", "controls"}
    if "solves" in parts:
        solves()
    if "prepare" in parts:
        prepares()
    if "steps" in parts:
        steps()
    if "controls" in parts:
        controls()
    if "stalls" in parts:
        stalls()


def solves():
    solves = []
    for config, seed, scale, index, keep in SOLVES:
        r = run("solve", *case_args(config, seed, scale, index), *(["--xy"] if keep else []))
        r.update(config=config, seed=seed, scale=scale, index=index)
        solves.append(r)
        print(config, seed, scale, index, r["status"], r["iterations"], r["objective"])
    (OUT / "reference_solves.json").write_text(json.dumps(solves, indent=1) + "\n")


def steps():
    steps = []
    for config, seed, scale, index, adaptive in [(2, 3, 1, 0, True), (0, 3, 20, 0, True), (0, 3, 20, 0, False)]:
        args = ["steps", *case_args(config, seed, scale, index), "--count", 10, "--chunks", 3,
                "--tau", 0.05, "--sigma", 0.05]
        if adaptive:
            args.append("--adaptive")
        r = run(*args)
        r.update(config=config, seed=seed, scale=scale, index=index, adaptive=adaptive, count=10, tau=0.05,
                 sigma=0.05)
        steps.append(r)
    (OUT / "reference_steps.json").write_text(json.dumps(steps) + "\n")


# The optional controls: anchored (Halpern, reflected, residual-restart)
# methods and the PID primal weight, as whole solves, raw backend chunks and
# controller sequences.
CONTROL_SOLVES = [
    # (config, seed, scale, index, method, primal_weight, adaptive)
    (0, 3, 20, 0, "halpern", "displacement", False),
    (0, 3, 20, 0, "rhpdhg", "displacement", False),
    (0, 3, 20, 0, "r2hpdhg", "displacement", False),
    (0, 3, 20, 0, "halpern", "pid", False),
    (0, 3, 20, 0, "rhpdhg", "pid", False),
    (0, 3, 20, 0, "r2hpdhg", "pid", False),
    (0, 3, 20, 0, "pdhg", "pid", True),
    (2, 3, 1, 1, "pdhg", "pid", True),
    (2, 3, 1, 0, "r2hpdhg", "displacement", False),
]
PID_SEQUENCES = [
    (1.0, [(100, 1), (0, 1)] + [(1e20, 1)] * 40),
    (0.37, [(2.5, 0.125), (1e-3, 4.0), (3.0, 3.0), (1e-30, 1.0), (7.0, 1e-9), (0.5, 0.25)] * 6),
    (250.0, [(1e-6, 1e6)] * 30 + [(1e6, 1e-6)] * 30),
]


def controls():
    out = {"solves": [], "steps": [], "pid": []}
    for config, seed, scale, index, method, weight, adaptive in CONTROL_SOLVES:
        args = ["solve", *case_args(config, seed, scale, index), "--method", method, "--primal-weight", weight,
                "--iterations", 30000]
        if not adaptive:
            args.append("--no-adaptive")
        r = run(*args)
        r.update(config=config, seed=seed, scale=scale, index=index, method=method, primal_weight=weight,
                 adaptive=adaptive)
        out["solves"].append(r)
        print("control", config, method, weight, r["status"], r["iterations"], r["restarts"], r["weight_updates"])
    for reflection, residual in ((1.0, False), (0.0, True), (1.0, True)):
        r = run("steps", *case_args(0, 3, 20, 0), "--count", 25, "--chunks", 8, "--tau", 0.05, "--sigma", 0.05,
                "--halpern", "--reflection", reflection,
                *(["--residual-restarts", "--reset-on-restart"] if residual else []))
        r.update(config=0, seed=3, scale=20, index=0, reflection=reflection, residual_restarts=residual,
                 count=25, tau=0.05, sigma=0.05)
        out["steps"].append(r)
    for w0, seq in PID_SEQUENCES:
        r = run("pid", "--weight", w0, "--updates", ";".join(f"{dx!r},{dy!r}" for dx, dy in seq))
        out["pid"].append({"weight": w0, "updates": seq, "weights": r["weights"]})
    (OUT / "reference_controls.json").write_text(json.dumps(out) + "\n")


# Where the reference's own PDHG stops converging within its 100,000-iteration
# default: the block-angular workload at shrinking scales, tol 1e-8, and the
# power-law workload at 1/50 scale, tol 1e-6 (status, iterations and rel-KKT
# reached; 4 OpenMP threads, so trajectories are not bit-exact pins).
STALLS = [(4, 42, 32, 1e-8), (4, 42, 16, 1e-8), (4, 42, 8, 1e-8), (4, 42, 4, 1e-8), (1, 42, 50, 1e-6)]


def stalls():
    out = []
    for config, seed, scale, tol in STALLS:
        r = run("solve", "--config", config, "--seed", seed, "--scale", scale, "--threads", 4, "--tol", tol,
                "--iterations", 100000)
        keep = {k: r[k] for k in ("status", "iterations", "restarts", "kkt", "primal", "dual", "gap", "objective",
                                  "planted_objective", "rows", "cols", "nnz", "seconds")}
        keep.update(config=config, seed=seed, scale=scale, tol=tol, threads=4)
        out.append(keep)
        print("stall", config, scale, tol, r["status"], r["iterations"], r["kkt"], flush=True)
    (OUT / "reference_stalls.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
