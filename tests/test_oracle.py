"""Pins the CPU oracle: against the reference's own outputs (golden fixtures
written by tests/golden/make_golden.py from the reference build, and the live
reference build when it is present), against closed forms from the
reference's tests, and against an independent LP solver (HiGHS via scipy).
"""
import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_DRIVER = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "ref_driver"


def load(name):
    return json.loads((GOLDEN / name).read_text())


def instance(engine, case):
    return engine.Instance(case["config"], case["seed"], case["scale"], case["index"])


@pytest.mark.parametrize("case", load("reference_solves.json"), ids=lambda c: f"c{c['config']}s{c['seed']}x{c['scale']}i{c['index']}")
def test_oracle_reproduces_reference_solve(engine, oracle, case):
    inst = instance(engine, case)
    assert inst.lp.rows == case["rows"] and inst.lp.cols == case["cols"] and inst.lp.nnz == case["nnz"]
    r = oracle.solve(inst.lp, threads=1, time_limit=1e9)
    assert r["status"] == case["status"]
    for key in ("iterations", "restarts", "rejected_steps", "weight_updates"):
        assert r[key] == case[key], key
    # One thread, same arithmetic: the restatement is bit-identical to the reference.
    assert r["objective"] == case["objective"]
    assert r["kkt"] == case["kkt"]
    if "x" in case:
        assert np.array_equal(r["x"], case["x"]) and np.array_equal(r["y"], case["y"])
    # and the reference itself agrees with the planted closed-form optimum
    assert abs(case["objective"] - inst.planted_objective) <= 1e-6 * (1 + abs(inst.planted_objective))


@pytest.mark.parametrize("case", load("reference_prepare.json"), ids=lambda c: f"c{c['config']}-{c['scaling']}")
def test_oracle_reproduces_reference_scaling(engine, oracle, case):
    inst = instance(engine, case)
    p = oracle.prepare(inst.lp, scaling=case["scaling"])
    assert p["failure"] == case["failure"]
    assert p["rows"] == case["rows"] and p["cols"] == case["cols"]
    assert p["objective_scale"] == case["objective_scale"]
    if "column_scale" in case:
        assert np.array_equal(p["column_scale"], case["column_scale"])
        assert np.array_equal(p["row_scale"], case["row_scale"])
        assert np.array_equal(p["col_map"], case["col_map"]) and np.array_equal(p["row_map"], case["row_map"])
    # the whole scaled model, bit for bit, through hashes of the reference's arrays
    q = oracle.prepared_model(inst.lp, scaling=case["scaling"])
    assert q["offset"] == case["offset"]
    arrays = dict(q, column_scale=p["column_scale"], row_scale=p["row_scale"], col_map=p["col_map"],
                  row_map=p["row_map"])
    for key, digest in case["model_sha256"].items():
        a = np.asarray(arrays[key])
        assert hashlib.sha256(a.astype(a.dtype.newbyteorder("<")).tobytes()).hexdigest() == digest, key


@pytest.mark.parametrize("case", load("reference_steps.json"), ids=lambda c: f"c{c['config']}-a{int(c['adaptive'])}")
def test_oracle_reproduces_reference_steps(engine, oracle, case):
    inst = instance(engine, case)
    lp = inst.lp
    x0 = np.clip(np.zeros(lp.cols), lp.lb, lp.ub)
    b = oracle.Backend(lp, x0, np.zeros(lp.rows), adaptive=case["adaptive"])
    for chunk in case["chunks"]:
        assert b.advance(case["count"], case["tau"], case["sigma"]) == chunk["accepted"]
        assert b.rejected() == chunk["rejected"]
        x, y, xa, ya = b.candidates()
        for got, key in ((x, "x"), (y, "y"), (xa, "xa"), (ya, "ya")):
            assert np.array_equal(got, chunk[key]), key


@pytest.mark.skipif(not REF_DRIVER.exists(), reason="reference build not present (built only where the sources are)")
@pytest.mark.parametrize("seed", [101, 202])
def test_oracle_matches_live_reference(engine, oracle, seed):
    for config, scale, index in ((2, 1, seed % 5), (0, 10, 0)):
        out = subprocess.run([str(REF_DRIVER), "solve", "--config", str(config), "--seed", str(seed), "--scale",
                              str(scale), "--index", str(index), "--threads", "1"], check=True,
                             capture_output=True, text=True).stdout
        ref = json.loads(out)
        inst = engine.Instance(config, seed, scale, index)
        r = oracle.solve(inst.lp, threads=1, time_limit=1e9)
        assert (r["status"], r["iterations"], r["restarts"]) == (ref["status"], ref["iterations"], ref["restarts"])
        assert r["objective"] == ref["objective"]


# ---------------------------------------------------------------- closed forms

def dense_lp(engine, A, c, lb, ub, rl, ru):
    return engine.LP.from_dense(np.array(A, float), c, lb, ub, rl, ru)


def test_planted_points_certify(engine, oracle):
    """The generator's planted (x0, -y0) is an exact KKT point of every workload."""
    for config, scale, index in ((0, 4, 0), (1, 400, 0), (2, 1, 3), (4, 200, 0)):
        inst = engine.Instance(config, 5, scale, index)
        acc = oracle.verify(inst.lp, inst.planted_x, inst.planted_y)
        assert acc["kkt"] <= 1e-9, (config, acc)
        assert abs(acc["objective"] - inst.planted_objective) <= 1e-9 * (1 + abs(inst.planted_objective))
        # a perturbed dual must be detected (a dropped term in the dual residual would hide it)
        y = inst.planted_y.copy()
        y[0] += 1.0
        assert oracle.verify(inst.lp, inst.planted_x, y)["kkt"] > 1e-6


def test_verifier_pins(engine, oracle):
    # tests/test_core.cpp:57-70, 113-122
    wide = dense_lp(engine, [[1]], [0], [0], [10], [1], [1e15])
    assert oracle.verify(wide, [0], [0])["primal"] >= .5
    contradiction = dense_lp(engine, [[1, 1], [1, 1]], [0, 0], [0, 0], [10, 10], [3, -np.inf], [np.inf, 1])
    assert oracle.infeasibility_bound(contradiction, [-1, 1]) > 1.9
    assert oracle.infeasibility_bound(contradiction, [1, -1]) <= 0
    assert oracle.infeasibility_bound(contradiction, [0, 0]) <= 0
    m = dense_lp(engine, [[1, 2]], [1, 1], [0, 0], [10, 10], [4], [np.inf])
    assert not oracle.verify(m, [np.nan, 0], [0])["finite"]
    assert oracle.verify(m, [0, 0], [0])["primal"] > 0
    overflow = dense_lp(engine, [[1e308, -1e308]], [0, 0], [0, 0], [10, 10], [-np.inf], [np.inf])
    assert not oracle.verify(overflow, [10, 10], [0])["finite"]
    assert oracle.verify(m, [0, 2], [-.6])["lower_bound"] <= 2


def test_one_adaptive_step_closed_form(engine, oracle):
    # tests/test_research.cpp:34-48: first trial has limit 1/2 and is rejected.
    lp = dense_lp(engine, [[1]], [-1], [-10], [10], [3], [3])
    b = oracle.Backend(lp, [0.0], [0.0], adaptive=True)
    assert b.advance(1, 2, 2) == 1
    assert b.rejected() > 0
    x, y, ax, ay = b.candidates()
    assert 0 < x[0] < 2
    assert (x[0] ** 2 + y[0] ** 2) / x[0] >= 2 * abs(x[0] * y[0])
    assert abs(ax[0] - x[0]) <= 1e-11 * (1 + abs(x[0])) and abs(ay[0] - y[0]) <= 1e-11 * (1 + abs(y[0]))


def test_halpern_scalar_recurrence(engine, oracle):
    # tests/test_research.cpp:65-92: min_x max_y (x-1) y
    lp = dense_lp(engine, [[1]], [0], [-np.inf], [np.inf], [1], [1])
    b = oracle.Backend(lp, [0.0], [0.0], adaptive=False, halpern=True)
    ex = ey = 0.0
    for k in range(20):
        tx = ex - .5 * ey
        ty = ey + .5 * (2 * tx - ex - 1)
        w = (k + 1) / (k + 2)
        ex, ey = w * tx, w * ty
        b.advance(1, .5, .5)
        x, y, rx, ry = b.candidates()
        for got, want in ((x[0], ex), (y[0], ey), (rx[0], tx), (ry[0], ty)):
            assert abs(got - want) <= 1e-11 * (1 + abs(want))
    b.reset([2.0], [3.0])
    b.advance(1, .5, .5)
    x, y, _, _ = b.candidates()
    assert abs(x[0] - 1.25) <= 1e-11 and abs(y[0] - 2.5) <= 1e-11


def test_zero_passes_is_identity(engine, oracle):
    inst = engine.Instance(2, 3, 1, 0)
    p = oracle.prepare(inst.lp, scaling_passes=0)
    assert np.all(p["column_scale"] == 1) and np.all(p["row_scale"] == 1) and p["objective_scale"] == 1


def test_oracle_matches_highs_on_random_lps(engine, oracle):
    """Independent solver cross-check on small random LPs with mixed row types."""
    linprog = pytest.importorskip("scipy.optimize").linprog
    rng = np.random.default_rng(3)
    for trial in range(4):
        m, n = 8 + trial, 16 + 2 * trial
        A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
        x0 = rng.uniform(0, 5, n) * (rng.random(n) < 0.7)
        b = A @ x0
        rl = np.where(np.arange(m) % 3 == 0, -np.inf, b)
        ru = b + np.where(np.arange(m) % 3 == 0, 1.0, 0.0)
        c = rng.standard_normal(n)
        lp = dense_lp(engine, A, c, np.zeros(n), np.full(n, 5.0), rl, ru)
        r = oracle.solve(lp, threads=1)
        eq = np.isfinite(rl)
        h = linprog(c, A_ub=A[~eq], b_ub=ru[~eq], A_eq=A[eq], b_eq=b[eq], bounds=[(0, 5)] * n, method="highs")
        assert h.status == 0 and r["status"] == "OPTIMAL"
        assert abs(r["objective"] - h.fun) <= 1e-5 * (1 + abs(h.fun))


# ------------------------------------------------------- optional controls

def test_pid_controller_closed_forms(oracle):
    """tests/test_research.cpp:1016-1022, plus two hand-derived steps.

    From a fresh controller, update(1, e^2, 1) has error = 0.5 * ln(e^2) = 1,
    integral 1, no derivative, correction 0.3 + 0.01 = 0.31: the weight is
    e^-0.31.  A second identical update has error 1 - 0.31 = 0.69, integral
    1.69, derivative -0.31, correction 0.207 + 0.0169 - 0.031 = 0.1929."""
    pid = oracle.Pid()
    w = pid.update(1, 100, 1)
    assert 0 < w < 1, "displacement feedback direction"
    assert pid.update(w, 0, 1) == w, "no update without a primal displacement"
    for _ in range(1000):
        w = pid.update(w, 1e20, 1)
    assert np.isfinite(w) and w >= 1e-4 * (1 - 1e-12), "saturation and anti-windup"
    fresh = oracle.Pid()
    w1 = fresh.update(1.0, np.exp(2.0), 1.0)
    assert abs(np.log(w1) + 0.31) <= 1e-14
    w2 = fresh.update(w1, np.exp(2.0), 1.0)
    assert abs(np.log(w2) - (np.log(w1) - 0.1929)) <= 1e-13
    # the per-update correction is clamped to +-ln 4
    big = oracle.Pid()
    assert abs(big.update(1.0, 1e12, 1e-12) - 0.25) <= 1e-15
    # anti-windup: while clamped at 1e-4 the integral stops growing, so a
    # reversal moves the weight back up by the full step at once
    wind = oracle.Pid()
    w = 1.0
    for _ in range(50):
        w = wind.update(w, 1e20, 1)
    assert abs(w - 1e-4) <= 1e-16
    assert wind.update(w, 1e-20, 1) > w * 3.99


def test_pid_controller_matches_reference(oracle):
    for case in load("reference_controls.json")["pid"]:
        pid, w = oracle.Pid(), case["weight"]
        for (dx, dy), want in zip(case["updates"], case["weights"]):
            w = pid.update(w, dx, dy)
            assert w == want


def test_reflected_halpern_closed_form(engine, oracle):
    # tests/test_research.cpp:1023-1041: reflected Halpern on min_x max_y (x-1) y
    lp = dense_lp(engine, [[1]], [0], [-np.inf], [np.inf], [1], [1])
    b = oracle.Backend(lp, [0.0], [0.0], adaptive=False, halpern=True, reflection=1.0)
    x = y = 0.0
    for k in range(20):
        tx = x - .5 * y
        ty = y + .5 * (2 * tx - x - 1)
        w = (k + 1) / (k + 2)
        x, y = w * (2 * tx - x), w * (2 * ty - y)
        b.advance(1, .5, .5)
        cx, cy, rx, ry = b.candidates()
        for got, want in ((cx[0], x), (cy[0], y), (rx[0], tx), (ry[0], ty)):
            assert abs(got - want) <= 1e-11 * (1 + abs(want))


def test_fixed_point_restart_interrupts_and_clears(engine, oracle):
    # tests/test_research.cpp:1042-1046
    lp = dense_lp(engine, [[1]], [0], [-np.inf], [np.inf], [1], [1])
    b = oracle.Backend(lp, [0.0], [0.0], adaptive=False, halpern=True, reflection=1.0, residual_restarts=True)
    assert b.advance(100, .5, .5) < 100 and b.restart_requested(), "fixed-point restart interrupts a chunk"
    b.reset([0.0], [0.0])
    assert not b.restart_requested(), "restart state cleared"
    # without the residual test the same chunk runs to the end
    plain = oracle.Backend(lp, [0.0], [0.0], adaptive=False, halpern=True, reflection=1.0)
    assert plain.advance(100, .5, .5) == 100 and not plain.restart_requested()


def test_anchored_steps_match_reference(engine, oracle):
    """Raw backend chunks of the reflected and residual-restart Halpern
    iterations, bit for bit against the reference build (a reset to the T(z)
    candidate after each chunk that requested a restart)."""
    for case in load("reference_controls.json")["steps"]:
        lp = instance(engine, case).lp
        b = oracle.Backend(lp, np.clip(np.zeros(lp.cols), lp.lb, lp.ub), np.zeros(lp.rows), adaptive=False,
                           halpern=True, reflection=case["reflection"], residual_restarts=case["residual_restarts"])
        for chunk in case["chunks"]:
            assert b.advance(case["count"], case["tau"], case["sigma"]) == chunk["accepted"]
            assert b.restart_requested() == bool(chunk["restart_requested"])
            x, y, xa, ya = b.candidates()
            for got, key in ((x, "x"), (y, "y"), (xa, "xa"), (ya, "ya")):
                assert np.array_equal(got, chunk[key]), key
            if case["residual_restarts"] and b.restart_requested():
                b.reset(xa, ya)


@pytest.mark.parametrize("case", load("reference_controls.json")["solves"],
                         ids=lambda c: f"c{c['config']}-{c['method']}-{c['primal_weight']}")
def test_control_solves_match_reference(engine, oracle, case):
    inst = instance(engine, case)
    r = oracle.solve(inst.lp, threads=1, time_limit=1e9, method=case["method"], primal_weight=case["primal_weight"],
                     adaptive=case["adaptive"], iteration_limit=30000)
    assert r["status"] == case["status"] == "OPTIMAL"
    for key in ("iterations", "restarts", "rejected_steps", "weight_updates"):
        assert r[key] == case[key], key
    assert r["objective"] == case["objective"] and r["kkt"] == case["kkt"]
    assert abs(case["objective"] - inst.planted_objective) <= 1e-6 * (1 + abs(inst.planted_objective))
