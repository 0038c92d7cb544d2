"""Drop-in boundary on the GPU: the CLI and the session C-ABI solving real
models end to end.

Netlib models are solved by `niryukti solve` and compared with the result
document the reference CLI produced for the same file (recorded by
tests/golden/make_netlib_golden.py): same status, objective within 1e-6
relative (both points are certified at rel-KKT <= 1e-6, which bounds their
objective difference by a few 1e-6), and the returned point passes the
oracle's FP64 verifier and `niryukti verify`.  Session scenarios follow the
reference's tests/test_extensions.py:20-38.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
NETLIB = ROOT / "tests" / "golden" / "netlib"


def run_cli(engine, *args, timeout=120):
    return subprocess.run([str(engine.CLI_PATH), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def linear_model():
    return {"name": "linear", "sense": "min",
            "variables": [{"name": "x", "ub": 10}],
            "objective": {"linear": [3]},
            "constraints": [{"name": "demand", "lb": 2, "coefficients": {"x": 1}}]}


@pytest.mark.parametrize("name", ("afiro", "adlittle", "israel"))
def test_netlib_solve_matches_reference(engine, name, tmp_path):
    ref = json.loads((NETLIB / f"{name}.reference_result.json").read_text())
    out_path = tmp_path / "result.json"
    out = run_cli(engine, "solve", NETLIB / f"{name}.mps", "--json-out", out_path)
    print(name, out.stderr[-400:])
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    mine = json.loads(out.stdout)
    print(name, mine["status"], mine["performance"]["iterations"], mine["objective"], ref["objective"])
    assert mine["status"] == "OPTIMAL"
    assert mine["model"]["fingerprint"] == ref["model"]["fingerprint"]
    assert mine["accuracy"]["kkt_error"] <= 1e-6
    assert abs(mine["objective"] - ref["objective"]) <= 1e-6 * (1 + abs(ref["objective"]))
    check = run_cli(engine, "verify", NETLIB / f"{name}.mps", out_path)
    assert check.returncode == 0 and json.loads(check.stdout)["status"] == "VERIFIED_OPTIMAL"


def test_netlib_iteration_limit_is_reported(engine):
    # e226 does not reach 1e-6 within the reference's default 100000 iterations;
    # a short limit must end in ITERATION_LIMIT with exit code 2 and a finite point.
    out = run_cli(engine, "solve", NETLIB / "e226.mps", "--iterations", "2000")
    assert out.returncode == 2
    r = json.loads(out.stdout)
    assert r["status"] == "ITERATION_LIMIT"
    assert r["performance"]["iterations"] <= 2000
    assert len(r["primal"]) == r["model"]["columns"]


def test_warm_start_from_result_file(engine, tmp_path):
    first = tmp_path / "first.json"
    out = run_cli(engine, "solve", NETLIB / "adlittle.mps", "--json-out", first)
    assert out.returncode == 0
    cold = json.loads(out.stdout)["performance"]["iterations"]
    out = run_cli(engine, "solve", NETLIB / "adlittle.mps", "--warm-start", first)
    assert out.returncode == 0
    warm = json.loads(out.stdout)
    assert warm["status"] == "OPTIMAL"
    assert warm["performance"]["iterations"] < cold


def test_session_updates_and_rollback(engine):
    with engine.NativeSession(linear_model()) as s:
        assert s.solve()["objective"] == pytest.approx(6, abs=1e-6 * 7)
        s.update_row_bounds({"demand": {"lb": 4}})
        assert s.solve()["objective"] == pytest.approx(12, abs=1e-6 * 13)
        with pytest.raises(ValueError):
            s.update_variable_bounds({"x": {"lb": 11}})
        assert s.solve()["objective"] == pytest.approx(12, abs=1e-6 * 13)
        with pytest.raises(ValueError):
            s._request("solve", unknown_option=True)
        s.update_objective({"x": 5})
        r = s.solve()
        assert r["status"] == "OPTIMAL" and r["objective"] == pytest.approx(20, abs=1e-6 * 21)
        assert r["session"]["warm_start_requested"] is True
    with pytest.raises(RuntimeError):
        s.solve()


def test_session_on_netlib_warm_start(engine):
    with engine.NativeSession(str(NETLIB / "afiro.mps")) as s:
        cold = s.solve()
        assert cold["status"] == "OPTIMAL"
        again = s.solve()
        assert again["status"] == "OPTIMAL"
        assert again["performance"]["iterations"] <= cold["performance"]["iterations"]
        fresh = s.solve(warm_start=False)
        assert fresh["performance"]["iterations"] == cold["performance"]["iterations"]


def test_power_iterations_estimate_and_solve(engine):
    """--power-iterations: the result reports the reference's estimate (the
    same bits, tests/golden/netlib/reference_power.json), the solve with the
    larger first step still certifies, and the option is range-checked like
    the reference's (src/solver.cpp:20-25)."""
    ref = json.loads((NETLIB / "reference_power.json").read_text())
    out = run_cli(engine, "solve", NETLIB / "afiro.mps", "--power-iterations", "30")
    assert out.returncode == 0, out.stderr
    doc = json.loads(out.stdout)
    assert doc["performance"]["operator_norm_estimate"] == ref["afiro/30"]
    assert doc["status"] == "OPTIMAL"
    assert run_cli(engine, "solve", NETLIB / "afiro.mps", "--power-iterations", "1001").returncode == 1
    m = engine.Model(NETLIB / "adlittle.mps")
    r = engine.solve_lp(m.lp, power_iterations=30)
    assert r["operator_norm_estimate"] == ref["adlittle/30"] and r["status"] == "OPTIMAL"


def test_polishing_cli_and_options_block(engine, tmp_path):
    """--polishing through the CLI (tests/test_cli.py:106-110): the solve still
    certifies, the options block records the request (app/main.cpp:610-630),
    and a quadratic model is refused with exit code 1."""
    out = run_cli(engine, "solve", NETLIB / "afiro.mps", "--polishing", "--power-iterations", "20")
    assert out.returncode == 0, out.stderr
    doc = json.loads(out.stdout)
    assert doc["status"] == "OPTIMAL" and doc["performance"]["operator_norm_estimate"] > 0
    assert doc["options"]["polishing"] is True and doc["options"]["power_iterations"] == 20
    assert doc["performance"]["polishing_iterations"] <= doc["performance"]["iterations"]
    assert run_cli(engine, "solve", NETLIB / "afiro.mps", "--power-iterations", "-1").returncode == 1
    qp = dict(linear_model(), objective={"linear": [3], "quadratic_diagonal": [0.2]})
    path = tmp_path / "dispatch.json"
    path.write_text(json.dumps(qp))
    assert run_cli(engine, "solve", path, "--polishing").returncode == 1


def test_polishing_shares_the_iteration_budget(engine, oracle):
    """A badly scaled LP started next to its optimum (tests/test_research.cpp:
    1084-1102): polishing runs, its iterations come out of the one iteration
    limit, and an OPTIMAL answer passes the original-space check."""
    lp = engine.LP(np.array([0, 1, 2], np.int64), np.array([0, 1], np.int64), np.array([1.0, 1e-3]),
                   np.array([1.0, 1e-3]), np.zeros(2), np.full(2, np.inf), np.array([1.0, 1e-3]),
                   np.array([1.0, 1e-3]), 0.0, 1.0)
    r = engine.solve_lp(lp, polishing=1, scaling_passes=0, tol=1e-10, iteration_limit=2000,
                        initial_x=np.array([0.999, 0.999]), initial_y=np.array([-1.0, -1.0]))
    print("slow lp", r["status"], r["iterations"], r["polishing_attempts"], r["polishing_iterations"], flush=True)
    assert r["polishing_attempts"] > 0
    assert r["iterations"] <= 2000 and r["polishing_iterations"] <= r["iterations"]
    if r["status"] == "OPTIMAL":
        assert oracle.verify(lp, r["x"], r["y"])["kkt"] <= 1e-10


@pytest.mark.parametrize("name", ("adlittle", "israel"))
def test_polishing_on_netlib(engine, oracle, name):
    """Polishing on real models: the answer still passes the original-space
    check at the requested tolerance within the same iteration limit, and its
    objective agrees with the reference CLI's result for the model."""
    ref = json.loads((NETLIB / f"{name}.reference_result.json").read_text())
    lp = engine.Model(NETLIB / f"{name}.mps").lp
    r = engine.solve_lp(lp, polishing=1)
    print(name, r["status"], r["iterations"], r["polishing_attempts"], r["polishing_iterations"], flush=True)
    assert r["status"] == "OPTIMAL" and r["polishing_attempts"] > 0
    assert r["polishing_iterations"] <= r["iterations"] <= 100000
    assert oracle.verify(lp, r["x"], r["y"])["kkt"] <= 1e-6
    assert abs(r["objective"] - ref["objective"]) <= 1e-6 * (1 + abs(ref["objective"]))
