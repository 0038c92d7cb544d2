"""Device-side scaling and transpose (include/niryukti_prepare.h).

npdhg_scale_lp runs the Ruiz passes, the objective scale and the explicit
transpose of a presolved LP on the device.  Its output must equal, bit for bit,
the oracle's prepared model (which tests/test_oracle.py pins to the
reference's own prepare() output through hashes of every array), and its
transpose must equal a stable numpy transpose (Sparse::transpose keeps each
column's entries in row order, src/model.cpp:39-58).  Solves with the scaling
on the device must be identical to solves with it on the host.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NETLIB = Path(__file__).resolve().parent / "golden" / "netlib"


def hazard_lp(engine):
    """Rows and columns whose Ruiz factor must be refused (src/preprocess.cpp:6-65):
    row 0 holds 1e300 next to the smallest subnormal (scaling it down would
    erase that entry), row 1 has a bound that scaling up would overflow, column
    1 a cost that would overflow, column 3 a lower bound that would underflow."""
    A = np.array([[1e300, 5e-324, 0.0, 0.0, 1.0],
                  [0.0, 1e-4, 0.0, 2e-4, 0.0],
                  [1.0, -0.02, 0.5, 0.0, 3.0],
                  [0.0, 0.0, 4.0, 1e-3, 0.0]])
    rows, cols = np.nonzero(A)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=4))]).astype(np.int64)
    c = np.array([1.0, 1.79e308, -1.0, 2.0, 0.5])
    lb = np.array([0.0, 0.0, 0.0, 5e-324, -np.inf])
    ub = np.full(5, np.inf)
    rl = np.array([-1.0, -np.inf, 1.0, -np.inf])
    ru = np.array([1.0, 1e307, 1.0, 2.0])
    return engine.LP(ptr, cols.astype(np.int64), A[rows, cols], c, lb, ub, rl, ru, 0.0, 1.0)


def cases(engine):
    out = {f"config{c}-{s}-{k}-{i}": engine.Instance(c, s, k, i).lp
           for c, s, k, i in [(0, 3, 20, 0), (4, 1, 200, 0), (1, 42, 50, 0), (2, 3, 1, 1), (3, 42, 1, 17)]}
    for name in ("afiro", "adlittle", "israel", "e226"):
        out[name] = engine.Model(NETLIB / f"{name}.mps").lp
    out["hazard"] = hazard_lp(engine)
    return out


def presolved(engine, oracle, lp):
    p = oracle.prepared_model(lp, scaling_passes=0)
    assert p["failure"] == ""
    return engine.LP(p["ptr"], p["index"], p["value"], p["c"], p["lb"], p["ub"], p["rl"], p["ru"], p["offset"], 1.0), p


def stable_transpose(ptr, index, value, cols):
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    order = np.argsort(index, kind="stable")
    t_ptr = np.concatenate([[0], np.cumsum(np.bincount(index, minlength=cols))])
    return t_ptr, rows[order], value[order]


@pytest.mark.parametrize("name", ["config0-3-20-0", "config4-1-200-0", "config1-42-50-0", "config2-3-1-1",
                                  "config3-42-1-17", "afiro", "adlittle", "israel", "e226", "hazard"])
def test_device_scaling_matches_oracle_bitwise(engine, oracle, name):
    lp = cases(engine)[name]
    lp0, pre = presolved(engine, oracle, lp)
    print("case", name, lp0.rows, lp0.cols, lp0.nnz, flush=True)
    dev = engine.scale_lp(lp0, 5)
    ref = oracle.prepared_model(lp)
    scales = oracle.prepare(lp)
    for key in ("value", "c", "lb", "ub", "rl", "ru"):
        assert np.array_equal(dev[key], ref[key]), key
    assert np.array_equal(dev["row_scale"], scales["row_scale"])
    assert np.array_equal(dev["column_scale"], scales["column_scale"])
    assert dev["objective_scale"] == scales["objective_scale"]
    assert pre["offset"] / dev["objective_scale"] == ref["offset"]
    t_ptr, t_index, t_value = stable_transpose(ref["ptr"], ref["index"], ref["value"], ref["cols"])
    assert np.array_equal(dev["t_ptr"], t_ptr)
    assert np.array_equal(dev["t_index"], t_index)
    assert np.array_equal(dev["t_value"], t_value)
    if name == "hazard":
        # the refusals really happened: these factors stayed at 1
        assert dev["row_scale"][0] == dev["row_scale"][1] == 1.0 and dev["column_scale"][3] == 1.0
        # after one pass the cost of column 1 has refused its factor (> 1) too
        one = engine.scale_lp(lp0, 1)
        ref1 = oracle.prepared_model(lp, scaling_passes=1)
        for key in ("value", "c", "lb", "ub", "rl", "ru"):
            assert np.array_equal(one[key], ref1[key]), key
        assert one["column_scale"][1] == 1.0 and np.array_equal(one["column_scale"],
                                                                 oracle.prepare(lp, scaling_passes=1)["column_scale"])


def test_transpose_only_and_empty_columns(engine):
    rng = np.random.default_rng(4)
    m, n = 300, 700
    A = (rng.random((m, n)) < 0.01) * rng.standard_normal((m, n))
    A[:, 5] = 0  # an empty column
    A[7, :] = rng.standard_normal(n)  # a dense row
    rows, cols = np.nonzero(A)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]).astype(np.int64)
    lp = engine.LP(ptr, cols.astype(np.int64), A[rows, cols], rng.standard_normal(n), np.zeros(n), np.ones(n),
                   -np.ones(m), np.ones(m), 0.0, 1.0)
    dev = engine.scale_lp(lp, 0)
    assert np.array_equal(dev["value"], lp.value) and dev["objective_scale"] == 1.0
    assert np.all(dev["row_scale"] == 1) and np.all(dev["column_scale"] == 1)
    t_ptr, t_index, t_value = stable_transpose(lp.ptr, lp.index, lp.value, n)
    assert np.array_equal(dev["t_ptr"], t_ptr) and np.array_equal(dev["t_index"], t_index)
    assert np.array_equal(dev["t_value"], t_value)


@pytest.mark.parametrize("case", [(0, 3, 20, 0), (1, 42, 100, 0)])
def test_solve_with_device_prepare_is_identical(engine, case):
    lp = engine.Instance(*case).lp
    print("case", case, lp.nnz, flush=True)
    host = engine.solve_lp(lp, device_prepare=0, iteration_limit=3000)
    dev = engine.solve_lp(lp, device_prepare=1, iteration_limit=3000)
    assert host["device_prepare"] == 0 and dev["device_prepare"] == 1
    assert host["status"] == dev["status"] and host["iterations"] == dev["iterations"]
    assert np.array_equal(host["x"], dev["x"]) and np.array_equal(host["y"], dev["y"])


def test_cli_device_prepare_flag(engine):
    outs = {}
    for flag in ("on", "off"):
        p = subprocess.run([str(engine.CLI_PATH), "solve", str(NETLIB / "afiro.mps"), "--device-prepare", flag],
                           capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
        outs[flag] = json.loads(p.stdout)
    assert outs["on"]["gpu_execution"]["device_prepare"] is True
    assert outs["off"]["gpu_execution"]["device_prepare"] is False
    assert outs["on"]["objective"] == outs["off"]["objective"]
    assert outs["on"]["primal"] == outs["off"]["primal"]
