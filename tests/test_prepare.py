"""The engine's host prepare step (presolve, scaling, objective scale) against
the oracle's, bit for bit: the prepared arrays, scales and row/column maps
(src/preprocess.cpp:133-346).  The oracle's prepare is itself pinned to the
reference's output by hashes (tests/test_oracle.py).  Runs on the CPU.
"""
import json
from pathlib import Path

import numpy as np
import pytest

NETLIB = Path(__file__).resolve().parent / "golden" / "netlib"
KEYS = ("ptr", "index", "value", "c", "lb", "ub", "rl", "ru")


def presolve_lp(engine):
    """Every presolve rule at once: a fixed column (x2), a column in no row
    (x5), a row left empty by substitution (r3, feasible: dropped), a row the
    box makes redundant (r4), and an exactly repeated row with tighter bounds
    (r5 repeats r0: intersected into r0)."""
    A = np.array([[1.0, 2.0, 0.0, 1.0, 0.0, 0.0],
                  [0.0, 1.0, 3.0, 0.0, -1.0, 0.0],
                  [2.0, 0.0, 1.0, 0.0, 1.0, 0.0],
                  [0.0, 0.0, 1.0, 0.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 1e-3, 1e-3, 0.0],
                  [1.0, 2.0, 0.0, 1.0, 0.0, 0.0]])
    c = np.array([1.0, -2.0, 0.5, 3.0, 1.0, -1.0])
    lb = np.array([0.0, 0.0, 4.0, -1.0, 0.0, 0.0])
    ub = np.array([10.0, 5.0, 4.0, 2.0, 1.0, 7.0])
    rl = np.array([-np.inf, 12.0, 10.0, 3.5, -10.0, 0.5])
    ru = np.array([8.0, 14.0, np.inf, 4.5, 10.0, 6.0])
    return engine.LP.from_dense(A, c, lb, ub, rl, ru)


def cases(engine):
    out = {f"config{c}-{s}-{k}-{i}": engine.Instance(c, s, k, i).lp
           for c, s, k, i in [(0, 3, 20, 0), (4, 1, 200, 0), (1, 42, 50, 0), (2, 3, 1, 1), (3, 42, 1, 17)]}
    for name in ("afiro", "adlittle", "israel", "e226"):
        out[name] = engine.Model(NETLIB / f"{name}.mps").lp
    out["presolve"] = presolve_lp(engine)
    return out


@pytest.mark.parametrize("scaling", ["ruiz", "combined"])
@pytest.mark.parametrize("name", ["config0-3-20-0", "config4-1-200-0", "config1-42-50-0", "config2-3-1-1",
                                  "config3-42-1-17", "afiro", "adlittle", "israel", "e226", "presolve"])
def test_host_prepare_matches_oracle_bitwise(engine, oracle, name, scaling):
    lp = cases(engine)[name]
    ours = engine.prepare_lp(lp, scaling=scaling)
    ref = oracle.prepared_model(lp, scaling=scaling)
    maps = oracle.prepare(lp, scaling=scaling)
    assert ours["failure"] == "" and ref["failure"] == ""
    for key in KEYS:
        assert np.array_equal(ours[key], ref[key]), key
    for key in ("column_scale", "row_scale", "col_map", "row_map"):
        assert np.array_equal(ours[key], maps[key]), key
    assert ours["objective_scale"] == maps["objective_scale"] and ours["offset"] == ref["offset"]
    if name == "presolve":
        assert list(ours["col_map"]) == [0, 1, 3, 4]  # x2 fixed, x5 in no row
        assert list(ours["row_map"]) == [0, 1, 2]     # r3 empty, r4 redundant, r5 merged into r0


def test_presolve_infeasibility_matches_oracle(engine, oracle):
    base = presolve_lp(engine)
    # repeated row with disjoint bounds; a row left empty but required nonzero;
    # activity bounds disjoint from the row bounds
    variants = []
    rl, ru = base.rl.copy(), base.ru.copy()
    rl[5], ru[5] = 9.0, 12.0
    variants.append(engine.LP(base.ptr, base.index, base.value, base.c, base.lb, base.ub, rl, ru))
    rl, ru = base.rl.copy(), base.ru.copy()
    rl[3], ru[3] = 5.0, 6.0
    variants.append(engine.LP(base.ptr, base.index, base.value, base.c, base.lb, base.ub, rl, ru))
    rl, ru = base.rl.copy(), base.ru.copy()
    rl[4], ru[4] = 1.0, 2.0
    variants.append(engine.LP(base.ptr, base.index, base.value, base.c, base.lb, base.ub, rl, ru))
    for lp in variants:
        ours = engine.prepare_lp(lp)
        ref = oracle.prepared_model(lp)
        assert ours["failure"].startswith("INFEASIBLE") and ref["failure"] == "INFEASIBLE"


def prepared_lp(engine, lp):
    p = engine.prepare_lp(lp)
    return engine.LP(p["ptr"], p["index"], p["value"], p["c"], p["lb"], p["ub"], p["rl"], p["ru"], p["offset"], 1.0)


@pytest.mark.parametrize("name", ["afiro", "adlittle", "israel", "e226"])
def test_power_norm_matches_reference(engine, name):
    """--power-iterations (src/research.cpp:6-50, src/solver.cpp:315-321): the
    estimate on the prepared model equals the reference CLI's, bit for bit."""
    ref = json.loads((NETLIB / "reference_power.json").read_text())
    lp = prepared_lp(engine, engine.Model(NETLIB / f"{name}.mps").lp)
    for its in (1, 30):
        assert engine.power_norm(lp, its) == ref[f"{name}/{its}"], its


def test_power_norm_closed_forms(engine):
    # diagonal: the largest |entry|; rank one u v^T: |u| |v| once the start vector is in range
    diag = engine.LP.from_dense(np.diag([3.0, -1.0, 2.0]), np.zeros(3), np.zeros(3), np.ones(3), np.zeros(3), np.ones(3))
    assert abs(engine.power_norm(diag, 200) - 3.0) <= 1e-9
    u, v = np.array([1.0, -2.0, 2.0]), np.array([0.5, 0.0, 1.5, -1.0])
    rank1 = engine.LP.from_dense(np.outer(u, v), np.zeros(4), np.zeros(4), np.ones(4), np.zeros(3), np.ones(3))
    assert abs(engine.power_norm(rank1, 3) - np.linalg.norm(u) * np.linalg.norm(v)) <= 1e-12 * 10
    assert engine.power_norm(diag, 0) == 0.0
