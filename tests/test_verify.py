"""The engine's certifying host verifier against the oracle's, bit for bit (CPU).

The oracle's verifier restates src/verify.cpp and is pinned to the reference
build through the golden solves (objective and rel-KKT of the returned points
equal bit for bit).  The engine's verifier is written differently (a row and
a column pass feeding shared accumulators, an outward-rounded interval for
the certificate bound), so these tests require every figure to agree exactly
on points that exercise each term: the planted optimum, perturbed and random
points, bound and row violations, prices on rows whose active side is
infinite, and both signs of every reduced cost.
"""
from pathlib import Path

import numpy as np
import pytest

NETLIB = Path(__file__).resolve().parent / "golden" / "netlib"
KEYS = ("objective", "primal", "dual", "gap", "kkt", "complementarity", "lower_bound", "primal_absolute",
        "dual_absolute", "finite")


def points(inst_lp, x_star, y_star, rng):
    n, m = inst_lp.cols, inst_lp.rows
    yield x_star, y_star
    yield x_star, np.zeros(m)
    yield np.zeros(n), y_star
    for scale in (1e-9, 1e-6, 1e-3, 1.0):
        yield x_star + scale * rng.standard_normal(n), y_star + scale * rng.standard_normal(m)
    # outside the box on purpose, and prices of either sign everywhere
    yield x_star + 5 * rng.standard_normal(n), 10 * rng.standard_normal(m)
    lo = np.where(np.isfinite(inst_lp.lb), inst_lp.lb, -3.0)
    hi = np.where(np.isfinite(inst_lp.ub), inst_lp.ub, 3.0)
    yield rng.uniform(lo, hi), rng.standard_normal(m) * (rng.random(m) < 0.5)


def compare(engine, oracle, lp, x, y):
    for safe in (True, False):
        mine = engine.verify_lp(lp, x, y, safe_bound=safe)
        ref = oracle.verify(lp, x, y, safe=safe)
        for key in KEYS:
            a, b = mine[key], ref[key]
            if key == "objective" and not ref["finite"]:
                continue  # reported as NaN by the engine when not finite
            assert (a == b) or (np.isnan(a) and np.isnan(b)), (key, safe, a, b)


@pytest.mark.parametrize("case", [(0, 3, 20, 0), (1, 42, 400, 0), (2, 42, 1, 5), (4, 1, 200, 0), (0, 7, 4, 0)],
                         ids=str)
def test_verifier_matches_oracle_bitwise(engine, oracle, case):
    inst = engine.Instance(*case)
    rng = np.random.default_rng(sum(case))
    for x, y in points(inst.lp, inst.planted_x, inst.planted_y, rng):
        compare(engine, oracle, inst.lp, x, y)


@pytest.mark.parametrize("name", ("afiro", "adlittle", "israel"))
def test_verifier_matches_oracle_on_netlib(engine, oracle, name):
    """Netlib models have ranged, free and one-sided rows and free columns."""
    lp = engine.Model(NETLIB / f"{name}.mps").lp
    rng = np.random.default_rng(len(name))
    x0 = np.clip(np.zeros(lp.cols), lp.lb, lp.ub)
    for x, y in points(lp, x0, np.zeros(lp.rows), rng):
        compare(engine, oracle, lp, x, y)


def test_verifier_special_points(engine, oracle):
    lp = engine.LP.from_dense([[1.0, 2.0], [1.0, -1.0]], [1, 1], [0, -np.inf], [10, np.inf], [4, -np.inf],
                              [np.inf, 1])
    for x, y in (([0, 2], [-0.6, 0]), ([1, 1.5], [0.5, -0.25]), ([np.nan, 0], [0, 0]), ([0, 0], [0, 0]),
                 ([1e300, 1e300], [1e300, -1e300]), ([2, 1], [-1e-300, 1e-300])):
        compare(engine, oracle, lp, np.array(x, float), np.array(y, float))
    # a Farkas ray certifies infeasibility through the same bound
    contradiction = engine.LP.from_dense([[1, 1], [1, 1]], [0, 0], [0, 0], [10, 10], [3, -np.inf], [np.inf, 1])
    assert oracle.infeasibility_bound(contradiction, [-1, 1]) > 1.9
