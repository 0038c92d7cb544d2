"""Device-side candidate checks of the single-LP engine.

npdhg_evaluate restores both candidates to the original space on the device
and evaluates them with the verifier's formulas (src/verify.cpp:103-240,
without the outward-rounded bound).  Here every figure is compared with the
oracle's FP64 verifier on the same restored points, the restored points with a
numpy restatement of restore_x / restore_y (src/preprocess.cpp:121-132), and
the epoch displacement with a numpy sum.  End-to-end solves with the checks on
the device must reach the same certified answers as with host checks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / (1 + abs(b))


def sparse_lp(engine, rng, m, n, per_col, long_row=0, long_col=0):
    rows, cols, vals = [], [], []
    for j in range(n):
        r = rng.choice(m, per_col, replace=False)
        rows += r.tolist()
        cols += [j] * per_col
        vals += rng.standard_normal(per_col).tolist()
    if long_row:
        c = rng.choice(n, long_row, replace=False)
        rows += [1] * long_row
        cols += c.tolist()
        vals += rng.standard_normal(long_row).tolist()
    if long_col:
        r = rng.choice(m, long_col, replace=False)
        rows += r.tolist()
        cols += [2] * long_col
        vals += rng.standard_normal(long_col).tolist()
    A = np.zeros((m, n))
    np.add.at(A, (np.array(rows), np.array(cols)), np.array(vals))
    return A


def restore(x, y, cols_map, lower, upper, fixed, dc, dr, sobj, m0):
    ox = fixed.copy()
    ox[cols_map] = x * dc
    oy = np.zeros(m0)
    for i in range(len(y)):
        oy[upper[i] if y[i] >= 0 else lower[i]] += y[i] * dr[i] * sobj
    return ox, oy


def test_evaluate_matches_oracle_verifier(engine, oracle):
    rng = np.random.default_rng(3)
    m0, n0 = 300, 700
    A0 = sparse_lp(engine, rng, m0, n0, 4, long_row=600, long_col=250)
    kind = rng.integers(0, 4, m0)
    b = rng.uniform(-1, 1, m0)
    rl = np.where(kind == 1, -np.inf, b - (kind == 3) * 0.5)
    ru = np.where(kind == 2, np.inf, b + (kind == 3) * 0.5)
    rl[kind == 0] = ru[kind == 0] = b[kind == 0]
    lb = np.where(rng.random(n0) < 0.2, -np.inf, 0.0)
    ub = np.where(rng.random(n0) < 0.3, np.inf, rng.uniform(1, 5, n0))
    lp0 = engine.LP.from_dense(A0, rng.standard_normal(n0), lb, ub, rl, ru, offset=3.5)

    m, n = m0 - 20, n0 - 30
    As = sparse_lp(engine, rng, m, n, 3)
    lps = engine.LP.from_dense(As, rng.standard_normal(n), -5 * np.ones(n), 5 * np.ones(n), -np.ones(m), np.ones(m))
    cols_map = rng.permutation(n0)[:n]
    upper = rng.permutation(m0)[:m]
    lower = np.where(rng.random(m) < 0.5, upper, rng.permutation(m0)[:m])
    # lower must be one-to-one on its own: rebuild it where the mix collided
    if len(set(lower.tolist())) != m:
        lower = upper.copy()
    fixed = rng.uniform(-1, 1, n0)
    dc, dr, sobj = rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, m), 1.7

    x0, y0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, m)
    gpu = engine.Backend(lps, x0, y0, adaptive=True)
    gpu.attach_original(lp0, cols_map, lower, upper, fixed, dc, dr, sobj)
    for _ in range(2):
        assert gpu.advance(20, 0.05, 0.05) == 20
    cur, avg = gpu.evaluate()
    x, y, xa, ya = gpu.candidates()
    for acc, (px, py), which in ((cur, (x, y), 0), (avg, (xa, ya), 1)):
        ox, oy = restore(px, py, cols_map, lower, upper, fixed, dc, dr, sobj, m0)
        gx, gy = gpu.original_point(which)
        assert np.array_equal(gx, ox) and np.array_equal(gy, oy)
        ref = oracle.verify(lp0, ox, oy, safe=False)
        assert acc["finite"] == 1 and ref["finite"] == 1
        print(which, acc, {k: ref[k] for k in ("objective", "primal", "dual", "gap", "kkt")})
        assert rel(acc["objective"], ref["objective"]) <= 1e-12
        for key in ("primal", "dual", "gap", "kkt"):
            assert rel(acc[key], ref[key]) <= 1e-9, key
    gpu.keep_best(True)
    bx, by = gpu.original_point(2)
    ox, oy = restore(xa, ya, cols_map, lower, upper, fixed, dc, dr, sobj, m0)
    assert np.array_equal(bx, ox) and np.array_equal(by, oy)
    dx2, dy2 = gpu.epoch_displacement(False)
    assert rel(dx2, float(np.sum((x - x0) ** 2))) <= 1e-12
    assert rel(dy2, float(np.sum((y - y0) ** 2))) <= 1e-12
    gpu.set_epoch(True)
    dx2, dy2 = gpu.epoch_displacement(True)
    assert dx2 == 0 and dy2 == 0
    gpu.close()


@pytest.mark.parametrize("case", [(0, 4, 0), (1, 100, 0), (4, 200, 0), (2, 1, 3)])
def test_device_checks_reach_certified_optimum(engine, oracle, case):
    config, scale, index = case
    inst = engine.Instance(config, 42, scale, index)
    host = engine.solve_lp(inst.lp, device_checks=0)
    dev = engine.solve_lp(inst.lp, device_checks=1)
    print(case, "host", host["iterations"], host["objective"], "device", dev["iterations"], dev["objective"],
          dev["device_candidate_checks"], dev["host_candidate_checks"])
    for r in (host, dev):
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6
        assert rel(r["objective"], inst.planted_objective) <= 1e-6
    assert dev["device_candidate_checks"] > 0
    assert dev["host_candidate_checks"] <= 3
    assert oracle.verify(inst.lp, dev["x"], dev["y"])["kkt"] <= 1e-6


def test_device_checks_find_farkas_certificate(engine, oracle):
    # x in [0,1]^2, x1 + x2 >= 1.5 and x1 + 2 x2 <= 1.2: no single row is
    # infeasible, so presolve cannot see it; PDHG's diverging dual gives the ray.
    lp = engine.LP.from_dense([[1, 1], [1, 2]], [1, 1], [0, 0], [1, 1], [1.5, -np.inf], [np.inf, 1.2])
    for checks in (0, 1):
        r = engine.solve_lp(lp, device_checks=checks, presolve=False)
        print(checks, r["status"], r["iterations"], r["message"])
        assert r["status"] == "INFEASIBLE"
