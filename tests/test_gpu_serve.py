"""Serving loop (continuous batching) on the GPU.

A mixed stream of LPs from the serving workload (nonzeros log-uniform over
[1e3, 5e6], batch generator) plus warm-started re-solves with the right-hand
side perturbed by up to 5% is pushed through one Server while earlier requests
are still running.  Every result must be certified OPTIMAL and be the same
(bitwise) as the same LP solved alone; the planted instances must also hit
their planted optimum, the perturbed ones the oracle's optimum.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = [2, 3, 6, 10, 11, 13, 15, 16, 18, 21, 22, 23, 25]
MID = [17, 20, 26]  # 33k-45k nonzeros: routed to the grid engine below


def rel(a, b):
    return abs(a - b) / (1 + abs(b))


def perturbed(engine, lp, seed):
    rng = np.random.default_rng(seed)
    f = 1 + rng.uniform(-0.05, 0.05, lp.rows)
    rl = np.where(np.isfinite(lp.rl), lp.rl * f, lp.rl)
    ru = np.where(np.isfinite(lp.ru), lp.ru * f, lp.ru)
    return engine.LP(lp.ptr, lp.index, lp.value, lp.c, lp.lb, lp.ub, rl, ru, lp.offset, lp.sense)


def alone_batch(engine, lp, x0=None, y0=None):
    b = engine.Batch(arena_bytes=1 << 30, max_slots=4)
    t = b.submit(lp, x0, y0)
    b.run(1)
    r = b.result(t)
    b.close()
    return r


def test_stream_matches_alone(engine, oracle):
    insts = {i: engine.Instance(3, 42, 1, i) for i in SMALL + MID}
    with engine.Server(hbm_budget=8 << 30, large_nonzeros=30_000, grid_workers=2) as server:
        first = {i: server.submit(insts[i].lp) for i in SMALL + MID}
        # warm re-solves arrive while the first wave is running
        base = {i: first[i].wait() for i in (SMALL[0], SMALL[5], MID[0])}
        redo = {}
        for k, i in enumerate(base):
            lp2 = perturbed(engine, insts[i].lp, 100 + k)
            redo[i] = (lp2, base[i]["x"], base[i]["y"], server.submit(lp2, base[i]["x"], base[i]["y"]))
        results = {i: r.wait() for i, r in first.items()}
        routes = {i: r.route for i, r in first.items()}
        lat = [r.latency for r in first.values()]
    assert all(routes[i] == "batch" for i in SMALL) and all(routes[i] == "grid" for i in MID)
    print("latency p50/max", np.percentile(lat, 50), max(lat))
    for i, r in results.items():
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6, (i, r["status"])
        assert rel(r["objective"], insts[i].planted_objective) <= 1e-6
        alone = alone_batch(engine, insts[i].lp) if routes[i] == "batch" else engine.solve_lp(insts[i].lp)
        assert np.array_equal(alone["x"], r["x"]) and np.array_equal(alone["y"], r["y"]), i
        assert alone["iterations"] == r["iterations"]
    for i, (lp2, x0, y0, req) in redo.items():
        r = req.wait()
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6
        ref = oracle.solve(lp2)
        assert ref["status"] == "OPTIMAL"
        assert rel(r["objective"], ref["objective"]) <= 1e-6
        assert oracle.verify(lp2, r["x"], r["y"])["kkt"] <= 1e-6
        if req.route == "batch":
            alone = alone_batch(engine, lp2, x0, y0)
        else:
            alone = engine.solve_lp(lp2, initial_x=x0, initial_y=y0)
        assert np.array_equal(alone["x"], r["x"])


def test_solve_request_shape(engine):
    model = {"name": "linear", "sense": "min", "variables": [{"name": "x", "ub": 10}],
             "objective": {"linear": [3]}, "constraints": [{"name": "demand", "lb": 2, "coefficients": {"x": 1}}]}
    with engine.Server(hbm_budget=1 << 30) as server:
        code, body = server.solve_request({"model": model, "options": {"tol": 1e-6}})
        assert code == 200 and body["status"] == "OPTIMAL"
        assert abs(body["objective"] - 6) <= 1e-6 * 7
        assert body["serving"]["route"] == "batch"
        code, body = server.solve_request({"model": model, "options": {"iterations": 50000}})
        assert code == 200 and body["serving"]["route"] == "grid" and body["status"] == "OPTIMAL"
        assert server.solve_request({"model": model, "extra": 1})[0] == 400
        assert server.solve_request({"model": model, "options": {"threads": 0}})[0] == 400
        assert server.solve_request({"model": {"variables": "bad"}})[0] == 400


def test_lps_beyond_the_lds_classes_take_the_grid_route(engine):
    """An LP with more than 8,000 rows + columns would run in the batched
    engine's global class (one workgroup walking its CSR from memory, 2-80x
    slower per iteration than the grid engine), so the server routes it to the
    grid engine by size even when its nonzeros are below large_nonzeros."""
    small, big = engine.Instance(3, 42, 1, 3), engine.Instance(3, 42, 1, 17)
    assert small.lp.rows + small.lp.cols <= 8000 < big.lp.rows + big.lp.cols
    with engine.Server(hbm_budget=8 << 30, large_nonzeros=10 ** 9, grid_workers=1) as server:
        rs, rb = server.submit(small.lp), server.submit(big.lp)
        a, b = rs.wait(), rb.wait()
    assert rs.route == "batch" and rb.route == "grid"
    assert a["status"] == b["status"] == "OPTIMAL"
    assert np.array_equal(b["x"], engine.solve_lp(big.lp)["x"])
