"""End-to-end GPU solves against the CPU oracle and the planted optimum.

Both paths must reach OPTIMAL at rel-KKT <= 1e-6 (verified in FP64 in the
original space), and the GPU objective must match the oracle's and the
closed-form planted optimum to 1e-6 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(0, 4, 0), (0, 1, 0), (2, 1, 0), (2, 1, 7), (4, 200, 0), (1, 400, 0), (1, 200, 0)]


def rel(a, b):
    return abs(a - b) / (1 + abs(b))


@pytest.mark.parametrize("case", CASES)
def test_solve_matches_oracle(engine, oracle, case):
    config, scale, index = case
    inst = engine.Instance(config, 42, scale, index)
    cpu = oracle.solve(inst.lp)
    gpu = engine.solve_lp(inst.lp)
    print(case, "gpu", gpu["status"], gpu["iterations"], gpu["objective"], gpu["kkt_error"],
          "cpu", cpu["status"], cpu["iterations"], cpu["objective"], cpu["kkt"])
    assert cpu["status"] == "OPTIMAL"
    assert gpu["status"] == "OPTIMAL"
    assert gpu["kkt_error"] <= 1e-6
    assert rel(gpu["objective"], cpu["objective"]) <= 1e-6
    assert rel(gpu["objective"], inst.planted_objective) <= 1e-6
    # Independent FP64 verification of the returned point by the oracle's verifier.
    acc = oracle.verify(inst.lp, gpu["x"], gpu["y"])
    assert acc["kkt"] <= 1e-6


def test_deterministic_and_graph_invariant(engine):
    inst = engine.Instance(0, 3, 4)
    a = engine.solve_lp(inst.lp, graphs=True)
    b = engine.solve_lp(inst.lp, graphs=True)
    c = engine.solve_lp(inst.lp, graphs=False)
    assert a["iterations"] == b["iterations"] == c["iterations"]
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["x"], c["x"])
    assert np.array_equal(a["y"], b["y"]) and np.array_equal(a["y"], c["y"])


def long_row_lp(engine, seed=11, m=2000, n=6000, per_col=3, long_rows=24, long_len=3000):
    """Planted LP whose A has rows far above the lane-group limit, so the
    step kernels sum them by segments (and one of the segment workgroups,
    whichever finishes last, completes each such row)."""
    rng = np.random.default_rng(seed)
    rows = [rng.choice(m, per_col, replace=False) for _ in range(n)]
    r = np.concatenate(rows)
    c_ = np.repeat(np.arange(n), per_col)
    lr = np.repeat(rng.choice(m, long_rows, replace=False), long_len)
    lc = np.concatenate([rng.choice(n, long_len, replace=False) for _ in range(long_rows)])
    A = np.zeros((m, n))
    np.add.at(A, (np.concatenate([r, lr]), np.concatenate([c_, lc])), rng.standard_normal(len(r) + len(lr)))
    x0 = rng.uniform(0, 10, n) * (rng.random(n) > 0.3)
    y0 = rng.standard_normal(m)
    c = A.T @ y0 + np.where(x0 == 0, rng.uniform(0, 1, n), 0.0)
    b = A @ x0
    return engine.LP.from_dense(A, c, np.zeros(n), np.full(n, 10.0), b, b)


def test_long_row_solves_are_deterministic(engine):
    lp = long_row_lp(engine)
    for checks in (0, 1):  # host checks, device checks
        a, b = (engine.solve_lp(lp, iteration_limit=600, device_checks=checks) for _ in range(2))
        assert a["iterations"] == b["iterations"] and a["rejected_steps"] == b["rejected_steps"]
        assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])


def test_index_width_invariant(engine):
    inst = engine.Instance(2, 5, 1, 3)
    a = engine.solve_lp(inst.lp, gpu_indices="32")
    b = engine.solve_lp(inst.lp, gpu_indices="64")
    assert a["index_bits"] == 32 and b["index_bits"] == 64
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["x"], b["x"])


def test_mixed_precision_verified_in_fp64(engine, oracle):
    """FP32 matrix values without refinement (the reference's "mixed")
    converge to the FP32-rounded LP, so on this instance they stall at a
    rel-KKT of the original LP just above 1e-6 (7.95e-6) and stop at the
    iteration limit, never claiming OPTIMAL; the fp32 mode's FP64 refinement
    phase certifies the same LP, and the oracle's FP64 verifier agrees."""
    inst = engine.Instance(2, 42, 1, 11)
    mixed = engine.solve_lp(inst.lp, matrix_precision="mixed")
    print("mixed", mixed["status"], mixed["kkt_error"], flush=True)
    assert mixed["status"] == "ITERATION_LIMIT" and 1e-6 < mixed["kkt_error"] < 1e-4
    refined = engine.solve_lp(inst.lp, matrix_precision="fp32")
    assert refined["status"] == "OPTIMAL" and refined["fp32_iterations"] < refined["iterations"]
    assert oracle.verify(inst.lp, refined["x"], refined["y"])["kkt"] <= 1e-6
    assert rel(refined["objective"], inst.planted_objective) <= 1e-6


@pytest.mark.parametrize("method", ["halpern", "rhpdhg", "r2hpdhg"])
def test_anchored_methods(engine, oracle, method):
    # Halpern LP with both finite bounds and ranged rows (tests/test_research.cpp:94-108): optimum 2.
    lp = engine.LP.from_dense([[1, 1], [1, -1]], [1, 2], [0, 0], [4, 4], [2, -1], [3, 2])
    g = engine.solve_lp(lp, method=method, adaptive=False)
    c = oracle.solve(lp, method=method, adaptive=False)
    assert c["status"] == "OPTIMAL" and g["status"] == "OPTIMAL"
    assert abs(g["objective"] - 2) <= 1e-5
    assert oracle.verify(lp, g["x"], g["y"])["kkt"] <= 1e-6


def test_analytical_lps(engine):
    # tests/test_core.cpp:43-97 analytical optima
    lp = engine.LP.from_dense([[1, 2]], [1, 1], [0, 0], [10, 10], [4], [np.inf])
    assert abs(engine.solve_lp(lp)["objective"] - 2) < 1e-4 * 3
    fixed = engine.LP.from_dense([[1, 2]], [1, 1], [1, 0], [1, 10], [4], [np.inf])
    for presolve in (True, False):
        r = engine.solve_lp(fixed, presolve=presolve)
        assert r["status"] == "OPTIMAL" and abs(r["objective"] - 2.5) < 1e-4 * 3.5
    scaled = engine.LP.from_dense([[1e8, 2e8]], [1, 1], [0, 0], [10, 10], [4e8], [np.inf])
    assert abs(engine.solve_lp(scaled)["objective"] - 2) < 1e-4 * 3
    free = engine.LP.from_dense([[1]], [1], [-np.inf], [np.inf], [-2], [-2])
    assert abs(engine.solve_lp(free)["objective"] + 2) < 1e-4 * 3
    infeasible = engine.LP(np.array([0, 0]), np.zeros(0), np.zeros(0), [0.0], [0.0], [10.0], [1.0], [np.inf])
    assert engine.solve_lp(infeasible)["status"] == "INFEASIBLE"
    unbounded = engine.LP(np.array([0]), np.zeros(0), np.zeros(0), [-1.0], [0.0], [np.inf], [], [])
    assert engine.solve_lp(unbounded)["status"] == "UNBOUNDED"
    m = engine.LP.from_dense([[1, 2]], [1, 1], [0, 0], [10, 10], [4], [np.inf])
    assert engine.solve_lp(m, iteration_limit=0)["status"] == "ITERATION_LIMIT"
    assert engine.solve_lp(m, iteration_limit=0, time_limit=0)["status"] == "TIME_LIMIT"
    contradiction = engine.LP.from_dense([[1, 1], [1, 1]], [0, 0], [0, 0], [10, 10], [3, -np.inf], [np.inf, 1])
    assert engine.solve_lp(contradiction)["status"] == "INFEASIBLE"


def test_concurrent_solves_from_threads(engine):
    """Single-LP solves on several host threads at once (each captures its own
    hipGraph on its own stream) while a batch runs on the main thread: every
    result must equal the same solve run alone."""
    from concurrent.futures import ThreadPoolExecutor

    insts = [engine.Instance(2, 42, 1, i) for i in (421, 542, 7, 11, 19, 23)]
    alone = [engine.solve_lp(i.lp) for i in insts]
    batch_lps = [engine.Instance(2, 42, 1, i).lp for i in range(40, 56)]
    with ThreadPoolExecutor(4) as pool:
        futures = [pool.submit(engine.solve_lp, i.lp) for i in insts]
        b = engine.Batch(arena_bytes=1 << 30, max_slots=64)
        tickets = [b.submit(lp) for lp in batch_lps]
        b.run(1)
        batch = [b.result(t) for t in tickets]
        b.close()
        results = [f.result() for f in futures]
    for a, r in zip(alone, results):
        assert r["status"] == a["status"] == "OPTIMAL", r["message"]
        assert np.array_equal(a["x"], r["x"]) and a["iterations"] == r["iterations"]
    assert all(r["status"] == "OPTIMAL" for r in batch)
