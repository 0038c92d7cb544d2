"""GPU tests of the batched / served engine.

Every batched result must be certified OPTIMAL at rel-KKT <= 1e-6 by the FP64
verifier, match the planted closed-form optimum and the CPU oracle's objective
to 1e-6, and be bit-identical whether the LP runs alone or among others
(each LP is reduced by its own workgroup in a fixed order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / (1 + abs(b))


def solve_batch(engine, lps, checks=5, arena=4 << 30, slots=4096, warm=None, **opts):
    b = engine.Batch(arena_bytes=arena, max_slots=slots, **opts)
    tickets = []
    for k, lp in enumerate(lps):
        x0, y0 = warm[k] if warm else (None, None)
        tickets.append(b.submit(lp, x0, y0))
    b.run(checks)
    out = [b.result(t) for t in tickets]
    stats = b.stats()
    b.close()
    return out, stats


def test_batch_certified_against_planted_and_oracle(engine, oracle):
    insts = [engine.Instance(2, 42, 1, i) for i in range(24)]
    res, stats = solve_batch(engine, [i.lp for i in insts])
    assert stats["finished"] == len(insts)
    for inst, r in zip(insts, res):
        assert r is not None and r["status"] == "OPTIMAL", r and (r["status"], r["message"])
        assert r["kkt_error"] <= 1e-6
        assert rel(r["objective"], inst.planted_objective) <= 1e-6
        acc = oracle.verify(inst.lp, r["x"], r["y"])
        assert acc["kkt"] <= 1e-6
    # the CPU oracle reaches the same optimum on a few of them
    for inst, r in list(zip(insts, res))[:4]:
        cpu = oracle.solve(inst.lp)
        assert cpu["status"] == "OPTIMAL"
        assert rel(r["objective"], cpu["objective"]) <= 1e-6
        print(inst.lp.rows, inst.lp.cols, "iterations batch", r["iterations"], "oracle", cpu["iterations"])


def test_batch_alone_equals_in_mix(engine):
    insts = [engine.Instance(2, 7, 1, i) for i in range(40)]
    mix, _ = solve_batch(engine, [i.lp for i in insts])
    for k in (0, 13, 39):
        alone, _ = solve_batch(engine, [insts[k].lp])
        assert alone[0]["status"] == mix[k]["status"] == "OPTIMAL"
        assert alone[0]["iterations"] == mix[k]["iterations"]
        assert np.array_equal(alone[0]["x"], mix[k]["x"]) and np.array_equal(alone[0]["y"], mix[k]["y"])


def test_global_class_beside_lds_classes(engine, oracle):
    """An LP too large for the LDS classes (m + n = 15,000) runs in the global
    class, which walks the CSR instead of the 16-bit sliced copies, next to
    LDS-class LPs; all are certified, and the large one matches its solve
    alone in the batch bit for bit."""
    big = engine.Instance(1, 42, 200)
    small = [engine.Instance(2, 7, 1, i) for i in range(3)]
    assert big.lp.rows + big.lp.cols > 8000
    mix, _ = solve_batch(engine, [big.lp] + [i.lp for i in small])
    alone, _ = solve_batch(engine, [big.lp])
    print("global class", mix[0]["status"], mix[0]["iterations"], flush=True)
    for inst, r in zip([big] + small, mix):
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6
        assert rel(r["objective"], inst.planted_objective) <= 1e-6
        assert oracle.verify(inst.lp, r["x"], r["y"])["kkt"] <= 1e-6
    assert alone[0]["iterations"] == mix[0]["iterations"]
    assert np.array_equal(alone[0]["x"], mix[0]["x"]) and np.array_equal(alone[0]["y"], mix[0]["y"])


def test_continuous_admission_small_arena(engine):
    """An arena that holds only a few LPs forces many admission rounds while
    other LPs keep iterating; results are unchanged."""
    insts = [engine.Instance(2, 11, 1, i) for i in range(16)]
    big, _ = solve_batch(engine, [i.lp for i in insts])
    small, stats = solve_batch(engine, [i.lp for i in insts], checks=1, arena=24 << 20, slots=4)
    assert stats["launches"] > 4
    for a, b in zip(big, small):
        assert a["status"] == b["status"] == "OPTIMAL"
        assert np.array_equal(a["x"], b["x"]) and a["iterations"] == b["iterations"]


def test_batch_matches_single_engine_accuracy(engine):
    insts = [engine.Instance(0, 42, 4), engine.Instance(4, 42, 200), engine.Instance(1, 42, 400)]
    res, _ = solve_batch(engine, [i.lp for i in insts])
    for inst, r in zip(insts, res):
        single = engine.solve_lp(inst.lp)
        assert r["status"] == single["status"] == "OPTIMAL"
        assert rel(r["objective"], single["objective"]) <= 1e-6
        assert rel(r["objective"], inst.planted_objective) <= 1e-6


def test_warm_start_resolve_of_perturbed_rhs(engine, oracle):
    inst = engine.Instance(2, 5, 1, 2)
    first, _ = solve_batch(engine, [inst.lp])
    rng = np.random.default_rng(0)
    lp = inst.lp
    b = lp.rl * (1 + rng.uniform(-0.05, 0.05, lp.rows))
    pert = engine.LP(lp.ptr, lp.index, lp.value, lp.c, lp.lb, lp.ub, b, b)
    cold, _ = solve_batch(engine, [pert])
    warm, _ = solve_batch(engine, [pert], warm=[(first[0]["x"], first[0]["y"])])
    cpu = oracle.solve(pert)
    for r in (cold[0], warm[0]):
        assert r["status"] == cpu["status"]
        if r["status"] == "OPTIMAL":
            assert rel(r["objective"], cpu["objective"]) <= 1e-6
    print("cold", cold[0]["iterations"], "warm", warm[0]["iterations"], "cpu", cpu["iterations"])


def test_batch_reports_infeasible_and_limits(engine):
    infeasible = engine.LP(np.array([0, 0]), np.zeros(0), np.zeros(0), [0.0], [0.0], [10.0], [1.0], [np.inf])
    contradiction = engine.LP.from_dense([[1, 1], [1, 1]], [0, 0], [0, 0], [10, 10], [3, -np.inf], [np.inf, 1])
    inst = engine.Instance(2, 3, 1, 1)
    res, _ = solve_batch(engine, [infeasible, contradiction])
    assert [r["status"] for r in res] == ["INFEASIBLE", "INFEASIBLE"]
    res, _ = solve_batch(engine, [inst.lp], iteration_limit=200)
    assert res[0]["status"] == "ITERATION_LIMIT" and res[0]["iterations"] == 200


def test_handoff_to_grid_engine(engine, oracle):
    """LPs still running after the handoff count continue on the grid-wide
    engine from their last checked candidate.  Results must be certified, at
    the planted optimum, and the same alone and in the mix, since the move
    depends only on the LP's own iteration count."""
    insts = [engine.Instance(2, 42, 1, i) for i in range(10)]

    def run(lps, checks):
        b = engine.Batch(arena_bytes=4 << 30, max_slots=64)
        b.set_handoff(300, 2)
        tickets = [b.submit(lp) for lp in lps]
        b.run(checks)
        out = [b.result(t) for t in tickets]
        stats = b.stats()
        b.close()
        return out, stats

    mix, stats = run([i.lp for i in insts], 2)
    assert stats["handoffs"] > 0 and stats["handoffs_running"] == 0
    moved = [k for k, r in enumerate(mix) if r["iterations"] > 300]
    assert len(moved) == stats["handoffs"]
    for k, inst in enumerate(insts):
        r = mix[k]
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6, (k, r["status"])
        assert rel(r["objective"], inst.planted_objective) <= 1e-6
        assert oracle.verify(inst.lp, r["x"], r["y"])["kkt"] <= 1e-6
    for k in moved[:2]:
        alone, _ = run([insts[k].lp], 5)
        assert np.array_equal(alone[0]["x"], mix[k]["x"]) and alone[0]["iterations"] == mix[k]["iterations"]


def test_launch_shape_does_not_change_results(engine, monkeypatch):
    """A crowded class runs 256-thread workgroups with several rows per thread
    (1,024 threads, one row each, for the two larger classes) and a separate
    check kernel; once few LPs are left it switches to
    the fused tail kernel (1024-thread workgroups that also run the checks).
    The per-tile sums keep every step's arithmetic independent of that shape,
    and the fused check forms its sums on 256 threads as the check kernel
    does, so an LP that changes shape mid-solve matches the same LP run alone
    throughout in the tail shape, and the rows-per-thread setting changes
    nothing."""
    insts = [engine.Instance(2, 5, 1, i) for i in range(900)]
    mix, stats = solve_batch(engine, [i.lp for i in insts], checks=1)
    assert stats["launches"] > 2
    # the smallest class starts with well over 256 LPs
    assert sum(i.lp.rows + i.lp.cols <= 2048 for i in insts) > 400
    picks = [0, 101, 257, 511, 899]
    alone, _ = solve_batch(engine, [insts[k].lp for k in picks])
    for k, a in zip(picks, alone):
        assert a["status"] == mix[k]["status"] == "OPTIMAL"
        assert a["iterations"] == mix[k]["iterations"]
        # the checks decided alike: the same restarts and primal weight updates
        assert a["restarts"] == mix[k]["restarts"] and a["weight_updates"] == mix[k]["weight_updates"]
        assert np.array_equal(a["x"], mix[k]["x"]) and np.array_equal(a["y"], mix[k]["y"])
    # 2 or 4 rows per thread in the smallest class (read when a batch is created)
    for rows in ("2", "4"):
        monkeypatch.setenv("NIRYUKTI_BATCH_ROWS_PER_THREAD", rows)
        other, _ = solve_batch(engine, [i.lp for i in insts[:300]], checks=1)
        for k in (0, 101, 257):
            assert other[k]["iterations"] == mix[k]["iterations"], rows
            assert np.array_equal(other[k]["x"], mix[k]["x"]) and np.array_equal(other[k]["y"], mix[k]["y"]), rows


def test_power_iterations_in_batch(engine):
    """--power-iterations in the batched engine: the same host estimate of
    ||A||_2 as the single-LP engine (the same seeded power iteration on the
    same prepared matrix), the larger first step, and a certified answer that
    is the same alone and in a mix."""
    insts = [engine.Instance(2, 7, 1, i) for i in range(4)]
    mix, _ = solve_batch(engine, [i.lp for i in insts], power_iterations=20)
    alone, _ = solve_batch(engine, [insts[1].lp], power_iterations=20)
    single = engine.solve_lp(insts[1].lp, power_iterations=20)
    print("estimates", mix[1]["operator_norm_estimate"], single["operator_norm_estimate"], flush=True)
    assert mix[1]["operator_norm_estimate"] == single["operator_norm_estimate"] > 0
    for inst, r in zip(insts, mix):
        assert r["status"] == "OPTIMAL" and rel(r["objective"], inst.planted_objective) <= 1e-6
    assert alone[0]["iterations"] == mix[1]["iterations"] and np.array_equal(alone[0]["x"], mix[1]["x"])


def test_autotuner_keeps_results(engine, monkeypatch):
    """The launch-shape autotuner tries every shape four times on the first
    full rounds of a crowded class and keeps the start shape unless another is
    clearly faster; results are bit-identical to a run with the tuner off
    (shapes never change the arithmetic)."""
    insts = [engine.Instance(2, 5, 1, i) for i in range(2000)]
    lps = [i.lp for i in insts]
    # every LP resident from the first round, and trials counted from 128 LPs
    # in the class (512 by default): the smallest class (over 256 LPs, so not
    # in the tail group) reaches a decision within this batch, after 2
    # settling and 16 trial rounds
    assert sum(i.lp.rows + i.lp.cols <= 2048 for i in insts) > 1200
    monkeypatch.setenv("NIRYUKTI_BATCH_TUNE_MIN_LPS", "128")
    b = engine.Batch(arena_bytes=8 << 30, max_slots=4096)
    tickets = [b.submit(lp) for lp in lps]
    assert b.admit() == len(lps)
    b.run(1)
    tuned, stats = [b.result(t) for t in tickets], b.stats()
    b.close()
    monkeypatch.delenv("NIRYUKTI_BATCH_TUNE_MIN_LPS")
    print("tuned shapes", stats["class_threads"], stats["class_rows"], stats["class_tuned"], flush=True)
    assert stats["class_tuned"][0] == 1
    monkeypatch.setenv("NIRYUKTI_BATCH_TUNE", "0")
    fixed, fstats = solve_batch(engine, lps, checks=1, arena=8 << 30)
    assert list(fstats["class_threads"]) == [256, 1024, 1024] and list(fstats["class_rows"]) == [1, 1, 1]
    monkeypatch.delenv("NIRYUKTI_BATCH_TUNE")
    monkeypatch.setenv("NIRYUKTI_BATCH_SHAPES", "2,1,0")
    other, ostats = solve_batch(engine, lps[:600], checks=1)
    assert list(ostats["class_threads"]) == [256, 256, 256] and list(ostats["class_rows"]) == [4, 2, 1]
    for a, b in zip(tuned, fixed):
        assert a["status"] == b["status"] and a["iterations"] == b["iterations"]
        assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    for a, b in zip(other, fixed):
        assert a["iterations"] == b["iterations"] and np.array_equal(a["x"], b["x"])


def test_submit_from_another_thread_beside_steps(engine):
    """niryukti_batch_submit on a second thread while the scheduling thread
    steps (the streamed-pass pattern): every LP finishes once, under the ticket
    submit returned, bit-identical to the same LP solved alone."""
    import threading

    lps = [engine.Instance(2, 42, 1, i).lp for i in range(48)]
    b = engine.Batch(arena_bytes=4 << 30, max_slots=64)
    tickets, errors = [None] * len(lps), []

    def submitter():
        try:
            for k, lp in enumerate(lps):
                tickets[k] = b.submit(lp)
                if k % 8 == 7:
                    threading.Event().wait(0.005)  # arrive over several rounds
        except BaseException as e:
            errors.append(e)

    th = threading.Thread(target=submitter)
    th.start()
    import time

    out, t0 = {}, time.perf_counter()
    while True:
        running = b.step(2)
        for t in b.finished(1024):
            assert t not in out
            out[t] = b.result(t)
        if not th.is_alive() and running == 0 and len(out) == len(lps):
            break
        assert time.perf_counter() - t0 < 60
    th.join()
    b.close()
    assert not errors
    assert sorted(tickets) == list(range(len(lps)))
    for k in (0, 17, 47):
        alone, _ = solve_batch(engine, [lps[k]])
        r = out[tickets[k]]
        assert r["status"] == alone[0]["status"] == "OPTIMAL"
        assert r["iterations"] == alone[0]["iterations"]
        assert np.array_equal(r["x"], alone[0]["x"]) and np.array_equal(r["y"], alone[0]["y"])


def test_collect_returns_every_result_once(engine):
    """niryukti_batch_collect: every finished ticket once, with the fields
    niryukti_batch_result gives, and released when asked."""
    lps = [engine.Instance(2, 42, 1, i).lp for i in range(24)]
    ref, _ = solve_batch(engine, lps)
    b = engine.Batch(arena_bytes=4 << 30, max_slots=64)
    tickets = [b.submit(lp) for lp in lps]
    seen = {}
    while True:
        running = b.step(2)
        fin, rows = b.collect(capacity=5)  # small capacity: several calls per round
        for t, r in zip(fin.tolist(), rows):
            assert t not in seen
            seen[t] = r
        if running == 0 and len(seen) == len(lps):
            break
    assert sorted(seen) == tickets
    for t, r in zip(tickets, ref):
        row = seen[t]
        assert row["status"].decode() == r["status"] == "OPTIMAL"
        assert row["iterations"] == r["iterations"] and row["objective"] == r["objective"]
        assert row["kkt_error"] == r["kkt_error"] and row["x"] == 0
    with pytest.raises(RuntimeError):
        b.result(tickets[0])  # released by collect
    b.close()


def test_mixed_precision_batch_certified_against_oracle(engine, oracle):
    """Mixed precision in the batch (FP32 matrix values, FP64 products, sums
    and iterates, FP64 values once a candidate reaches 1e-5 or the kkt stalls):
    every LP is certified by the host's FP64 verifier at 1e-6 and its objective
    lies within 1e-6 of the planted optimum and of the FP64 oracle's; the
    objective error against the oracle is printed per LP.  Every LP ran on
    FP32 values first, and most finished on FP64 values (an LP whose candidate
    falls from above 1e-5 to 1e-6 in one check period needs none)."""
    insts = [engine.Instance(2, 42, 1, i) for i in range(16)]
    res, _ = solve_batch(engine, [i.lp for i in insts], matrix_precision="mixed")
    assert sum(r["fp32_iterations"] < r["iterations"] for r in res) >= len(res) // 2
    for k, (inst, r) in enumerate(zip(insts, res)):
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6, (r["status"], r["message"])
        assert 0 < r["fp32_iterations"] <= r["iterations"], (r["fp32_iterations"], r["iterations"])
        assert rel(r["objective"], inst.planted_objective) <= 1e-6
        assert oracle.verify(inst.lp, r["x"], r["y"])["kkt"] <= 1e-6
        if k < 4:
            cpu = oracle.solve(inst.lp)
            err = rel(r["objective"], cpu["objective"])
            print(f"mixed LP {k}: fp32 iterations {r['fp32_iterations']} of {r['iterations']}, "
                  f"objective error vs FP64 oracle {err:.2e}", flush=True)
            assert err <= 1e-6


def test_mixed_precision_batch_alone_equals_in_mix(engine):
    """An LP's switch from FP32 to FP64 values depends only on its own checks.
    Among 600 LPs the smallest class runs crowded (more than 256 LPs), each
    period launching the step in FP32 values over its LPs still on them and
    in FP64 values over those that switched; alone an LP runs in the tail
    launch, whose workgroups pick the value type per LP.  Either way an LP
    follows the same trajectory bit for bit; the LPs compared are small ones
    that switched earliest in the mix (while their class was crowded) and
    three others."""
    insts = [engine.Instance(2, 7, 1, i) for i in range(600)]
    mix, stats = solve_batch(engine, [i.lp for i in insts], matrix_precision="mixed")
    assert stats["class_launches"][0] > 0
    small = [k for k, i in enumerate(insts) if i.lp.rows + i.lp.cols <= 1500 and mix[k]["status"] == "OPTIMAL"
             and mix[k]["fp32_iterations"] < mix[k]["iterations"]]
    assert len(small) > 256
    early = sorted(small, key=lambda k: mix[k]["fp32_iterations"])[:3]
    for k in early + [0, 13, 39]:
        alone, _ = solve_batch(engine, [insts[k].lp], matrix_precision="mixed")
        assert alone[0]["status"] == mix[k]["status"] == "OPTIMAL"
        assert alone[0]["iterations"] == mix[k]["iterations"]
        assert alone[0]["fp32_iterations"] == mix[k]["fp32_iterations"]
        assert np.array_equal(alone[0]["x"], mix[k]["x"]) and np.array_equal(alone[0]["y"], mix[k]["y"])


def test_time_limit_stops_only_the_slow_lp(engine):
    """A time limit stops a running LP at the round that finds it past the
    limit (TIME_LIMIT, with its best candidate), is left out of the next round
    and frees its slot; the LPs around it finish OPTIMAL and equal to their
    solves alone."""
    slow = engine.Instance(2, 42, 1, 222)  # runs to the 100,000-iteration limit (~1 s alone)
    quick = [engine.Instance(2, 42, 1, i) for i in (0, 1, 2)]
    res, stats = solve_batch(engine, [slow.lp] + [q.lp for q in quick], time_limit=0.3)
    assert stats["finished"] == 4 and stats["active"] == 0
    assert res[0]["status"] == "TIME_LIMIT"
    assert 0 < res[0]["iterations"] < 100_000
    assert res[0]["x"] is not None and np.all(np.isfinite(res[0]["x"]))
    alone = [solve_batch(engine, [q.lp])[0][0] for q in quick]
    for r, a in zip(res[1:], alone):
        assert r["status"] == a["status"] == "OPTIMAL"
        assert r["iterations"] == a["iterations"] and np.array_equal(r["x"], a["x"])


def _long_row_lp(engine, m, n, per_row, seed, dense_col=False, dense_row=False):
    """A sparse LP with per_row random entries a row, optionally one fixed
    column in every row or one row over every column; a ranged row around the
    activity of a random point in the box keeps it feasible."""
    rng = np.random.default_rng(seed)
    rows = [np.sort(rng.choice(n, size=per_row, replace=False)) for _ in range(m)]
    if dense_row:
        rows[0] = np.arange(n)
    if dense_col:
        rows = [np.append(r, n) for r in rows]
    cols = n + (1 if dense_col else 0)
    ptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    index = np.concatenate(rows)
    value = rng.normal(size=len(index))
    lb, ub = np.zeros(cols), rng.uniform(1, 10, size=cols)
    if dense_col:
        lb[n] = ub[n] = 1.0
    point = rng.uniform(0, 1, size=cols) * ub
    point[lb == ub] = lb[lb == ub]
    act = np.array([value[ptr[i]:ptr[i + 1]] @ point[index[ptr[i]:ptr[i + 1]]] for i in range(m)])
    rl, ru = act - rng.uniform(0.5, 2, size=m), act + rng.uniform(0.5, 2, size=m)
    return engine.LP(ptr, index, value, rng.normal(size=cols), lb, ub, rl, ru)


def test_long_rows_in_checks(engine, oracle):
    """Rows longer than 2,048 entries are summed by the checks' workgroup tree
    (the others one thread per row).  A fixed column in all 2,100 rows lies
    in the original A^T only (presolve removes it, so the LP runs in an LDS
    class, here in the tail launch); a row over all 2,200 columns stays in the
    scaled A too (the LP runs in the global class, checked by the crowded check
    kernel).  Both are certified, reach the oracle's objective, and equal their
    solves alone bit for bit."""
    lps = [_long_row_lp(engine, 2100, 2000, 2, 1, dense_col=True),
           _long_row_lp(engine, 600, 2200, 3, 5, dense_row=True)]
    assert max(np.diff(lps[0].ptr)) <= 3 and lps[0].rows > 2048
    assert max(np.diff(lps[1].ptr)) == 2200
    mix, stats = solve_batch(engine, lps)
    assert stats["class_launches"][4] > 0 and stats["class_launches"][3] > 0
    for lp, r in zip(lps, mix):
        assert r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6, (r["status"], r["message"])
        assert oracle.verify(lp, r["x"], r["y"])["kkt"] <= 1e-6
        cpu = oracle.solve(lp)
        assert cpu["status"] == "OPTIMAL"
        assert rel(r["objective"], cpu["objective"]) <= 1e-6
        alone, _ = solve_batch(engine, [lp])
        assert alone[0]["iterations"] == r["iterations"]
        assert np.array_equal(alone[0]["x"], r["x"]) and np.array_equal(alone[0]["y"], r["y"])
