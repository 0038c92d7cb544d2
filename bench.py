"""Benchmark: LPs solved per second to rel-KKT <= 1e-6 on one MI355X.

Workload: the batched workload of 4,096 independent LPs (m log-uniform in
[50, 2000], n uniform in [2m, 3m], 3-8 nonzeros per column at uniform rows,
N(0,1) values, box [0, U] with U ~ U(1, 50), planted optimum; seed 42), solved
by the batched engine (one workgroup per LP, FP64 iterates and matrices).  A
solve counts only when the engine reports OPTIMAL, i.e. after the host's FP64
original-space verifier with the outward-rounded bound certified kkt <= 1e-6.

One step = one pass of the engine over the whole batch.  The passes stream
through one engine the way a server takes arriving work: pass k+1 is submitted
once fewer than --stream-threshold LPs of the earlier passes are still
running, so the next pass's prepares and its crowded phase overlap the few
slow LPs that end a pass.  The timed region runs from the first submit of the
first timed pass to the last certified result of the last one: host presolve
and scaling (on the engine's prepare workers, beside the device rounds),
upload, iteration, checks and certification are all inside, for every LP.
Also reported: p50/p99 latency (each LP's own submit to its certified
result), one pass alone on a fresh engine (`one_pass_alone`), the solve phase
alone (`solve_phase`: LPs prepared and uploaded before the clock starts), the
plain one-LP-at-a-time GPU engine on a bounded sample, the step kernel's
roofline, and the CPU baseline (the reference's own CPU PDHG, or the oracle
when that build is absent) on the host cores of this machine on a bounded
sample, at one thread and on all cores.  --no-stream runs the passes one
after another, each on a fresh engine.

    python bench.py [--gpus 1] [--steps 4] [--warmup 1] [--count 4096]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "LPs solved/sec to rel-KKT≤1e-6 + p50/p99 latency, served vs plain PDHG, 1 GPU"
HBM_PEAK_GBS = 8000.0
# launch shapes (threads, rows per thread) of the crowded classes' step kernel,
# in the order of the engine's per-shape statistics
LDS_SHAPES = ((256, 1), (256, 2), (256, 4), (1024, 1))
SEED = 42


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_instances(engine, count):
    return [engine.Instance(2, SEED, 1, i) for i in range(count)]


def batch_step(engine, insts, checks, handoff=0, prepared=False, precision="fp64"):
    """One step over the whole batch.  Timed from the first submit to the last
    certified result; with `prepared`, every LP is prepared and uploaded
    before the clock starts (the solve phase alone)."""
    b = engine.Batch(arena_bytes=96 << 30, max_slots=len(insts), time_limit=1e9, matrix_precision=precision)
    b.set_handoff(handoff, 4)
    if prepared:
        tickets = [b.submit(i.lp) for i in insts]
        assert b.admit() == len(insts), "arena could not hold the whole batch"
    t0 = time.perf_counter()
    if not prepared:
        tickets = [b.submit(i.lp) for i in insts]
    b.run(checks)
    elapsed = time.perf_counter() - t0
    results = [b.result(t) for t in tickets]
    stats = b.stats()
    b.close()
    solved, latencies = 0, []
    for inst, r in zip(insts, results):
        ok = r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6 and \
            abs(r["objective"] - inst.planted_objective) <= 1e-6 * (1 + abs(inst.planted_objective))
        solved += ok
        if ok:
            # an LP's `seconds` run from its own submit to its certified result
            latencies.append(r["seconds"])
    return elapsed, solved, results, stats, latencies


def batch_stream(b, insts, passes, checks, threshold):
    """`passes` passes of the batch streamed through one engine `b`: pass 0 is
    submitted at once, and pass k+1 as soon as fewer than `threshold` LPs of
    the earlier passes are still running, so its prepares and its crowded
    phase overlap the slowest LPs of the pass before (continuous batching: the
    engine admits the new LPs beside the old ones).  A submitter thread hands
    the LPs to the engine (niryukti_batch_submit is safe beside a running
    step) while this thread runs the scheduling rounds and collects results.
    Timed from the first submit to the last certified result; every LP's
    latency runs from its own submit.  Returns (elapsed, per-pass results)."""
    import threading

    lock = threading.Condition()
    tickets = [[None] * len(insts) for _ in range(passes)]
    state = {"submitted": 0, "finished": 0, "done_submitting": False, "error": None, "abort": False}

    def submitter():
        try:
            for k in range(passes):
                with lock:
                    lock.wait_for(lambda: state["abort"] or k == 0
                                  or state["submitted"] - state["finished"] < threshold)
                    if state["abort"]:
                        return
                for i, inst in enumerate(insts):
                    t = b.submit(inst.lp)
                    tickets[k][i] = t
                    with lock:
                        state["submitted"] += 1
        except BaseException as e:  # reported by the scheduling thread
            state["error"] = e
        finally:
            with lock:
                state["done_submitting"] = True

    out, collected = {}, []
    switch = sys.getswitchinterval()
    sys.setswitchinterval(1e-4)  # the scheduling thread's bookkeeping should not wait on the submitter
    t0 = time.perf_counter()
    th = threading.Thread(target=submitter, daemon=True)
    th.start()
    in_step = in_results = 0.0
    try:
        while True:
            ts = time.perf_counter()
            running = b.step(checks)
            tr = time.perf_counter()
            fin, rows = b.collect()  # results and release of every LP that finished, in one call
            if len(fin):
                collected.append((fin, rows))
            in_step += tr - ts
            in_results += time.perf_counter() - tr
            with lock:
                state["finished"] += len(fin)
                lock.notify_all()
                finished_all = state["done_submitting"] and state["finished"] == state["submitted"]
            if state["error"] is not None:
                raise state["error"]
            if finished_all:
                break
            if running == 0 and not len(fin):
                time.sleep(1e-4)  # the submitter is between passes
        elapsed = time.perf_counter() - t0
        log(f"stream: {elapsed:.3f} s, inside step {in_step:.3f} s, reading results {in_results:.3f} s")
    finally:
        sys.setswitchinterval(switch)
        with lock:
            state["abort"] = True
            lock.notify_all()
        th.join()
    for fin, rows in collected:
        for t, r in zip(fin.tolist(), rows):
            out[t] = {"status": r["status"].decode(), "kkt_error": float(r["kkt_error"]),
                      "objective": float(r["objective"]), "seconds": float(r["seconds"]),
                      "iterations": int(r["iterations"]), "message": r["message"].decode()}
    return elapsed, [[out[t] for t in row] for row in tickets]


def score(insts, results):
    """Certified solves of one pass (OPTIMAL, kkt <= 1e-6, objective within
    1e-6 of the planted optimum) and their latencies."""
    solved, latencies = 0, []
    for inst, r in zip(insts, results):
        ok = r["status"] == "OPTIMAL" and r["kkt_error"] <= 1e-6 and \
            abs(r["objective"] - inst.planted_objective) <= 1e-6 * (1 + abs(inst.planted_objective))
        solved += ok
        if ok:
            latencies.append(r["seconds"])
    return solved, latencies


def plain_sample(engine, insts, budget_s=15.0):
    """One LP at a time through the single-LP GPU engine (bounded sample):
    each solve_lp call prepares, uploads, iterates and certifies one LP."""
    solved, t0, n = 0, time.perf_counter(), 0
    for inst in insts:
        r = engine.solve_lp(inst.lp, time_limit=1e9)
        n += 1
        solved += r["status"] == "OPTIMAL"
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"lps": n, "solved": solved, "seconds": round(dt, 3), "lps_per_s": round(solved / dt, 2)}


def machine():
    """Host CPU description: nproc, the cores this process may use, lscpu."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": _cpu_model()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Thread(s) per core", "Core(s) per socket", "Socket(s)", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def physical_cores(info):
    """Physical cores this process may use (SMT siblings counted once), at most
    16 (the host share of one GPU on this machine class)."""
    try:
        tpc = max(1, int(info.get("Thread(s) per core", "1")))
    except ValueError:
        tpc = 1
    return max(1, min(16, info["affinity"] // tpc))


def cpu_baseline(count, workers, budget_s):
    """Reference CPU PDHG (oracle/_ref, built from the reference's sources) on
    `workers` host cores over the same batch: worker w solves LPs w, w+W, ...
    one after another in one single-threaded process until its summed solve()
    time (presolve, scaling, iterations, checks, certification; instance
    generation excluded, as in the GPU figure) reaches `budget_s`.  Rate =
    LPs solved / the longest worker's summed solve time."""
    ref = ROOT / "oracle" / "_ref" / "ref_driver"
    if not ref.exists():
        return None
    cmd = [str(ref), "solve-range", "--config", "2", "--seed", str(SEED), "--threads", "1",
           "--stride", str(workers), "--total", str(count), "--budget", str(budget_s)]
    procs = [subprocess.Popen(cmd + ["--first", str(w)], stdout=subprocess.PIPE, text=True)
             for w in range(workers)]
    solved, lps, spent = 0, 0, []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"reference CPU baseline failed ({p.returncode})")
        rows = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
        lps += len(rows)
        solved += sum(r["status"] == "OPTIMAL" for r in rows)
        spent.append(sum(r["seconds"] for r in rows))
    wall = max(spent)
    return {"value": round(solved / wall, 2), "unit": "LPs/s", "cores": workers, "kind": "reference",
            "sample": f"{lps} LPs of the same 4,096-LP batch (seed {SEED}), {workers} single-threaded reference "
                      f"processes, each solving its stride of LPs for ~{budget_s:.0f} s of solve() time "
                      f"(presolve+scaling+iterations+certification); {solved} OPTIMAL, longest worker "
                      f"{wall:.1f} s"}


def serve_stream(engine, qps, duration, seed=SEED, warm_fraction=0.2, hbm_budget=64 << 30):
    """Open-loop serving run: Poisson arrivals at `qps` for `duration` seconds
    of the serving workload (nonzeros log-uniform over [1e3, 5e6], batch
    generator), `warm_fraction` of them warm-started re-solves of an earlier,
    finished request with every row bound scaled by 1 + U(-5%, 5%).
    Instances are generated before the clock starts."""
    from niryukti_amd.serve import Server

    rng = np.random.default_rng(seed)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / qps)
        if t >= duration:
            break
        arrivals.append(t)
    warm_flags = rng.random(len(arrivals)) < warm_fraction
    fresh = int((~warm_flags).sum()) + 1
    log(f"serve: {len(arrivals)} arrivals, generating up to {fresh} instances")
    insts = []
    for i in range(fresh):
        insts.append(engine.Instance(3, seed, 1, i))
        if i % 20 == 19:
            log(f"serve: generated {i + 1}")
    reqs, kinds, next_fresh = [], [], 0
    with Server(hbm_budget=hbm_budget, grid_workers=3) as server:
        t0 = time.perf_counter()
        last_log = t0
        for a, warm in zip(arrivals, warm_flags):
            if time.perf_counter() - last_log > 5:
                last_log = time.perf_counter()
                log(f"serve: t={last_log - t0:.1f}s submitted {len(reqs)}, done {sum(r.done() for r in reqs)}")
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            done = [k for k, (r, kind) in enumerate(zip(reqs, kinds)) if kind[0] == "fresh" and r.done()
                    and not isinstance(r.result, BaseException) and r.result["status"] == "OPTIMAL"]
            if warm and done:
                k = done[int(rng.integers(len(done)))]
                lp = insts[kinds[k][1]].lp
                f = 1 + rng.uniform(-0.05, 0.05, lp.rows)
                lp2 = engine.LP(lp.ptr, lp.index, lp.value, lp.c, lp.lb, lp.ub, lp.rl * f, lp.ru * f)
                reqs.append(server.submit(lp2, reqs[k].result["x"], reqs[k].result["y"]))
                kinds.append(("warm", kinds[k][1]))
            else:
                reqs.append(server.submit(insts[next_fresh].lp))
                kinds.append(("fresh", next_fresh))
                next_fresh += 1
        while not all(r.done() for r in reqs):
            time.sleep(1.0)
            if time.perf_counter() - last_log > 5:
                last_log = time.perf_counter()
                log(f"serve: t={last_log - t0:.1f}s draining, done {sum(r.done() for r in reqs)}/{len(reqs)}")
        results = [r.wait() for r in reqs]
        wall = max(r.finish for r in reqs) - t0
    lat = np.array([r.latency for r in reqs])
    ok = np.array([res["status"] == "OPTIMAL" and res["kkt_error"] <= 1e-6 for res in results])
    planted_ok = sum(1 for res, kind in zip(results, kinds) if kind[0] == "fresh" and res["status"] == "OPTIMAL"
                     and abs(res["objective"] - insts[kind[1]].planted_objective)
                     <= 1e-6 * (1 + abs(insts[kind[1]].planted_objective)))
    statuses, routes = {}, {}
    for res, r in zip(results, reqs):
        statuses[res["status"]] = statuses.get(res["status"], 0) + 1
        routes[r.route] = routes.get(r.route, 0) + 1
    nnz = np.array([r.lp.nnz for r in reqs])
    backlog = sum(1 for r in reqs if r.finish is None or r.finish - t0 > duration)
    in_window = sum(1 for r, res in zip(reqs, results) if r.finish - t0 <= duration and res["status"] == "OPTIMAL"
                    and res["kkt_error"] <= 1e-6)
    return {"mode": "serve", "qps": qps, "duration_s": duration, "requests": len(reqs),
            "arrival_rate_per_s": round(len(reqs) / duration, 3),
            "finished_after_arrivals_end": backlog,
            "warm_resolves": int(sum(k[0] == "warm" for k in kinds)), "routes": routes, "statuses": statuses,
            "certified": int(ok.sum()), "fresh_at_planted_optimum": planted_ok,
            "solves_per_s": round(float(ok.sum()) / wall, 3), "wall_s": round(wall, 2),
            # certified inside the arrival window, per second of it (the sustained rate while
            # requests arrive; solves_per_s above includes the drain after the last arrival)
            "certified_in_window": in_window, "in_window_solves_per_s": round(in_window / duration, 3),
            "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 1),
                           "p99": round(1e3 * float(np.percentile(lat, 99)), 1),
                           "p50_warm": round(1e3 * float(np.percentile(lat[[k[0] == "warm" for k in kinds]], 50)), 1)
                           if any(k[0] == "warm" for k in kinds) else None},
            "nonzeros": {"median": int(np.median(nnz)), "mean": int(nnz.mean()), "max": int(nnz.max())},
            "hbm_budget_bytes": hbm_budget}


def single_lp(engine, config, scale, iterations, tol, seed=SEED, cpu_threads=16, cpu_iterations=None,
              matrix_precision="fp64", verbose=False, index=0):
    """One LP of config 0 / 1 / 4 on the grid-wide single-LP engine for a fixed
    iteration budget, beside the reference's CPU PDHG on the same instance
    (oracle/_ref/ref_driver, `cpu_threads` OpenMP threads, `cpu_iterations`
    iterations, by default the same budget)."""
    inst = engine.Instance(config, seed, scale, index)
    lp = inst.lp
    m, n, nnz = lp.rows, lp.cols, lp.nnz
    log(f"single: config {config} scale {scale} index {index}: {m} x {n}, {nnz} nonzeros")
    r = engine.solve_lp(lp, iteration_limit=iterations, time_limit=1e9, tol=tol, matrix_precision=matrix_precision,
                        verbose=int(verbose))
    attempts = r["iterations"] + r["rejected_steps"]
    s_p = 4 if nnz < 2 ** 31 else 8
    s_v = 4 if matrix_precision in ("mixed", "fp32") else 8
    s_g = 4 if matrix_precision == "fp32" else 8  # the gathered y and dx (FP32 shadows in the FP32 phase)
    # algorithmic bytes of one attempt (s_v-byte values, int32 indices), original sizes
    b_attempt = 2 * nnz * (s_v + 4) + (n + m + 2) * s_p + m * s_g + n * s_g + 64 * (n + m)
    achieved = b_attempt * attempts / max(r["iteration_seconds"], 1e-12) / 1e9
    out = {"mode": "single", "config": config, "scale": scale, "index": index, "rows": m, "cols": n, "nonzeros": nnz,
           "matrix_precision": matrix_precision,
           "tol": tol, "status": r["status"], "iterations": r["iterations"], "attempts": attempts,
           "kkt": r["kkt_error"], "objective": r["objective"], "planted_objective": inst.planted_objective,
           "seconds": round(r["seconds"], 3), "iteration_seconds": round(r["iteration_seconds"], 3),
           "preprocess_seconds": round(r["preprocess_seconds"], 3),
           "verification_seconds": round(r["verification_seconds"], 3),
           "ms_per_attempt": round(1e3 * r["iteration_seconds"] / max(attempts, 1), 4),
           "fp32_iterations": r["fp32_iterations"],
           "algorithmic_bytes_per_attempt": b_attempt,
           "achieved_GBs": round(achieved, 1), "hbm_frac": round(achieved / HBM_PEAK_GBS, 4)}
    ref = ROOT / "oracle" / "_ref" / "ref_driver"
    if cpu_threads and ref.exists():
        cmd = [str(ref), "solve", "--config", str(config), "--seed", str(seed), "--scale", str(scale), "--index", str(index),
               "--iterations", str(cpu_iterations or iterations), "--threads", str(cpu_threads), "--tol", str(tol)]
        c = json.loads(subprocess.run(cmd, capture_output=True, text=True, check=True).stdout)
        out["cpu_baseline"] = {"kind": "reference", "threads": cpu_threads, "status": c["status"],
                               "iterations": c["iterations"], "kkt": c["kkt"],
                               "seconds": round(c["seconds"], 3), "iteration_seconds": round(c["iteration_seconds"], 3),
                               "ms_per_iteration": round(1e3 * c["iteration_seconds"] / max(c["iterations"], 1), 3)}
        out["speedup_per_iteration"] = round(out["cpu_baseline"]["ms_per_iteration"] /
                                             max(1e3 * r["iteration_seconds"] / max(r["iterations"], 1), 1e-9), 1)
        out["speedup_wall"] = round(c["seconds"] / max(r["seconds"], 1e-9), 1)
    return out


def _csv_rows(directory, suffix):
    import csv
    import glob

    rows = []
    for path in sorted(glob.glob(f"{directory}/**/*{suffix}", recursive=True)):
        with open(path, newline="") as f:
            rows += list(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *{suffix} under {directory} (run rocprofv3 with --output-format csv)")
    return rows


def _col(row, *names):
    lower = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in lower:
            return row[lower[n.lower()]]
    raise SystemExit(f"rocprofv3 CSV lacks {names}; columns are {sorted(row)}")


def _short(name):
    return name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()


def summarize_trace(directory):
    """Per (kernel, LDS bytes) dispatch count and mean duration from a
    `rocprofv3 --kernel-trace --stats --output-format csv` directory, plus the
    tool's own kernel_stats table."""
    out = ["# rocprofv3 kernel_stats.csv (durations in ns)"]
    for r in _csv_rows(directory, "kernel_stats.csv")[:12]:
        out.append(f"{_short(_col(r, 'Name'))[:60]:60s} calls={_col(r, 'Calls'):>7s} "
                   f"total_ns={_col(r, 'TotalDurationNs'):>14s} avg_ns={float(_col(r, 'AverageNs')):14.1f} "
                   f"pct={float(_col(r, 'Percentage')):6.2f}")
    groups = {}
    for r in _csv_rows(directory, "kernel_trace.csv"):
        key = (_short(_col(r, "Kernel_Name")), int(_col(r, "LDS_Block_Size", "Lds_Block_Size", "LDS_Size")))
        dur = int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))
        g = groups.setdefault(key, [0, 0])
        g[0] += 1
        g[1] += dur
    out.append("# per kernel and LDS allocation (from kernel_trace.csv)")
    for (name, lds), (n, tot) in sorted(groups.items(), key=lambda kv: -kv[1][1]):
        out.append(json.dumps({"kernel": name, "lds_bytes": lds, "dispatches": n,
                               "mean_ms": round(tot / n / 1e6, 4), "total_s": round(tot / 1e9, 4)}))
    return "\n".join(out)


def _pmc_means(directory, counter):
    per = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        if _col(r, "Counter_Name") != counter:
            continue
        key = (_short(_col(r, "Kernel_Name")), int(_col(r, "LDS_Block_Size", "Lds_Block_Size", "LDS_Size")))
        per.setdefault(key, {}).setdefault(_col(r, "Dispatch_Id"), 0.0)
        per[key][_col(r, "Dispatch_Id")] += float(_col(r, "Counter_Value"))
    return {k: (len(v), sum(v.values()) / len(v)) for k, v in per.items()}


def summarize_counters(directory):
    """Per kernel: dispatches and the mean of every counter of a rocprofv3
    --pmc run per dispatch (one JSON line per kernel, the busiest first)."""
    sums, dispatches = {}, {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        k = _short(_col(r, "Kernel_Name"))
        c = _col(r, "Counter_Name")
        sums.setdefault(k, {}).setdefault(c, 0.0)
        sums[k][c] += float(_col(r, "Counter_Value"))
        dispatches.setdefault(k, set()).add(_col(r, "Dispatch_Id"))
    lines = []
    for k in sorted(sums, key=lambda k: -sums[k].get("SQ_BUSY_CU_CYCLES", 0.0)):
        n = len(dispatches[k])
        lines.append(json.dumps({"kernel": k, "dispatches": n,
                                 **{c: round(v / n) for c, v in sorted(sums[k].items())}}))
    return "\n".join(lines)


def summarize_traffic(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes):
    """HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE runs (KiB per
    dispatch), each scaled by what that counter reads for a plain copy of
    `copy_bytes` bytes (the largest dispatch of the copy-probe run)."""
    cf = max(v[1] for v in _pmc_means(copy_fetch_dir, "FETCH_SIZE").values())
    cw = max(v[1] for v in _pmc_means(copy_write_dir, "WRITE_SIZE").values())
    fetch_scale = copy_bytes / (cf * 1024.0)
    write_scale = copy_bytes / (cw * 1024.0)
    out = [f"# copy probe of {copy_bytes} bytes: FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB",
           f"# scale factors: fetch x{fetch_scale:.4f}, write x{write_scale:.4f}"]
    fetch, write = _pmc_means(fetch_dir, "FETCH_SIZE"), _pmc_means(write_dir, "WRITE_SIZE")
    for key in sorted(fetch, key=lambda k: -fetch[k][0] * fetch[k][1]):
        n, f = fetch[key]
        w = write.get(key, (0, 0.0))[1]
        out.append(json.dumps({"kernel": key[0], "lds_bytes": key[1], "dispatches": n,
                               "fetch_KiB": round(f, 1), "write_KiB": round(w, 1),
                               "hbm_bytes_per_dispatch": round(f * 1024 * fetch_scale + w * 1024 * write_scale)}))
    return "\n".join(out)


TRAFFIC_PROFILE = ROOT / "profiles" / "batch_advance_hbm_traffic_4096lp.txt"


def committed_traffic(kernel):
    """HBM bytes per launch of `kernel` (a kernel symbol, whichever classes
    launched it) from the committed FETCH/WRITE summary, or None."""
    if not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("{"):
            e = json.loads(line)
            if e["kernel"].replace(" ", "") == kernel.replace(" ", ""):
                return e["hbm_bytes_per_dispatch"]
    return None


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--checks", type=int, default=5, help="check periods per launch")
    ap.add_argument("--stream-threshold", type=int, default=2048,
                    help="submit the next pass once fewer than this many LPs of earlier passes run")
    ap.add_argument("--no-stream", action="store_true",
                    help="passes one after another, each on a fresh engine (no overlap between passes)")
    ap.add_argument("--handoff", type=int, default=0,
                    help="iterations after which an LP moves to the grid-wide engine (0: never)")
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--summarize-trace", metavar="DIR", help="summarise a rocprofv3 --kernel-trace --stats CSV run")
    ap.add_argument("--summarize-counters", metavar="DIR", help="summarise a rocprofv3 --pmc CSV run per kernel")
    ap.add_argument("--summarize-traffic", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="summarise FETCH_SIZE / WRITE_SIZE CSV runs of the bench and of --copy-probe")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--serve", action="store_true", help="serving stream instead of the batch workload")
    ap.add_argument("--single", type=int, default=None, metavar="CONFIG",
                    help="one LP of config 0, 1 or 4 on the single-LP engine beside the reference CPU")
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--index", type=int, default=0, help="single-LP mode: instance index (configs 2 and 3)")
    ap.add_argument("--iterations", type=int, default=1000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-iterations", type=int, default=None)
    ap.add_argument("--verbose", action="store_true", help="single-LP mode: one progress line per check period")
    ap.add_argument("--matrix-precision", default="fp64", choices=("fp64", "mixed", "fp32"),
                    help="single-LP mode: FP64; mixed = FP32 matrix values (FP64 accumulation and iterates); "
                         "fp32 = FP32 SpMV arithmetic over FP64 iterates, then FP64 refinement")
    ap.add_argument("--batch-precision", default="fp64", choices=("fp64", "mixed"),
                    help="batched engine: FP64 matrix values, or FP32 values (FP64 arithmetic and iterates) "
                         "until an LP stalls, then FP64")
    ap.add_argument("--qps", type=float, default=1.0, help="serving stream: Poisson arrival rate")
    ap.add_argument("--duration", type=float, default=30.0, help="serving stream: arrival window in seconds")
    ap.add_argument("--no-serve", action="store_true", help="leave the serving stream out of the default run")
    ap.add_argument("--no-config1", action="store_true",
                    help="leave the capped config-1 single-LP comparison out of the default run")
    ap.add_argument("--config1-iterations", type=int, default=3000,
                    help="iteration cap of the config-1 single-LP comparison (GPU and CPU alike)")
    ap.add_argument("--no-solve-phase", action="store_true",
                    help="leave the solve-phase step (LPs prepared before the clock) out of the default run")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace))
        return
    if args.summarize_traffic:
        print(summarize_traffic(*args.summarize_traffic, args.copy_bytes))
        return
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters))
        return
    if args.copy_probe:
        import torch

        src = torch.ones(args.copy_probe // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    import niryukti_amd as engine

    if engine.device_count() < 1:
        raise SystemExit("no HIP device visible")
    log("device:", engine.device_name(0))
    if args.serve:
        print(json.dumps(serve_stream(engine, args.qps, args.duration)), flush=True)
        return
    if args.single is not None:
        print(json.dumps(single_lp(engine, args.single, args.scale, args.iterations, args.tol,
                                   cpu_threads=args.cpu_threads, cpu_iterations=args.cpu_iterations,
                                   matrix_precision=args.matrix_precision, verbose=args.verbose,
                                   index=args.index)), flush=True)
        return
    insts = make_instances(engine, args.count)
    nnz = sum(i.lp.nnz for i in insts)
    log(f"instances: {len(insts)} LPs, {nnz} nonzeros")
    total_t, total_solved, lat, stats_all, last = 0.0, 0, [], [], None
    if args.no_stream:
        for w in range(args.warmup):
            el, solved, _, _, _ = batch_step(engine, insts, args.checks, args.handoff, precision=args.batch_precision)
            log(f"warmup {w}: {el:.3f} s, {solved} solved")
        for s in range(args.steps):
            el, solved, results, stats, latencies = batch_step(engine, insts, args.checks, args.handoff,
                                                               precision=args.batch_precision)
            log(f"step {s}: {el:.3f} s, {solved}/{len(insts)} solved, stats {stats}")
            total_t += el
            total_solved += solved
            lat += latencies
            stats_all.append(stats)
            last = results
    else:
        # one engine for every pass, warm-up passes included (its launch-shape
        # tuner settles there); the timed passes start once they have drained
        b = engine.Batch(arena_bytes=96 << 30, max_slots=2 * len(insts), time_limit=1e9,
                         matrix_precision=args.batch_precision)
        b.set_handoff(args.handoff, 4)
        if args.warmup:
            el, passes = batch_stream(b, insts, args.warmup, args.checks, args.stream_threshold)
            log(f"warmup: {args.warmup} passes streamed in {el:.3f} s, "
                f"{sum(score(insts, r)[0] for r in passes)} solved")
        before = b.stats()
        total_t, passes = batch_stream(b, insts, args.steps, args.checks, args.stream_threshold)
        after = b.stats()
        b.close()
        for k, results in enumerate(passes):
            solved, latencies = score(insts, results)
            total_solved += solved
            lat += latencies
            log(f"pass {k}: {solved}/{len(insts)} solved")
        log(f"{args.steps} passes streamed in {total_t:.3f} s, stats {after}")
        d = {k: after[k] - before[k] for k in ("device_seconds", "check_seconds", "advance_seconds", "launches",
                                                "certify_wait_seconds", "settle_seconds", "admit_seconds")}
        log(f"timed region host work in step: waiting for certification {d['certify_wait_seconds']:.3f} s, "
            f"settling {d['settle_seconds']:.3f} s, admitting {d['admit_seconds']:.3f} s")
        log(f"timed region: {d['launches']} rounds, host waiting on the device {d['device_seconds']:.3f} s of "
            f"{total_t:.3f} s; summed event time: advance {d['advance_seconds']:.3f} s, checks {d['check_seconds']:.3f} s; "
            "per class " + ", ".join(f"{a - c:.3f}" for a, c in zip(after["class_seconds"], before["class_seconds"])))
        stats_all.append({key: [a - c for a, c in zip(after[key], before[key])]
                          for key in ("class_bytes", "class_seconds", "class_launches",
                                      "shape_bytes", "shape_seconds", "shape_launches")})
        stats_all[0].update({key: after[key] for key in ("class_threads", "class_rows", "class_tuned", "handoffs")})
        last = passes[-1]
    value = total_solved / total_t
    # Dominant kernel: the launch class (one step-kernel instantiation) that
    # moved the most algorithmic bytes over the timed region, whichever
    # classes launched it (a kernel trace reports per symbol, too); its rate
    # = bytes / its HIP-event time.
    def total(key, k):
        return sum(st[key][k] for st in stats_all)

    last_stats = stats_all[-1]
    # the crowded step kernel's value type: FP64, or FP32 values in mixed precision
    vt = "float" if args.batch_precision == "mixed" else "double"
    kernels = [(f"batch_advance_lds_kernel<{t}, {r}, {vt}>", total("shape_bytes", k), total("shape_seconds", k),
                total("shape_launches", k)) for k, (t, r) in enumerate(LDS_SHAPES)]
    kernels += [("batch_advance_kernel", total("class_bytes", 3), total("class_seconds", 3), total("class_launches", 3)),
                ("batch_tail_kernel<1024>", total("class_bytes", 4), total("class_seconds", 4),
                 total("class_launches", 4))]
    symbol, adv_b, adv_s, launches = max(kernels, key=lambda k: k[1])
    device_b = sum(total("class_bytes", k) for k in range(len(stats_all[0]["class_bytes"])))
    achieved = adv_b / adv_s / 1e9 if adv_s > 0 else 0.0
    users = [name for c, name in enumerate(("m+n <= 2048", "m+n <= 4096", "m+n <= 8000"))
             if (last_stats["class_threads"][c], last_stats["class_rows"][c]) in
             [(t, r) for (t, r) in LDS_SHAPES if symbol == f"batch_advance_lds_kernel<{t}, {r}, {vt}>"]]
    if symbol.startswith("batch_advance_lds_kernel"):
        kernel_name = f"{symbol} (crowded LDS classes; tuned choice of class {', '.join(users) or 'none'})"
        if vt == "float":
            # the shape's FP64-value launches (LPs that switched) are counted with it
            kernel_name += " with the same shape's FP64-value launches"
    elif symbol == "batch_advance_kernel":
        kernel_name = symbol + " (global class)"
    else:
        kernel_name = symbol + " (tail group: attempts and checks in one launch)"
    traffic = committed_traffic(symbol)
    statuses = {}
    for r in last:
        statuses[r["status"]] = statuses.get(r["status"], 0) + 1
    its = np.array([r["iterations"] for r in last])
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "LPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total_t / args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.batch_precision == "fp64" else "f64 (f32 matrix values until 1e-5)",
        "data": "synthetic",
        "config": {"workload": f"batch of {args.count} independent LPs: m log-uniform [50,2000], n uniform "
                               f"[2m,3m], 3-8 nnz/col, N(0,1), box [0,U], U~U(1,50), planted optimum, seed {SEED}",
                   "lps": args.count, "nonzeros": int(nnz), "tol": 1e-6, "checks_per_launch": args.checks,
                   "handoff_iterations": args.handoff,
                   "passes": "one engine, pass k+1 submitted once fewer than "
                             f"{args.stream_threshold} LPs of earlier passes run" if not args.no_stream
                             else "one engine per pass, passes one after another",
                   "engine": "batched (one workgroup per LP), " +
                             ("FP64" if args.batch_precision == "fp64" else "FP32 matrix values, FP64 arithmetic")},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 1) if lat else None,
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 1) if lat else None},
        "statuses": statuses,
        "unsolved_indices": [k for k, r in enumerate(last) if r["status"] != "OPTIMAL"][:20],
        "unsolved_messages": sorted({r["message"][:120] for r in last if r["status"] != "OPTIMAL" and r["message"]})[:5],
        "iterations": {"mean": float(its.mean()), "max": int(its.max())},
        "handoffs": int(stats_all[-1].get("handoffs", 0)),
        "launch_shapes": {f"class{c}": {"threads": last_stats["class_threads"][c], "rows": last_stats["class_rows"][c],
                                        "tuned": bool(last_stats["class_tuned"][c])} for c in range(3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": TRAFFIC_PROFILE.name if traffic is not None else None,
                     "kernel": kernel_name,
                     "algorithmic_bytes_per_launch": round(adv_b / max(1, launches), 1),
                     "avg_launch_ms": round(1e3 * adv_s / max(1, launches), 3),
                     "launches": launches,
                     # the whole device: every step launch's algorithmic bytes over the timed
                     # region (the classes' launches run side by side and share HBM)
                     "device": {"achieved": round(device_b / total_t / 1e9, 1),
                                "frac": round(device_b / total_t / 1e9 / HBM_PEAK_GBS, 4),
                                "algorithmic_bytes": device_b, "seconds": round(total_t, 3)}},
    }
    if not args.no_stream and not args.no_solve_phase:
        # one pass alone on a fresh engine, first submit to last result
        el, solved, _, _, latencies = batch_step(engine, insts, args.checks, args.handoff, precision=args.batch_precision)
        line["one_pass_alone"] = {"lps_per_s": round(solved / el, 2), "seconds": round(el, 3), "solved": solved,
                                  "p50_ms": round(1e3 * float(np.percentile(latencies, 50)), 1),
                                  "p99_ms": round(1e3 * float(np.percentile(latencies, 99)), 1)}
        log(f"one pass alone: {el:.3f} s, {solved} solved")
    if not args.no_solve_phase:
        # the solve phase alone: every LP prepared and uploaded before the clock starts
        el, solved, _, _, _ = batch_step(engine, insts, args.checks, args.handoff, prepared=True, precision=args.batch_precision)
        line["solve_phase"] = {"lps_per_s": round(solved / el, 2), "seconds": round(el, 3), "solved": solved,
                               "note": "host prepare and upload before the timed region (not the headline)"}
        log(f"solve phase: {el:.3f} s, {solved} solved")
    if not args.no_baselines:
        line["plain_gpu"] = plain_sample(engine, insts)
        line["served_vs_plain"] = round(value / max(1e-9, line["plain_gpu"]["lps_per_s"]), 2)
        host = machine()
        cores = physical_cores(host)
        log(f"host: {host}; CPU baseline on 1 and {cores} cores")
        one = cpu_baseline(args.count, 1, 10.0)
        line["cpu_baseline"] = cpu_baseline(args.count, cores, 20.0)
        if line["cpu_baseline"] is not None:
            line["cpu_baseline"]["host"] = host
            line["cpu_baseline"]["one_thread"] = one
            line["vs_cpu_all_cores"] = round(value / max(1e-9, line["cpu_baseline"]["value"]), 2)
    if not args.no_serve:
        # the serving stream (Poisson arrivals of LPs of 1e3..5e6 nonzeros, 20 %
        # warm re-solves): an extra key beside the batch headline
        line["serving"] = serve_stream(engine, args.qps, args.duration)
        log(f"serving: {line['serving']}")
    if not args.no_baselines and not args.no_config1:
        # the single large LP (1,000,000 x 2,000,000, power-law rows) on the
        # grid-wide single-LP engine beside the reference's CPU PDHG on all
        # cores, both capped at the same iteration count: neither converges
        # within the 100,000-iteration default, so the record is "not
        # converged at the cap" with the KKT each reached, and the time per
        # iteration (an extra key, not the headline)
        c1 = single_lp(engine, 1, 1, args.config1_iterations, 1e-6, cpu_threads=cores)
        c1["note"] = (f"capped at {args.config1_iterations} iterations: GPU {c1['status']} at kkt {c1['kkt']:.3g}, "
                      f"CPU {c1.get('cpu_baseline', {}).get('status')} at kkt "
                      f"{c1.get('cpu_baseline', {}).get('kkt', float('nan')):.3g}")
        line["config1_single_lp"] = c1
        log(f"config 1: {c1}")
        # the same LP and cap in the FP32-bulk form (FP32 values, gathered
        # vectors and lane sums over FP64 iterates; FP64 checks, and FP64
        # refinement once a candidate reaches the switch point, which a capped
        # run of this LP does not), timed against the same CPU run
        c1f = single_lp(engine, 1, 1, args.config1_iterations, 1e-6, cpu_threads=0, matrix_precision="fp32")
        cpu = c1.get("cpu_baseline")
        if cpu:
            c1f["cpu_baseline"] = dict(cpu, note="the run of config1_single_lp")
            c1f["speedup_per_iteration"] = round(cpu["ms_per_iteration"] /
                                                 max(1e3 * c1f["iteration_seconds"] / max(c1f["iterations"], 1), 1e-9), 1)
            c1f["speedup_wall"] = round(cpu["seconds"] / max(c1f["seconds"], 1e-9), 1)
        c1f["note"] = (f"capped at {args.config1_iterations} iterations: GPU {c1f['status']} at kkt {c1f['kkt']:.3g}, "
                       f"{c1f['fp32_iterations']} iterations in the FP32 phase")
        line["config1_single_lp_fp32"] = c1f
        log(f"config 1, FP32 bulk: {c1f}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
