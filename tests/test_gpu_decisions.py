"""The GPU driver's decisions against the oracle's (src/solver.cpp:299-333, 511-726).

An objective check alone would still pass with a wrong initial step, restart
rule or primal-weight update, so these tests compare the decisions
themselves, through the C-ABI, on seeded config-0 and config-2 instances:

* the first step and primal weight (host sums over the prepared model, which
  is bit-identical to the oracle's): exactly equal;
* the iteration of the first restart: exactly equal (the restart rule compares
  rel-KKT figures at multiples of check_every, far from any tie);
* the primal weight right after the first restart: within 1e-12 relative when
  that restart comes after 4 steps (check_every 2), where the device's
  double-double displacement sums and the host's long-double sums see
  iterates that differ only by the rounding of 4 steps; within 1e-5 after
  the default 200 steps, by which point the trajectories have drifted apart
  (device and host sum each step's energy and coupling in different orders,
  and a borderline step acceptance can flip);
* total iterations, restarts and weight updates: within a band, since that
  drift can move the stopping check by one or a few check periods.

Band: |iterations - oracle| <= max(300, 20 % of the oracle's), restarts and
weight updates within 2 + 20 %.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIG0 = [(0, s, sc, 0) for s, sc in ((3, 20), (7, 20), (11, 10), (13, 10), (17, 8), (19, 8), (23, 5), (29, 4),
                                        (31, 20), (37, 10))]
CONFIG2 = [(2, 42, 1, i) for i in (0, 1, 2, 3, 5, 7, 11, 13, 17, 19)] + [(2, 3, 1, i) for i in (0, 1, 2)]


def within_band(gpu, cpu):
    it_ok = abs(gpu["iterations"] - cpu["iterations"]) <= max(300, 0.2 * cpu["iterations"])
    rs_ok = abs(gpu["restarts"] - cpu["restarts"]) <= 2 + 0.2 * cpu["restarts"]
    wu_ok = abs(gpu["weight_updates"] - cpu["weight_updates"]) <= 2 + 0.2 * cpu["weight_updates"]
    return it_ok and rs_ok and wu_ok


def check_decisions(name, gpu, cpu, weight_tol=1e-5):
    print(name, "gpu", gpu["status"], gpu["iterations"], gpu["restarts"], gpu["weight_updates"],
          gpu["first_restart_iteration"], gpu["first_restart_weight"], "| cpu", cpu["status"], cpu["iterations"],
          cpu["restarts"], cpu["weight_updates"], cpu["first_restart_iteration"], cpu["first_restart_weight"],
          flush=True)
    assert gpu["status"] == cpu["status"] == "OPTIMAL"
    assert gpu["initial_step"] == cpu["initial_step"]
    assert gpu["initial_weight"] == cpu["initial_weight"]
    assert gpu["first_restart_iteration"] == cpu["first_restart_iteration"]
    if cpu["first_restart_iteration"] > 0:
        err = abs(gpu["first_restart_weight"] - cpu["first_restart_weight"])
        assert err <= weight_tol * abs(cpu["first_restart_weight"])
    assert within_band(gpu, cpu), name


@pytest.mark.parametrize("case", CONFIG0 + CONFIG2, ids=lambda c: f"c{c[0]}s{c[1]}x{c[2]}i{c[3]}")
def test_single_lp_driver_decisions_match_oracle(engine, oracle, case):
    inst = engine.Instance(*case)
    cpu = oracle.solve(inst.lp, threads=1)
    gpu = engine.solve_lp(inst.lp)
    check_decisions(case, gpu, cpu)


@pytest.mark.parametrize("case", [(0, 3, 20, 0), (0, 7, 10, 0), (2, 42, 1, 0), (2, 42, 1, 3), (2, 3, 1, 1)], ids=str)
def test_early_restart_weight_matches_oracle(engine, oracle, case):
    """check_every 2: the first restart can come after 4 steps, before the
    device and host trajectories have drifted, so the restart rule and the
    displacement weight update (src/solver.cpp:692-716) are compared almost
    exactly."""
    inst = engine.Instance(*case)
    cpu = oracle.solve(inst.lp, threads=1, check_every=2, iteration_limit=400)
    gpu = engine.solve_lp(inst.lp, check_every=2, iteration_limit=400)
    print(case, gpu["first_restart_iteration"], gpu["first_restart_weight"], cpu["first_restart_iteration"],
          cpu["first_restart_weight"], flush=True)
    assert 0 < cpu["first_restart_iteration"] <= 20
    assert gpu["initial_step"] == cpu["initial_step"] and gpu["initial_weight"] == cpu["initial_weight"]
    assert gpu["first_restart_iteration"] == cpu["first_restart_iteration"]
    assert abs(gpu["first_restart_weight"] - cpu["first_restart_weight"]) <= 1e-12 * cpu["first_restart_weight"]
    assert gpu["first_restart_weight"] != gpu["initial_weight"]


def test_batched_driver_decisions_match_oracle(engine, oracle):
    """The batched engine runs the same driver per workgroup on the device."""
    insts = [engine.Instance(*c) for c in CONFIG2]
    b = engine.Batch(arena_bytes=2 << 30, max_slots=64)
    tickets = [b.submit(i.lp) for i in insts]
    b.run(5)
    results = [b.result(t) for t in tickets]
    b.close()
    for case, inst, gpu in zip(CONFIG2, insts, results):
        check_decisions(("batch",) + case, gpu, oracle.solve(inst.lp, threads=1))


@pytest.mark.parametrize("case", [(0, 3, 20, 0), (2, 3, 1, 1), (2, 42, 1, 5)], ids=str)
def test_pid_weight_matches_oracle(engine, oracle, case):
    """--primal-weight pid (src/research.cpp:52-70) with adaptive steps: the
    controller starts from weight 1 (src/solver.cpp:323) and its updates
    drive the same restarts as the oracle's."""
    inst = engine.Instance(*case)
    cpu = oracle.solve(inst.lp, threads=1, primal_weight="pid")
    gpu = engine.solve_lp(inst.lp, primal_weight="pid")
    check_decisions(("pid",) + case, gpu, cpu)
    assert gpu["initial_weight"] == 1.0
    rel = abs(gpu["objective"] - inst.planted_objective) / (1 + abs(inst.planted_objective))
    assert rel <= 1e-6
    assert oracle.verify(inst.lp, gpu["x"], gpu["y"])["kkt"] <= 1e-6


@pytest.mark.parametrize("method", ["halpern", "rhpdhg", "r2hpdhg"])
@pytest.mark.parametrize("weight", ["displacement", "pid"])
def test_anchored_methods_match_oracle(engine, oracle, method, weight):
    """Halpern, reflected Halpern and the residual-restart variants
    (src/cpu.cpp:174-202) on a 100 x 250 planted LP, fixed steps: the same
    decisions as the oracle (whose solves equal the reference build's bit for
    bit on this instance) and the planted optimum to 1e-6."""
    inst = engine.Instance(0, 3, 20, 0)
    opts = dict(method=method, adaptive=False, primal_weight=weight, iteration_limit=30000)
    cpu = oracle.solve(inst.lp, threads=1, **opts)
    gpu = engine.solve_lp(inst.lp, **opts)
    check_decisions((method, weight), gpu, cpu)
    rel = abs(gpu["objective"] - inst.planted_objective) / (1 + abs(inst.planted_objective))
    assert rel <= 1e-6
    assert oracle.verify(inst.lp, gpu["x"], gpu["y"])["kkt"] <= 1e-6


def test_device_checks_take_the_host_checks_decisions(engine):
    """Candidate checks on the device (npdhg_evaluate) and on the host make
    the same decisions: same iterations, restarts, weight updates."""
    for case in ((0, 7, 20, 0), (2, 42, 1, 3)):
        inst = engine.Instance(*case)
        host = engine.solve_lp(inst.lp, device_checks=0)
        dev = engine.solve_lp(inst.lp, device_checks=1)
        print(case, host["iterations"], dev["iterations"], host["restarts"], dev["restarts"], flush=True)
        assert host["status"] == dev["status"] == "OPTIMAL"
        assert host["first_restart_iteration"] == dev["first_restart_iteration"]
        assert within_band(dev, host)
