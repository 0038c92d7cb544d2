"""Checkpoint / resume of a continuous solve (src/solver.cpp:384-509) on the GPU.

A solve stopped at an iteration limit leaves its last checkpoint (written at
the top of a chunk every `checkpoint_nodes` iterations); resuming from it
must end exactly where the uninterrupted solve ends: same iterate, same
counters, bit for bit (the backend snapshot holds the iterates, averages,
activity, anchors, step factor and counters, and the file the driver's loop
state).
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("host checks", (0, 7, 4, 0), {}),
    ("device checks", (2, 42, 1, 3), {"device_checks": 1}),
    ("pid weight", (2, 3, 1, 1), {"primal_weight": "pid"}),
    ("reflected halpern", (0, 3, 20, 0), {"method": "r2hpdhg", "adaptive": False}),
]
SAME = ("status", "iterations", "restarts", "rejected_steps", "weight_updates", "objective", "kkt_error",
        "primal_residual", "dual_residual", "relative_gap", "first_restart_iteration", "first_restart_weight",
        "host_candidate_checks", "device_candidate_checks")


@pytest.mark.parametrize("name,case,opts", CASES, ids=[c[0] for c in CASES])
def test_resumed_solve_equals_uninterrupted(engine, tmp_path, name, case, opts):
    lp = engine.Instance(*case).lp
    whole = engine.solve_lp(lp, **opts)
    path = tmp_path / "run.ckpt"
    cut = max(300, whole["iterations"] - 350)
    first = engine.solve_lp(lp, checkpoint_path=str(path), checkpoint_nodes=200, iteration_limit=cut, **opts)
    assert first["status"] == "ITERATION_LIMIT" and path.exists()
    saved = json.loads(path.read_text())
    assert saved["schema"] == "vantage-continuous-1" and 0 < saved["iterations"] < cut
    resumed = engine.solve_lp(lp, resume_path=str(path), **opts)
    print(name, whole["iterations"], cut, saved["iterations"], resumed["iterations"], flush=True)
    for key in SAME:
        assert resumed[key] == whole[key], key
    assert np.array_equal(resumed["x"], whole["x"]) and np.array_equal(resumed["y"], whole["y"])


def test_resume_rejects_damaged_state(engine, tmp_path):
    lp = engine.Instance(0, 7, 4, 0).lp
    path = tmp_path / "run.ckpt"
    engine.solve_lp(lp, checkpoint_path=str(path), checkpoint_nodes=100, iteration_limit=200)
    good = json.loads(path.read_text())

    def resume_with(change):
        doc = json.loads(json.dumps(good))
        change(doc)
        bad = tmp_path / "bad.ckpt"
        bad.write_text(json.dumps(doc))
        with pytest.raises(RuntimeError) as err:
            engine.solve_lp(lp, resume_path=str(bad))
        return str(err.value)

    assert "dimensions" in resume_with(lambda d: d["best_x"].pop())
    assert "dimensions" in resume_with(lambda d: d["backend_state"][0].append(1.0))
    assert "mismatch" in resume_with(lambda d: d["configuration"].update(device_checks=True))
    assert "counters" in resume_with(lambda d: d["backend_state"][8].__setitem__(0, -1.0))
    assert "controller" in resume_with(lambda d: d["control"].__setitem__(0, 0.0))
    # the undamaged file still resumes
    assert engine.solve_lp(lp, resume_path=str(path))["status"] == "OPTIMAL"
