"""Checkpoint files that do not belong to the model or the settings are
refused before any device work (CPU): the schema, the model fingerprint and
the solver configuration must match (src/solver.cpp:430-437)."""
import json

import pytest


def linear_model():
    # native JSON model (the reader of src/io.cpp:84-140): min 3x, 2 <= x <= 10
    return {"name": "linear", "sense": "min", "variables": [{"name": "x", "ub": 10}],
            "objective": {"linear": [3]}, "constraints": [{"name": "demand", "lb": 2, "coefficients": {"x": 1}}]}


def header(engine, model):
    return {"schema": "vantage-continuous-1", "fingerprint": model.fingerprint,
            "configuration": {"backend": "hip", "method": "pdhg", "tol": 1e-6, "presolve": True, "scaling": "ruiz",
                              "scaling_passes": 5, "adaptive": True, "restart": True,
                              "primal_weight": "displacement", "power_iterations": 0, "polishing": False,
                              "check_every": 100, "graphs": True, "precision": "fp64", "indices": "auto",
                              "device_checks": False}}


def run(engine, *args):
    import subprocess

    return subprocess.run([str(engine.CLI_PATH), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_mismatched_checkpoints_are_refused(engine, tmp_path):
    mp = tmp_path / "m.json"
    mp.write_text(json.dumps(linear_model()))
    model = engine.Model(mp)
    out = run(engine, "solve", mp, "--resume", tmp_path / "missing.ckpt")
    assert out.returncode == 1 and "Cannot open continuous checkpoint" in out.stderr + out.stdout
    for change in (lambda d: d.update(schema="vantage-continuous-0"), lambda d: d.update(fingerprint="0" * 16),
                   lambda d: d["configuration"].update(tol=1e-8), lambda d: d["configuration"].pop("method"),
                   lambda d: d.pop("configuration")):
        doc = header(engine, model)
        change(doc)
        path = tmp_path / "c.ckpt"
        path.write_text(json.dumps(doc))
        out = run(engine, "solve", mp, "--resume", path)
        assert out.returncode == 1, out.stdout + out.stderr
        assert "mismatch" in out.stderr + out.stdout
    path.write_text("{not json")
    assert run(engine, "solve", mp, "--resume", path).returncode == 1
