"""Drop-in boundary on the CPU: model readers, result/solution files, the
session C-ABI and the CLI, checked against the reference's own outputs
(fingerprints and result documents recorded from the reference build by
tests/golden/make_netlib_golden.py) and the reference's boundary tests
(tests/test_cli.py:71-110, tests/test_extensions.py:20-38).
"""
import ctypes
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
NETLIB = ROOT / "tests" / "golden" / "netlib"
NAMES = ("afiro", "adlittle", "israel", "e226")


def run_cli(engine, *args):
    return subprocess.run([str(engine.CLI_PATH), *map(str, args)], capture_output=True, text=True)


def linear_model():
    # tests/test_extensions.py:17: x <= 10, x >= 2 (demand), min 3x
    return {"name": "linear", "sense": "min",
            "variables": [{"name": "x", "ub": 10}],
            "objective": {"linear": [3]},
            "constraints": [{"name": "demand", "lb": 2, "coefficients": {"x": 1}}]}


def test_library_exports_every_declared_symbol(engine):
    lib = ctypes.CDLL(str(engine.LIBRARY_PATH))
    declared = set()
    for header in (ROOT / "include").rglob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        declared |= set(re.findall(r"\b((?:npdhg|niryukti|vantage)_[a-z_0-9]+)\s*\(", text))
    assert len(declared) > 30
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.parametrize("name", NAMES)
def test_mps_fingerprint_matches_reference(engine, name):
    ref = json.loads((NETLIB / f"{name}.reference_result.json").read_text())
    out = run_cli(engine, "inspect", NETLIB / f"{name}.mps")
    assert out.returncode == 0, out.stderr
    mine = json.loads(out.stdout)
    for key in ("fingerprint", "rows", "columns", "nonzeros", "name"):
        assert mine[key] == ref["reference_inspect"][key], key


@pytest.mark.parametrize("name", ("afiro", "adlittle", "israel"))
def test_verify_accepts_reference_result_and_rejects_tampering(engine, name, tmp_path):
    ref = json.loads((NETLIB / f"{name}.reference_result.json").read_text())
    assert ref["status"] == "OPTIMAL"
    good = tmp_path / "good.json"
    good.write_text(json.dumps(ref))
    out = run_cli(engine, "verify", NETLIB / f"{name}.mps", good)
    assert out.returncode == 0, out.stdout + out.stderr
    assert json.loads(out.stdout)["status"] == "VERIFIED_OPTIMAL"
    bad = dict(ref)
    bad["objective"] = ref["objective"] + 1.0 + abs(ref["objective"]) * 1e-3
    tampered = tmp_path / "bad.json"
    tampered.write_text(json.dumps(bad))
    out = run_cli(engine, "verify", NETLIB / f"{name}.mps", tampered)
    assert out.returncode == 2
    assert json.loads(out.stdout)["status"] == "VERIFICATION_FAILED"
    # a dual perturbed away from optimality is no longer optimal
    worse = dict(ref)
    worse["dual"] = [v + 1.0 for v in ref["dual"]]
    worse_path = tmp_path / "worse.json"
    worse_path.write_text(json.dumps(worse))
    out = run_cli(engine, "verify", NETLIB / f"{name}.mps", worse_path)
    assert out.returncode == 2


def test_verifier_matches_reference_accuracy(engine, oracle):
    """The engine's host verifier (via CLI verify) and the oracle reproduce the
    accuracy numbers the reference reported for its own solution."""
    ref = json.loads((NETLIB / "afiro.reference_result.json").read_text())
    out = json.loads(run_cli(engine, "verify", NETLIB / "afiro.mps", NETLIB / "afiro.reference_result.json").stdout)
    for key in ("primal_residual", "dual_residual", "relative_gap", "kkt_error"):
        assert out["accuracy"][key] == pytest.approx(ref["accuracy"][key], rel=1e-9, abs=1e-15)


def test_result_document_has_reference_keys(engine):
    ref = json.loads((NETLIB / "afiro.reference_result.json").read_text())
    out = json.loads(run_cli(engine, "solve", NETLIB / "afiro.mps").stdout)
    for section in ("model", "accuracy", "performance", "hardware", "presolve", "selection", "gpu_execution"):
        missing = set(ref[section]) - set(out[section]) - {"end_to_end_seconds", "parse_seconds"}
        assert not missing, (section, missing)
    for key in ("solver", "version", "status", "message", "problem_type", "objective", "primal", "dual"):
        assert key in out


def test_json_model_reader_matches_reference_fingerprint(engine, tmp_path):
    path = tmp_path / "m.json"
    path.write_text(json.dumps(linear_model()))
    out = json.loads(run_cli(engine, "inspect", path).stdout)
    ref_cli = ROOT / "oracle" / "_ref" / "niryukti_ref"
    if ref_cli.exists():
        ref = json.loads(subprocess.run([str(ref_cli), "inspect", str(path)], capture_output=True, text=True).stdout)
        assert out["fingerprint"] == ref["fingerprint"]
    assert out["rows"] == 1 and out["columns"] == 1


def test_cli_errors_and_exit_codes(engine, tmp_path):
    assert run_cli(engine, "inspect", tmp_path / "missing.mps").returncode == 1
    assert run_cli(engine, "solve", NETLIB / "afiro.mps", "--bogus").returncode == 1
    assert run_cli(engine, "solve", NETLIB / "afiro.mps", "--device", "quantum").returncode == 1
    out = run_cli(engine, "solve", NETLIB / "afiro.mps", "--device", "cpu")
    assert out.returncode == 2 and json.loads(out.stdout)["status"] == "UNSUPPORTED"
    assert run_cli(engine, "devices").returncode == 0
    qp = tmp_path / "q.json"
    m = linear_model()
    m["objective"]["quadratic_diagonal"] = [2.0]
    qp.write_text(json.dumps(m))
    assert "UNSUPPORTED" in run_cli(engine, "inspect", qp).stderr


def test_session_protocol(engine):
    with engine.NativeSession(linear_model()) as s:
        with pytest.raises(ValueError):
            s._request("solve", unknown_option=True)
        with pytest.raises(ValueError):
            s.update_variable_bounds({"x": {"lb": 11}})  # lb > ub is rejected, model unchanged
        with pytest.raises(ValueError):
            s.update_row_bounds({"nope": {"lb": 1}})
        up = s.update_row_bounds({"demand": {"lb": 4}})
        assert up["status"] == "UPDATED"
        r = s.solve(device="cpu")
        assert r["status"] == "UNSUPPORTED"
        assert r["session"]["model_retained"] is True
    with pytest.raises(RuntimeError):
        s.solve()
    with pytest.raises(ValueError):
        engine.NativeSession("/missing/model.json")
    lib = engine.load_library()
    lib.vantage_request.restype = ctypes.c_char_p
    assert lib.vantage_request(None, b"{}") is None
    with engine.NativeSession(linear_model()) as s:
        assert s._lib.vantage_request(s._handle, b"not json") is None
        assert lib.niryukti_abi_version() == 1


def test_serving_request_validation(engine):
    """`/v1/solve` body checks of the reference (python/niryukti/service.py:54-69)."""
    from niryukti_amd.serve import parse_request

    model, extra = parse_request({"model": linear_model(), "options": {"tol": 1e-7, "threads": 4}})
    assert model.lp.rows == 1 and model.lp.cols == 1 and extra == {"tol": 1e-7}
    assert parse_request({"model": linear_model()})[1] == {}
    assert parse_request({"model": linear_model(), "options": {"method": "halpern"}})[1] == {"method": "halpern"}
    for bad in ({"model": linear_model(), "extra": 1}, {"model": "x"}, {"options": {}},
                {"model": linear_model(), "options": {"gpu": 1}},
                {"model": linear_model(), "options": {"time_limit": 0}},
                {"model": linear_model(), "options": {"time_limit": 7200}},
                {"model": linear_model(), "options": {"threads": 65}},
                {"model": {"variables": "bad"}}, [1, 2]):
        with pytest.raises((ValueError, TypeError)):
            parse_request(bad)


def test_model_handle_and_result_document(engine):
    m = engine.Model(NETLIB / "afiro.mps")
    ref = json.loads((NETLIB / "afiro.reference_result.json").read_text())
    assert m.fingerprint == ref["model"]["fingerprint"]
    doc = m.result_json({"status": "OPTIMAL", "message": "", "objective": ref["objective"], "finite": 1,
                         "kkt_error": 0.0, "x": ref["primal"], "y": ref["dual"], "iterations": 7})
    assert doc["model"]["fingerprint"] == ref["model"]["fingerprint"]
    assert doc["objective"] == pytest.approx(ref["objective"], rel=1e-12)
    assert doc["primal"] == ref["primal"] and doc["performance"]["iterations"] == 7
    with pytest.raises(ValueError):
        engine.Model(json_model={"variables": [{"name": "x", "type": "integer"}]})


REF_CLI = ROOT / "oracle" / "_ref" / "niryukti_ref"


@pytest.mark.parametrize("name", NAMES)
def test_convert_round_trips_to_reference_fingerprint(engine, tmp_path, name):
    """`convert` (app/main.cpp:423-428, src/io.cpp:815-901): the written MPS and
    JSON read back to the model the reference fingerprinted."""
    ref = json.loads((NETLIB / f"{name}.reference_result.json").read_text())["reference_inspect"]
    for ext in ("mps", "json"):
        out = tmp_path / f"{name}.{ext}"
        assert run_cli(engine, "convert", NETLIB / f"{name}.mps", out).returncode == 0
        back = json.loads(run_cli(engine, "inspect", out).stdout)
        assert back["fingerprint"] == ref["fingerprint"], ext
        if REF_CLI.exists():
            theirs = tmp_path / f"{name}.ref.{ext}"
            subprocess.run([str(REF_CLI), "convert", str(NETLIB / f"{name}.mps"), str(theirs)], check=True)
            if ext == "mps":
                assert out.read_bytes() == theirs.read_bytes()
            else:
                a, b = json.loads(out.read_text()), json.loads(theirs.read_text())
                assert a == b and list(a) == list(b)


def test_convert_mps_layout_matches_reference_file(engine, tmp_path):
    out = tmp_path / "afiro.mps"
    assert run_cli(engine, "convert", NETLIB / "afiro.mps", out).returncode == 0
    assert out.read_bytes() == (NETLIB / "afiro.reference_convert.mps").read_bytes()


def test_convert_sections_and_errors(engine, tmp_path):
    # max sense, objective offset, a ranged row, >= and = rows, free, fixed and
    # one-sided columns: every MPS section the writer has
    model = {"name": "sections", "sense": "max",
             "variables": [{"name": "a", "lb": -2, "ub": 4}, {"name": "b", "lb": None, "ub": 3},
                           {"name": "c", "lb": 1.5, "ub": 1.5}, {"name": "d", "lb": None}],
             "objective": {"linear": [1, -2, 0.25, 0.1], "offset": 7},
             "constraints": [{"name": "range", "lb": -1, "ub": 5, "coefficients": {"a": 1, "b": 2}},
                             {"name": "ge", "lb": 0.5, "coefficients": {"b": 1, "d": -1}},
                             {"name": "eq", "lb": 2, "ub": 2, "coefficients": {"a": 1, "c": 1, "d": 3}},
                             {"name": "le", "ub": 9, "coefficients": {"d": 1}}]}
    src = tmp_path / "sections.json"
    src.write_text(json.dumps(model))
    fp = json.loads(run_cli(engine, "inspect", src).stdout)["fingerprint"]
    for ext in ("mps", "json"):
        out = tmp_path / f"out.{ext}"
        assert run_cli(engine, "convert", src, out).returncode == 0
        assert json.loads(run_cli(engine, "inspect", out).stdout)["fingerprint"] == fp
        if REF_CLI.exists():
            theirs = tmp_path / f"ref.{ext}"
            subprocess.run([str(REF_CLI), "convert", str(src), str(theirs)], check=True)
            assert out.read_bytes() == theirs.read_bytes() if ext == "mps" else \
                json.loads(out.read_text()) == json.loads(theirs.read_text())
    text = (tmp_path / "out.mps").read_text()
    for section in ("OBJSENSE\n MAX", "RANGES\n RNG1 range 6", " MI BND b", " FX BND c 1.5", " MI BND d", "RHS1 OBJ -7"):  # MPS objective RHS is -offset
        assert section in text, section
    bad = run_cli(engine, "convert", src, tmp_path / "out.lp")
    assert bad.returncode == 1 and "Convert output must be .json or .mps" in bad.stderr
    assert run_cli(engine, "convert", src).returncode == 1
    free_row = dict(model, constraints=[{"name": "free", "coefficients": {"a": 1}}])
    src.write_text(json.dumps(free_row))
    assert run_cli(engine, "convert", src, tmp_path / "f.mps").returncode == 1
    assert run_cli(engine, "convert", src, tmp_path / "f.json").returncode == 0
    spaced = dict(model, variables=[dict(v, name=v["name"] + " x") for v in model["variables"]],
                  constraints=[])
    src.write_text(json.dumps(spaced))
    out = run_cli(engine, "convert", src, tmp_path / "s.mps")
    assert out.returncode == 1 and "whitespace" in out.stderr


def test_model_handle_writes_reference_layout(engine, tmp_path):
    m = engine.Model(NETLIB / "afiro.mps")
    m.write(tmp_path / "afiro.mps")
    assert (tmp_path / "afiro.mps").read_bytes() == (NETLIB / "afiro.reference_convert.mps").read_bytes()
    m.write(tmp_path / "afiro.json")
    assert engine.Model(tmp_path / "afiro.json").fingerprint == m.fingerprint
    with pytest.raises(ValueError, match="Convert output"):
        m.write(tmp_path / "afiro.lp")


def test_batch_refuses_polishing(engine):
    """Feasibility polishing is a single-LP driver feature; the batched engine
    refuses it at creation (before it looks for a device)."""
    with pytest.raises(RuntimeError, match="polishing"):
        engine.Batch(arena_bytes=1 << 20, max_slots=1, polishing=1)
