"""Records the reference CLI's own results on the Netlib fixtures.

Runs the reference build (oracle/_ref/niryukti_ref, from oracle/ref.mk) with
`solve MODEL --method pdhg --device cpu` and `inspect MODEL`, and stores the
result documents under tests/golden/netlib/ as data; also keeps the reference's
`convert afiro.mps afiro.reference_convert.mps` output and its --power-iterations
estimates (reference_power.json).  Run from the repository
root after `make -f oracle/ref.mk REF=/root/reference`.
"""
import json
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
CLI = ROOT / "oracle" / "_ref" / "niryukti_ref"
OUT = ROOT / "tests" / "golden" / "netlib"

for name in ("afiro", "adlittle", "israel", "e226"):
    model = OUT / f"{name}.mps"
    inspect = json.loads(subprocess.run([str(CLI), "inspect", str(model)], capture_output=True, text=True,
                                        check=True).stdout)
    res = subprocess.run([str(CLI), "solve", str(model), "--method", "pdhg", "--device", "cpu"],
                         capture_output=True, text=True)
    doc = json.loads(res.stdout)
    doc["reference_exit_code"] = res.returncode
    doc["reference_inspect"] = inspect
    (OUT / f"{name}.reference_result.json").write_text(json.dumps(doc) + "\n")
    print(name, doc["status"], doc["objective"], doc["performance"]["iterations"], inspect["fingerprint"])

subprocess.run([str(CLI), "convert", str(OUT / "afiro.mps"), str(OUT / "afiro.reference_convert.mps")], check=True)

# operator-norm estimates of the prepared models (--power-iterations, src/research.cpp:6-50)
power = {}
for name in ("afiro", "adlittle", "israel", "e226"):
    for its in (1, 30):
        res = subprocess.run([str(CLI), "solve", str(OUT / f"{name}.mps"), "--method", "pdhg", "--device", "cpu",
                              "--power-iterations", str(its), "--iterations", "1"], capture_output=True, text=True)
        power[f"{name}/{its}"] = json.loads(res.stdout)["performance"]["operator_norm_estimate"]
(OUT / "reference_power.json").write_text(json.dumps(power, indent=1) + "\n")
