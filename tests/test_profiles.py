"""The committed rocprofv3 summaries and the bench's roofline fields agree.

bench.py takes the dominant kernel's HBM traffic per launch from
profiles/batch_advance_hbm_traffic_4096lp.txt; the kernel trace summary states
the bench's own HIP-event average launch time beside rocprof's mean for the
same kernel.  These tests keep the two files usable and consistent: every
summary names the command that made it, the traffic file has an entry for the
dominant kernel, and the two launch-time figures agree within 2 %.
"""
import json
import re
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PROFILES = ROOT / "profiles"


def test_every_summary_names_its_command():
    for path in PROFILES.glob("*.txt"):
        text = path.read_text()
        assert "rocprofv3" in text and "bench.py" in text, path.name


def dominant_kernel():
    """The bench's dominant kernel and its HIP-event launch time, as the trace
    summary's header states them."""
    text = (PROFILES / "batch_advance_trace_4096lp.txt").read_text()
    bench_ms = float(re.search(r"avg_launch_ms ([0-9.]+)", text).group(1))
    kernel = re.search(r"^#\s+kernel (\S+(?:, \S+)*) \(", text, re.M).group(1)
    return text, kernel, bench_ms


def test_traffic_entry_for_the_dominant_kernel():
    sys.path.insert(0, str(ROOT))
    import bench

    _, kernel, _ = dominant_kernel()
    traffic = bench.committed_traffic(kernel)
    assert traffic is not None and traffic > 0, kernel
    entries = [json.loads(line) for line in (PROFILES / "batch_advance_hbm_traffic_4096lp.txt").read_text().splitlines()
               if line.startswith("{")]
    assert all(e["hbm_bytes_per_dispatch"] > 0 and e["dispatches"] > 0 for e in entries)


def test_trace_agrees_with_the_bench_launch_time():
    text, kernel, bench_ms = dominant_kernel()
    stats = [line for line in text.splitlines() if line.startswith(kernel + " ") and "avg_ns=" in line]
    assert stats, kernel
    rocprof_ms = float(re.search(r"avg_ns=\s*([0-9.]+)", stats[0]).group(1)) / 1e6
    assert abs(bench_ms - rocprof_ms) <= 0.02 * rocprof_ms, (bench_ms, rocprof_ms)
