"""bench.py's host logic: the committed traffic figure is used only for the
workload, kernel variant and kernel sources it was measured on, and more
than one GPU is refused."""
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parent.parent


def test_traffic_needs_matching_workload_variant_and_sources(monkeypatch):
    text = bench.TRAFFIC_FILE.read_text()
    fields = dict(l.split(":", 1) for l in text.splitlines()
                  if l.startswith(("primal_variant:", "kernel_source_sha256:")))
    v = int(fields["primal_variant"])
    recorded = fields["kernel_source_sha256"].strip()
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: recorded)
    assert bench.measured_traffic(66, v) > 0
    assert bench.measured_traffic(20, v) is None
    assert bench.measured_traffic(66, v + 1) is None
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: "0" * 16)
    assert bench.measured_traffic(66, v) is None


def test_valu_issue_needs_matching_workload_variant_and_sources(monkeypatch):
    text = bench.SQ_FILE.read_text()
    fields = dict(l.split(":", 1) for l in text.splitlines()
                  if l.startswith(("primal_variant:", "kernel_source_sha256:",
                                   "valu_instructions_per_launch:")))
    v = int(fields["primal_variant"])
    recorded = fields["kernel_source_sha256"].strip()
    instr = float(fields["valu_instructions_per_launch"])
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: recorded)
    r = bench.valu_issue(66, v, 200.0)
    # 4 cycles per wave64 instruction over 1,024 SIMDs at 2.4 GHz
    assert abs(r["issue_limit_us"] - instr * 4 / (1024 * 2.4e3)) < 1e-9
    assert abs(r["frac"] - r["issue_limit_us"] / 200.0) < 1e-12
    assert bench.valu_issue(20, v, 200.0) is None
    assert bench.valu_issue(66, v + 1, 200.0) is None
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: "0" * 16)
    assert bench.valu_issue(66, v, 200.0) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
