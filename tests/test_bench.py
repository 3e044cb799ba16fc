"""Host logic of bench.py (no GPU): the multi-GPU refusal, the algorithmic
byte count of a BMV call, and the provenance of the stored traffic figure."""
import hashlib
import importlib.util
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_bmv_bytes_counts_each_array_once(bench):
    # 3 block rows, 5 tiles at T = 8: rowptr 4*4, colind 4*5, words 5*8*1,
    # packed x 3 words of 1 byte rounded up to 8, counts 4*3*8
    assert bench.bmv_bytes(8, 3, 5) == 16 + 20 + 40 + 8 + 96
    # T = 32: 4-byte words; packed x 1 block row rounded up to 32 words
    assert bench.bmv_bytes(32, 1, 2) == 8 + 8 + 2 * 32 * 4 + 32 * 4 + 4 * 32


def test_traffic_source_flags_a_stale_figure(bench, tmp_path, monkeypatch):
    src = tmp_path / "bgb_bmv.hip"
    src.write_bytes(b"kernel source v1\n")
    h1 = hashlib.sha256(b"kernel source v1\n").hexdigest()
    table = {"rmat20": {"4": {"bytes_per_call": 123, "kernel": "k4", "source_sha256": h1},
                        "8": {"bytes_per_call": 456, "kernel": "k8"}}}
    tf = tmp_path / "traffic.json"
    tf.write_text(json.dumps(table))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(tf))
    monkeypatch.setattr(bench, "BMV_SOURCE", str(src))
    assert bench.measured_traffic("rmat20", 4) == 123
    assert bench.traffic_source("rmat20", 4)["measured_on_this_source"] is True
    assert bench.traffic_source("rmat20", 4)["kernel"] == "k4"
    # no hash recorded: unknown, not current
    assert bench.traffic_source("rmat20", 8)["measured_on_this_source"] is None
    # not measured at all
    assert bench.measured_traffic("rmat22", 4) is None
    assert bench.traffic_source("rmat22", 4) is None
    # the kernel source changes after the measurement
    src.write_bytes(b"kernel source v2\n")
    assert bench.traffic_source("rmat20", 4)["measured_on_this_source"] is False


def test_committed_traffic_figures_name_their_kernel(bench):
    with open(bench.TRAFFIC_FILE) as f:
        table = json.load(f)
    for workload in ("rmat20", "rmat22"):
        for T in ("4", "8"):
            e = table[workload][T]
            assert e["bytes_per_call"] > 0 and "bmv_piece_kernel" in e["kernel"]
            assert len(e.get("source_sha256", "")) == 64
