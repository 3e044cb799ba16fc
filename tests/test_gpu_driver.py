"""Runs the reference-shaped C++ driver (tests/cpp/bmv_driver.cpp): the
reference's test flow (CSR2HSBSF -> packRandVec -> 5 accumulating launches ->
verifyResult) through the bgb_ref C++ mirror of the reference's operator
names, verified against the CPU oracle inside the driver."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "bmv_driver")


@pytest.mark.parametrize("op", ["bin_bin_full", "bin_bin_bin", "bin_full_full"])
@pytest.mark.parametrize("T", [4, 8, 16, 32, 64])
def test_reference_shaped_driver(op, T):
    assert os.path.exists(DRIVER), "driver not built (run __graft_entry__.build())"
    for graph in ("er:4096:16", "rmat:14:16", "grid:200:150"):
        print(f"[case] driver {op} T={T} {graph}", flush=True)
        sys.stdout.flush()
        r = subprocess.run([DRIVER, op, str(T), graph], capture_output=True, text=True, timeout=120)
        print(r.stdout.strip(), r.stderr.strip(), flush=True)
        assert r.returncode == 0 and "success: 1" in r.stdout, (op, T, graph, r.stdout, r.stderr)
