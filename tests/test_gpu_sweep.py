"""The sweep over local MatrixMarket files (bgb.sweep, the reference's
testbmv.sh output format) and the reference-shaped C++ driver reading .mtx
files with and without the transpose flag."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from mtx_files import write_mtx, write_mycielskian

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "bmv_driver")
NUM = r"\d+\.\d{6}"


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("mtx")
    out = []
    for k in (4, 8, 11):
        p = str(d / f"mycielskian{k}.mtx")
        write_mycielskian(p, k)
        out.append(p)
    rng = np.random.default_rng(5)
    n = 5000
    src, dst = rng.integers(0, n, 40000), rng.integers(0, n, 40000)
    p = str(d / "directed5000.mtx")
    write_mtx(p, n, n, list(zip(src, dst)), sym="general")
    out.append(p)
    return out


@pytest.mark.parametrize("op", ["bin_bin_full", "bin_full_full", "bin_bin_bin", "baseline"])
def test_sweep_csv_lines(cuda, files, tmp_path, op):
    from bgb import sweep
    out = str(tmp_path / f"output_{op}.csv")
    log(f"[case] sweep {op}")
    assert sweep.main([op, *files, "--tiles", "4,8,16,32,64", "--out", out]) == 0
    lines = open(out).read().split("\n")
    assert lines[-1] == "" and len(lines) == len(files) + 1
    nvals = 1 if op == "baseline" else 5
    for line, path in zip(lines, files):
        name = os.path.splitext(os.path.basename(path))[0]
        assert re.fullmatch(re.escape(name) + " " + f"(?:{NUM} ){{{nvals}}}", line), line
        assert all(float(v) > 0 for v in line.split()[1:])


def test_sweep_transpose(cuda, files):
    from bgb import sweep
    log("[case] sweep transpose")
    line = sweep.sweep_line("bin_bin_full", files[-1], [8, 32], transpose=1)
    assert re.fullmatch(f"directed5000 (?:{NUM} ){{2}}", line), line


@pytest.mark.parametrize("transpose", [0, 1])
@pytest.mark.parametrize("T", [8, 32, 64])
def test_driver_reads_mtx(files, T, transpose):
    assert os.path.exists(DRIVER), "driver not built (run __graft_entry__.build())"
    for path in (files[2], files[3]):
        for op in ("bin_bin_full", "bin_full_full", "bin_bin_bin"):
            log(f"[case] driver {op} T={T} {os.path.basename(path)} transpose={transpose}")
            r = subprocess.run([DRIVER, op, str(T), "mtx:" + path, str(transpose)], capture_output=True,
                               text=True, timeout=120)
            print(r.stdout.strip(), r.stderr.strip(), flush=True)
            assert r.returncode == 0 and "success: 1" in r.stdout, (op, T, path, r.stdout, r.stderr)
