"""SuiteSparse-style sweep over local MatrixMarket files, in the reference's
CSV format.

The reference's sweep (testbmv.sh:1011-1110) downloads each SuiteSparse
matrix, runs one driver binary per tile dim on it (bmv{T}_<op> <file> 0,
test_bin_bin_full.cu:420-460) and appends what the drivers print to
output_<op>.csv: one line per matrix, "<name> <ms T=4> <ms T=8> <ms T=16>
<ms T=32> " -- each value the mean time of one BMV over TEST_TIMES (5)
launches (test_bin_bin_full.cu:311-353). Mode "baseline" prints one value per
matrix, the vendor CSR SpMV time (the BASELINE build, cuSPARSE csrmv there,
rocSPARSE csrmv here).

This module does the same over .mtx files already on disk (nothing is
downloaded):

    PYTHONPATH=bit-graphblas_amd python -m bgb.sweep bin_bin_full a.mtx b.mtx --tiles 4,8,16,32 [--transpose 1]

Inputs follow the drivers: readMtx with directed = 0 (bgb_read_mtx), an
optional transpose (argv[2] of the drivers; csr2csc on the GPU), a random 0/1
vector (genRandVec, test_bin_bin_full.cu:223-233) packed with ToBit{T}Row, and
the output zeroed before the timed launches, which accumulate as the
reference's do.
"""
import argparse
import ctypes
import os
import sys

import numpy as np
import torch

from . import b2sr as B
from . import graph

OPS = ("bin_bin_full", "bin_full_full", "bin_bin_bin", "baseline")
TEST_TIMES = 5


def _device_csr(csr, dev, transpose):
    rp = torch.from_numpy(csr.rowptr).to(dev)
    ci = torch.from_numpy(csr.colind if csr.nnz else np.zeros(1, np.int32)).to(dev)
    n, m = csr.nrows, csr.ncols
    if transpose:
        rp, ci, _ = B.csr2csc(n, m, csr.nnz, rp, ci)
        n, m = m, n
    return n, m, rp, ci


def _time_ms(fn, times):
    fn()  # first launch (partition pass, workspace) outside the timed region
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(times):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / times


def bmv_time(op, T, n, m, nnz, d_rp, d_ci, x, dev, times=TEST_TIMES):
    """Mean ms of one `op` BMV at tile dim T (accumulating launches, as the drivers)."""
    A = B.B2SR.from_device_csr(T, n, m, nnz, d_rp, d_ci)
    nbr = A.nblockrows
    # x has one entry per column; padded to whole block columns (and at least NBR*T)
    k = max(nbr, (m + T - 1) // T)
    xpad = torch.zeros(k * T, dtype=torch.float32, device=dev)
    xpad[: x.numel()] = x
    if op == "bin_full_full":
        out = torch.zeros(nbr * T, dtype=torch.float32, device=dev)
        return _time_ms(lambda: B.bmv_bin_full_full(A, xpad, out=out, accumulate=True), times)
    xb = B.pack_vec(xpad, T, k)
    if op == "bin_bin_full":
        out = torch.zeros(nbr * T, dtype=torch.float32, device=dev)
        return _time_ms(lambda: B.bmv_bin_bin_full(A, xb, out=out, accumulate=True, dtype=torch.float32),
                        times)
    out = B._bits_out(None, A)
    return _time_ms(lambda: B.bmv_bin_bin_bin(A, xb, out=out, accumulate=True), times)


def baseline_time(n, nnz, d_rp, d_ci, x, times=TEST_TIMES):
    """rocSPARSE csrmv (FP32 CSR, all values 1) mean ms, from lib/libbgb_baseline.so.
    Square matrices only, as the reference's baseline driver."""
    if x.numel() != n:
        raise ValueError("the baseline times square matrices only")
    path = os.path.join(os.path.dirname(B.__file__), "..", "lib", "libbgb_baseline.so")
    L = ctypes.CDLL(os.path.abspath(path))
    f = L.bgbx_rocsparse_csrmv
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f.restype = ci
    f.argtypes = [ci, ci, vp, vp, vp, vp, vp, ci, ci, vp, vp, vp]
    val = torch.ones(max(nnz, 1), dtype=torch.float32, device=d_rp.device)
    y = torch.zeros(max(n, 1), dtype=torch.float32, device=d_rp.device)
    ms, pre = ctypes.c_double(), ctypes.c_double()
    st = f(n, nnz, d_rp.data_ptr(), d_ci.data_ptr(), val.data_ptr(), x.data_ptr(), y.data_ptr(), 0, times,
           ctypes.byref(ms), ctypes.byref(pre), torch.cuda.current_stream().cuda_stream)
    if st != 0:
        raise RuntimeError(f"rocSPARSE csrmv failed ({st})")
    return ms.value


def sweep_line(op, path, tiles, transpose=0, seed=0, times=TEST_TIMES, dev="cuda"):
    """One output line of the sweep for the matrix in `path` (without the newline)."""
    dev = torch.device(dev)
    csr = graph.read_mtx(path, directed=0)
    n, m, d_rp, d_ci = _device_csr(csr, dev, transpose)
    rng = np.random.default_rng(seed)
    x = torch.from_numpy((rng.random(m) > 0.5).astype(np.float32)).to(dev)
    name = os.path.splitext(os.path.basename(path))[0]
    if op == "baseline":
        vals = [baseline_time(n, csr.nnz, d_rp, d_ci, x, times)]
    else:
        vals = [bmv_time(op, T, n, m, csr.nnz, d_rp, d_ci, x, dev, times) for T in tiles]
    return name + " " + "".join(f"{v:f} " for v in vals)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("op", choices=OPS)
    ap.add_argument("files", nargs="+", help="MatrixMarket coordinate files")
    ap.add_argument("--tiles", default="4,8,16,32")
    ap.add_argument("--transpose", type=int, default=0, choices=(0, 1))
    ap.add_argument("--times", type=int, default=TEST_TIMES, help="timed launches per value (TEST_TIMES)")
    ap.add_argument("--out", default=None, help="CSV path (default output_<op>.csv)")
    args = ap.parse_args(argv)
    tiles = [int(t) for t in args.tiles.split(",")]
    out = args.out or f"output_{args.op}.csv"
    with open(out, "w") as f:
        for p in args.files:
            line = sweep_line(args.op, p, tiles, args.transpose, times=args.times)
            print(line, flush=True)
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
