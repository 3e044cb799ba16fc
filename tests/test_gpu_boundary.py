"""GPU checks of the C-ABI contract around the BMV operators.

* Bit outputs are exactly NBR words (the reference caller's allocation,
  test_bin_bin_bin.cu:329): the last block rows -- whose words share a dword
  with no word after them -- carry pieces of hub rows across many tile chunks,
  and the bytes behind the output must come back unchanged.
* The caching entry points refuse to run while their stream is being
  captured (they would leave the graph pointing at library-owned scratch).
* A wide matrix (more block columns than block rows) reads x past
  ceil(NBR/T)*T words on every x-staging path.
Results are compared with the CPU oracle bit-exactly.
"""
import sys

import numpy as np
import pytest

from test_gpu_bmv import check_all_ops, check_guard, guarded_bits, to_np

pytestmark = pytest.mark.gpu

TS = (4, 8, 16, 32, 64)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


@pytest.mark.parametrize("T", TS)
def test_bit_outputs_exact_size_tail_rows(orc, dev, T):
    """NBR = 4k + r block rows (r = 0..3, the last one partial), hubs in the
    last three block rows (each spans tens of chunks, so the carries of the
    words in the last partial dword come from more than 64 chunks at T = 4):
    bin_bin_bin (overwrite and accumulate) and masked match the oracle, and
    the guard bytes behind the NBR words are unchanged."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(300 + T)
    for r in (1, 2, 3, 0):
        nbr = 4 * (150000 // (4 * T)) + r
        n = nbr * T - T // 2
        hubs = [(nbr - 1 - i) * T + i for i in range(3)] + [n - 1]
        src = np.concatenate([np.full(n, h) for h in hubs] + [rng.integers(0, n, 2 * n)])
        dst = np.concatenate([np.arange(n) for _ in hubs] + [rng.integers(0, n, 2 * n)])
        csr = graph.from_coo(n, n, src, dst, False, True)
        A = B.B2SR.from_csr(csr, T, dev)
        for dens in (0.5, 0.0005):
            log(f"[case] tail rows T={T} NBR={nbr} (mod 4 = {r}) x density {dens}")
            x01 = (rng.random(n) < dens).astype(np.float32)
            xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
            counts = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
            want = orc.pack_u8(T, (counts > 0).astype(np.uint8), nbr)
            raw, C = guarded_bits(A, dev)
            np.testing.assert_array_equal(to_np(B.bmv_bin_bin_bin(A, xb, out=C), T), want)
            check_guard(raw, A)
            # accumulate onto random words
            raw, C = guarded_bits(A, dev)
            pre = rng.integers(0, 1 << min(T, 62), nbr, dtype=np.uint64).astype(B.WORD_NUMPY[T])
            if T == 4:
                pre &= 0xF
            C.copy_(torch.from_numpy(pre.view(to_np(C).dtype)).to(dev))
            np.testing.assert_array_equal(to_np(B.bmv_bin_bin_bin(A, xb, out=C, accumulate=True), T), want | pre)
            check_guard(raw, A)
            mbits = (rng.random(n) < 0.5).astype(np.uint8)
            mask = torch.from_numpy(orc.pack_u8(T, mbits, nbr).view(to_np(C).dtype)).to(dev)
            raw, C = guarded_bits(A, dev)
            got = to_np(B.bmv_bin_bin_bin_masked(A, xb, mask, out=C), T)
            np.testing.assert_array_equal(got, orc.pack_u8(T, ((counts > 0) & (mbits == 0)).astype(np.uint8), nbr))
            check_guard(raw, A)


def test_bmv_rejects_stream_capture(dev):
    """A BMV entry point called while its stream is being captured returns
    BGB_ERR_ARG before it enqueues anything; the stream and the library work
    as before afterwards."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    from bgb._lib import BgbError
    csr = graph.rmat(10, 8, seed=1)
    A = B.B2SR.from_csr(csr, 8, dev)
    x01 = (np.arange(csr.nrows) % 3 == 0).astype(np.float32)
    xb = B.pack_vec(B.padded_vector(x01, 8, dev), 8)
    C = torch.zeros(A.nblockrows * 8, dtype=torch.int32, device=dev)
    want = to_np(B.bmv_bin_bin_full(A, xb)).copy()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    codes = []
    with torch.cuda.stream(s):
        g.capture_begin()
        try:
            for fn in (lambda: B.bmv_bin_bin_full(A, xb, out=C), lambda: B.bmv_bin_bin_bin(A, xb)):
                try:
                    fn()
                    codes.append(0)
                except BgbError as e:
                    codes.append(e.status)
            C.add_(0)  # one node, so the graph is not empty
        finally:
            g.capture_end()
    assert codes == [1, 1]
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(B.bmv_bin_bin_full(A, xb)), want)


@pytest.mark.parametrize("T", TS)
def test_wide_matrix_every_x_staging(orc, dev, T):
    """ncols > nrows: the packed x holds a word per block column, more than
    ceil(NBR/T)*T; gathered from memory, staged in LDS, or automatic, the
    results match the oracle. Also the B2SR transpose of a tall matrix."""
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(500 + T)
    n = 3000
    m = 4 * n + 37
    e = 12 * m
    wide = graph.from_coo(n, m, rng.integers(0, n, e), rng.integers(0, m, e), False, False)
    tall = graph.from_coo(m, n, rng.integers(0, m, e), rng.integers(0, n, e), False, False)
    try:
        for mode in (0, 1, 2):
            bgb.call("bgb_set_x_staging", T, mode)
            check_all_ops(orc, wide, T, rng, dev, f"wide {n}x{m} x staging {mode}", xdensity=0.2)
            At = B.B2SR.from_csr(tall, T, dev).transpose()  # n x m
            x01 = (rng.random(m) < 0.2).astype(np.float32)
            xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
            # counts of the transpose: column sums of the tall matrix
            rows = np.repeat(np.arange(m), np.diff(tall.rowptr))
            want = np.zeros(n, np.int64)
            np.add.at(want, tall.colind, x01[rows].astype(np.int64))
            log(f"[case] transpose of tall {m}x{n} T={T} x staging {mode}")
            got = to_np(B.bmv_bin_bin_full(At, xb))[:n]
            np.testing.assert_array_equal(got, want)
    finally:
        bgb.call("bgb_set_x_staging", T, 0)
