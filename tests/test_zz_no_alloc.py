"""The allocation-free bin_bin_full entry point (bgb_bmv_bin_bin_full_i32_noalloc)
through the C ABI, against the allocating one (bgb_bmv_bin_bin_full_i32) and
the oracle: bit-exact int32 counts on two calls in a row, the second on a
smaller matrix (the static scratch keeps the first call's contents), with
accumulate, at every tile dim; a matrix needing more than the 64 MiB static
scratch is refused before anything is launched; and the call captured into a
graph gives the same counts on every replay, with other work in between.
"""
import ctypes

import numpy as np
import pytest

from test_gpu_bmv import dirty, to_np

pytestmark = pytest.mark.gpu

TS = (4, 8, 16, 32, 64)


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


def _case(orc, csr, T, dev, seed):
    from bgb import b2sr as B
    A = B.B2SR.from_csr(csr, T, dev)
    x01 = (np.random.default_rng(seed).random(csr.nrows) < 0.3).astype(np.float32)
    xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
    return A, xb, orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))


@pytest.mark.parametrize("T", TS)
def test_noalloc_matches_allocating_entry_point(orc, dev, T):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    for k, csr in enumerate((graph.rmat(16, 16, seed=21), graph.rmat(11, 8, seed=22))):
        A, xb, counts = _case(orc, csr, T, dev, 40 + k)
        n = A.nblockrows * T
        want = to_np(B.bmv_bin_bin_full(A, xb)).copy()
        np.testing.assert_array_equal(want[: csr.nrows], counts)
        got = B.bmv_bin_bin_full(A, xb, out=dirty(n, torch.int32, dev), noalloc=True)
        np.testing.assert_array_equal(to_np(got), want)
        B.bmv_bin_bin_full(A, xb, out=got, accumulate=True, noalloc=True)
        np.testing.assert_array_equal(to_np(got), 2 * want)


def test_noalloc_refuses_matrices_beyond_the_scratch(dev):
    """nblocks past what 64 MiB of carry records hold at T = 32 (1 M chunks of
    64 tiles): BGB_ERR_SIZE from the host-side check, before any launch."""
    import torch
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(10, 8, seed=1)
    A = B.B2SR.from_csr(csr, 32, dev)
    xb = B.pack_vec(B.padded_vector(np.ones(csr.nrows, np.float32), 32, dev), 32)
    C = torch.zeros(A.nblockrows * 32, dtype=torch.int32, device=dev)
    st = bgb.lib().bgb_bmv_bin_bin_full_i32_noalloc(32, A.val.data_ptr(), xb.data_ptr(), C.data_ptr(),
                                                     A.rowptr.data_ptr(), A.colind.data_ptr(), A.nblockrows,
                                                     1 << 26, 0, torch.cuda.current_stream().cuda_stream)
    assert st == 5  # BGB_ERR_SIZE
    assert b"scratch" in bgb.lib().bgb_last_error()
    torch.cuda.synchronize()
    assert not C.any()


def test_noalloc_graph_replays(orc, dev):
    """Captured once, replayed three times with the output cleared and other
    BMVs run eagerly between replays: every replay gives the counts."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    A, xb, counts = _case(orc, graph.rmat(14, 8, seed=31), 8, dev, 50)
    A2, xb2, counts2 = _case(orc, graph.rmat(12, 16, seed=32), 8, dev, 51)
    C = torch.empty(A.nblockrows * 8, dtype=torch.int32, device=dev)
    B.bmv_bin_bin_full(A, xb, out=C, noalloc=True)  # first call on the device: scratch lookup
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g.capture_begin()
        B.bmv_bin_bin_full(A, xb, out=C, noalloc=True)
        g.capture_end()
    for r in range(3):
        C.fill_(-1)
        other = B.bmv_bin_bin_full(A2, xb2, noalloc=(r % 2 == 0))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(to_np(other)[: A2.nrows], counts2)
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(to_np(C)[: A.nrows], counts)
