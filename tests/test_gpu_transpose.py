"""GPU parity of the transposes (bgb_csr2csc, bgb_b2sr_transpose) and the tile
statistics (bgb_tile_row_stats, bgb_estimate_tiles) against the CPU oracle.

All of it is index / bit work: results must match bit-exactly. At full size
(R-MAT scale 20, symmetric) the transpose is checked through a property that
needs no oracle: a symmetric matrix is its own transpose, so the transposed
B2SR must equal the packed one word for word.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TS = (4, 8, 16, 32, 64)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


def directed_csr(orc, rng, n, m, nedges):
    from bgb import graph
    src = rng.integers(0, n, nedges).astype(np.int32)
    dst = rng.integers(0, m, nedges).astype(np.int32)
    rp, ci = orc.normalize(n, m, src, dst, undirected=False, drop_self=False)
    return graph.HostCSR(n, m, rp, ci)


def words(t, T):
    from bgb.b2sr import WORD_NUMPY
    return t.cpu().numpy().view(WORD_NUMPY[T])


@pytest.mark.parametrize("shape", [(1, 1), (300, 170), (2000, 3100), (1 << 14, 1 << 14)])
def test_csr2csc_matches_oracle(orc, dev, shape):
    import torch
    from bgb import b2sr as B
    n, m = shape
    rng = np.random.default_rng(n + m)
    csr = directed_csr(orc, rng, n, m, 8 * max(n, m))
    log(f"[case] csr2csc {n}x{m} nnz={csr.nnz}")
    vals = rng.standard_normal(csr.nnz).astype(np.float32)
    d_rp = torch.from_numpy(csr.rowptr).to(dev)
    d_ci = torch.from_numpy(csr.colind).to(dev)
    d_v = torch.from_numpy(vals).to(dev)
    cp, ri, cv = B.csr2csc(n, m, csr.nnz, d_rp, d_ci, d_v)
    wcp, wri, wcv = orc.csr2csc(n, m, csr.rowptr, csr.colind, vals)
    np.testing.assert_array_equal(cp.cpu().numpy(), wcp)
    np.testing.assert_array_equal(ri.cpu().numpy(), wri)
    np.testing.assert_array_equal(cv.cpu().numpy(), wcv)
    # structure only
    cp2, ri2, cv2 = B.csr2csc(n, m, csr.nnz, d_rp, d_ci)
    assert cv2 is None
    np.testing.assert_array_equal(ri2.cpu().numpy(), wri)


def test_csr2csc_unsorted_rows_and_empty(orc, dev):
    import torch
    from bgb import b2sr as B
    log("[case] csr2csc unsorted rows")
    rng = np.random.default_rng(3)
    csr = directed_csr(orc, rng, 500, 400, 3000)
    ci = csr.colind.copy()
    for r in range(csr.nrows):  # shuffle every row: csr2csc does not need sorted rows
        s, e = csr.rowptr[r], csr.rowptr[r + 1]
        ci[s:e] = rng.permutation(ci[s:e])
    d_rp = torch.from_numpy(csr.rowptr).to(dev)
    cp, ri, _ = B.csr2csc(500, 400, csr.nnz, d_rp, torch.from_numpy(ci).to(dev))
    wcp, wri, _ = orc.csr2csc(500, 400, csr.rowptr, ci)
    np.testing.assert_array_equal(cp.cpu().numpy(), wcp)
    np.testing.assert_array_equal(ri.cpu().numpy(), wri)
    log("[case] csr2csc empty")
    rp0 = torch.zeros(6, dtype=torch.int32, device=dev)
    cp, ri, _ = B.csr2csc(5, 9, 0, rp0, torch.zeros(1, dtype=torch.int32, device=dev))
    assert cp.cpu().tolist() == [0] * 10 and ri.numel() == 0


def test_csr2csc_rejects_bad_input(dev):
    import torch
    from bgb import b2sr as B
    from bgb._lib import BgbError
    log("[case] csr2csc bad input")
    rp = torch.tensor([0, 2, 3], dtype=torch.int32, device=dev)
    with pytest.raises(BgbError, match="ERR_FORMAT"):
        B.csr2csc(2, 4, 3, rp, torch.tensor([0, 4, 1], dtype=torch.int32, device=dev))  # column 4 of 4
    with pytest.raises(BgbError, match="ERR_FORMAT"):
        B.csr2csc(2, 4, 3, torch.tensor([0, 3, 2], dtype=torch.int32, device=dev),
                  torch.tensor([0, 1, 1], dtype=torch.int32, device=dev))  # rowptr decreases


@pytest.mark.parametrize("T", TS)
def test_b2sr_transpose_matches_oracle(orc, dev, T):
    from bgb import b2sr as B
    from bgb import graph
    for n, m in ((37, 53), (1500, 900), (4096, 4096)):
        rng = np.random.default_rng(T * 7 + n)
        csr = directed_csr(orc, rng, n, m, 6 * max(n, m))
        log(f"[case] b2sr_transpose T={T} {n}x{m} nnz={csr.nnz}")
        A = B.B2SR.from_csr(csr, T, dev)
        At = A.transpose()
        assert (At.nrows, At.ncols, At.nblocks) == (m, n, A.nblocks)
        rp, ci, val = At.to_host()
        # oracle: pack the CSR of A^T
        cp, ri, _ = orc.csr2csc(n, m, csr.rowptr, csr.colind)
        wrp, wci, wval = orc.b2sr(T, m, n, cp, ri)
        np.testing.assert_array_equal(rp, wrp)
        np.testing.assert_array_equal(ci, wci)
        np.testing.assert_array_equal(val, wval)
        # and the oracle's own tile transpose
        orp, oci, oval = orc.b2sr_transpose(T, -(-m // T), *A.to_host())
        np.testing.assert_array_equal(val, oval)
        # packing A^T's CSR on the GPU gives the same words; transposing back gives A
        Bt = B.B2SR.from_csr(graph.HostCSR(m, n, cp, ri), T, dev)
        np.testing.assert_array_equal(Bt.to_host()[2], val)
        Att = At.transpose()
        for a, b in zip(Att.to_host(), A.to_host()):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("T", (8, 32, 64))
def test_b2sr_transpose_symmetric_rmat20_is_identity(dev, T):
    """Full-size property: R-MAT 20 is symmetrised, so A^T = A word for word."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(20, 16, seed=1)
    log(f"[case] symmetric rmat20 T={T} nnz={csr.nnz}")
    A = B.B2SR.from_csr(csr, T, dev)
    At = A.transpose()
    assert At.nblocks == A.nblocks
    assert torch.equal(At.rowptr, A.rowptr)
    assert torch.equal(At.colind[: A.nblocks], A.colind[: A.nblocks])
    assert torch.equal(At.val[: A.nblocks * T], A.val[: A.nblocks * T])


def test_transposed_bmv_is_the_push_product(orc, dev):
    """bin_bin_full over the transposed B2SR equals the CSR count product with A^T."""
    from bgb import b2sr as B
    rng = np.random.default_rng(11)
    n = 3000
    csr = directed_csr(orc, rng, n, n, 10 * n)
    log(f"[case] transposed bmv n={n} nnz={csr.nnz}")
    for T in (8, 32):
        At = B.B2SR.from_csr(csr, T, dev).transpose()
        x01 = (rng.random(n) < 0.3).astype(np.float32)
        xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
        got = B.bmv_bin_bin_full(At, xb).cpu().numpy()[:n]
        cp, ri, _ = orc.csr2csc(n, n, csr.rowptr, csr.colind)
        np.testing.assert_array_equal(got, orc.csr_count(cp, ri, x01.astype(np.uint8)))


@pytest.mark.parametrize("T", TS)
def test_tile_stats_match_oracle(orc, dev, T):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    for csr in (graph.rmat(14, 16, seed=5), directed_csr(orc, np.random.default_rng(T), 777, 1234, 5000),
                graph.HostCSR(3, 3, np.zeros(4, np.int32), np.zeros(0, np.int32))):
        log(f"[case] tile stats T={T} n={csr.nrows} nnz={csr.nnz}")
        A = B.B2SR.from_csr(csr, T, dev)
        tpr, nzr = A.row_stats()
        wtpr, wnzr = orc.tile_row_stats(T, *A.to_host()[::2])
        np.testing.assert_array_equal(tpr.cpu().numpy(), wtpr)
        np.testing.assert_array_equal(nzr.cpu().numpy(), wnzr)
        d_rp = torch.from_numpy(csr.rowptr).to(dev)
        d_ci = torch.from_numpy(csr.colind if csr.nnz else np.zeros(1, np.int32)).to(dev)
        est = B.estimate_tiles(T, csr.nrows, d_rp, d_ci)
        assert est == orc.estimate_tiles(csr.rowptr, csr.colind, T)
        # sorted rows: the estimate counts the non-zero row words exactly
        assert est == int(wnzr.sum())
        assert est >= A.nblocks


@pytest.mark.parametrize("T", (4, 8, 16, 32))
def test_directed_bfs_pulls_over_the_transpose(orc, dev, T):
    """Directed graphs: the batched BFS pulls over in-edges, i.e. the rows of A^T,
    so BFS over the transposed B2SR gives the forward (out-edge) levels of A --
    the oracle's queue BFS over A's CSR. The BFS over A itself gives the levels
    of the reversed graph (the oracle's BFS over A^T)."""
    from bgb import b2sr as B
    rng = np.random.default_rng(100 + T)
    n = 2500
    csr = directed_csr(orc, rng, n, n, 4 * n)
    log(f"[case] directed bfs T={T} nnz={csr.nnz}")
    A = B.B2SR.from_csr(csr, T, dev)
    At = A.transpose()
    src = rng.integers(0, n, 64)
    lev, _ = B.bfs_multi(At, src)
    want = np.stack([orc.bfs(csr.rowptr, csr.colind, int(s))[0] for s in src])
    np.testing.assert_array_equal(lev.cpu().numpy(), want)
    cp, ri, _ = orc.csr2csc(n, n, csr.rowptr, csr.colind)
    lev_rev, _ = B.bfs_multi(A, src)
    want_rev = np.stack([orc.bfs(cp, ri, int(s))[0] for s in src])
    np.testing.assert_array_equal(lev_rev.cpu().numpy(), want_rev)
    # direction switching pushes along the rows of the pulled matrix's transpose
    lev_d, _ = B.bfs_multi(At, src, push=A)
    np.testing.assert_array_equal(lev_d.cpu().numpy(), want)
    lev_rd, _ = B.bfs_multi(A, src, push=At)
    np.testing.assert_array_equal(lev_rd.cpu().numpy(), want_rev)
