"""Pins the CPU oracle before it is trusted as the parity reference.

Each oracle function is checked against something that does not restate it:
the hand-derived fixture tests/golden/b2sr_5x5.json, dense numpy products,
scipy.sparse.csgraph BFS, the reference's own readMtx.hpp (oracle/_ref) and
closed forms.
"""
import json
import os

import numpy as np
import pytest

from oracle_lib import REF_SO, WORD_NP, RefReadMtx, trailing_entry_erased

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "b2sr_5x5.json")
TS = (4, 8, 16, 32, 64)


@pytest.fixture(scope="module")
def gold():
    with open(GOLDEN) as f:
        return json.load(f)


def rand_csr(rng, n, m, density, directed=True):
    """Random 0/1 matrix as (dense, rowptr, colind)."""
    D = (rng.random((n, m)) < density).astype(np.int64)
    if not directed:
        D = np.triu(D, 1)
        D = D + D.T
    rp = np.concatenate([[0], np.cumsum(D.sum(1))]).astype(np.int32)
    ci = np.nonzero(D)[1].astype(np.int32)
    return D, rp, ci


def dense_tiles(D, T):
    """Independent B2SR construction: numpy reshape + packbits (MSB first)."""
    n, m = D.shape
    nbr, nbc = -(-n // T), -(-m // T)
    P = np.zeros((nbr * T, nbc * T), np.uint8)
    P[:n, :m] = D
    tiles = P.reshape(nbr, T, nbc, T).transpose(0, 2, 1, 3)  # [I, J, r, c]
    rowptr, colind, words = [0], [], []
    for I in range(nbr):
        for J in range(nbc):
            t = tiles[I, J]
            if t.any():
                colind.append(J)
                # row r -> T-bit integer with column 0 as the most significant bit
                words.extend(int("".join(map(str, t[r])), 2) for r in range(T))
        rowptr.append(len(colind))
    return (np.array(rowptr, np.int32), np.array(colind, np.int32),
            np.array(words, dtype=np.uint64).astype(WORD_NP[T]))


# --------------------------------------------------------------------------
def test_golden_b2sr_layout(orc, gold):
    rp = np.array(gold["csr_rowptr"], np.int32)
    ci = np.array(gold["csr_colind"], np.int32)
    n = gold["n"]
    for T in TS:
        g = gold["tiles"][str(T)]
        brp, bci, val = orc.b2sr(T, n, n, rp, ci)
        assert brp.tolist() == g["rowptr"], T
        assert bci.tolist() == g["colind"], T
        assert val.tolist() == g["val"], T


def test_golden_vector_and_ops(orc, gold):
    rp = np.array(gold["csr_rowptr"], np.int32)
    ci = np.array(gold["csr_colind"], np.int32)
    n = gold["n"]
    xbits = np.array(gold["x_bits"], np.float32)
    xf = np.array(gold["x_float"], np.float32)
    assert orc.csr_count(rp, ci, xbits.astype(np.uint8)).tolist() == gold["counts"]
    np.testing.assert_array_equal(orc.csr_f32(rp, ci, xf), np.array(gold["full_full"], np.float32))
    for T in TS:
        g = gold["tiles"][str(T)]
        nbr = -(-n // T)
        brp, bci, val = orc.b2sr(T, n, n, rp, ci)
        xb = orc.pack_vec(T, xbits, nbr)
        assert xb.tolist() == g["xb"], T
        assert orc.b2sr_bin_bin_full(T, val, xb, brp, bci).tolist() == g["bin_bin_full"]
        assert orc.b2sr_bin_bin_bin(T, val, xb, brp, bci).tolist() == g["bin_bin_bin"]
        mask = np.array(g["mask"], dtype=np.uint64).astype(WORD_NP[T])
        assert orc.b2sr_masked(T, val, xb, brp, bci, mask).tolist() == g["masked"]
        # the mask is the packed form of the visited rows
        mbits = np.zeros(n, np.uint8)
        mbits[gold["mask_rows"]] = 1
        assert orc.pack_u8(T, mbits, nbr).tolist() == g["mask"]
        xpad = np.zeros(nbr * T, np.float32)
        xpad[:n] = xf
        ff = orc.b2sr_bin_full_full(T, val, xpad, brp, bci)
        np.testing.assert_array_equal(ff[:n], np.array(gold["full_full"], np.float32))
        assert not ff[n:].any()


@pytest.mark.parametrize("T", TS)
@pytest.mark.parametrize("shape", [(1, 1), (5, 7), (33, 33), (70, 45), (129, 200)])
def test_b2sr_matches_dense_construction(orc, T, shape):
    rng = np.random.default_rng(shape[0] * 1000 + shape[1] + T)
    for density in (0.0, 0.03, 0.3):
        D, rp, ci = rand_csr(rng, shape[0], shape[1], density)
        brp, bci, val = orc.b2sr(T, shape[0], shape[1], rp, ci)
        erp, eci, ev = dense_tiles(D, T)
        np.testing.assert_array_equal(brp, erp)
        np.testing.assert_array_equal(bci, eci)
        np.testing.assert_array_equal(val, ev)
        # invariants: total popcount == nnz
        assert sum(bin(int(w)).count("1") for w in val) == rp[-1]


@pytest.mark.parametrize("T", TS)
def test_b2sr_ops_match_dense_products(orc, T):
    rng = np.random.default_rng(7 + T)
    for n in (1, 3, T, T + 1, 3 * T + 5, 150):
        D, rp, ci = rand_csr(rng, n, n, 0.1)
        nbr = -(-n // T)
        brp, bci, val = orc.b2sr(T, n, n, rp, ci)
        x01 = (rng.random(n) < 0.4).astype(np.float32)
        xf = rng.random(n).astype(np.float32)
        xb = orc.pack_vec(T, x01, nbr)
        want = D @ x01.astype(np.int64)
        got = orc.b2sr_bin_bin_full(T, val, xb, brp, bci)
        np.testing.assert_array_equal(got[:n], want)
        assert not got[n:].any()
        bits = orc.b2sr_bin_bin_bin(T, val, xb, brp, bci)
        np.testing.assert_array_equal(bits, orc.pack_u8(T, (want > 0).astype(np.uint8), nbr))
        # accumulate semantics: C |= and C +=
        bits2 = orc.b2sr_bin_bin_bin(T, val, xb, brp, bci, C=bits)
        np.testing.assert_array_equal(bits2, bits)
        got2 = orc.b2sr_bin_bin_full(T, val, xb, brp, bci, C=got)
        np.testing.assert_array_equal(got2, 2 * got)
        xpad = np.zeros(nbr * T, np.float32)
        xpad[:n] = xf
        ff = orc.b2sr_bin_full_full(T, val, xpad, brp, bci)
        np.testing.assert_allclose(ff[:n], D @ xf.astype(np.float64), rtol=1e-6, atol=1e-6)
        # B2SR tile loop and CSR loop sum in the same column order: bit-identical
        np.testing.assert_array_equal(ff[:n], orc.csr_f32(rp, ci, xf))
        mbits = (rng.random(n) < 0.5).astype(np.uint8)
        mask = orc.pack_u8(T, mbits, nbr)
        got_m = orc.b2sr_masked(T, val, xb, brp, bci, mask)
        want_m = ((want > 0) & (mbits == 0)).astype(np.uint8)
        np.testing.assert_array_equal(got_m, orc.pack_u8(T, want_m, nbr))


def test_csr_semirings_vs_dense(orc):
    rng = np.random.default_rng(3)
    D, rp, ci = rand_csr(rng, 300, 300, 0.05)
    x01 = (rng.random(300) < 0.3).astype(np.uint8)
    np.testing.assert_array_equal(orc.csr_count(rp, ci, x01), D @ x01.astype(np.int64))
    np.testing.assert_array_equal(orc.csr_or(rp, ci, x01), (D @ x01.astype(np.int64) > 0))
    xd = rng.standard_normal(300)
    np.testing.assert_allclose(orc.csr_f64(rp, ci, xd), D @ xd, rtol=1e-12)
    X = rng.standard_normal((300, 16))
    np.testing.assert_allclose(orc.csr_spmm_f64(rp, ci, X), D @ X, rtol=1e-12, atol=1e-12)


def test_pack_vec_bit_order(orc):
    # x[J*T + c] > 0 lands at bit T-1-c of word J; negative and zero are 0
    for T in TS:
        nbr = 3
        x = np.zeros(nbr * T, np.float32)
        x[0] = 1.0
        x[T + T - 1] = 2.0
        x[2 * T + 1] = -1.0
        xb = orc.pack_vec(T, x, nbr)
        assert int(xb[0]) == 1 << (T - 1)
        assert int(xb[1]) == 1
        assert int(xb[2]) == 0
        assert xb.size == -(-nbr // T) * T


# --------------------------------------------------------------------------
def _edges(rng, n, m, loops=True):
    src = rng.integers(0, n, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    if loops:
        src[: m // 10] = dst[: m // 10]
    return src, dst


def test_normalize_vs_dense(orc):
    rng = np.random.default_rng(11)
    n = 60
    src, dst = _edges(rng, n, 500)
    for undirected in (True, False):
        for drop in (True, False):
            rp, ci = orc.normalize(n, n, src, dst, undirected, drop)
            D = np.zeros((n, n), bool)
            D[src, dst] = True
            if undirected:
                D |= D.T
            if drop:
                np.fill_diagonal(D, False)
            _, erp, eci = rand_csr_from_dense(D)
            np.testing.assert_array_equal(rp, erp)
            np.testing.assert_array_equal(ci, eci)


def rand_csr_from_dense(D):
    rp = np.concatenate([[0], np.cumsum(D.sum(1))]).astype(np.int32)
    return D, rp, np.nonzero(D)[1].astype(np.int32)


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build oracle/_ref not present")
def test_normalize_vs_reference_readmtx(orc):
    """The oracle's normalisation equals the reference's own readMtx.hpp code."""
    ref = RefReadMtx()
    rng = np.random.default_rng(5)
    for n, m in ((10, 40), (200, 3000), (1000, 1000)):
        src, dst = _edges(rng, n, m)
        for undirected in (True, False):
            if trailing_entry_erased(list(zip(src.tolist(), dst.tolist())), undirected):
                # a trailing self-loop or duplicate costs the reference two
                # more entries (test_mtx.py pins that defect)
                keep = (src != n - 1) & (dst != n - 1)
                src, dst = np.append(src[keep], n - 1), np.append(dst[keep], n - 2)
            assert not trailing_entry_erased(list(zip(src.tolist(), dst.tolist())), undirected)
            rp, ci = orc.normalize(n, n, src, dst, undirected, True)
            rrp, rci = ref.normalize(n, n, src, dst, undirected)
            np.testing.assert_array_equal(rp, rrp)
            np.testing.assert_array_equal(ci, rci)


# --------------------------------------------------------------------------
def test_bfs_vs_scipy(orc):
    sp = pytest.importorskip("scipy.sparse")
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    rng = np.random.default_rng(2)
    n = 400
    src, dst = _edges(rng, n, 700, loops=False)
    rp, ci = orc.normalize(n, n, src, dst, True, True)
    A = sp.csr_matrix((np.ones(ci.size), ci, rp), shape=(n, n))
    for s in (0, 17, 399):
        lev, nl = orc.bfs(rp, ci, s)
        dist = csgraph.shortest_path(A, unweighted=True, indices=s)
        want = np.where(np.isinf(dist), -1, dist).astype(np.int32)
        np.testing.assert_array_equal(lev, want)
        assert nl == lev.max() + 1


def test_bfs_multi_vs_scipy(orc):
    """The sources-in-parallel BFS equals scipy's unweighted shortest paths
    from every source (and the one-source BFS), on a graph with isolated
    vertices and two components."""
    sp = pytest.importorskip("scipy.sparse")
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    rng = np.random.default_rng(12)
    n = 600
    src = np.concatenate([rng.integers(0, 250, 500), rng.integers(300, 580, 500)]).astype(np.int32)
    dst = np.concatenate([rng.integers(0, 250, 500), rng.integers(300, 580, 500)]).astype(np.int32)
    rp, ci = orc.normalize(n, n, src, dst, True, True)
    A = sp.csr_matrix((np.ones(ci.size), ci, rp), shape=(n, n))
    srcs = np.array([0, 5, 299, 310, 599, 420, 5], np.int32)
    lev, nl = orc.bfs_multi(rp, ci, srcs)
    dist = csgraph.shortest_path(A, unweighted=True, indices=srcs)
    want = np.where(np.isinf(dist), -1, dist).astype(np.int32)
    np.testing.assert_array_equal(lev, want)
    assert nl == want.max() + 1
    for i, s in enumerate(srcs):
        np.testing.assert_array_equal(orc.bfs(rp, ci, s)[0], lev[i])


def test_pagerank_regular_graph_closed_form(orc):
    # on a regular graph the uniform vector is the fixed point
    rp, ci = orc.normalize(64, 64, np.arange(64), (np.arange(64) + 1) % 64, True, True)
    pr, res = orc.pagerank(rp, ci, 0.85, 20)
    np.testing.assert_allclose(pr, 1 / 64, rtol=1e-14)
    assert res.max() < 1e-15


def test_pagerank_vs_dense_power_iteration(orc):
    rng = np.random.default_rng(9)
    n = 120
    src, dst = _edges(rng, n, 200, loops=False)
    rp, ci = orc.normalize(n, n, src, dst, True, True)
    D = np.zeros((n, n))
    for i in range(n):
        D[i, ci[rp[i]:rp[i + 1]]] = 1
    deg = D.sum(1)
    p = np.full(n, 1 / n)
    for _ in range(30):
        dang = p[deg == 0].sum()
        p = (1 - 0.85) / n + 0.85 * (D.T @ np.where(deg > 0, p / np.maximum(deg, 1), 0) + dang / n)
    pr, _ = orc.pagerank(rp, ci, 0.85, 30)
    np.testing.assert_allclose(pr, p, rtol=1e-12)
    assert abs(pr.sum() - 1) < 1e-12


# -- transposes and tile statistics -------------------------------------------
@pytest.mark.parametrize("shape", [(1, 1), (6, 9), (40, 33), (130, 70)])
def test_csr2csc_vs_dense_transpose(orc, shape):
    rng = np.random.default_rng(sum(shape))
    D, rp, ci = rand_csr(rng, *shape, 0.2)
    vals = rng.standard_normal(ci.size).astype(np.float32)
    W = np.zeros(shape, np.float32)
    W[np.repeat(np.arange(shape[0]), np.diff(rp)), ci] = vals
    cp, ri, cv = orc.csr2csc(shape[0], shape[1], rp, ci, vals)
    Dt = D.T
    assert cp.tolist() == [0] + np.cumsum(Dt.sum(1)).tolist()
    assert ri.tolist() == np.nonzero(Dt)[1].tolist()
    assert cv.tolist() == W.T[np.nonzero(Dt)].tolist()


def test_csr2csc_hand_case(orc):
    # 3 x 4: rows {1, 3}, {}, {0, 1}  ->  columns {2}, {0, 2}, {}, {0}
    cp, ri, _ = orc.csr2csc(3, 4, [0, 2, 2, 4], [1, 3, 0, 1])
    assert cp.tolist() == [0, 1, 3, 3, 4]
    assert ri.tolist() == [2, 0, 2, 0]


@pytest.mark.parametrize("T", TS)
@pytest.mark.parametrize("shape", [(1, 1), (5, 7), (70, 45), (129, 200)])
def test_b2sr_transpose_vs_dense(orc, T, shape):
    rng = np.random.default_rng(T * 1000 + shape[0])
    D, _, _ = rand_csr(rng, *shape, 0.1)
    brp, bci, bval = dense_tiles(D, T)
    nbc = -(-shape[1] // T)
    trp, tci, tval = orc.b2sr_transpose(T, nbc, brp, bci, bval)
    erp, eci, ev = dense_tiles(np.ascontiguousarray(D.T), T)
    assert trp.tolist() == erp.tolist() and tci.tolist() == eci.tolist()
    assert tval.tolist() == ev.tolist()
    # transposing twice gives the original back
    rrp, rci, rval = orc.b2sr_transpose(T, brp.size - 1, trp, tci, tval)
    assert rrp.tolist() == brp.tolist() and rci.tolist() == bci.tolist() and rval.tolist() == bval.tolist()


@pytest.mark.parametrize("T", TS)
def test_tile_row_stats_vs_dense(orc, T):
    rng = np.random.default_rng(T)
    n = 3 * T + 5
    D, _, _ = rand_csr(rng, n, 2 * n, 0.03)
    brp, bci, bval = dense_tiles(D, T)
    tpr, nzr = orc.tile_row_stats(T, brp, bval)
    nbr = brp.size - 1
    P = np.zeros((nbr * T, 2 * n + T), np.int64)
    P[:n, : 2 * n] = D
    for I in range(nbr):
        cols = bci[brp[I]:brp[I + 1]]
        for i in range(T):
            assert tpr[I * T + i] == cols.size
            row = P[I * T + i]
            assert nzr[I * T + i] == sum(int(row[J * T:(J + 1) * T].any()) for J in cols)


@pytest.mark.parametrize("T", TS)
def test_estimate_tiles_vs_numpy(orc, T):
    rng = np.random.default_rng(7 + T)
    D, rp, ci = rand_csr(rng, 150, 300, 0.05)
    want = sum(np.unique(ci[rp[r]:rp[r + 1]] // T).size for r in range(150))
    assert orc.estimate_tiles(rp, ci, T) == want
    # an upper bound on the exact tile count (a block column counted once per row)
    assert want >= dense_tiles(D, T)[1].size
    assert orc.estimate_tiles([0, 3, 3], [0, 1, 9], 4) == 2  # columns 0, 1 share block 0
