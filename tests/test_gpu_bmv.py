"""GPU parity of CSR->B2SR packing and the BMV operators against the CPU oracle.

Bit and count results must match bit-exactly; bin_full_full (FP32) within
1e-6 relative of the FP64 CSR sum (plus an absolute 1e-6 floor for values
near zero, the reference's checkResult tolerance, utility.cu:137-149).
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TS = (4, 8, 16, 32, 64)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "b2sr_5x5.json")
FP_RTOL, FP_ATOL = 1e-6, 1e-6


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


def to_np(t, T=None):
    a = t.cpu().numpy()
    if T is not None:
        from bgb.b2sr import WORD_NUMPY
        a = a.view(WORD_NUMPY[T])
    return a


GUARD_BYTES = 64
GUARD_VALUE = 0xA5


def dirty(n, dtype, dev):
    """An output buffer of n elements with every byte 0xFF (int32 -1, float NaN):
    an overwrite (accumulate=0) must replace every element the op defines."""
    import torch
    t = torch.empty(max(n, 1), dtype=dtype, device=dev)
    t.view(torch.uint8).fill_(0xFF)
    return t[:n]


def guarded_bits(A, dev):
    """(raw bytes, C): C is exactly NBR words of 0xFF bytes, the reference's
    allocation (test_bin_bin_bin.cu:329), followed in the same allocation by
    GUARD_BYTES bytes of GUARD_VALUE that no BMV may change."""
    import torch
    from bgb import b2sr as B
    nb = A.nblockrows * B.lib().bgb_word_bytes(A.T)
    raw = torch.full((nb + GUARD_BYTES,), GUARD_VALUE, dtype=torch.uint8, device=dev)
    raw[:nb] = 0xFF
    return raw, raw[:nb].view(B.WORD_TORCH[A.T])


def check_guard(raw, A):
    from bgb import b2sr as B
    nb = A.nblockrows * B.lib().bgb_word_bytes(A.T)
    tail = raw[nb:].cpu().numpy()
    assert (tail == GUARD_VALUE).all(), "bytes past the bit output changed"


def check_all_ops(orc, csr, T, rng, dev, label, xdensity=0.3):
    """Packs csr on the GPU and checks every operator against the oracle.
    Outputs start dirty (0xFF bytes) and are overwritten (accumulate=0); bit
    outputs are exactly NBR words with a guard region behind them. x covers
    every column (a wide matrix reads more block columns than it has block
    rows)."""
    import torch
    from bgb import b2sr as B
    log(f"[case] {label} T={T} n={csr.nrows}x{csr.ncols} nnz={csr.nnz}")
    n = csr.nrows
    nx = max(csr.nrows, csr.ncols)
    A = B.B2SR.from_csr(csr, T, dev)
    brp, bci, val = orc.b2sr(T, n, csr.ncols, csr.rowptr, csr.colind)
    g_rp, g_ci, g_val = A.to_host()
    np.testing.assert_array_equal(g_rp, brp)
    np.testing.assert_array_equal(g_ci, bci)
    np.testing.assert_array_equal(g_val, val)
    nbr = A.nblockrows
    x01 = (rng.random(nx) < xdensity).astype(np.float32)
    xpad = B.padded_vector(x01, T, dev)
    xb = B.pack_vec(xpad, T)
    want_xb = orc.pack_vec(T, x01, -(-nx // T))
    np.testing.assert_array_equal(to_np(xb, T)[: want_xb.size], want_xb)

    counts = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
    got = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(nbr * T, torch.int32, dev)))
    np.testing.assert_array_equal(got[:n], counts)
    assert not got[n:].any()
    gotf = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(nbr * T, torch.float32, dev), dtype=torch.float32))
    np.testing.assert_array_equal(gotf[:n], counts.astype(np.float32))
    assert not gotf[n:].any()

    want_bits = orc.pack_u8(T, (counts > 0).astype(np.uint8), nbr)
    raw, C = guarded_bits(A, dev)
    bits = to_np(B.bmv_bin_bin_bin(A, xb, out=C), T)
    np.testing.assert_array_equal(bits, want_bits)
    check_guard(raw, A)

    mbits = (rng.random(n) < 0.5).astype(np.uint8)
    mask = torch.from_numpy(orc.pack_u8(T, mbits, nbr).view(to_np(xb).dtype)).to(dev)
    raw, C = guarded_bits(A, dev)
    got_m = to_np(B.bmv_bin_bin_bin_masked(A, xb, mask, out=C), T)
    want_m = orc.pack_u8(T, ((counts > 0) & (mbits == 0)).astype(np.uint8), nbr)
    np.testing.assert_array_equal(got_m, want_m)
    check_guard(raw, A)

    xf = rng.random(nx).astype(np.float32)
    xfp = B.padded_vector(xf, T, dev)
    gff = to_np(B.bmv_bin_full_full(A, xfp, out=dirty(nbr * T, torch.float32, dev)))
    ref64 = orc.csr_f64(csr.rowptr, csr.colind, xf.astype(np.float64))
    np.testing.assert_allclose(gff[:n], ref64, rtol=FP_RTOL, atol=FP_ATOL)
    assert not gff[n:].any()
    xh = xfp.to(torch.bfloat16)
    want_h = orc.csr_f64(csr.rowptr, csr.colind, xh.float().cpu().numpy()[:nx].astype(np.float64))
    gh = to_np(B.bmv_bin_full_full_bf16(A, xh, out=dirty(nbr * T, torch.float32, dev)))
    np.testing.assert_allclose(gh[:n], want_h, rtol=FP_RTOL, atol=FP_ATOL)
    assert not gh[n:].any()
    return A, xb, xfp


def test_golden_fixture(orc, dev):
    import torch
    from bgb import b2sr as B
    from bgb.graph import HostCSR
    with open(GOLDEN) as f:
        gold = json.load(f)
    n = gold["n"]
    csr = HostCSR(n, n, gold["csr_rowptr"], gold["csr_colind"])
    for T in TS:
        log(f"[case] golden T={T}")
        g = gold["tiles"][str(T)]
        A = B.B2SR.from_csr(csr, T, dev)
        rp, ci, val = A.to_host()
        assert rp.tolist() == g["rowptr"] and ci.tolist() == g["colind"]
        assert val.tolist() == g["val"]
        xb = B.pack_vec(B.padded_vector(np.array(gold["x_bits"], np.float32), T, dev), T)
        assert to_np(xb, T).tolist() == g["xb"]
        assert to_np(B.bmv_bin_bin_full(A, xb)).tolist() == g["bin_bin_full"]
        assert to_np(B.bmv_bin_bin_bin(A, xb), T).tolist() == g["bin_bin_bin"]
        m = np.array(g["mask"], dtype=np.uint64).astype(B.WORD_NUMPY[T]).view(to_np(xb).dtype)
        got_m = B.bmv_bin_bin_bin_masked(A, xb, torch.from_numpy(m).to(dev))
        assert to_np(got_m, T).tolist() == g["masked"]
        ff = to_np(B.bmv_bin_full_full(A, B.padded_vector(np.array(gold["x_float"], np.float32), T, dev)))
        assert ff[:n].tolist() == gold["full_full"]


@pytest.mark.parametrize("T", TS)
def test_random_and_ragged(orc, dev, T):
    from bgb import graph
    rng = np.random.default_rng(100 + T)
    cases = [(1, 0), (2, 1), (T - 1, 3), (T + 1, 10), (5 * T + 3, 200), (1000, 3000), (4099, 20000)]
    for n, m in cases:
        src = rng.integers(0, n, m).astype(np.int32)
        dst = rng.integers(0, n, m).astype(np.int32)
        for undirected in (True, False):
            csr = graph.from_coo(n, n, src, dst, undirected, True)
            check_all_ops(orc, csr, T, rng, dev, f"random n={n} m={m} und={undirected}")


@pytest.mark.parametrize("T", TS)
def test_empty_rows_and_hubs(orc, dev, T):
    """Long runs of empty block rows and rows spanning many chunks."""
    from bgb import graph
    rng = np.random.default_rng(7 * T)
    n = 20000
    # star: vertex 0 adjacent to every vertex -> block row 0 spans many chunks
    src = np.zeros(n - 1, np.int32)
    dst = np.arange(1, n, dtype=np.int32)
    # isolated stretch [5000, 15000) and a few random edges elsewhere
    extra_s = rng.integers(15000, n, 3000).astype(np.int32)
    extra_d = rng.integers(15000, n, 3000).astype(np.int32)
    s = np.concatenate([src[:4999], src[15000:], extra_s])
    d = np.concatenate([dst[:4999], dst[15000:], extra_d])
    csr = graph.from_coo(n, n, s, d, True, True)
    check_all_ops(orc, csr, T, rng, dev, "star+holes", xdensity=0.5)
    # all-empty matrix
    csr0 = graph.from_coo(100, 100, np.zeros(0, np.int32), np.zeros(0, np.int32))
    check_all_ops(orc, csr0, T, rng, dev, "empty")
    # dense block (every tile full)
    k = 3 * T
    ii, jj = np.meshgrid(np.arange(k), np.arange(k))
    csr_d = graph.from_coo(k, k, ii.ravel(), jj.ravel(), False, False)
    check_all_ops(orc, csr_d, T, rng, dev, "dense")


@pytest.mark.parametrize("T", TS)
def test_accumulate_semantics(orc, dev, T):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(T)
    csr = graph.rmat(12, 8, seed=5)
    log(f"[case] accumulate T={T}")
    A = B.B2SR.from_csr(csr, T, dev)
    x01 = (rng.random(csr.nrows) < 0.2).astype(np.float32)
    xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
    counts = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
    C = B.bmv_bin_bin_full(A, xb)
    for _ in range(4):  # the drivers' 5 accumulating launches (test_bin_bin_full.cu:335-346)
        B.bmv_bin_bin_full(A, xb, out=C, accumulate=True)
    np.testing.assert_array_equal(to_np(C)[: csr.nrows], 5 * counts)
    Cb = torch.zeros(A.nblockrows, dtype=B.WORD_TORCH[T], device=dev)
    for _ in range(5):
        B.bmv_bin_bin_bin(A, xb, out=Cb, accumulate=True)
    np.testing.assert_array_equal(to_np(Cb, T), orc.pack_u8(T, (counts > 0).astype(np.uint8), A.nblockrows))
    xf = rng.random(csr.nrows).astype(np.float32)
    xfp = B.padded_vector(xf, T, dev)
    Cf = B.bmv_bin_full_full(A, xfp)
    B.bmv_bin_full_full(A, xfp, out=Cf, accumulate=True)
    ref = 2 * orc.csr_f64(csr.rowptr, csr.colind, xf.astype(np.float64))
    np.testing.assert_allclose(to_np(Cf)[: csr.nrows], ref, rtol=FP_RTOL, atol=FP_ATOL)


def test_config_er4096_bin_bin_bin(orc, dev):
    """ER n=4096, average degree 16, seed 1, T=32, x density 0.1: bit-exact."""
    from bgb import graph
    csr = graph.erdos_renyi(4096, 16, seed=1)
    assert csr.nnz == 65536
    check_all_ops(orc, csr, 32, np.random.default_rng(1), dev, "ER4096", xdensity=0.1)


@pytest.mark.parametrize("T", TS)
def test_config_rmat20_bin_bin_full(orc, dev, T):
    """R-MAT scale 20, edge factor 16, symmetrised: exact int32 counts at every T."""
    from bgb import b2sr as B
    from bgb import graph
    csr = _rmat20()
    log(f"[case] rmat20 T={T} nnz={csr.nnz}")
    A = B.B2SR.from_csr(csr, T, dev)
    x01 = graph.bernoulli(csr.nrows, 0.05, 2)
    xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
    counts = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
    got = to_np(B.bmv_bin_bin_full(A, xb))
    np.testing.assert_array_equal(got[: csr.nrows], counts)
    # invariant: total popcount of the packed tiles equals nnz
    _, _, val = A.to_host()
    assert int(np.unpackbits(val.view(np.uint8)).sum()) == csr.nnz


_RMAT20 = None


def _rmat20():
    global _RMAT20
    if _RMAT20 is None:
        from bgb import graph
        _RMAT20 = graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=1)
    return _RMAT20


@pytest.mark.parametrize("T", (4, 8))
def test_row_pieces_modes(orc, dev, T):
    """The T <= 8 count BMVs cut long block rows into 256-tile pieces aligned to
    the row's first 16-byte unit (one wave each, the last piece writing the row,
    the others leaving carry records) and give short rows a lane each. The
    piece path and the chunk walk must both give the oracle's counts: rows of
    31..34 tiles (the short/long threshold), 255..257, 511..513 and 1025
    tiles (piece edges), row starts at every alignment, an R-MAT graph, accumulate=1."""
    import torch
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(70 + T)
    n = 12000 * T // 4 + 3
    lens = [31, 32, 33, 34, 1, 255, 256, 257, 511, 512, 513, 2, 1025, 3, 700, 0, 5, 64, 65, 40]
    src, dst = [], []
    for i, k in enumerate(lens):
        cols = rng.choice(n // T, k, replace=False) * T + rng.integers(0, T, k)
        src.append(np.full(k, (i * 3 + 1) * T + i % T))
        dst.append(cols)
    src.append(rng.integers(0, n, 20000))
    dst.append(rng.integers(0, n, 20000))
    cases = (("piece edges", graph.from_coo(n, n, np.concatenate(src), np.concatenate(dst), False, False)),
             ("rmat14", graph.rmat(14, 8, seed=31)))
    prev = bgb.lib().bgb_get_row_pieces()
    try:
        for mode in (2, 1):
            bgb.call("bgb_set_row_pieces", mode)
            for label, csr in cases:
                log(f"[case] row pieces mode {mode} {label} T={T}")
                A = B.B2SR.from_csr(csr, T, dev)
                x01 = (rng.random(csr.nrows) < 0.4).astype(np.float32)
                xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
                want = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
                nbr = A.nblockrows
                got = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(nbr * T, torch.int32, dev)))
                np.testing.assert_array_equal(got[: csr.nrows], want)
                assert not got[csr.nrows:].any()
                gotf = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(nbr * T, torch.float32, dev), dtype=torch.float32))
                np.testing.assert_array_equal(gotf[: csr.nrows], want.astype(np.float32))
                twice = B.bmv_bin_bin_full(A, xb, out=torch.from_numpy(got).to(dev), accumulate=True)
                np.testing.assert_array_equal(to_np(twice)[: csr.nrows], 2 * want)
    finally:
        bgb.call("bgb_set_row_pieces", prev)


@pytest.mark.timeout(600)
def test_row_pieces_x_placements(orc, dev):
    """The piece kernel's other packed-x placements, pieces forced: x read
    from memory (x staging off, and a packed x at an address that is not
    16-byte aligned), and its first words staged with the rest read from
    memory (T = 4, 6 M vertices: x of 1.5 M words is too large even for the
    non-zero bitmap). Counts against the oracle, dirty output and accumulate."""
    import torch
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(140)

    def check(csr, T, label, misalign=False):
        log(f"[case] x placement: {label} T={T}")
        A = B.B2SR.from_csr(csr, T, dev)
        x01 = (rng.random(csr.ncols) < 0.05).astype(np.float32)
        xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
        if misalign:  # the same words one byte into a larger buffer
            buf = torch.zeros(xb.numel() + 16, dtype=torch.uint8, device=dev)
            buf[1 : 1 + xb.numel()] = xb
            xb = buf[1 : 1 + xb.numel()]
        want = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
        got = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(A.nblockrows * T, torch.int32, dev)))
        np.testing.assert_array_equal(got[: csr.nrows], want)
        twice = B.bmv_bin_bin_full(A, xb, out=torch.from_numpy(got).to(dev), accumulate=True)
        np.testing.assert_array_equal(to_np(twice)[: csr.nrows], 2 * want)

    prev = bgb.lib().bgb_get_row_pieces()
    try:
        bgb.call("bgb_set_row_pieces", 2)
        small = graph.rmat(16, 16, seed=41)
        for T in (4, 8):
            bgb.call("bgb_set_x_staging", T, 1)
            check(small, T, "staging off (x from memory)")
            bgb.call("bgb_set_x_staging", T, 0)
            check(small, T, "x not 16-byte aligned (x from memory)", misalign=True)
        n, m = 6000000, 6000000
        src = rng.integers(0, n, m)
        dst = np.where(rng.random(m) < 0.6, rng.integers(0, n // 8, m), rng.integers(0, n, m))
        hubs = np.repeat(np.array([3, n // 3, n - 5]), 40000)
        big = graph.from_coo(n, n, np.concatenate([src, hubs]).astype(np.int32),
                             np.concatenate([dst, rng.integers(0, n, hubs.size)]).astype(np.int32), False, True)
        check(big, 4, "first words staged, the rest from memory (6 M vertices)")
    finally:
        bgb.call("bgb_set_row_pieces", prev)
        for T in (4, 8):
            bgb.call("bgb_set_x_staging", T, 0)


def test_row_piece_plans_evicted_and_remade(orc, dev):
    """Piece plans are cached per (device, rowptr, NBR, nblocks, T), eight at a
    time: ten matrices in turn (T = 4 and 8, pieces forced) evict the oldest
    plans, and the second pass over all of them remakes those and reuses the
    rest, every result equal to the oracle's counts."""
    import torch
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(123)
    mats = []
    for i in range(10):
        T = 4 if i % 2 == 0 else 8
        csr = graph.rmat(11 + i % 3, 8 + i, seed=200 + i)
        x01 = (rng.random(csr.nrows) < 0.3).astype(np.float32)
        A = B.B2SR.from_csr(csr, T, dev)
        xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
        mats.append((csr, A, xb, orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))))
    prev = bgb.lib().bgb_get_row_pieces()
    try:
        bgb.call("bgb_set_row_pieces", 2)
        for rnd in range(2):
            for i, (csr, A, xb, want) in enumerate(mats):
                log(f"[case] piece plan cache round {rnd} matrix {i} T={A.T}")
                got = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(A.nblockrows * A.T, torch.int32, dev)))
                np.testing.assert_array_equal(got[: csr.nrows], want)
    finally:
        bgb.call("bgb_set_row_pieces", prev)


@pytest.mark.parametrize("T", (4, 8))
def test_row_pieces_bitmap_x(orc, dev, T):
    """A packed x larger than LDS on the row-piece path: LDS holds the first
    words of x and a bitmap of which of the others are non-zero; only those
    are read from memory. 3 M vertices (x: 375 KB at T = 4 and 8), columns
    skewed to low ids plus uniform ones, hub rows of up to 40 000 tiles, a
    square and a wide matrix; x of density 0.05, all ones (every unstaged
    word read) and all zeros (none read); overwrite on a dirty output and
    accumulate; pieces forced and automatic, against the oracle."""
    import torch
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(110 + T)
    n = 3000000
    m = 4000000
    src = rng.integers(0, n, m)
    dst = np.where(rng.random(m) < 0.6, rng.integers(0, n // 6, m), rng.integers(0, n, m))
    hubs = np.repeat(np.array([5, n // 2 + 1, n - 2]), 60000)
    far = rng.integers(0, n, hubs.size)
    src = np.concatenate([src, hubs]).astype(np.int32)
    dst = np.concatenate([dst, far]).astype(np.int32)
    cases = (("square", graph.from_coo(n, n, src, dst, False, True)),
             ("wide 1000000 x 3000000", graph.from_coo(1000000, n, src % 1000000, dst, False, True)))
    prev = bgb.lib().bgb_get_row_pieces()
    try:
        for mode in (2, 0):
            bgb.call("bgb_set_row_pieces", mode)
            for label, csr in cases:
                A = B.B2SR.from_csr(csr, T, dev)
                nbr = A.nblockrows
                for xlabel, x01 in (("density 0.05", (rng.random(csr.ncols) < 0.05).astype(np.float32)),
                                    ("all ones", np.ones(csr.ncols, np.float32)),
                                    ("all zeros", np.zeros(csr.ncols, np.float32))):
                    log(f"[case] bitmap x, pieces mode {mode}, {label}, x {xlabel}, T={T}")
                    xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
                    want = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
                    got = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(nbr * T, torch.int32, dev)))
                    np.testing.assert_array_equal(got[: csr.nrows], want)
                    assert not got[csr.nrows:].any()
                    twice = B.bmv_bin_bin_full(A, xb, out=torch.from_numpy(got).to(dev), accumulate=True)
                    np.testing.assert_array_equal(to_np(twice)[: csr.nrows], 2 * want)
    finally:
        bgb.call("bgb_set_row_pieces", prev)


@pytest.mark.parametrize("T", TS)
def test_rowptr_rewritten_in_place(orc, dev, T):
    """The chunk partition is cached per (stream, rowptr address, sizes): a caller
    that rewrites rowptr/colind/val in place with another matrix of the same
    NBR and nblocks (here: block rows permuted) must still get that matrix's
    results, on the first call after the rewrite and on the next."""
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(50 + T)
    csr1 = graph.rmat(13, 8, seed=8)
    n = csr1.nrows
    nbr = -(-n // T)
    perm = rng.permutation(nbr)  # block row I -> perm[I]
    rows = np.repeat(np.arange(n), np.diff(csr1.rowptr))
    new_rows = perm[rows // T] * T + rows % T
    keep = new_rows < n
    csr2 = graph.from_coo(n, n, new_rows[keep], csr1.colind[keep], False, False)
    A1 = B.B2SR.from_csr(csr1, T, dev)
    A2 = B.B2SR.from_csr(csr2, T, dev)
    if A2.nblocks != A1.nblocks:
        pytest.skip("permutation changed the tile count (a partial last block row moved)")
    x01 = (rng.random(n) < 0.3).astype(np.float32)
    xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
    want1 = orc.csr_count(csr1.rowptr, csr1.colind, x01.astype(np.uint8))
    want2 = orc.csr_count(csr2.rowptr, csr2.colind, x01.astype(np.uint8))
    log(f"[case] rowptr rewrite T={T} nblocks={A1.nblocks}")
    np.testing.assert_array_equal(to_np(B.bmv_bin_bin_full(A1, xb))[:n], want1)
    A1.rowptr.copy_(A2.rowptr)
    A1.colind.copy_(A2.colind)
    A1.val.copy_(A2.val)
    for _ in range(2):
        np.testing.assert_array_equal(to_np(B.bmv_bin_bin_full(A1, xb))[:n], want2)


@pytest.mark.parametrize("T", (4, 8))
def test_short_row_path_modes(orc, dev, T):
    """T <= 8 BMVs run the chunked kernels or one lane per block row (taken
    automatically when no block row holds more than 32 tiles). Both must give
    the oracle's results on short rows (a grid), on rows far longer than one
    batch of tiles (R-MAT hubs, forced onto the short-row path), on a ragged n
    and on runs of empty block rows; accumulate=1 adds onto the output."""
    import torch
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(90 + T)
    src = np.concatenate([rng.integers(0, 100, 400), rng.integers(200, 300, 400)])
    dst = np.concatenate([rng.integers(0, 100, 400), rng.integers(200, 300, 400)])
    star = graph.from_coo(30000, 30000, np.zeros(29999, np.int32), np.arange(1, 30000, dtype=np.int32))
    cases = (("grid 70x33", graph.grid2d(70, 33)), ("rmat14", graph.rmat(14, 8, seed=12)),
             ("ragged n=37", graph.erdos_renyi(37, 6, seed=4)),
             ("empty block rows", graph.from_coo(300, 300, src, dst)),
             # a 30000-entry row walked by one lane when forced: its FP32 sum is
             # folded per batch / 16 batches / row, not one chain of 30000 adds
             ("star n=30000", star))
    try:
        for mode in (1, 2):
            bgb.call("bgb_set_short_row_path", mode)
            for label, csr in cases:
                A, xb, xfp = check_all_ops(orc, csr, T, rng, dev, f"short rows mode {mode} {label}")
                n = csr.nrows
                once = B.bmv_bin_bin_full(A, xb)
                twice = B.bmv_bin_bin_full(A, xb, out=once.clone(), accumulate=True)
                np.testing.assert_array_equal(to_np(twice), 2 * to_np(once))
                bits = B.bmv_bin_bin_bin(A, xb)
                np.testing.assert_array_equal(to_np(B.bmv_bin_bin_bin(A, xb, out=bits.clone(), accumulate=True), T),
                                              to_np(bits, T))
                f1 = B.bmv_bin_full_full(A, xfp)
                f2 = B.bmv_bin_full_full(A, xfp, out=f1.clone(), accumulate=True)
                np.testing.assert_allclose(to_np(f2), 2 * to_np(f1), rtol=1e-6, atol=1e-6)
                xh = xfp.to(torch.bfloat16)
                want = orc.csr_f64(csr.rowptr, csr.colind, xh.float().cpu().numpy()[:n].astype(np.float64))
                got = to_np(B.bmv_bin_full_full_bf16(A, xh))[:n]
                np.testing.assert_allclose(got, want, rtol=FP_RTOL, atol=FP_ATOL)
    finally:
        bgb.call("bgb_set_short_row_path", 0)
    # automatic mode: the batch of tiles per load round follows the longest
    # block row (4, 5, 6 or 8 tiles, several batches past 8, the chunked
    # kernels past 32): block rows of exactly k tiles, k columns drawn per row
    nbr = 300
    for k in (1, 3, 5, 6, 7, 9, 17, 32, 33):
        rows, cols = [], []
        for I in range(nbr):
            J = rng.choice(nbr, k, replace=False)
            for j in J:
                rows.append(I * T + rng.integers(0, T, 2))
                cols.append(j * T + rng.integers(0, T, 2))
        csr = graph.from_coo(nbr * T, nbr * T, np.concatenate(rows), np.concatenate(cols), False, False)
        check_all_ops(orc, csr, T, rng, dev, f"auto path, {k} tiles per block row")


@pytest.mark.parametrize("T", TS)
def test_x_staging_modes(orc, dev, T):
    """The bit BMVs read the packed x either from memory or from a copy staged
    in LDS by persistent workgroups (T=4 packs two nibbles per LDS byte): both
    paths must give the oracle's results, including packed-x lengths that are
    not a multiple of 16 bytes (the staging copies a tail), and chunks that
    lie inside one block row (a star's hub row spans many chunks; such chunks
    publish a wave sum as their carry)."""
    import bgb
    from bgb import graph
    rng = np.random.default_rng(70 + T)
    n = 40000
    star = graph.from_coo(n, n, np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32), True, True)
    try:
        for mode in (1, 2):
            bgb.call("bgb_set_x_staging", T, mode)
            for label, csr in (("rmat14", graph.rmat(14, 8, seed=12)),
                               ("ragged n=1001", graph.erdos_renyi(1001, 12, seed=3)),
                               ("ragged n=37", graph.erdos_renyi(37, 6, seed=4)),
                               ("star n=40000", star)):
                check_all_ops(orc, csr, T, rng, dev, f"x staging mode {mode} {label}")
    finally:
        bgb.call("bgb_set_x_staging", T, 0)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("T", (4, 8, 16))
def test_partial_x_staging(orc, dev, T):
    """A packed x larger than LDS (1.5 M vertices): the LDS kernels stage its
    first words and read the rest from memory. Columns on both sides of the
    staged prefix -- an R-MAT-like skew toward low ids plus uniform ones, a
    hub row and a hub column, and a wide matrix whose extra block columns lie
    past the square part -- against the oracle, staging forced and automatic."""
    import bgb
    from bgb import graph
    rng = np.random.default_rng(90 + T)
    n = 1500000
    m = 3000000
    src = rng.integers(0, n, m)
    dst = np.where(rng.random(m) < 0.6, rng.integers(0, n // 5, m), rng.integers(0, n, m))
    hub = np.full(20000, 7)
    far = rng.integers(0, n, 20000)
    src = np.concatenate([src, hub, far]).astype(np.int32)
    dst = np.concatenate([dst, far, np.full(20000, n - 3)]).astype(np.int32)
    sq = graph.from_coo(n, n, src, dst, False, True)
    wide = graph.from_coo(400000, n, src % 400000, dst, False, True)
    try:
        for mode in (2, 0):
            bgb.call("bgb_set_x_staging", T, mode)
            for label, csr in (("square", sq), ("wide 400000 x 1500000", wide)):
                check_all_ops(orc, csr, T, rng, dev, f"partial x staging, mode {mode}, {label}")
    finally:
        bgb.call("bgb_set_x_staging", T, 0)
