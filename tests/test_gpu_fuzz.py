"""Seeded randomized parity against the oracle: every BMV operator across the
kernel paths a matrix can take (chunked walk with general, one-row and
two-row chunks; short-row kernel; row pieces; packed x in LDS or gathered
from memory),
the batched BFS, and the K=16 MFMA path single and batched.

Each case draws a graph shape (hub rows spanning many chunks, runs of empty
block rows, ragged n, non-square, near-regular rows), a tile dim and the
knob settings from a fixed seed, then checks counts (int32 and float),
bits, masked bits and FP32 sums exactly / within 1e-6 relative, and that
accumulate=1 adds onto the output. Bit and integer results must be exact.
"""
import numpy as np
import pytest

from test_gpu_bmv import check_all_ops, dirty, to_np

pytestmark = pytest.mark.gpu

TS = (4, 8, 16, 32, 64)


def random_graph(rng):
    from bgb import graph
    kind = rng.integers(0, 5)
    n = int(rng.integers(20, 30000))
    if kind == 0:  # hubs: a few vertices adjacent to a large share of the rest
        hubs = rng.integers(0, n, int(rng.integers(1, 4)))
        src = np.concatenate([np.full(n // 2, h) for h in hubs] + [rng.integers(0, n, 2 * n)])
        dst = np.concatenate([rng.integers(0, n, n // 2) for _ in hubs] + [rng.integers(0, n, 2 * n)])
        return "hubs", graph.from_coo(n, n, src, dst, True, True)
    if kind == 1:  # empty stretches between two random blocks
        a, b = sorted(rng.integers(0, n, 2))
        m = 3 * n
        src = np.concatenate([rng.integers(0, max(a, 1), m), rng.integers(b, n, m)])
        dst = np.concatenate([rng.integers(0, max(a, 1), m), rng.integers(b, n, m)])
        return "holes", graph.from_coo(n, n, src, dst, True, True)
    if kind == 2:  # non-square, tall or wide (x covers every column), directed
        m = int(rng.integers(min(20, n), 2 * n + 1))
        e = int(rng.integers(1, 12)) * max(n, m)
        return "rect", graph.from_coo(n, m, rng.integers(0, n, e), rng.integers(0, m, e), False, False)
    if kind == 3:  # near-regular short rows (a grid)
        w = int(rng.integers(2, 300))
        return "grid", graph.grid2d(w, max(2, n // w))
    scale = int(rng.integers(8, 15))
    return "rmat", graph.rmat(scale, int(rng.integers(2, 17)), seed=int(rng.integers(1, 1000)))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("seed", range(20))
def test_random_paths(orc, cuda, seed):
    import bgb
    import bgb.b2sr  # noqa: F401
    from bgb import b2sr as B
    rng = np.random.default_rng(1000 + seed)
    prng = np.random.default_rng(5000 + seed)  # row-piece mode (a stream of its own)
    try:
        for case in range(6):
            label, csr = random_graph(rng)
            T = int(rng.choice(TS))
            staging = int(rng.integers(0, 3))
            short = int(rng.integers(0, 3)) if T <= 8 else 0
            pieces = int(prng.integers(0, 3))
            bgb.call("bgb_set_x_staging", T, staging)
            bgb.call("bgb_set_short_row_path", short)
            bgb.call("bgb_set_row_pieces", pieces)
            A, xb, xfp = check_all_ops(orc, csr, T, rng, cuda,
                                       f"seed {seed}.{case} {label} staging={staging} short={short} pieces={pieces}",
                                       xdensity=float(rng.uniform(0.02, 0.9)))
            once = B.bmv_bin_bin_full(A, xb)
            twice = B.bmv_bin_bin_full(A, xb, out=once.clone(), accumulate=True)
            np.testing.assert_array_equal(to_np(twice), 2 * to_np(once))
            bgb.call("bgb_set_x_staging", T, 0)
    finally:
        for T in TS:
            bgb.call("bgb_set_x_staging", T, 0)
        bgb.call("bgb_set_short_row_path", 0)
        bgb.call("bgb_set_row_pieces", 0)


@pytest.mark.timeout(600)
def test_random_bfs(orc, cuda):
    """Batched BFS over random symmetric graphs: tile dim, source count (1-64),
    hipGraph or eager level loop, pull-only or direction switching; level
    arrays identical to the oracle's queue BFS from every source."""
    from bgb import b2sr as B
    done = 0
    for seed in range(7000, 7040):
        rng = np.random.default_rng(seed)
        label, csr = random_graph(rng)
        if csr.nrows != csr.ncols:  # symmetric shapes only (push = A)
            continue
        T = int(rng.choice((4, 8, 16, 32)))
        A = B.B2SR.from_csr(csr, T, cuda)
        src = rng.integers(0, csr.nrows, int(rng.integers(1, 65)))
        use_graph = bool(rng.integers(0, 2))
        push = A if rng.integers(0, 2) else None
        lev, nl = B.bfs_multi(A, src, use_graph, push=push)
        want = np.stack([orc.bfs(csr.rowptr, csr.colind, int(s))[0] for s in src])
        np.testing.assert_array_equal(lev.cpu().numpy(), want, err_msg=f"seed {seed} {label} T={T}")
        assert nl == want.max() + 1
        done += 1
    assert done >= 25


@pytest.mark.timeout(600)
def test_random_k16(orc, cuda):
    """K=16 MFMA bin_full_full over random graphs at T = 16 / 32, one launch per
    graph and all of them in one persistent batch: within the FP32-accumulation
    bound of the FP64 SpMM (1e-6 of sum |A||X| per entry, plus 1e-7)."""
    import torch
    from bgb import b2sr as B

    def check(csr, Y, X64, label):
        want = orc.csr_spmm_f64(csr.rowptr, csr.colind, X64)
        bound = orc.csr_spmm_f64(csr.rowptr, csr.colind, np.abs(X64))
        err = np.abs(Y[: csr.nrows].cpu().numpy().astype(np.float64) - want)
        assert np.all(err <= 1e-6 * bound + 1e-7), label

    mats, xbs, refs = [], [], []
    for seed in range(8000, 8030):
        rng = np.random.default_rng(seed)
        label, csr = random_graph(rng)
        if csr.nrows != csr.ncols:
            continue
        T = int(rng.choice((16, 32)))
        A = B.B2SR.from_csr(csr, T, cuda)
        X = torch.from_numpy(rng.uniform(-1, 1, (csr.nrows, 16)).astype(np.float16)).to(cuda)
        X64 = X.cpu().numpy().astype(np.float64)
        Xb = B.k16_pack_rhs(X, A.nblockrows * T)
        Y = B.bmv_k16(A, Xb, out=dirty(A.nblockrows * T * 16, torch.float32, cuda))
        check(csr, Y, X64, f"seed {seed} {label} T={T}")
        mats.append(A)
        xbs.append(Xb)
        refs.append((csr, X64))
    assert len(refs) >= 20
    batch = B.K16Batch(mats, xbs, nworkgroups=96)
    batch.y.view(torch.uint8).fill_(0xFF)
    batch.run()
    for i, (csr, X64) in enumerate(refs):
        check(csr, batch.result(i), X64, f"batch graph {i}")
