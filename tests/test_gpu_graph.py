"""GPU parity of the BFS drivers and PageRank against the CPU oracle.

BFS levels must be identical to the oracle's queue BFS for every source.
PageRank is compared with the oracle's FP64 power iteration; the tolerances
(max relative error per vertex) are stated per precision mode below.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TS = (4, 8, 16, 32)
# PageRank tolerances vs the FP64 oracle: FP32 storage gives ~1e-7 per step,
# damped by 1/(1-d); bf16 x carries a 2^-9 relative rounding per element.
PR_RTOL = {"fp32": 2e-5, "fp64sum": 2e-5, "bf16": 2e-2}
# total rank mass |sum(pr) - 1|: bf16 rounding of x biases the mass by up to ~2^-9
PR_MASS = {"fp32": 1e-4, "fp64sum": 1e-4, "bf16": 5e-3}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


def oracle_levels(orc, csr, sources):
    return np.stack([orc.bfs(csr.rowptr, csr.colind, s)[0] for s in sources])


@pytest.mark.parametrize("T", TS)
def test_bfs_multi_random(orc, dev, T):
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(T)
    # two ER components plus isolated vertices: unreached pairs stay -1
    n = 3001
    s1 = rng.integers(0, 1400, 3000)
    d1 = rng.integers(0, 1400, 3000)
    s2 = rng.integers(1500, 2900, 2500)
    d2 = rng.integers(1500, 2900, 2500)
    csr = graph.from_coo(n, n, np.concatenate([s1, s2]), np.concatenate([d1, d2]))
    A = B.B2SR.from_csr(csr, T, dev)
    for nsrc, use_graph, push in ((64, True, None), (64, False, None), (5, True, None), (1, False, None),
                                  (64, True, A), (64, False, A), (3, True, A)):
        log(f"[case] bfs_multi T={T} nsrc={nsrc} graph={use_graph} push={push is not None}")
        src = rng.integers(0, n, nsrc)
        src[0] = 2950  # an isolated vertex
        lev, nl = B.bfs_multi(A, src, use_graph, push=push)
        want = oracle_levels(orc, csr, src)
        np.testing.assert_array_equal(lev.cpu().numpy(), want)
        assert nl == want.max() + 1


@pytest.mark.parametrize("T", TS)
def test_bfs_plain_matches_oracle(orc, dev, T):
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(13, 8, seed=4)
    A = B.B2SR.from_csr(csr, T, dev)
    for s in (0, 7, 4000):
        log(f"[case] bfs_plain T={T} src={s}")
        lev, nl = B.bfs_plain(A, s)
        want, wl = orc.bfs(csr.rowptr, csr.colind, s)
        np.testing.assert_array_equal(lev.cpu().numpy(), want)
        assert nl == wl


@pytest.mark.parametrize("T", (4, 8))
@pytest.mark.parametrize("mode", (0, 1))
def test_bfs_gather_modes(orc, dev, T, mode):
    """Both frontier-gather walks of the pull levels (per set bit of a row, per
    set column of a tile with the group's first columns gathered ahead) give the
    oracle's levels: hub rows that saturate early, a hub in the last block row,
    isolated vertices and n mod T != 0."""
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(40 + T + mode)
    n = 20003
    src_e = rng.integers(0, n, 60000)
    dst_e = rng.integers(0, n, 60000)
    hub = np.concatenate([np.zeros(4000, np.int64), np.full(3000, n - 2)])
    nb = rng.integers(0, n - 50, hub.size)
    csr = graph.from_coo(n, n, np.concatenate([src_e, hub]), np.concatenate([dst_e, nb]))
    A = B.B2SR.from_csr(csr, T, dev)
    src = rng.integers(0, n, 64)
    src[:2] = (0, n - 2)
    want = oracle_levels(orc, csr, src)
    prev = bgb.lib().bgb_get_bfs_gather_mode()
    bgb.call("bgb_set_bfs_gather_mode", mode)
    try:
        for push in (None, A):
            log(f"[case] gather mode {mode} T={T} push={push is not None}")
            lev, nl = B.bfs_multi(A, src, push=push)
            np.testing.assert_array_equal(lev.cpu().numpy(), want)
            assert nl == want.max() + 1
    finally:
        bgb.call("bgb_set_bfs_gather_mode", prev)


def test_bfs_rmat_batched_equals_plain_and_oracle(orc, dev):
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(16, 16, seed=9)
    deg = csr.degrees()
    rng = np.random.default_rng(3)
    src = rng.choice(np.nonzero(deg)[0], 64, replace=False)
    want = oracle_levels(orc, csr, src)
    for T in TS:
        log(f"[case] rmat16 batched T={T}")
        A = B.B2SR.from_csr(csr, T, dev)
        lev, nl = B.bfs_multi(A, src)
        np.testing.assert_array_equal(lev.cpu().numpy(), want)
        # direction switching (symmetric: A is its own transpose) gives the same levels
        levd, nld = B.bfs_multi(A, src, push=A)
        np.testing.assert_array_equal(levd.cpu().numpy(), want)
        assert nld == nl
        if T == 8:
            for i in (0, 31, 63):
                pl, _ = B.bfs_plain(A, int(src[i]))
                np.testing.assert_array_equal(pl.cpu().numpy(), want[i])


def pr_errors(pr, ref):
    return np.max(np.abs(pr.astype(np.float64) - ref) / ref)


@pytest.mark.parametrize("T", TS + (64,))
def test_pagerank_grid_modes(orc, dev, T):
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.grid2d(97, 61)
    ref, rres = orc.pagerank(csr.rowptr, csr.colind, 0.85, 50)
    A = B.B2SR.from_csr(csr, T, dev)
    for mode in ("fp32", "bf16", "fp64sum"):
        log(f"[case] pagerank grid T={T} mode={mode}")
        pr, res = B.pagerank(A, 0.85, 50, mode)
        err = pr_errors(pr.cpu().numpy(), ref)
        assert err < PR_RTOL[mode], (mode, err)
        assert abs(float(pr.double().sum()) - 1.0) < PR_MASS[mode]
        # residual history follows the FP64 run (same decay); bf16 quantisation
        # noise (~2^-9 of each value) dominates its residuals, so only the
        # FP32 / FP64-sum modes are compared
        if mode != "bf16":
            np.testing.assert_allclose(res[:10], rres[:10], rtol=1e-3)


def test_pagerank_grid_stride(orc, dev):
    """More vertices than one pass of the capped update grid covers (1024
    workgroups x 256), R-MAT with isolated vertices so the dangling sum is
    non-zero: ranks and residuals match the oracle, and a second run matches
    the first to FP32 rounding (the partial sums are folded in index order;
    the BMV adds the FP32 carries of rows spanning more than 64 chunks with
    atomics, whose order can move a sum by an ulp)."""
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(19, 8, seed=9)
    assert csr.nrows > 1024 * 256 and (csr.degrees() == 0).any()
    ref, rres = orc.pagerank(csr.rowptr, csr.colind, 0.85, 20)
    A = B.B2SR.from_csr(csr, 8, dev)
    for mode in ("fp32", "fp64sum"):
        log(f"[case] pagerank rmat19 mode={mode}")
        pr, res = B.pagerank(A, 0.85, 20, mode)
        assert pr_errors(pr.cpu().numpy(), ref) < PR_RTOL[mode]
        # FP32 rounding moves an FP32 replay of the iteration's residuals by
        # < 2e-5 relative over the first 8 steps (numpy check of this graph)
        np.testing.assert_allclose(res[:8], rres[:8], rtol=1e-3)
        pr2, res2 = B.pagerank(A, 0.85, 20, mode)
        np.testing.assert_allclose(pr2.cpu().numpy(), pr.cpu().numpy(), rtol=1e-5, atol=0)
        np.testing.assert_allclose(res2[:8], res[:8], rtol=1e-3)


def test_pagerank_fused_short_rows(orc, dev):
    """A grid has 5 tiles per block row at T = 4 and 8, so each iteration is one
    fused kernel (short-row BMV + update, x double-buffered); 2048 x 2100
    vertices give more block rows than one pass of its capped grid covers
    (4096 workgroups x 256) at T = 4. Checked against the oracle and against
    the unfused iteration (chunked BMV, then the update kernel)."""
    import bgb
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.grid2d(2048, 2100)
    assert -(-csr.nrows // 4) > 4096 * 256
    ref, rres = orc.pagerank(csr.rowptr, csr.colind, 0.85, 20)
    for T in (4, 8):
        A = B.B2SR.from_csr(csr, T, dev)
        for mode in ("fp32", "bf16", "fp64sum"):
            log(f"[case] pagerank fused grid T={T} mode={mode}")
            pr, res = B.pagerank(A, 0.85, 20, mode)
            assert pr_errors(pr.cpu().numpy(), ref) < PR_RTOL[mode]
            if mode != "bf16":
                np.testing.assert_allclose(res[:8], rres[:8], rtol=1e-3)
            try:
                bgb.call("bgb_set_short_row_path", 1)
                pr_u, _ = B.pagerank(A, 0.85, 20, mode)
            finally:
                bgb.call("bgb_set_short_row_path", 0)
            np.testing.assert_allclose(pr.cpu().numpy(), pr_u.cpu().numpy(),
                                       rtol=1e-5 if mode != "bf16" else 2e-2)


def test_pagerank_rmat_with_dangling(orc, dev):
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(14, 8, seed=2)
    assert (csr.degrees() == 0).any()  # isolated vertices carry dangling mass
    ref, _ = orc.pagerank(csr.rowptr, csr.colind, 0.85, 30)
    for T in (4, 32):
        A = B.B2SR.from_csr(csr, T, dev)
        for mode in ("fp32", "fp64sum", "bf16"):
            log(f"[case] pagerank rmat14 T={T} mode={mode}")
            pr, _ = B.pagerank(A, 0.85, 30, mode)
            assert pr_errors(pr.cpu().numpy(), ref) < PR_RTOL[mode]


def test_bf16_bmv_matches_rounded_oracle(orc, dev):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    csr = graph.rmat(12, 8, seed=6)
    rng = np.random.default_rng(1)
    x = rng.random(csr.nrows).astype(np.float32)
    xb = torch.from_numpy(x).to(dev).to(torch.bfloat16)
    xr = xb.float().cpu().numpy().astype(np.float64)  # the values the kernel sees
    want = orc.csr_f64(csr.rowptr, csr.colind, xr)
    for T in TS:
        A = B.B2SR.from_csr(csr, T, dev)
        xpad = torch.zeros(A.nblockrows * T, dtype=torch.bfloat16, device=dev)
        xpad[: csr.nrows] = xb
        got = B.bmv_bin_full_full_bf16(A, xpad).cpu().numpy()[: csr.nrows]
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)
