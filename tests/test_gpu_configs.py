"""The benchmark workloads at their full sizes, gated against the CPU oracle.

* R-MAT scale 22, edge factor 16 (a=.57 b=.19 c=.19), symmetrised: the
  bin_bin_full BMV (exact int32 neighbour counts, x Bernoulli(0.05)) at every
  tile dim, and the 64-source batched BFS (pull only and direction
  switching) against 64 queue BFS runs of the oracle -- the reference's eval
  path is test_bin_bin_full.cu:311-353.
* The 4096 x 4096 4-neighbour grid: bin_full_full (x uniform in [0, 1)) at
  every tile dim within 1e-6 relative of the FP64 CSR sum
  (test_bin_full_full.cu:274-305), and PageRank (d = 0.85, 50 iterations) in
  the three precision modes against the oracle's FP64 power iteration.
* 1000 graphs of uneven size (n log-uniform in [1e3, 1e6], half Erdos-Renyi
  with degree U{8..32}, half R-MAT with edge factor 8), K = 16 FP16
  right-hand sides on MFMA, tile dim from the autotuner, streamed through
  the persistent multi-graph launch in 8 batches of 125 graphs (one test
  each): every graph within the FP32-accumulation bound of the FP64 SpMM.
Inputs use the benchmark's seeds. Outputs start dirty (0xFF bytes).
"""
import sys
import time

import numpy as np
import pytest

from test_gpu_bmv import dirty, to_np
from test_gpu_graph import PR_MASS, PR_RTOL

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


_CACHE = {}


def rmat22():
    if "rmat22" not in _CACHE:
        from bgb import graph
        t0 = time.perf_counter()
        _CACHE["rmat22"] = graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1)
        log(f"[gen] R-MAT 22: nnz {_CACHE['rmat22'].nnz} in {time.perf_counter() - t0:.1f} s")
    return _CACHE["rmat22"]


def grid4096():
    if "grid" not in _CACHE:
        from bgb import graph
        _CACHE["grid"] = graph.grid2d(4096, 4096)
    return _CACHE["grid"]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("T", (4, 8, 16, 32))
def test_rmat22_bin_bin_full_exact(orc, dev, T):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    csr = rmat22()
    x01 = graph.bernoulli(csr.nrows, 0.05, 2)
    if "counts22" not in _CACHE:
        _CACHE["counts22"] = orc.csr_count(csr.rowptr, csr.colind, x01.astype(np.uint8))
    counts = _CACHE["counts22"]
    A = B.B2SR.from_csr(csr, T, dev)
    log(f"[case] R-MAT 22 bin_bin_full T={T} nblocks={A.nblocks}")
    xb = B.pack_vec(B.padded_vector(x01, T, dev), T)
    got = to_np(B.bmv_bin_bin_full(A, xb, out=dirty(A.nblockrows * T, torch.int32, dev)))
    np.testing.assert_array_equal(got[: csr.nrows], counts)
    assert not got[csr.nrows:].any()
    # invariant: the tiles hold every edge once
    _, _, val = A.to_host()
    assert int(np.unpackbits(val.view(np.uint8)).sum()) == csr.nnz


@pytest.mark.timeout(600)
def test_rmat22_batched_bfs_64_sources(orc, dev):
    """64 sources among the non-isolated vertices (the benchmark's draw), tile
    dim from the autotuner and T = 8; level arrays identical to the oracle's
    64 queue BFS runs, pulling every level and with direction switching."""
    from bgb import b2sr as B
    from bgb import graph
    csr = rmat22()
    rng = np.random.default_rng(64)
    src = rng.choice(np.nonzero(csr.degrees())[0], 64, replace=False).astype(np.int32)
    t0 = time.perf_counter()
    want, wl = orc.bfs_multi(csr.rowptr, csr.colind, src)
    log(f"[case] R-MAT 22 oracle BFS x64: {wl} levels in {time.perf_counter() - t0:.1f} s")
    tuned = graph.autotune_tiledim(csr, "bfs")[0]
    for T in sorted({tuned, 8}):
        A = B.B2SR.from_csr(csr, T, dev)
        for push in (None, A):  # symmetric: A is its own transpose
            log(f"[case] R-MAT 22 batched BFS T={T} push={push is not None}")
            lev, nl = B.bfs_multi(A, src, True, push=push)
            np.testing.assert_array_equal(lev.cpu().numpy(), want)
            assert nl == wl
        del A


@pytest.mark.timeout(600)
@pytest.mark.parametrize("T", (4, 8, 16, 32))
def test_grid4096_bin_full_full(orc, dev, T):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    csr = grid4096()
    assert csr.nnz == 67092480
    x = graph.uniform(csr.nrows, 0.0, 1.0, 3)
    if "ref_grid" not in _CACHE:
        _CACHE["ref_grid"] = orc.csr_f64(csr.rowptr, csr.colind, x.astype(np.float64))
    ref = _CACHE["ref_grid"]
    A = B.B2SR.from_csr(csr, T, dev)
    log(f"[case] 4096^2 grid bin_full_full T={T} nblocks={A.nblocks}")
    got = to_np(B.bmv_bin_full_full(A, B.padded_vector(x, T, dev), out=dirty(A.nblockrows * T, torch.float32, dev)))
    np.testing.assert_allclose(got[: csr.nrows], ref, rtol=1e-6, atol=1e-6)
    assert not got[csr.nrows:].any()


@pytest.mark.timeout(600)
def test_grid4096_pagerank_three_modes(orc, dev):
    """d = 0.85, 50 iterations at the autotuned tile dim: max relative error
    per vertex against the FP64 oracle within the per-mode bound (FP32 and
    FP64-sum 2e-5, BF16 x 2e-2) and total mass 1."""
    from bgb import b2sr as B
    from bgb import graph
    csr = grid4096()
    t0 = time.perf_counter()
    ref, _ = orc.pagerank(csr.rowptr, csr.colind, 0.85, 50)
    log(f"[case] 4096^2 grid oracle PageRank in {time.perf_counter() - t0:.1f} s")
    T = graph.autotune_tiledim(csr, "bin_full_full")[0]
    A = B.B2SR.from_csr(csr, T, dev)
    for mode in ("fp32", "bf16", "fp64sum"):
        pr, _ = B.pagerank(A, 0.85, 50, mode)
        p = pr.cpu().numpy().astype(np.float64)
        err = float(np.max(np.abs(p - ref) / ref))
        log(f"[case] 4096^2 grid PageRank T={T} {mode}: max rel err {err:.3e}")
        assert err < PR_RTOL[mode], (mode, err)
        assert abs(p.sum() - 1.0) < PR_MASS[mode]


NPARTS = 8


@pytest.mark.timeout(600)
@pytest.mark.parametrize("part", range(NPARTS))
def test_uneven_graphs_k16_persistent(orc, dev, part):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    specs = graph.uneven_graph_specs(1000)
    per = len(specs) // NPARTS
    mats, xbs, refs = [], [], []
    t0 = time.perf_counter()
    for i in range(part * per, (part + 1) * per):
        csr = graph.uneven_graph(specs[i])
        T = graph.autotune_tiledim(csr, "k16", allowed=(16, 32))[0]
        A = B.B2SR.from_csr(csr, T, dev)
        rng = np.random.default_rng(7000 + i)
        X = torch.from_numpy(rng.uniform(-1, 1, (csr.nrows, 16)).astype(np.float16)).to(dev)
        xbs.append(B.k16_pack_rhs(X, A.nblockrows * T))
        mats.append(A)
        refs.append((i, csr, X.cpu().numpy().astype(np.float64), T))
    log(f"[case] uneven graphs part {part}: {per} graphs generated and packed in {time.perf_counter() - t0:.1f} s")
    batch = B.K16Batch(mats, xbs)
    batch.y.view(torch.uint8).fill_(0xFF)  # the launch overwrites every output row
    batch.run()
    worst = 0.0
    for k, (i, csr, X64, T) in enumerate(refs):
        want = orc.csr_spmm_f64(csr.rowptr, csr.colind, X64)
        bound = orc.csr_spmm_f64(csr.rowptr, csr.colind, np.abs(X64))
        Y = batch.result(k)
        got = Y[: csr.nrows].cpu().numpy().astype(np.float64)
        err = np.abs(got - want)
        assert np.all(err <= 1e-6 * bound + 1e-7), f"graph {i} {specs[i]} T={T}"
        assert not Y[csr.nrows:].abs().sum().item()
        worst = max(worst, float((err / np.maximum(bound, 1e-30)).max()))
    log(f"[case] uneven graphs part {part}: max error / bound {worst:.3e}, {time.perf_counter() - t0:.1f} s")
