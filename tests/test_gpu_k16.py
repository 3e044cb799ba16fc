"""GPU parity of the K=16 MFMA bin_full_full (single graph and the persistent
multi-graph launch) against the oracle's FP64 CSR SpMM.

X is FP16; the oracle multiplies the same FP16 values in FP64, so the only
difference is FP32 accumulation: |Y - Y64| <= 1e-6 * sum_u |X[u]| per entry
(plus 1e-7 absolute) is the tolerance, stated here and reported per graph.
"""
import sys

import numpy as np
import pytest

from test_gpu_bmv import dirty

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(cuda):
    import bgb.b2sr  # noqa: F401
    return cuda


def rhs(n, seed, dev):
    import torch
    rng = np.random.default_rng(seed)
    X = torch.from_numpy(rng.uniform(-1, 1, (n, 16)).astype(np.float16)).to(dev)
    return X, X.cpu().numpy().astype(np.float64)


def check(orc, csr, Y, X64, label):
    want = orc.csr_spmm_f64(csr.rowptr, csr.colind, X64)
    bound = orc.csr_spmm_f64(csr.rowptr, csr.colind, np.abs(X64))
    got = Y[: csr.nrows].cpu().numpy().astype(np.float64)
    err = np.abs(got - want)
    ok = np.all(err <= 1e-6 * bound + 1e-7)
    log(f"[k16] {label}: max abs err {err.max():.3e}, max rel (to L1 bound) "
        f"{(err / np.maximum(bound, 1e-30)).max():.3e}")
    assert ok, label
    assert not Y[csr.nrows:].abs().sum().item()


@pytest.mark.parametrize("T", (16, 32))
def test_k16_single(orc, dev, T):
    import torch
    from bgb import b2sr as B
    from bgb import graph
    for label, csr in (("rmat13", graph.rmat(13, 8, seed=3)), ("er3000", graph.erdos_renyi(3000, 20, 5)),
                       ("grid", graph.grid2d(70, 33))):
        A = B.B2SR.from_csr(csr, T, dev)
        X, X64 = rhs(csr.nrows, T, dev)
        Xb = B.k16_pack_rhs(X, A.nblockrows * T)
        Y = B.bmv_k16(A, Xb, out=dirty(A.nblockrows * T * 16, torch.float32, dev))  # overwritten
        check(orc, csr, Y, X64, f"{label} T={T}")
        Y2 = B.bmv_k16(A, Xb, out=Y.clone(), accumulate=True)
        np.testing.assert_allclose(Y2.cpu().numpy(), 2 * Y.cpu().numpy(), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("T", (16, 32))
def test_k16_empty_block_rows(orc, dev, T):
    """Block rows without tiles -- leading, scattered and a long trailing run
    (200 000 vertices, edges only among 20 000 of them) -- are zero in the
    output (which starts dirty), single graph and batched."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(60 + T)
    n = 200000
    src = rng.integers(1000, 21000, 150000)
    dst = rng.integers(1000, 21000, 150000)
    keep = ((src // 64) % 3 != 0) & ((dst // 64) % 3 != 0)  # isolated stretches of 64 vertices
    csr = graph.from_coo(n, n, src[keep], dst[keep], True, True)
    A = B.B2SR.from_csr(csr, T, dev)
    X, X64 = rhs(n, 70 + T, dev)
    Xb = B.k16_pack_rhs(X, A.nblockrows * T)
    Y = B.bmv_k16(A, Xb, out=dirty(A.nblockrows * T * 16, torch.float32, dev))
    check(orc, csr, Y, X64, f"empty block rows T={T}")
    batch = B.K16Batch([A, A], [Xb, Xb])
    batch.y.view(torch.uint8).fill_(0xFF)
    batch.run()
    for i in range(2):
        check(orc, csr, batch.result(i), X64, f"empty block rows batched #{i} T={T}")


@pytest.mark.timeout(120)
def test_k16_persistent_batch(orc, dev):
    """Graphs of uneven sizes (log-uniform n), half ER, half R-MAT, tile dim from
    the autotuner, one launch."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    rng = np.random.default_rng(11)
    mats, xbs, refs = [], [], []
    for i in range(40):
        n = int(np.exp(rng.uniform(np.log(100), np.log(20000))))
        if i % 2 == 0:
            csr = graph.erdos_renyi(max(n, 64), int(rng.integers(8, 33)), seed=i)
        else:
            csr = graph.rmat(max(7, int(np.log2(max(n, 128)))), 8, seed=i)
        T = graph.autotune_tiledim(csr, "k16", allowed=(16, 32))[0]
        A = B.B2SR.from_csr(csr, T, dev)
        X, X64 = rhs(csr.nrows, 100 + i, dev)
        mats.append(A)
        xbs.append(B.k16_pack_rhs(X, A.nblockrows * T))
        refs.append((csr, X64, T))
    batch = B.K16Batch(mats, xbs, nworkgroups=64)
    log(f"[case] k16 persistent batch: {len(mats)} graphs, 64 workgroups")
    batch.y.view(torch.uint8).fill_(0xFF)  # the launch overwrites every output row
    batch.run()
    for i, (csr, X64, T) in enumerate(refs):
        check(orc, csr, batch.result(i), X64, f"batch graph {i} n={csr.nrows} T={T}")


@pytest.mark.timeout(600)
def test_k16_batch_rmat22(orc, dev):
    """Three R-MAT scale-22 graphs at T=16 in one persistent launch (about a
    million 256-tile chunks; hub rows carry across many chunks and across the
    fix-up's 64-chunk waves): every graph within the FP32 bound."""
    import torch
    from bgb import b2sr as B
    from bgb import graph
    mats, xbs, refs = [], [], []
    chunks = 0
    for i in range(3):
        csr = graph.rmat(22, 16, seed=40 + i)
        A = B.B2SR.from_csr(csr, 16, dev)
        X, X64 = rhs(csr.nrows, 200 + i, dev)
        mats.append(A)
        xbs.append(B.k16_pack_rhs(X, A.nblockrows * 16))
        refs.append((csr, X64))
        chunks += -(-A.nblocks // 256)
    log(f"[case] k16 batch of 3 R-MAT 22 graphs: {chunks} chunks")
    batch = B.K16Batch(mats, xbs)
    batch.y.view(torch.uint8).fill_(0xFF)
    batch.run()
    for i, (csr, X64) in enumerate(refs):
        check(orc, csr, batch.result(i), X64, f"rmat22 #{i}")
