"""Seeded randomized parity of every BMV operator against the oracle, across
the kernel paths a matrix can take: chunked walk (general, one-row and
two-row chunks), short-row kernel, packed x in LDS or gathered from memory.

Each case draws a graph shape (hub rows spanning many chunks, runs of empty
block rows, ragged n, non-square, near-regular rows), a tile dim and the
knob settings from a fixed seed, then checks counts (int32 and float),
bits, masked bits and FP32 sums exactly / within 1e-6 relative, and that
accumulate=1 adds onto the output. Bit and integer results must be exact.
"""
import numpy as np
import pytest

from test_gpu_bmv import check_all_ops, to_np

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
    if kind == 2:  # non-square (m <= n: x is drawn per row and covers every column), directed
        m = int(rng.integers(min(20, n), n + 1))
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
    try:
        for case in range(6):
            label, csr = random_graph(rng)
            T = int(rng.choice(TS))
            staging = int(rng.integers(0, 3))
            short = int(rng.integers(0, 3)) if T <= 8 else 0
            bgb.call("bgb_set_x_staging", T, staging)
            bgb.call("bgb_set_short_row_path", short)
            A, xb, xfp = check_all_ops(orc, csr, T, rng, cuda,
                                       f"seed {seed}.{case} {label} staging={staging} short={short}",
                                       xdensity=float(rng.uniform(0.02, 0.9)))
            once = B.bmv_bin_bin_full(A, xb)
            twice = B.bmv_bin_bin_full(A, xb, out=once.clone(), accumulate=True)
            np.testing.assert_array_equal(to_np(twice), 2 * to_np(once))
            bgb.call("bgb_set_x_staging", T, 0)
    finally:
        for T in TS:
            bgb.call("bgb_set_x_staging", T, 0)
        bgb.call("bgb_set_short_row_path", 0)
