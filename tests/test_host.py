"""CPU tests of libbgb: exports, generators, normalisation, argument checks.

No compute call reaches the GPU here: the BMV / packing entry points are only
called with arguments that fail validation before any HIP call.
"""
import ctypes
import re

import numpy as np
import pytest

import bgb
from bgb import graph

HEADER_FN = re.compile(r"^\s*(?:const\s+)?[a-z_0-9 ]+\s*\*?\s*(bgb_[a-z0-9_]+)\s*\(", re.M)


def header_functions():
    with open(bgb.HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(HEADER_FN.findall(text)))


def test_library_exports_every_header_symbol():
    names = header_functions()
    assert len(names) >= 25
    L = bgb.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    from bgb._lib import SIGNATURES
    assert sorted(set(names) - set(SIGNATURES)) == []


def test_info_calls():
    L = bgb.lib()
    assert L.bgb_version() >= 100
    assert L.bgb_status_string(0) == b"ok"
    assert [L.bgb_word_bytes(t) for t in (4, 8, 16, 32, 64, 5)] == [1, 1, 2, 4, 8, 0]
    assert L.bgb_packed_vec_words(32, 128) == 128
    assert L.bgb_packed_vec_words(4, 5) == 8
    assert L.bgb_packed_vec_words(3, 5) == -1


def test_er_generator_exact_and_deterministic():
    g = graph.erdos_renyi(4096, 16, seed=1)
    assert g.nrows == 4096 and g.nnz == 4096 * 16
    deg = g.degrees()
    assert deg.sum() == 65536
    # symmetric, no loops, sorted unique rows
    rows = np.repeat(np.arange(g.nrows), deg)
    assert not np.any(rows == g.colind)
    key = rows.astype(np.int64) * g.nrows + g.colind
    assert np.all(np.diff(key) > 0)
    tkey = g.colind.astype(np.int64) * g.nrows + rows
    assert np.array_equal(np.sort(tkey), key)
    g2 = graph.erdos_renyi(4096, 16, seed=1)
    assert np.array_equal(g.colind, g2.colind)
    assert not np.array_equal(g.colind, graph.erdos_renyi(4096, 16, seed=2).colind)


def test_rmat_properties(orc):
    g = graph.rmat(12, 8, seed=3)
    n = 1 << 12
    assert g.nrows == n
    rows = np.repeat(np.arange(n), g.degrees())
    assert not np.any(rows == g.colind)
    key = rows.astype(np.int64) * n + g.colind
    assert np.all(np.diff(key) > 0)
    assert np.array_equal(np.sort(g.colind.astype(np.int64) * n + rows), key)
    # skew: vertex 0 (all quadrant-a bits) is the highest-degree vertex
    assert g.degrees()[0] == g.degrees().max()
    assert 0.5 * 2 * 8 * n < g.nnz <= 2 * 8 * n


def test_grid_closed_form():
    g = graph.grid2d(7, 5)
    n = 35
    assert g.nnz == 4 * n - 2 * 7 - 2 * 5
    deg = g.degrees().reshape(5, 7)
    assert deg[0, 0] == 2 and deg[2, 3] == 4 and deg[0, 3] == 3
    assert list(g.colind[g.rowptr[8]:g.rowptr[9]]) == [1, 7, 9, 15]


def test_csr_from_coo_matches_oracle(orc):
    rng = np.random.default_rng(4)
    for undirected in (True, False):
        for drop in (True, False):
            src = rng.integers(0, 500, 4000).astype(np.int32)
            dst = rng.integers(0, 500, 4000).astype(np.int32)
            src[:300] = dst[:300]
            g = graph.from_coo(500, 500, src, dst, undirected, drop)
            rp, ci = orc.normalize(500, 500, src, dst, undirected, drop)
            np.testing.assert_array_equal(g.rowptr, rp)
            np.testing.assert_array_equal(g.colind, ci)


def test_csr_from_coo_rejects_bad_index():
    with pytest.raises(bgb.BgbError) as e:
        graph.from_coo(4, 4, np.array([0, 4]), np.array([1, 1]))
    assert e.value.status == 1


def test_vectors_deterministic():
    a = graph.bernoulli(10000, 0.1, 7)
    assert set(np.unique(a)) <= {0.0, 1.0}
    assert 800 < a.sum() < 1200
    assert np.array_equal(a, graph.bernoulli(10000, 0.1, 7))
    u = graph.uniform(10000, -1, 1, 3)
    assert u.min() >= -1 and u.max() < 1


def test_entry_points_validate_before_touching_the_gpu():
    L = bgb.lib()
    vp = ctypes.c_void_p
    # bad tile dim
    assert L.bgb_bmv_bin_bin_full_i32(5, vp(16), vp(16), vp(16), vp(16), vp(16), 4, 4, 0, None) == 2
    # null output
    assert L.bgb_bmv_bin_bin_bin(32, vp(16), vp(16), None, vp(16), vp(16), 4, 4, 0, None) == 1
    # bad accumulate flag
    assert L.bgb_bmv_bin_full_full(8, vp(16), vp(16), vp(16), vp(16), vp(16), 4, 4, 2, None) == 1
    # misaligned tiles
    assert L.bgb_bmv_bin_bin_full_f32(8, vp(17), vp(16), vp(16), vp(16), vp(16), 4, 4, 0, None) == 1
    # masked needs a mask
    assert L.bgb_bmv_bin_bin_bin_masked(16, vp(16), vp(16), vp(16), vp(16), vp(16), 4, 4, None, None) == 1
    # packing
    assert L.bgb_pack_vec(7, vp(16), 10, vp(16), None) == 2
    assert L.bgb_csr2b2sr(4, 8, 8, 4, vp(16), vp(16), vp(16), 9, vp(16), vp(16), None) == 1
    nb = ctypes.c_int()
    assert L.bgb_b2sr_nnzb(4, -1, 8, 0, vp(16), vp(16), vp(16), ctypes.byref(nb), None) == 1
    assert b"tile dim" in L.bgb_last_error() or L.bgb_last_error() != b""


def test_tile_stats_match_oracle_packing(orc):
    rng = np.random.default_rng(8)
    for n, m in ((1, 0), (37, 50), (1000, 8000)):
        g = graph.from_coo(n, n, rng.integers(0, n, m), rng.integers(0, n, m))
        nb, ne = graph.tile_stats(g)
        for T in (4, 8, 16, 32):
            brp, _, _ = orc.b2sr(T, n, n, g.rowptr, g.colind)
            assert nb[T] == brp[-1]
            assert ne[T] == int(np.count_nonzero(np.diff(brp)))


def test_autotuner_prefers_small_tiles_on_sparse_random_graphs():
    g = graph.erdos_renyi(20000, 8, seed=3)
    T, cost = graph.autotune_tiledim(g, "bin_bin")
    assert T in (4, 8) and cost[T] == min(cost.values())
    T16, _ = graph.autotune_tiledim(g, "k16")
    assert T16 in (16, 32)
    # the count-only form agrees on the small-tile choice
    assert graph.autotune_counts(g, "bin_bin")[0] in (4, 8)


def test_tile_profile_matches_the_oracle_statistics(orc):
    """Tiles and non-empty block rows per T equal the oracle's B2SR; the row
    words equal estimatevec's count; the longest block row is the largest
    rowptr difference."""
    g = graph.rmat(12, 8, seed=4)
    p = graph.tile_profile(g).as_dict()
    for T in (4, 8, 16, 32):
        brp, _, _ = orc.b2sr(T, g.nrows, g.ncols, g.rowptr, g.colind)
        assert p["tiles"][T] == brp[-1]
        assert p["nonempty_rows"][T] == int(np.count_nonzero(np.diff(brp)))
        assert p["max_row_tiles"][T] == int(np.diff(brp).max())
        assert p["row_words"][T] == orc.estimate_tiles(g.rowptr, g.colind, T)


def test_autotuner_model_paths():
    """A 2-D grid (5 tiles per block row at T <= 8) takes the short-row kernel
    at T = 8; one million dense 32 x 32 blocks (a profile built by hand: tiles
    hold 4 x the bytes per edge at T = 4 than at T = 32) favour T = 32."""
    c = graph.autotune(graph.grid2d(1024, 1024), "bin_bin")
    assert c["T"] in (4, 8) and c["x_staging"] == 1
    p = graph.TileProfile()
    nb = 1 << 20
    p.nrows = p.ncols = 32 * nb
    p.nnz = 1024 * nb
    for i, T in enumerate((4, 8, 16, 32)):
        k = 32 // T
        p.tiles[i] = nb * k * k
        p.nonempty_rows[i] = nb * k
        p.max_row_tiles[i] = k
        p.row_words[i] = nb * k * k * T
    c = graph.autotune_profile(None, "bin_bin", profile=p)
    assert c["T"] == 32, c


def test_tuning_knobs_validate_and_round_trip():
    L = bgb.lib()
    for T in (4, 8, 16, 32):
        assert L.bgb_get_run_length(T) in (8, 16)
        assert L.bgb_get_x_staging(T) == 0
    assert L.bgb_get_run_length(5) == -1 and L.bgb_get_x_staging(5) == -1
    assert L.bgb_set_run_length(8, 12) == 1      # BGB_ERR_ARG
    assert L.bgb_set_x_staging(8, 3) == 1        # BGB_ERR_ARG
    assert L.bgb_set_x_staging(6, 0) == 2        # BGB_ERR_TILEDIM
    assert L.bgb_get_short_row_path() == 0
    assert L.bgb_set_short_row_path(3) == 1      # BGB_ERR_ARG
    assert L.bgb_get_row_pieces() == 0
    assert L.bgb_set_row_pieces(3) == 1          # BGB_ERR_ARG
    assert L.bgb_set_row_pieces(-1) == 1
    try:
        assert L.bgb_set_x_staging(16, 2) == 0 and L.bgb_get_x_staging(16) == 2
        assert L.bgb_set_run_length(16, 16) == 0 and L.bgb_get_run_length(16) == 16
    finally:
        L.bgb_set_x_staging(16, 0)
        L.bgb_set_run_length(16, 8)


def test_autotune_reports_the_whole_configuration():
    # R-MAT 16: the packed x fits in LDS at every T <= 16, so the bit ops stage it
    csr = graph.rmat(16, 8, seed=3)
    c = graph.autotune(csr, "bin_bin")
    assert c["T"] in (4, 8, 16, 32) and c["run_length"] == 8
    assert c["x_staging"] == (2 if c["T"] <= 16 else 1)
    assert c["T"] == graph.autotune_tiledim(csr, "bin_bin")[0]
    # FP ops never stage a bit vector; K=16 picks from 16 / 32 only
    assert graph.autotune(csr, "bin_full_full")["x_staging"] == 1
    assert graph.autotune(csr, "k16")["T"] in (16, 32)
