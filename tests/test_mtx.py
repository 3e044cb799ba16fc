"""MatrixMarket ingestion (bgb_read_mtx) on the CPU.

Pinned against the reference's own reader -- readMtx + coo2csr from
bmv_eval/backend/readMtx.hpp compiled into oracle/_ref, the drivers' path
(test_bin_bin_full.cu:139-160) -- on pattern files written here, and against
the oracle's normalisation of the same entries for every field. The reference
reads integer / real values through a format string it leaves unterminated
(`char type_str[3]`, readMtx.hpp:93-98), so its parse of those files is
undefined; they are pinned by the normalisation check only. The reference
calls exit() on the malformed files of the error tests, so those are checked
against the documented status codes only.
"""
import os

import numpy as np
import pytest

import bgb
from bgb import graph
from oracle_lib import REF_SO, RefReadMtx, trailing_entry_erased

BGB_ERR_ARG, BGB_ERR_FORMAT = 1, 6


def write_mtx(path, M, N, entries, field="pattern", sym="general", comments=("% written by the test",),
              values=None):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {sym}\n")
        for c in comments:
            f.write(c + "\n")
        f.write(f"{M} {N} {len(entries)}\n")
        for k, (i, j) in enumerate(entries):
            if field == "pattern":
                f.write(f"{i + 1} {j + 1}\n")
            elif field == "integer":
                f.write(f"{i + 1} {j + 1} {values[k] if values is not None else 3}\n")
            else:
                f.write(f"{i + 1} {j + 1} {values[k] if values is not None else 0.25:.6g}\n")


def cases(rng):
    out = []
    for n, m in ((7, 12), (100, 600), (1000, 4000)):
        ij = rng.integers(0, n, (m, 2))
        ij = np.concatenate([ij, ij[: m // 10], np.stack([np.arange(5) % n] * 2, 1)])  # duplicates, loops
        for field in ("pattern", "integer", "real"):
            for sym in ("general", "symmetric"):
                ent = ij
                if sym == "symmetric":  # a symmetric file stores one triangle
                    ent = np.stack([ij.max(1), ij.min(1)], 1)
                out.append((n, n, [tuple(map(int, e)) for e in ent], field, sym))
    return out


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build oracle/_ref not present")
def test_read_mtx_matches_reference_reader(tmp_path):
    ref = RefReadMtx()
    rng = np.random.default_rng(9)
    compared = 0
    for k, (M, N, ent, field, sym) in enumerate(cases(rng)):
        if field != "pattern":
            continue  # the reference's value parse is undefined (see the module docstring)
        path = str(tmp_path / f"g{k}.mtx")
        write_mtx(path, M, N, ent, field, sym)
        for directed in (0, 1, 2):
            undirected = directed == 2 or (directed == 0 and sym == "symmetric")
            if trailing_entry_erased(ent, undirected):
                continue
            compared += 1
            g = graph.read_mtx(path, directed)
            rM, rN, rrp, rci = ref.read_mtx(path, directed)
            assert (g.nrows, g.ncols) == (rM, rN)
            np.testing.assert_array_equal(g.rowptr, rrp, err_msg=f"{field} {sym} directed={directed}")
            np.testing.assert_array_equal(g.colind, rci, err_msg=f"{field} {sym} directed={directed}")
    assert compared >= 8


def test_reference_reader_drops_two_after_a_trailing_erased_entry(orc, tmp_path):
    """The reference defect the comparison above steps around, pinned so the
    cross-check stays deterministic: a trailing duplicate costs the reference
    two valid entries, the library reader none."""
    ref = RefReadMtx()
    ent = [(0, 1), (1, 2), (2, 3), (3, 4), (3, 4)]
    assert trailing_entry_erased(ent, False)
    path = str(tmp_path / "trailing_dup.mtx")
    write_mtx(path, 5, 5, ent)
    g = graph.read_mtx(path, 1)
    _, _, rrp, _ = ref.read_mtx(path, 1)
    assert g.nnz == 4 and rrp[-1] == 2


def test_read_mtx_matches_normalisation_of_its_entries(orc, tmp_path):
    rng = np.random.default_rng(10)
    for k, (M, N, ent, field, sym) in enumerate(cases(rng)):
        path = str(tmp_path / f"h{k}.mtx")
        write_mtx(path, M, N, ent, field, sym, comments=("%", "% blank line next", ""))
        e = np.array(ent, np.int32)
        g = graph.read_mtx(path)
        rp, ci = orc.normalize(M, N, e[:, 0], e[:, 1], sym == "symmetric", True)
        np.testing.assert_array_equal(g.rowptr, rp)
        np.testing.assert_array_equal(g.colind, ci)


def test_read_mtx_rectangular_directed(tmp_path):
    path = str(tmp_path / "rect.mtx")
    write_mtx(path, 3, 5, [(0, 4), (2, 0), (0, 4), (1, 1)], "real", "general")
    g = graph.read_mtx(path)
    assert (g.nrows, g.ncols) == (3, 5)
    assert g.rowptr.tolist() == [0, 1, 1, 2] and g.colind.tolist() == [4, 0]   # loop (1,1) and duplicate dropped
    assert bgb.lib().bgb_read_mtx(path.encode(), 2, None) != 0   # undirected needs a square matrix


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", BGB_ERR_FORMAT),      # array format
    ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 0\n", BGB_ERR_FORMAT),
    ("%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n", BGB_ERR_FORMAT),
    ("not a banner\n2 2 1\n1 1\n", BGB_ERR_FORMAT),
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 3\n1 2\n2 1\n", BGB_ERR_FORMAT),  # short
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n3 1\n", BGB_ERR_FORMAT),      # out of range
    ("%%MatrixMarket matrix coordinate pattern general\n", BGB_ERR_FORMAT),                  # no size line
    # an entry count no file of this size can hold: the reservation follows the
    # file, not the count (an allocation of the count aborted the process)
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 999999999999999\n1 1\n", BGB_ERR_FORMAT),
])
def test_read_mtx_rejects_malformed_files(tmp_path, text, code):
    import ctypes
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    h = ctypes.c_void_p()
    assert bgb.lib().bgb_read_mtx(str(path).encode(), 0, ctypes.byref(h)) == code
    assert bgb.lib().bgb_last_error()


def test_read_mtx_missing_file():
    import ctypes
    h = ctypes.c_void_p()
    assert bgb.lib().bgb_read_mtx(b"/nonexistent/x.mtx", 0, ctypes.byref(h)) == BGB_ERR_ARG


def test_mycielskian_inputs(tmp_path):
    """The sweep's Mycielski inputs: known sizes (M4 is the Groetzsch graph,
    11 vertices and 20 edges), the size recurrence, and a symmetric read."""
    from mtx_files import mycielskian, write_mycielskian
    sizes = {k: (mycielskian(k)[0], len(mycielskian(k)[1])) for k in range(2, 9)}
    assert sizes[2] == (2, 1) and sizes[3] == (5, 5) and sizes[4] == (11, 20) and sizes[5] == (23, 71)
    for k in range(2, 8):
        n, e = sizes[k]
        assert sizes[k + 1] == (2 * n + 1, 3 * e + n)
    # triangle-free: no edge (a, b) with a common neighbour
    n, e = mycielskian(6)
    adj = np.zeros((n, n), bool)
    adj[e[:, 0], e[:, 1]] = adj[e[:, 1], e[:, 0]] = True
    assert not (adj & (adj.astype(np.int64) @ adj.astype(np.int64) > 0)).any()
    p = str(tmp_path / "mycielskian6.mtx")
    write_mycielskian(p, 6)
    csr = graph.read_mtx(p)
    assert csr.nrows == n and csr.nnz == 2 * len(e)
    D = np.zeros((n, n), bool)
    D[np.repeat(np.arange(n), np.diff(csr.rowptr)), csr.colind] = True
    assert (D == adj).all()
