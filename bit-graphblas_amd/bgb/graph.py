"""Host-side CSR graphs from libbgb's generators (numpy views of int32 CSR)."""
import ctypes
import os

import numpy as np

from ._lib import call, lib


class HostCSR:
    """An int32 CSR matrix on the host (sorted, duplicate-free rows)."""

    def __init__(self, nrows, ncols, rowptr, colind):
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colind = np.ascontiguousarray(colind, dtype=np.int32)
        assert self.rowptr.shape == (self.nrows + 1,)

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    @classmethod
    def _from_handle(cls, h):
        L = lib()
        try:
            n, m, nnz = L.bgb_csr_nrows(h), L.bgb_csr_ncols(h), L.bgb_csr_nnz(h)
            rp = np.ctypeslib.as_array(ctypes.cast(L.bgb_csr_rowptr(h), ctypes.POINTER(ctypes.c_int32)),
                                       shape=(n + 1,)).copy()
            if nnz > 0:
                ci = np.ctypeslib.as_array(ctypes.cast(L.bgb_csr_colind(h), ctypes.POINTER(ctypes.c_int32)),
                                           shape=(nnz,)).copy()
            else:
                ci = np.zeros(0, np.int32)
        finally:
            L.bgb_csr_free(h)
        return cls(n, m, rp, ci)

    def degrees(self):
        return np.diff(self.rowptr)


def _gen(fn, *args):
    h = ctypes.c_void_p()
    call(fn, *args, ctypes.byref(h))
    return HostCSR._from_handle(h)


def from_coo(nrows, ncols, src, dst, undirected=True, drop_self_loops=True):
    """Normalised CSR from an edge list (the reference's readMtx.hpp:152-306 rules)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    assert src.shape == dst.shape
    return _gen("bgb_csr_from_coo", nrows, ncols, src.size, src.ctypes.data, dst.ctypes.data,
                int(undirected), int(drop_self_loops))


def read_mtx(path, directed=0):
    """Normalised CSR from a MatrixMarket coordinate file (the reference's readMtx
    rules, readMtx.hpp:353-422): directed 0 = undirected iff the banner says
    symmetric, 1 = directed, 2 = undirected."""
    return _gen("bgb_read_mtx", os.fsencode(path), int(directed))


def erdos_renyi(n, avg_degree, seed):
    """G(n, m) with m = n*avg_degree/2 distinct undirected edges, no self-loops."""
    return _gen("bgb_gen_er", n, int(n * avg_degree) // 2, seed)


def rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1):
    return _gen("bgb_gen_rmat", scale, edge_factor, a, b, c, seed)


def grid2d(width, height):
    return _gen("bgb_gen_grid2d", width, height)


def bernoulli(n, p, seed):
    out = np.empty(n, np.float32)
    call("bgb_gen_bernoulli_f32", n, p, seed, out.ctypes.data)
    return out


def uniform(n, lo, hi, seed):
    out = np.empty(n, np.float32)
    call("bgb_gen_uniform_f32", n, lo, hi, seed, out.ctypes.data)
    return out


def tile_stats(csr):
    """Exact tile counts (and non-empty block rows) for T = 4, 8, 16, 32."""
    nb = np.zeros(4, np.int64)
    ne = np.zeros(4, np.int64)
    call("bgb_tile_stats", csr.nrows, csr.ncols, csr.rowptr.ctypes.data,
         csr.colind.ctypes.data if csr.nnz else None, nb.ctypes.data, ne.ctypes.data)
    return {T: int(nb[i]) for i, T in enumerate((4, 8, 16, 32))}, {T: int(ne[i]) for i, T in enumerate((4, 8, 16, 32))}


OP_FAMILY = {"bin_bin": 0, "bin_full_full": 1, "k16": 2, "bfs": 3}

TILE_DIMS4 = (4, 8, 16, 32)


class TileProfile(ctypes.Structure):
    """bgb_tile_profile: per T = 4, 8, 16, 32 the tiles, non-empty block rows,
    the longest block row and the non-zero tile row words."""
    _fields_ = [("tiles", ctypes.c_longlong * 4), ("nonempty_rows", ctypes.c_longlong * 4),
                ("max_row_tiles", ctypes.c_longlong * 4), ("row_words", ctypes.c_longlong * 4),
                ("nnz", ctypes.c_longlong), ("nrows", ctypes.c_int), ("ncols", ctypes.c_int)]

    def as_dict(self):
        return {f: {T: int(getattr(self, f)[i]) for i, T in enumerate(TILE_DIMS4)}
                for f in ("tiles", "nonempty_rows", "max_row_tiles", "row_words")}


def tile_profile(csr):
    """The autotuner's features of a host CSR (bgb_tile_profile_csr)."""
    p = TileProfile()
    call("bgb_tile_profile_csr", csr.nrows, csr.ncols, csr.rowptr.ctypes.data,
         csr.colind.ctypes.data if csr.nnz else None, ctypes.byref(p))
    return p


def autotune_profile(csr, op="bin_bin", allowed=TILE_DIMS4, profile=None):
    """Per-graph configuration from the tile profile: {"T", "run_length",
    "x_staging", "cost_us"} (bgb_autotune_profile)."""
    p = profile if profile is not None else tile_profile(csr)
    cost = np.zeros(4, np.float64)
    mask = sum(1 << i for i, T in enumerate(TILE_DIMS4) if T in allowed)
    t, rl, xs = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    call("bgb_autotune_profile", OP_FAMILY[op], ctypes.byref(p), mask, ctypes.byref(t), ctypes.byref(rl),
         ctypes.byref(xs), cost.ctypes.data)
    return {"T": t.value, "run_length": rl.value, "x_staging": xs.value,
            "cost_us": {T: float(cost[i]) for i, T in enumerate(TILE_DIMS4)}}


def autotune(csr, op="bin_bin", allowed=(4, 8, 16, 32), profile=None):
    """Per-graph configuration from the tile profile: {"T", "run_length",
    "x_staging", "cost_us"} (bgb_autotune_profile)."""
    return autotune_profile(csr, op, allowed, profile)


def autotune_tiledim(csr, op="bin_bin", allowed=(4, 8, 16, 32), profile=None):
    """(best T, predicted microseconds per T) from the tile profile."""
    c = autotune_profile(csr, op, allowed, profile)
    return c["T"], c["cost_us"]


def autotune_counts(csr, op="bin_bin", allowed=(4, 8, 16, 32)):
    """The count-only form (bgb_autotune_tiledim: tiles per T, no row
    statistics): (best T, predicted microseconds per T)."""
    nbd, _ = tile_stats(csr)
    nb = np.array([nbd[T] for T in (4, 8, 16, 32)], np.int64)
    cost = np.zeros(4, np.float64)
    mask = sum(1 << i for i, T in enumerate((4, 8, 16, 32)) if T in allowed)
    t = ctypes.c_int(0)
    call("bgb_autotune_tiledim", OP_FAMILY[op], csr.nrows, nb.ctypes.data, mask, ctypes.byref(t), cost.ctypes.data)
    return t.value, {T: float(cost[i]) for i, T in enumerate((4, 8, 16, 32))}


def uneven_graph_specs(ngraphs=1000, nmin=1e3, nmax=1e6, seed=5):
    """Specs of a batch of graphs of uneven size: n log-uniform in [nmin, nmax];
    even indices Erdos-Renyi with average degree U{8..32}, odd ones R-MAT with
    edge factor 8. Returns [(kind, n, degree or edge factor, graph seed)]."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(ngraphs):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        if i % 2 == 0:
            specs.append(("er", n, int(rng.integers(8, 33)), 1000 + i))
        else:
            specs.append(("rmat", n, 8, 1000 + i))
    return specs


def uneven_graph(spec):
    """The graph of one uneven_graph_specs entry (R-MAT at scale round(log2 n),
    at least 10, so its n is the nearest power of two)."""
    kind, n, deg, seed = spec
    if kind == "er":
        return erdos_renyi(n, deg, seed=seed)
    return rmat(max(10, int(round(np.log2(n)))), deg, seed=seed)
