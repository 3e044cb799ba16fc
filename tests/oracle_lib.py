"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Built by ``make -C oracle`` (``__graft_entry__.build()`` does it). The
reference-built cross-check oracle/_ref/libreadmtx_ref.so is optional.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libreadmtx_ref.so")

WORD_NP = {4: np.uint8, 8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def _p(a):
    return a.ctypes.data if a is not None else None


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        self.L = ctypes.CDLL(path)
        vp, ci, ll, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
        sig = {
            "orc_csr_normalize": (ci, [ci, ci, ll, vp, vp, ci, ci, vp, vp, vp]),
            "orc_b2sr_nnzb": (ci, [ci, ci, ci, vp, vp, vp]),
            "orc_b2sr_fill": (None, [ci, ci, ci, vp, vp, vp, vp, vp]),
            "orc_pack_vec": (None, [ci, vp, ci, ll, vp]),
            "orc_pack_u8": (None, [ci, vp, ci, ll, vp]),
            "orc_b2sr_bmv_bin_bin_bin": (None, [ci, vp, vp, vp, vp, vp, ci]),
            "orc_b2sr_bmv_bin_bin_full": (None, [ci, vp, vp, vp, vp, vp, ci]),
            "orc_b2sr_bmv_bin_full_full": (None, [ci, vp, vp, vp, vp, vp, ci]),
            "orc_b2sr_bmv_masked": (None, [ci, vp, vp, vp, vp, vp, ci, vp]),
            "orc_csr_spmv_count": (None, [ci, vp, vp, vp, vp]),
            "orc_csr_spmv_or": (None, [ci, vp, vp, vp, vp]),
            "orc_csr_spmv_f32": (None, [ci, vp, vp, vp, vp]),
            "orc_csr_spmv_f64": (None, [ci, vp, vp, vp, vp]),
            "orc_csr_spmm_f64": (None, [ci, vp, vp, ci, vp, vp]),
            "orc_bfs": (ci, [ci, vp, vp, ci, vp]),
            "orc_bfs_multi": (ci, [ci, vp, vp, ci, vp, vp]),
            "orc_pagerank_f64": (None, [ci, vp, vp, dbl, ci, vp, vp]),
            "orc_csr2csc": (None, [ci, ci, ll, vp, vp, vp, vp, vp, vp]),
            "orc_b2sr_transpose": (None, [ci, ci, ci, vp, vp, vp, vp, vp, vp]),
            "orc_tile_row_stats": (None, [ci, ci, vp, vp, vp, vp]),
            "orc_estimate_tiles": (ll, [ci, vp, vp, ci]),
            "orc_num_threads": (ci, []),
            "orc_set_threads": (None, [ci]),
        }
        for name, (res, args) in sig.items():
            f = getattr(self.L, name)
            f.restype, f.argtypes = res, args

    # -- normalisation / packing -------------------------------------------
    def normalize(self, nrows, ncols, src, dst, undirected=True, drop_self=True):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        rp = np.zeros(nrows + 1, np.int32)
        ci = np.zeros(max(2 * src.size, 1), np.int32)
        nnz = ctypes.c_longlong()
        st = self.L.orc_csr_normalize(nrows, ncols, src.size, _p(src), _p(dst), int(undirected),
                                      int(drop_self), _p(rp), _p(ci), ctypes.byref(nnz))
        assert st == 0, "edge index out of range"
        return rp, ci[: nnz.value].copy()

    def b2sr(self, T, nrows, ncols, rowptr, colind):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        colind = np.ascontiguousarray(colind if len(colind) else np.zeros(1), np.int32)
        nbr = (nrows + T - 1) // T
        brp = np.zeros(nbr + 1, np.int32)
        nb = self.L.orc_b2sr_nnzb(T, nrows, ncols, _p(rowptr), _p(colind), _p(brp))
        bci = np.zeros(max(nb, 1), np.int32)
        val = np.zeros(max(nb, 1) * T, WORD_NP[T])
        self.L.orc_b2sr_fill(T, nrows, ncols, _p(rowptr), _p(colind), _p(brp), _p(bci), _p(val))
        return brp, bci[:nb], val[: nb * T]

    @staticmethod
    def nwords(T, nbr):
        return -(-nbr // T) * T

    def pack_vec(self, T, x, nbr):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(max(self.nwords(T, nbr), 1), WORD_NP[T])
        self.L.orc_pack_vec(T, _p(x), x.size, self.nwords(T, nbr), _p(out))
        return out[: self.nwords(T, nbr)]

    def pack_u8(self, T, bits, nwords):
        bits = np.ascontiguousarray(bits, np.uint8)
        out = np.zeros(max(nwords, 1), WORD_NP[T])
        self.L.orc_pack_u8(T, _p(bits), bits.size, nwords, _p(out))
        return out[:nwords]

    # -- B2SR tile loops -----------------------------------------------------
    def b2sr_bin_bin_bin(self, T, val, xb, rowptr, colind, C=None):
        nbr = rowptr.size - 1
        C = np.zeros(nbr, WORD_NP[T]) if C is None else C.copy()
        self.L.orc_b2sr_bmv_bin_bin_bin(T, _p(val), _p(xb), _p(C), _p(rowptr), _p(colind), nbr)
        return C

    def b2sr_bin_bin_full(self, T, val, xb, rowptr, colind, C=None):
        nbr = rowptr.size - 1
        C = np.zeros(nbr * T, np.int32) if C is None else C.astype(np.int32).copy()
        self.L.orc_b2sr_bmv_bin_bin_full(T, _p(val), _p(xb), _p(C), _p(rowptr), _p(colind), nbr)
        return C

    def b2sr_bin_full_full(self, T, val, x, rowptr, colind, C=None):
        nbr = rowptr.size - 1
        x = np.ascontiguousarray(x, np.float32)
        C = np.zeros(nbr * T, np.float32) if C is None else C.astype(np.float32).copy()
        self.L.orc_b2sr_bmv_bin_full_full(T, _p(val), _p(x), _p(C), _p(rowptr), _p(colind), nbr)
        return C

    def b2sr_masked(self, T, val, xb, rowptr, colind, mask):
        nbr = rowptr.size - 1
        C = np.zeros(nbr, WORD_NP[T])
        mask = np.ascontiguousarray(mask, WORD_NP[T])
        self.L.orc_b2sr_bmv_masked(T, _p(val), _p(xb), _p(C), _p(rowptr), _p(colind), nbr, _p(mask))
        return C

    # -- CSR semirings -------------------------------------------------------
    def csr_count(self, rowptr, colind, x01):
        n = rowptr.size - 1
        x01 = np.ascontiguousarray(x01, np.uint8)
        y = np.zeros(n, np.int32)
        self.L.orc_csr_spmv_count(n, _p(rowptr), _p(colind), _p(x01), _p(y))
        return y

    def csr_or(self, rowptr, colind, x01):
        n = rowptr.size - 1
        x01 = np.ascontiguousarray(x01, np.uint8)
        y = np.zeros(n, np.uint8)
        self.L.orc_csr_spmv_or(n, _p(rowptr), _p(colind), _p(x01), _p(y))
        return y

    def csr_f32(self, rowptr, colind, x):
        n = rowptr.size - 1
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(n, np.float32)
        self.L.orc_csr_spmv_f32(n, _p(rowptr), _p(colind), _p(x), _p(y))
        return y

    def csr_f64(self, rowptr, colind, x):
        n = rowptr.size - 1
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(n, np.float64)
        self.L.orc_csr_spmv_f64(n, _p(rowptr), _p(colind), _p(x), _p(y))
        return y

    def csr_spmm_f64(self, rowptr, colind, X):
        n = rowptr.size - 1
        X = np.ascontiguousarray(X, np.float64)
        Y = np.zeros((n, X.shape[1]), np.float64)
        self.L.orc_csr_spmm_f64(n, _p(rowptr), _p(colind), X.shape[1], _p(X), _p(Y))
        return Y

    def bfs(self, rowptr, colind, src):
        n = rowptr.size - 1
        lev = np.zeros(n, np.int32)
        nl = self.L.orc_bfs(n, _p(rowptr), _p(colind), int(src), _p(lev))
        return lev, nl

    def bfs_multi(self, rowptr, colind, srcs):
        """Queue BFS from every source (sources in parallel): levels[nsrc, n], max level count."""
        n = rowptr.size - 1
        srcs = np.ascontiguousarray(srcs, np.int32)
        lev = np.zeros((srcs.size, n), np.int32)
        nl = self.L.orc_bfs_multi(n, _p(rowptr), _p(colind), srcs.size, _p(srcs), _p(lev))
        return lev, nl

    def pagerank(self, rowptr, colind, damping, iters):
        n = rowptr.size - 1
        pr = np.zeros(n, np.float64)
        res = np.zeros(max(iters, 1), np.float64)
        self.L.orc_pagerank_f64(n, _p(rowptr), _p(colind), damping, iters, _p(pr), _p(res))
        return pr, res[:iters]

    # -- transposes / statistics -------------------------------------------
    def csr2csc(self, nrows, ncols, rowptr, colind, vals=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        nnz = int(rowptr[-1])
        colind = np.ascontiguousarray(colind if nnz else np.zeros(1), np.int32)
        vals = None if vals is None else np.ascontiguousarray(vals, np.float32)
        cp = np.zeros(ncols + 1, np.int32)
        ri = np.zeros(max(nnz, 1), np.int32)
        cv = None if vals is None else np.zeros(max(nnz, 1), np.float32)
        self.L.orc_csr2csc(nrows, ncols, nnz, _p(vals), _p(rowptr), _p(colind), _p(cv), _p(ri), _p(cp))
        return cp, ri[:nnz], (None if cv is None else cv[:nnz])

    def b2sr_transpose(self, T, nbc, rowptr, colind, val):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        nb = int(rowptr[-1])
        colind = np.ascontiguousarray(colind if nb else np.zeros(1), np.int32)
        val = np.ascontiguousarray(val if nb else np.zeros(T), WORD_NP[T])
        trp = np.zeros(nbc + 1, np.int32)
        tci = np.zeros(max(nb, 1), np.int32)
        tval = np.zeros(max(nb, 1) * T, WORD_NP[T])
        self.L.orc_b2sr_transpose(T, rowptr.size - 1, nbc, _p(rowptr), _p(colind), _p(val), _p(trp),
                                  _p(tci), _p(tval))
        return trp, tci[:nb], tval[: nb * T]

    def tile_row_stats(self, T, rowptr, val):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        nbr = rowptr.size - 1
        val = np.ascontiguousarray(val if rowptr[-1] else np.zeros(T), WORD_NP[T])
        tpr = np.zeros(max(nbr * T, 1), np.int32)
        nzr = np.zeros(max(nbr * T, 1), np.int32)
        self.L.orc_tile_row_stats(T, nbr, _p(rowptr), _p(val), _p(tpr), _p(nzr))
        return tpr[: nbr * T], nzr[: nbr * T]

    def estimate_tiles(self, rowptr, colind, T):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        colind = np.ascontiguousarray(colind if len(colind) else np.zeros(1), np.int32)
        return int(self.L.orc_estimate_tiles(rowptr.size - 1, _p(rowptr), _p(colind), T))

    def threads(self):
        return self.L.orc_num_threads()

    def set_threads(self, t):
        self.L.orc_set_threads(int(t))


class RefReadMtx:
    """The reference's own readMtx.hpp CSR construction, built into oracle/_ref."""

    def __init__(self, path=REF_SO):
        self.L = ctypes.CDLL(path)
        vp, ci, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        f = self.L.ref_normalize
        f.restype, f.argtypes = ci, [ci, ci, ll, vp, vp, ci, vp, vp, vp]

    def read_mtx(self, path, directed=0, cap=1 << 20):
        """The reference driver's path: readMtx (directed) + coo2csr."""
        f = self.L.ref_read_mtx
        vp, ci, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        f.restype = ci
        f.argtypes = [ctypes.c_char_p, ci, vp, vp, vp, vp, vp, ll]
        M, N, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        rp = np.zeros(cap + 1, np.int32)
        col = np.zeros(2 * cap + 1, np.int32)
        st = f(os.fsencode(path), int(directed), ctypes.byref(M), ctypes.byref(N), ctypes.byref(nnz), _p(rp),
               _p(col), cap)
        assert st == 0
        return M.value, N.value, rp[: M.value + 1].copy(), col[: nnz.value].copy()

    def normalize(self, nrows, ncols, src, dst, undirected=True):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        rp = np.zeros(nrows + 1, np.int32)
        ci = np.zeros(max(2 * src.size, 1), np.int32)
        nnz = ctypes.c_longlong()
        st = self.L.ref_normalize(nrows, ncols, src.size, _p(src), _p(dst), int(undirected), _p(rp),
                                  _p(ci), ctypes.byref(nnz))
        assert st == 0
        return rp, ci[: nnz.value].copy()


def trailing_entry_erased(ent, undirected):
    """True when the last entry of the reference's sorted list is a self-loop
    or a duplicate. removeSelfloop's compaction then scans two slots past the
    end of its column vector (readMtx.hpp:202-204) and drops two more entries
    than it erased (the oracle build fills those slots with -1, an erased
    entry, so the outcome is fixed); such inputs do not pin the reader."""
    full = list(ent) + ([(j, i) for i, j in ent if i != j] if undirected else [])
    full.sort()
    last = full[-1]
    return last[0] == last[1] or (len(full) > 1 and full[-2] == last)
