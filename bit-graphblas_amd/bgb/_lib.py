"""ctypes binding of libbgb's C ABI (include/bgb.h).

The shared library is built in-tree by ``make -C bit-graphblas_amd`` (or
``__graft_entry__.build()``) into ``bit-graphblas_amd/lib/libbgb.so``. Loading
fails loudly when it is missing: there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libbgb.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "bgb.h")

BGB_OK = 0
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_TILEDIM", 3: "ERR_HIP", 4: "ERR_ALLOC",
                5: "ERR_SIZE", 6: "ERR_FORMAT"}


class BgbError(RuntimeError):
    """A libbgb call returned a non-zero status."""

    def __init__(self, fn, status, detail):
        self.status = status
        super().__init__(f"{fn} -> {STATUS_NAMES.get(status, status)}: {detail}")


_c_int, _c_ll, _c_u64, _c_dbl = ctypes.c_int, ctypes.c_longlong, ctypes.c_uint64, ctypes.c_double
_vp = ctypes.c_void_p

# name -> (restype, argtypes). Pointers are passed as integers (device
# addresses from torch tensors or host addresses from numpy arrays).
SIGNATURES = {
    "bgb_version": (_c_int, []),
    "bgb_status_string": (ctypes.c_char_p, [_c_int]),
    "bgb_last_error": (ctypes.c_char_p, []),
    "bgb_packed_vec_words": (_c_ll, [_c_int, _c_int]),
    "bgb_word_bytes": (_c_int, [_c_int]),
    "bgb_b2sr_nnzb": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "bgb_csr2b2sr": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp]),
    "bgb_pack_vec": (_c_int, [_c_int, _vp, _c_int, _vp, _vp]),
    "bgb_csr2csc": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bgb_b2sr_transpose": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bgb_tile_row_stats": (_c_int, [_c_int, _vp, _vp, _c_int, _vp, _vp, _vp]),
    "bgb_estimate_tiles": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "bgb_set_run_length": (_c_int, [_c_int, _c_int]),
    "bgb_get_run_length": (_c_int, [_c_int]),
    "bgb_set_x_staging": (_c_int, [_c_int, _c_int]),
    "bgb_get_x_staging": (_c_int, [_c_int]),
    "bgb_set_short_row_path": (_c_int, [_c_int]),
    "bgb_set_row_pieces": (_c_int, [_c_int]),
    "bgb_get_row_pieces": (_c_int, []),
    "bgb_get_short_row_path": (_c_int, []),
    "bgb_bmv_bin_bin_bin": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_bmv_bin_bin_full_i32": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_bmv_bin_bin_full_i32_noalloc": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_bmv_bin_bin_full_f32": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_bmv_bin_full_full": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_bmv_bin_bin_bin_masked": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp]),
    "bgb_bmv_bin_full_full_bf16": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_pagerank": (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, ctypes.c_float, _c_int, _c_int,
                              _vp, _vp, _vp]),
    "bgb_k16_pack_rhs": (_c_int, [_vp, _c_int, _vp, _vp]),
    "bgb_bmv_bin_full_full_k16": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]),
    "bgb_k16_batched": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp]),
    "bgb_bfs_multi": (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int,
                               _vp, _vp]),
    "bgb_bfs_multi_dir": (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _c_int,
                                   _vp, _c_int, _vp, _vp]),
    "bgb_set_bfs_level_timing": (_c_int, [_c_int]),
    "bgb_set_bfs_gather_mode": (_c_int, [_c_int]),
    "bgb_get_bfs_gather_mode": (_c_int, []),
    "bgb_bfs_level_times": (_c_int, [_c_int, _vp, _vp, _vp]),
    "bgb_bfs_plain": (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "bgb_tile_stats": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "bgb_tile_profile_csr": (_c_int, [_c_int, _c_int, _vp, _vp, _vp]),
    "bgb_autotune_profile": (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp]),
    "bgb_autotune_measure": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                      _vp]),
    "bgb_autotune_tiledim": (_c_int, [_c_int, _c_int, _vp, _c_int, _vp, _vp]),
    "bgb_autotune": (_c_int, [_c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp]),
    "bgb_csr_from_coo": (_c_int, [_c_int, _c_int, _c_ll, _vp, _vp, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "bgb_read_mtx": (_c_int, [ctypes.c_char_p, _c_int, ctypes.POINTER(_vp)]),
    "bgb_gen_er": (_c_int, [_c_int, _c_ll, _c_u64, ctypes.POINTER(_vp)]),
    "bgb_gen_rmat": (_c_int, [_c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_u64, ctypes.POINTER(_vp)]),
    "bgb_gen_grid2d": (_c_int, [_c_int, _c_int, ctypes.POINTER(_vp)]),
    "bgb_csr_nrows": (_c_int, [_vp]),
    "bgb_csr_ncols": (_c_int, [_vp]),
    "bgb_csr_nnz": (_c_ll, [_vp]),
    "bgb_csr_rowptr": (_vp, [_vp]),
    "bgb_csr_colind": (_vp, [_vp]),
    "bgb_csr_free": (None, [_vp]),
    "bgb_gen_bernoulli_f32": (_c_int, [_c_ll, _c_dbl, _c_u64, _vp]),
    "bgb_gen_uniform_f32": (_c_int, [_c_ll, _c_dbl, _c_dbl, _c_u64, _vp]),
}

_lib = None


def lib():
    """The loaded libbgb (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libbgb not built: {LIB_PATH} is missing (run `make -C bit-graphblas_amd`)")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def call(name, *args):
    """Calls a status-returning entry point and raises BgbError on failure."""
    st = getattr(lib(), name)(*args)
    if st != BGB_OK:
        detail = lib().bgb_last_error()
        raise BgbError(name, st, detail.decode() if detail else "")
    return st
