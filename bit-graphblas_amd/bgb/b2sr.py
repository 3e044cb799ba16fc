"""Device-side B2SR matrices and BMV operators (torch tensors as device memory).

Torch only allocates device memory and supplies the current HIP stream; every
computation runs in libbgb's HIP kernels through the C ABI.
"""
import ctypes

import numpy as np
import torch

from ._lib import call, lib

# Raw storage dtypes for the B2SR words (bit patterns; view as unsigned on host).
WORD_TORCH = {4: torch.uint8, 8: torch.uint8, 16: torch.int16, 32: torch.int32, 64: torch.int64}
WORD_NUMPY = {4: np.uint8, 8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
TILE_DIMS = (4, 8, 16, 32, 64)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return t.data_ptr() if t is not None else 0


def packed_vec_words(T, nblockrows):
    return int(lib().bgb_packed_vec_words(T, nblockrows))


class B2SR:
    """Bit-packed block sparse row matrix with tile dim T (the reference's layout)."""

    def __init__(self, T, nrows, ncols, rowptr, colind, val, nblocks):
        assert T in TILE_DIMS
        self.T, self.nrows, self.ncols = T, int(nrows), int(ncols)
        self.rowptr, self.colind, self.val = rowptr, colind, val
        self.nblocks = int(nblocks)

    @property
    def nblockrows(self):
        return (self.nrows + self.T - 1) // self.T

    @property
    def device(self):
        return self.rowptr.device

    def nbytes(self):
        """Bytes of rowptr + colind + tile words."""
        wb = lib().bgb_word_bytes(self.T)
        return 4 * (self.nblockrows + 1) + 4 * self.nblocks + self.nblocks * self.T * wb

    @classmethod
    def from_csr(cls, csr, T, device="cuda"):
        """Packs a host CSR (bgb.graph.HostCSR) on the GPU: nnzb pass then fill."""
        dev = torch.device(device)
        d_rp = torch.from_numpy(csr.rowptr).to(dev)
        d_ci = torch.from_numpy(csr.colind if csr.nnz else np.zeros(1, np.int32)).to(dev)
        return cls.from_device_csr(T, csr.nrows, csr.ncols, csr.nnz, d_rp, d_ci)

    @classmethod
    def from_device_csr(cls, T, nrows, ncols, nnz, d_rowptr, d_colind):
        dev = d_rowptr.device
        nbr = (nrows + T - 1) // T
        bsr_rp = torch.empty(nbr + 1, dtype=torch.int32, device=dev)
        nb = ctypes.c_int(0)
        call("bgb_b2sr_nnzb", T, nrows, ncols, nnz, _ptr(d_rowptr), _ptr(d_colind), _ptr(bsr_rp),
             ctypes.byref(nb), _stream())
        nblocks = nb.value
        # 16-byte padding keeps vector loads of the last tiles inside the allocation
        colind = torch.empty(max(nblocks, 1) + 4, dtype=torch.int32, device=dev)
        val = torch.zeros(max(nblocks, 1) * T + 16, dtype=WORD_TORCH[T], device=dev)
        call("bgb_csr2b2sr", T, nrows, ncols, nnz, _ptr(d_rowptr), _ptr(d_colind), _ptr(bsr_rp),
             nblocks, _ptr(colind), _ptr(val), _stream())
        return cls(T, nrows, ncols, bsr_rp, colind, val, nblocks)

    def to_host(self):
        """(rowptr, colind, val) as numpy arrays (val as unsigned words)."""
        rp = self.rowptr.cpu().numpy()
        ci = self.colind[: self.nblocks].cpu().numpy()
        v = self.val[: self.nblocks * self.T].cpu().numpy().view(WORD_NUMPY[self.T])
        return rp, ci, v

    def transpose(self):
        """The B2SR of A^T, built on the GPU from this one (bgb_b2sr_transpose)."""
        T, dev = self.T, self.device
        nbc = (self.ncols + T - 1) // T
        rp = torch.empty(nbc + 1, dtype=torch.int32, device=dev)
        colind = torch.empty(max(self.nblocks, 1) + 4, dtype=torch.int32, device=dev)
        val = torch.zeros(max(self.nblocks, 1) * T + 16, dtype=WORD_TORCH[T], device=dev)
        call("bgb_b2sr_transpose", T, self.nblockrows, nbc, self.nblocks, _ptr(self.rowptr), _ptr(self.colind),
             _ptr(self.val), _ptr(rp), _ptr(colind), _ptr(val), _stream())
        return B2SR(T, self.ncols, self.nrows, rp, colind, val, self.nblocks)

    def row_stats(self):
        """printMatrixStats_T: (tile_per_row, nnz_tile_per_row), int32 [NBR*T] each."""
        n = self.nblockrows * self.T
        tpr = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        nzr = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        call("bgb_tile_row_stats", self.T, _ptr(self.val), _ptr(self.rowptr), self.nblockrows, _ptr(tpr),
             _ptr(nzr), _stream())
        return tpr[:n], nzr[:n]


def csr2csc(m, n, nnz, d_rowptr, d_colind, d_val=None):
    """cusparseScsr2csc on the GPU: (colptr[n+1], rowind[nnz], val[nnz] or None)."""
    dev = d_rowptr.device
    colptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
    rowind = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    cval = None if d_val is None else torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)
    call("bgb_csr2csc", m, n, nnz, _ptr(d_val), _ptr(d_rowptr), _ptr(d_colind), _ptr(cval), _ptr(rowind),
         _ptr(colptr), _stream())
    return colptr, rowind[:nnz], (None if cval is None else cval[:nnz])


def estimate_tiles(T, nrows, d_rowptr, d_colind):
    """estimatevec on the GPU: runs of equal colind / T summed over rows."""
    out = torch.empty(1, dtype=torch.int64, device=d_rowptr.device)
    call("bgb_estimate_tiles", T, nrows, _ptr(d_rowptr), _ptr(d_colind), _ptr(out), _stream())
    return int(out.item())


def pack_vec(x, T, nblockrows=None):
    """ToBit{T}Row: float vector -> packed words (length ceil(NBR/T)*T)."""
    assert x.dtype == torch.float32 and x.is_cuda
    n = x.numel()
    nbr = (n + T - 1) // T if nblockrows is None else nblockrows
    words = packed_vec_words(T, nbr)
    xb = torch.zeros(max(words, 1), dtype=WORD_TORCH[T], device=x.device)
    call("bgb_pack_vec", T, _ptr(x), n, _ptr(xb), _stream())
    return xb


def _out(out, n, dtype, device, accumulate):
    if out is None:
        return torch.zeros(n, dtype=dtype, device=device)
    assert out.numel() >= n and out.dtype == dtype
    return out


def _bits_out(out, A):
    """Bit-output words: exactly NBR of them (nothing past the last is touched)."""
    return _out(out, A.nblockrows, WORD_TORCH[A.T], A.device, True)


def bmv_bin_bin_bin(A, xb, out=None, accumulate=False):
    C = _bits_out(out, A)
    call("bgb_bmv_bin_bin_bin", A.T, _ptr(A.val), _ptr(xb), _ptr(C), _ptr(A.rowptr), _ptr(A.colind),
         A.nblockrows, A.nblocks, int(accumulate), _stream())
    return C


def bmv_bin_bin_full(A, xb, out=None, accumulate=False, dtype=torch.int32, noalloc=False):
    """Neighbour counts; noalloc=True calls the allocation-free int32 entry point."""
    C = _out(out, A.nblockrows * A.T, dtype, A.device, accumulate)
    fn = "bgb_bmv_bin_bin_full_i32" if dtype == torch.int32 else "bgb_bmv_bin_bin_full_f32"
    if noalloc:
        assert dtype == torch.int32
        fn = "bgb_bmv_bin_bin_full_i32_noalloc"
    call(fn, A.T, _ptr(A.val), _ptr(xb), _ptr(C), _ptr(A.rowptr), _ptr(A.colind), A.nblockrows,
         A.nblocks, int(accumulate), _stream())
    return C


def bmv_bin_full_full(A, x, out=None, accumulate=False):
    assert x.dtype == torch.float32 and x.numel() >= A.nblockrows * A.T
    C = _out(out, A.nblockrows * A.T, torch.float32, A.device, accumulate)
    call("bgb_bmv_bin_full_full", A.T, _ptr(A.val), _ptr(x), _ptr(C), _ptr(A.rowptr), _ptr(A.colind),
         A.nblockrows, A.nblocks, int(accumulate), _stream())
    return C


def bmv_bin_bin_bin_masked(A, xb, mask, out=None):
    C = _bits_out(out, A)
    call("bgb_bmv_bin_bin_bin_masked", A.T, _ptr(A.val), _ptr(xb), _ptr(C), _ptr(A.rowptr),
         _ptr(A.colind), A.nblockrows, A.nblocks, _ptr(mask), _stream())
    return C


def padded_vector(x, T, device="cuda"):
    """x (host, float32) zero-padded to NBR*T on the device (the drivers' fB)."""
    n = x.shape[0]
    nbr = (n + T - 1) // T
    out = torch.zeros(nbr * T, dtype=torch.float32, device=device)
    out[:n] = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    return out


def bmv_bin_full_full_bf16(A, x_bf16, out=None, accumulate=False):
    """bin_full_full with x given as raw bfloat16 bits (torch.bfloat16 or int16)."""
    assert x_bf16.element_size() == 2 and x_bf16.numel() >= A.nblockrows * A.T
    C = _out(out, A.nblockrows * A.T, torch.float32, A.device, accumulate)
    call("bgb_bmv_bin_full_full_bf16", A.T, _ptr(A.val), _ptr(x_bf16), _ptr(C), _ptr(A.rowptr),
         _ptr(A.colind), A.nblockrows, A.nblocks, int(accumulate), _stream())
    return C


def bfs_multi(A, sources, use_graph=True, push=None):
    """Batched BFS from up to 64 sources; returns (levels[nsrc, n] int32 tensor, nlevels).

    push: the B2SR of A^T (A itself when A is symmetric) to switch small-frontier
    levels to push (bgb_bfs_multi_dir); None pulls every level."""
    src = np.ascontiguousarray(sources, dtype=np.int32)
    n = A.nrows
    lev = torch.empty((src.size, n), dtype=torch.int32, device=A.device)
    nl = ctypes.c_int(0)
    if push is None:
        call("bgb_bfs_multi", A.T, _ptr(A.val), _ptr(A.rowptr), _ptr(A.colind), A.nblockrows, A.nblocks, n,
             src.ctypes.data, src.size, _ptr(lev), int(use_graph), ctypes.byref(nl), _stream())
    else:
        assert push.T == A.T and push.nblockrows == A.nblockrows
        call("bgb_bfs_multi_dir", A.T, _ptr(A.val), _ptr(A.rowptr), _ptr(A.colind), A.nblockrows, A.nblocks,
             _ptr(push.val), _ptr(push.rowptr), _ptr(push.colind), n, src.ctypes.data, src.size, _ptr(lev),
             int(use_graph), ctypes.byref(nl), _stream())
    return lev, nl.value


def autotune_measure(csr, op="bin_bin", allowed=(4, 8, 16, 32), reps=10, device="cuda"):
    """Measured tuning on the GPU (bgb_autotune_measure) for a host CSR:
    {"T", "run_length", "x_staging", "times_us": {(T, S, placement): us}}."""
    dev = torch.device(device)
    d_rp = torch.from_numpy(csr.rowptr).to(dev)
    d_ci = torch.from_numpy(csr.colind if csr.nnz else np.zeros(1, np.int32)).to(dev)
    times = np.zeros(16, np.float64)
    t, rl, xs = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    mask = sum(1 << i for i, T in enumerate((4, 8, 16, 32)) if T in allowed)
    call("bgb_autotune_measure", {"bin_bin": 0, "bin_full_full": 1}[op], csr.nrows, csr.ncols, csr.nnz, _ptr(d_rp),
         _ptr(d_ci), mask, reps, ctypes.byref(t), ctypes.byref(rl), ctypes.byref(xs), times.ctypes.data, _stream())
    out = {}
    for i, T in enumerate((4, 8, 16, 32)):
        for si, S in enumerate((8, 16)):
            for xi, place in enumerate(("memory", "lds")):
                v = times[i * 4 + si * 2 + xi]
                if v > 0:
                    out[(T, S, place)] = float(v)
    return {"T": t.value, "run_length": rl.value, "x_staging": xs.value, "times_us": out}


def bfs_level_times(A, sources, push=None):
    """Runs bfs_multi with the eager level loop, one level at a time between HIP
    events: returns (levels, nlevels, [(ms, pulled)] per level)."""
    lib().bgb_set_bfs_level_timing(1)
    try:
        lev, nl = bfs_multi(A, sources, use_graph=False, push=push)
    finally:
        lib().bgb_set_bfs_level_timing(0)
    cap = 1 << 16
    ms = np.zeros(cap, np.float32)
    pulled = np.zeros(cap, np.int32)
    n = ctypes.c_int(0)
    call("bgb_bfs_level_times", cap, ms.ctypes.data, pulled.ctypes.data, ctypes.byref(n))
    k = min(n.value, cap)
    return lev, nl, [(float(ms[i]), bool(pulled[i])) for i in range(k)]


def bfs_plain(A, source, out=None):
    """Single-source BFS as one masked BMV per level; returns (levels[n], nlevels)."""
    n = A.nrows
    lev = torch.empty(n, dtype=torch.int32, device=A.device) if out is None else out
    nl = ctypes.c_int(0)
    call("bgb_bfs_plain", A.T, _ptr(A.val), _ptr(A.rowptr), _ptr(A.colind), A.nblockrows, A.nblocks, n,
         int(source), _ptr(lev), ctypes.byref(nl), _stream())
    return lev, nl.value


PR_MODES = {"fp32": 0, "bf16": 1, "fp64sum": 2}


def pagerank(A, damping=0.85, iters=50, mode="fp32"):
    """PageRank via bin_full_full; returns (pr float32 tensor [n], residuals np.float64[iters])."""
    n = A.nrows
    pr = torch.empty(n, dtype=torch.float32, device=A.device)
    res = np.zeros(max(iters, 1), np.float64)
    call("bgb_pagerank", A.T, _ptr(A.val), _ptr(A.rowptr), _ptr(A.colind), A.nblockrows, A.nblocks, n,
         float(damping), int(iters), PR_MODES[mode], _ptr(pr), res.ctypes.data, _stream())
    return pr, res[:iters]


def k16_pack_rhs(X, nrows_pad=None):
    """X: torch.float16 [n, 16] on the device -> MFMA-ordered RHS for the K=16 kernels.

    nrows_pad (e.g. NBR*T of the matrix) sets how many rows the packed form covers."""
    assert X.dtype == torch.float16 and X.dim() == 2 and X.shape[1] == 16 and X.is_contiguous()
    n = X.shape[0]
    rows = max(n, nrows_pad or 0)
    Xb = torch.empty(-(-rows // 16) * 256, dtype=torch.float16, device=X.device)
    if rows > n:
        Xp = torch.zeros((rows, 16), dtype=torch.float16, device=X.device)
        Xp[:n] = X
        X, n = Xp, rows
    call("bgb_k16_pack_rhs", _ptr(X), n, _ptr(Xb), _stream())
    return Xb


def bmv_k16(A, Xb, out=None, accumulate=False):
    """Y[NBR*T, 16] (+)= A X with FP16 X (packed by k16_pack_rhs) on MFMA, FP32 accumulate."""
    assert A.T in (16, 32)
    n = A.nblockrows * A.T
    Y = _out(out, n * 16, torch.float32, A.device, accumulate).view(-1)
    call("bgb_bmv_bin_full_full_k16", A.T, _ptr(A.val), _ptr(Xb), _ptr(Y), _ptr(A.rowptr), _ptr(A.colind),
         A.nblockrows, A.nblocks, int(accumulate), _stream())
    return Y.view(n, 16)


class K16Graph(ctypes.Structure):
    _fields_ = [("T", ctypes.c_int), ("n", ctypes.c_int), ("nblocks", ctypes.c_int), ("reserved", ctypes.c_int),
                ("rowptr_off", ctypes.c_longlong), ("tile_off", ctypes.c_longlong),
                ("val_byte_off", ctypes.c_longlong), ("x_off", ctypes.c_longlong), ("y_off", ctypes.c_longlong)]


class K16Batch:
    """Many (B2SR, packed RHS) pairs concatenated for one persistent K=16 launch."""

    def __init__(self, mats, xbs, nworkgroups=0):
        assert len(mats) == len(xbs) and mats
        dev = mats[0].device
        descs = (K16Graph * len(mats))()
        rps, cis, vals, xs = [], [], [], []
        rp_off = ci_off = v_off = x_off = y_off = 0
        self.y_offsets = []
        for i, (A, Xb) in enumerate(zip(mats, xbs)):
            d = descs[i]
            d.T, d.n, d.nblocks = A.T, A.nrows, A.nblocks
            d.rowptr_off, d.tile_off, d.val_byte_off, d.x_off, d.y_off = rp_off, ci_off, v_off, x_off, y_off
            rps.append(A.rowptr)
            cis.append(A.colind[: max(A.nblocks, 1)])
            vb = A.val[: max(A.nblocks, 1) * A.T].view(torch.uint8)
            pad = (-vb.numel()) % 16
            vals.append(torch.cat([vb, torch.zeros(pad, dtype=torch.uint8, device=dev)]) if pad else vb)
            need = -(-(A.nblockrows * A.T) // 16) * 256
            assert Xb.numel() >= need
            xs.append(Xb[:need])
            self.y_offsets.append((y_off, A.nblockrows * A.T))
            rp_off += A.rowptr.numel()
            ci_off += cis[-1].numel()
            v_off += vals[-1].numel()
            x_off += need
            y_off += A.nblockrows * A.T * 16
        self.descs = descs
        self.rowptr = torch.cat(rps)
        self.colind = torch.cat(cis)
        self.val = torch.cat(vals)
        self.xb = torch.cat(xs)
        self.y = torch.empty(y_off, dtype=torch.float32, device=dev)
        self.nworkgroups = nworkgroups

    def run(self):
        call("bgb_k16_batched", len(self.descs), ctypes.addressof(self.descs), _ptr(self.rowptr),
             _ptr(self.colind), _ptr(self.val), _ptr(self.xb), _ptr(self.y), int(self.nworkgroups), _stream())
        return self.y

    def result(self, i):
        off, rows = self.y_offsets[i]
        return self.y[off: off + rows * 16].view(rows, 16)
