/*
 * bgb.h — C ABI of the MI355X-native B2SR bit-packed binary SpMV (BMV) library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (BGB_OK == 0) and, when it touches the GPU, runs asynchronously on
 * the caller's HIP stream (`bgb_stream_t`, a hipStream_t; NULL = default stream).
 * Device pointers are prefixed `d_`, host pointers `h_`.
 *
 * B2SR layout (identical to the reference's, bsrbmv.cu:28-272 and
 * test_bin_bin_full.cu:49-59): tile dim T in {4, 8, 16, 32, 64},
 *   NBR          = ceil(nrows / T) block rows
 *   bsrRowPtr    int32[NBR + 1]
 *   bsrColInd    int32[nblocks]            ascending block column J per block row
 *   bsrVal       W_T[nblocks * T]          word k*T + r holds row r of tile k,
 *                                          column c at bit (T-1-c)
 *   W_T = uint8 (T=4, 8), uint16 (T=16), uint32 (T=32), uint64 (T=64); for T=4 the
 *   high nibble is 0. The reference packs T=64 (csr2bsr_batch_64, ToBit64Row/Col,
 *   csr2bsr_batch_bsrbmv.cu:510-628, bsrbmv.cu:224-272) but has no T=64 BMV; here
 *   every BMV operator, the vector packer and PageRank take T=64 too (the BFS
 *   drivers take T <= 32, the K=16 MFMA path T = 16, 32).
 * Packed bit vector (ToBit{T}Row, bsrbmv.cu:65-222): W_T[ceil(NBR/T)*T] words,
 *   word J holds x[J*T + c] > 0 at bit (T-1-c); for T=4 one nibble per byte.
 * Bit output of bin_bin_bin: W_T[NBR], word I holds row I*T + r at bit (T-1-r).
 * Count / FP32 outputs: [NBR * T] values, element I*T + r.
 */
#ifndef BGB_H
#define BGB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *bgb_stream_t; /* hipStream_t */

enum {
  BGB_OK = 0,
  BGB_ERR_ARG = 1,      /* null pointer, negative size, bad flag */
  BGB_ERR_TILEDIM = 2,  /* T not in {4, 8, 16, 32, 64}, or not one the entry point takes */
  BGB_ERR_HIP = 3,      /* a HIP runtime call failed */
  BGB_ERR_ALLOC = 4,    /* device / host allocation failed */
  BGB_ERR_SIZE = 5,     /* a size exceeds the int32 index range or a stated limit */
  BGB_ERR_FORMAT = 6    /* input structure invalid (non-monotone rowptr, ...) */
};

/* Library information / errors. bgb_last_error() is per host thread. */
int bgb_version(void);
const char *bgb_status_string(int status);
const char *bgb_last_error(void);
/* Number of W_T words of a packed vector for NBR block rows: ceil(NBR/T)*T
 * (the reference's allocation, test_bin_bin_full.cu:246-266). */
long long bgb_packed_vec_words(int T, int nblockrows);
/* sizeof(W_T) in bytes; 0 if T is invalid. */
int bgb_word_bytes(int T);

/* ------------------------------------------------------------------------ *
 * CSR -> B2SR packing (GPU).                                                 *
 * ------------------------------------------------------------------------ */

/* Block-row tile counts: fills d_bsrRowPtr[NBR+1] and *h_nblocks.
 * Replaces cusparseXcsr2bsrNnz (test_bin_bin_full.cu:199-200).
 * Input CSR must have ascending, duplicate-free column indices per row.
 * Synchronises `stream` (nblocks is returned to the host). */
int bgb_b2sr_nnzb(int T, int nrows, int ncols, int nnz, const int *d_csrRowPtr,
                  const int *d_csrColInd, int *d_bsrRowPtr, int *h_nblocks,
                  bgb_stream_t stream);

/* Fills d_bsrColInd[nblocks] and d_bsrVal[nblocks*T] given the row pointer
 * from bgb_b2sr_nnzb. Replaces csr2bsr_batch_{4,8,16,32}
 * (csr2bsr_batch_bsrbmv.cu:16,140,263,386) and the cusparseScsr2bsr +
 * ToBit{T}Col pipeline it drives (csr2bsr_batch_bsrbmv.cu:100-107). */
int bgb_csr2b2sr(int T, int nrows, int ncols, int nnz, const int *d_csrRowPtr,
                 const int *d_csrColInd, const int *d_bsrRowPtr, int nblocks,
                 int *d_bsrColInd, void *d_bsrVal, bgb_stream_t stream);

/* Packs a float vector into B2SR bit words: word J bit (T-1-c) = x[J*T+c] > 0.
 * Reads x[0..n) (entries at or beyond n count as 0) and writes
 * bgb_packed_vec_words(T, ceil(n/T)) words. Replaces ToBit{T}Row
 * (bsrbmv.cu:65,117,163,206). */
int bgb_pack_vec(int T, const float *d_x, int n, void *d_xb, bgb_stream_t stream);

/* ------------------------------------------------------------------------ *
 * Transposes (GPU). The reference transposes a directed graph's CSR with   *
 * cusparseScsr2csc before packing it (transposeMtx,                        *
 * test_bin_bin_bin.cu:110-136). Both calls synchronise `stream` once (after *
 * the structure check) and are deterministic.                              *
 * ------------------------------------------------------------------------ */

/* CSR (m x n) -> CSC, i.e. the CSR of the transpose: d_cscColPtr[n+1],
 * d_cscRowInd[nnz] (rows ascending inside each column), and d_cscVal[nnz]
 * when values are given (d_csrVal and d_cscVal both NULL = structure only).
 * Argument order of cusparseScsr2csc (CUSPARSE_ACTION_NUMERIC,
 * CUSPARSE_INDEX_BASE_ZERO). Rows need not be sorted; BGB_ERR_FORMAT if
 * rowptr is not monotone from 0 to nnz or an index is outside [0, n). */
int bgb_csr2csc(int m, int n, int nnz, const float *d_csrVal, const int *d_csrRowPtr,
                const int *d_csrColInd, float *d_cscVal, int *d_cscRowInd, int *d_cscColPtr,
                bgb_stream_t stream);

/* B2SR of A^T from the B2SR of A (no reference counterpart): block (I, J)
 * becomes block (J, I) with its T x T bit tile transposed. nblockcols =
 * ceil(ncols / T) of A = the block rows of A^T; outputs d_t_rowptr
 * [nblockcols+1], d_t_colind[nblocks], d_t_val[nblocks*T] -- identical to
 * packing the CSR of A^T with bgb_csr2b2sr. */
int bgb_b2sr_transpose(int T, int nblockrows, int nblockcols, int nblocks, const int *d_rowptr,
                       const int *d_colind, const void *d_val, int *d_t_rowptr,
                       int *d_t_colind, void *d_t_val, bgb_stream_t stream);

/* ------------------------------------------------------------------------ *
 * BMV (GPU). A = bsrVal, rowptr/colind = bsrRowPtr/bsrColInd, nblocks =     *
 * rowptr[nblockrows] (passed so that no device read is needed).             *
 * accumulate = 1 reproduces the reference's in-place accumulation           *
 * (|= for bits, += for counts and sums); 0 overwrites C.                    *
 * Bit outputs are exactly W_T[NBR] (the reference's allocation,            *
 * test_bin_bin_bin.cu:329): no byte past C[NBR-1] is read or written.      *
 * Outputs are 4-byte aligned (count / FP32 outputs 16-byte, T=64 bit       *
 * outputs and masks 8-byte); hipMalloc gives 256.                          *
 * x: a packed word (or T floats) for every block column the matrix         *
 * references -- ceil(NBR/T)*T words for a square matrix, as the reference  *
 * allocates them; a wide matrix (more block columns than block rows, e.g.  *
 * the transpose of a tall one) needs one per block column.                 *
 * Calls on one stream are ordered; calls that overlap on different streams *
 * use separate per-stream carry workspaces. These entry points keep a      *
 * per-stream workspace and per-matrix facts (chunk partition, longest      *
 * block row, largest block column -- the latter two measured on the first  *
 * call for a matrix, which synchronises the stream), so they return        *
 * BGB_ERR_ARG while their stream is being captured into a graph;           *
 * bgb_bmv_bin_bin_full_i32_noalloc is the capturable form.                 *
 * ------------------------------------------------------------------------ */

/* Tiles each lane group walks per chunk (8 or 16) for tile dim T; process-wide
 * tuning knob of the BMV kernels (no reference counterpart). Default 8. */
int bgb_set_run_length(int T, int S);
int bgb_get_run_length(int T);

/* Where the bit-vector BMVs (bin_bin_bin, bin_bin_full, masked) read the
 * packed x for tile dim T: 0 = automatic (the whole packed x is staged in LDS
 * by one persistent workgroup per CU when it fits -- about 140 KB of packed
 * x, i.e. n up to ~1.1M -- and the tile stream is large next to it; a larger
 * x has its first words staged, and for the bin_bin_full row pieces also a
 * bitmap of which of the other words are non-zero, so that only those are
 * read from memory), 1 = always gathered from memory, 2 = staged whenever it
 * fits. Process-wide tuning knob (no reference counterpart). Default 0. */
int bgb_set_x_staging(int T, int mode);
int bgb_get_x_staging(int T);

/* Short-row path for T = 4 and 8, every BMV op: one lane per block row walks
 * that row's tiles and writes it once (no chunk partition, carries or
 * fix-up). 0 = automatic: taken when no block row holds more than 32 tiles
 * (regular matrices such as 2-D grids); the figure is measured on the first
 * call for a (rowptr, nblockrows, nblocks) -- one reduction and a stream
 * synchronisation -- and kept. 1 = never, 2 =
 * always (exact for any matrix, slow for long rows). Process-wide tuning
 * knob (no reference counterpart; the reference's kernels are one warp per
 * 1-8 block rows, bsrbmv.cu:298-303). Default 0. */
int bgb_set_short_row_path(int mode);
int bgb_get_short_row_path(void);

/* Row pieces for bgb_bmv_bin_bin_full_i32 / _f32 at T = 4 and 8: block rows of
 * at most 32 tiles are counted by one lane each, longer rows in pieces of 256
 * tiles by one wave each (whole 16-byte units per lane, a DPP wave sum per
 * piece), with no row ends inside a piece; the pieces that do not end a row
 * leave carry records for the fix-up. The plan (pieces per block row) is
 * made on the first call for a (rowptr, nblockrows, nblocks, T) -- a count,
 * a scan and one stream synchronisation -- and kept; the kernels read the
 * row extents from rowptr itself, so a rowptr rewritten in place still gives
 * exact results. 0 = automatic: taken when the whole packed x, or its first
 * words and the non-zero bitmap of the others, fit in LDS and the pieces are
 * well filled (otherwise the chunk walk); 1 = never,
 * 2 = always. Process-wide tuning knob (no reference counterpart). Default 0.
 * Same results on either path. */
int bgb_set_row_pieces(int mode);
int bgb_get_row_pieces(void);

/* C[I] (|)= bits(T-1-r) = OR_c A[I*T+r][c] & x[c].  bmv{T}_bin_bin_bin
 * (bsrbmv.cu:277,437,508,579). B: packed bit vector, C: W_T[NBR]. */
int bgb_bmv_bin_bin_bin(int T, const void *d_A, const void *d_B, void *d_C,
                        const int *d_rowptr, const int *d_colind, int nblockrows,
                        int nblocks, int accumulate, bgb_stream_t stream);

/* C[I*T+r] (+)= popcount(A row & x), exact int32.  bmv{T}_bin_bin_full
 * (bsrbmv.cu:618,678,738,789) with an int32 output. */
int bgb_bmv_bin_bin_full_i32(int T, const void *d_A, const void *d_B, int *d_C,
                             const int *d_rowptr, const int *d_colind,
                             int nblockrows, int nblocks, int accumulate,
                             bgb_stream_t stream);

/* bgb_bmv_bin_bin_full_i32 without any allocation (no reference counterpart;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results. Its scratch (carry records and the chunk partition,
 * rebuilt on every call) is a static 64 MiB device array per GPU, held from
 * the library's load until the process ends: BGB_ERR_SIZE when a matrix
 * needs more (about 2.6 M tiles at T=32, 12 M at T=16, 90 M at T=4/8).
 * It calls no hipMalloc / hipFree, never synchronises and caches nothing, so
 * it always takes the chunked kernels with x gathered from memory. Every call
 * on a GPU shares that GPU's scratch and it keeps the last call's contents:
 * calls that overlap on one GPU (different streams, or graph replays in
 * flight at once) race and must not be issued. */
int bgb_bmv_bin_bin_full_i32_noalloc(int T, const void *d_A, const void *d_B, int *d_C,
                                     const int *d_rowptr, const int *d_colind, int nblockrows,
                                     int nblocks, int accumulate, bgb_stream_t stream);

/* Same counts stored as float, the reference's output type
 * (test_bin_bin_full.cu:62,314,335-346). Exact below 2^24. */
int bgb_bmv_bin_bin_full_f32(int T, const void *d_A, const void *d_B, float *d_C,
                             const int *d_rowptr, const int *d_colind,
                             int nblockrows, int nblocks, int accumulate,
                             bgb_stream_t stream);

/* C[I*T+r] (+)= sum over tiles (I,J), columns c of A[I*T+r][J*T+c] * x[J*T+c],
 * FP32; B: float[NBR*T]
 * (padded with zeros beyond nrows).  bmv{T}_bin_full_full
 * (bsrbmv.cu:826,899,975,1040). Each row is summed in ascending column order
 * inside a 64-lane chunk; rows that cross chunks are combined by a fix-up
 * pass, in a fixed order within each 64 chunks and with FP32 atomics across
 * them (the reference uses atomicAdd, bsrbmv.cu:895,971,1036). On the
 * short-row path (bgb_set_short_row_path) every row is summed by one lane in
 * the CSR order. */
int bgb_bmv_bin_full_full(int T, const void *d_A, const float *d_B, float *d_C,
                          const int *d_rowptr, const int *d_colind, int nblockrows,
                          int nblocks, int accumulate, bgb_stream_t stream);

/* bin_full_full with x stored as bfloat16 (raw uint16 bits), FP32
 * accumulation. No reference counterpart (the reference's x is FP32); used by
 * the BF16 PageRank variant to halve the gathered bytes of x. */
int bgb_bmv_bin_full_full_bf16(int T, const void *d_A, const uint16_t *d_B, float *d_C,
                               const int *d_rowptr, const int *d_colind, int nblockrows,
                               int nblocks, int accumulate, bgb_stream_t stream);

/* C[I] = bits & ~mask[I] (always overwrites).  bmv{T}_bin_bin_bin_masked
 * (bsrbmv.cu:1086,1122,1158,1194), with the reference's T=8/16 shift and
 * nibble-mask defects corrected (the mask applies to all T bits). */
int bgb_bmv_bin_bin_bin_masked(int T, const void *d_A, const void *d_B, void *d_C,
                               const int *d_rowptr, const int *d_colind,
                               int nblockrows, int nblocks, const void *d_mask,
                               bgb_stream_t stream);

/* ------------------------------------------------------------------------ *
 * K = 16 right-hand sides on MFMA (no reference counterpart; the reference's *
 * bin_full_full takes one FP32 vector). Tile dims 16 and 32. X is FP16      *
 * (raw uint16 bits), 0/1 tiles are exact in FP16, accumulation is FP32.     *
 * X must be finite: a product runs over every column of a 4-column group    *
 * that holds a set bit, so an Inf/NaN there becomes 0*Inf in the tile's     *
 * other rows.                                                                *
 * ------------------------------------------------------------------------ */

/* Permutes X (row-major FP16 [n][16]) into the MFMA operand order used by
 * the K=16 kernels: Xb[(b*64 + l)*4 + j] = X[16b + 4(l/16) + j][l % 16],
 * ceil(n/16)*256 values, zero beyond n. */
int bgb_k16_pack_rhs(const uint16_t *d_X, int n, void *d_Xb, bgb_stream_t stream);

/* Y[v][0..16) (+)= sum_u A[v][u] X[u][0..16), Y FP32 row-major
 * [nblockrows*T][16]. Xb from bgb_k16_pack_rhs over nblockrows*T rows. */
int bgb_bmv_bin_full_full_k16(int T, const void *d_A, const void *d_Xb, float *d_Y,
                              const int *d_rowptr, const int *d_colind, int nblockrows,
                              int nblocks, int accumulate, bgb_stream_t stream);

/* One graph of a batched K=16 launch: offsets into the shared arrays. */
typedef struct {
  int T;                  /* 16 or 32 */
  int n;                  /* vertices (nblockrows = ceil(n/T)) */
  int nblocks;
  int reserved;
  long long rowptr_off;   /* elements into d_rowptr (graph-local values) */
  long long tile_off;     /* elements into d_colind (graph-local values) */
  long long val_byte_off; /* bytes into d_val, multiple of 16 */
  long long x_off;        /* FP16 elements into d_Xb, multiple of 256 */
  long long y_off;        /* floats into d_Y */
} bgb_k16_graph;

/* Y = A X for every graph of the batch in one persistent launch: waves take
 * tile chunks of all graphs from one queue until it is drained, so graphs of
 * very different sizes share the machine. nworkgroups <= 0 picks a default.
 * Overwrites Y. Synchronises. */
int bgb_k16_batched(int ngraphs, const bgb_k16_graph *h_graphs, const int *d_rowptr,
                    const int *d_colind, const void *d_val, const void *d_Xb, float *d_Y,
                    int nworkgroups, bgb_stream_t stream);

/* ------------------------------------------------------------------------ *
 * BFS over a square B2SR matrix (no reference counterpart; the reference's  *
 * building block is the masked BMV above). Vertex v is reached from the     *
 * vertices u with A[v][u] = 1 (its row: pull), so a symmetric A gives the   *
 * undirected BFS and the transpose of a directed graph's matrix its forward *
 * BFS. n = number of vertices, nblockrows = ceil(n/T). Levels are int32 hop *
 * counts, -1 = unreached.                                                   *
 * ------------------------------------------------------------------------ */

/* Up to 64 sources at once: one uint64 frontier word per vertex, one pull
 * pass over the tiles per level. d_levels[s*n + v] = level of v from
 * h_sources[s]. use_graph = 1 captures the level loop in a hipGraph.
 * *h_nlevels = number of non-empty levels (max level + 1). Synchronises. */
int bgb_bfs_multi(int T, const void *d_A, const int *d_rowptr, const int *d_colind,
                  int nblockrows, int nblocks, int n, const int *h_sources, int nsrc,
                  int *d_levels, int use_graph, int *h_nlevels, bgb_stream_t stream);

/* bgb_bfs_multi with direction switching: d_At / d_rowptrT / d_colindT are
 * the B2SR of A^T (pass A's own arrays when A is symmetric; all NULL = pull
 * only, as bgb_bfs_multi). Levels whose frontier is cheap to push (the
 * first, and those whose frontier vertices' A^T block rows hold fewer than
 * nblocks/16 tiles in total) push along the rows of A^T with atomics instead
 * of pulling over every tile; the others pull. Same results as
 * bgb_bfs_multi. */
int bgb_bfs_multi_dir(int T, const void *d_A, const int *d_rowptr, const int *d_colind,
                      int nblockrows, int nblocks, const void *d_At, const int *d_rowptrT,
                      const int *d_colindT, int n, const int *h_sources, int nsrc, int *d_levels,
                      int use_graph, int *h_nlevels, bgb_stream_t stream);

/* Tuning knob (no reference counterpart): frontier gathers of the pull levels
 * at T <= 8. 0 = one gather per set bit of each tile row in turn (the loop
 * ending once the row is saturated); 1 = one gather per set column of a tile
 * (over the rows not yet saturated), OR-ed into every row holding that column.
 * Same results either way. Process-wide; default 1. */
int bgb_set_bfs_gather_mode(int mode);
int bgb_get_bfs_gather_mode(void);

/* Measurement aid (no reference counterpart): with on = 1, calls of
 * bgb_bfs_multi / bgb_bfs_multi_dir with use_graph = 0 run one level at a
 * time between two HIP events (and read the level flags before each level,
 * which synchronises the stream per level). Process-wide; default 0. */
int bgb_set_bfs_level_timing(int on);
/* The levels of the last timed call on this host thread: h_ms[i] = GPU time
 * of level i in milliseconds, h_pulled[i] = 1 for a pull pass over the tiles,
 * 0 for a push; *h_n = number of levels (at most cap are written). */
int bgb_bfs_level_times(int cap, float *h_ms, int *h_pulled, int *h_n);

/* One source: one masked bin_bin_bin BMV per level (frontier & ~visited),
 * the per-source loop the batched BFS is compared with. Synchronises once
 * per level. */
int bgb_bfs_plain(int T, const void *d_A, const int *d_rowptr, const int *d_colind,
                  int nblockrows, int nblocks, int n, int source, int *d_level,
                  int *h_nlevels, bgb_stream_t stream);

/* PageRank with bin_full_full (no reference counterpart): pr_0 = 1/n,
 * pr_{t+1} = (1-d)/n + d*(A (pr_t/deg) + dangling_t/n), `iters` iterations.
 * mode 0: FP32 throughout; 1: x = pr/deg stored as bfloat16, FP32 accumulate;
 * 2: FP32 vector, FP64 residual and dangling/teleport sums. d_pr[n] receives
 * the result; h_residuals[iters] (may be NULL) the L1 change per iteration.
 * Synchronises. */
int bgb_pagerank(int T, const void *d_A, const int *d_rowptr, const int *d_colind,
                 int nblockrows, int nblocks, int n, float damping, int iters, int mode,
                 float *d_pr, double *h_residuals, bgb_stream_t stream);

/* ------------------------------------------------------------------------ *
 * Tile-dim autotuner (host). The reference fixes T at compile time         *
 * (Makefile:12-46); its tile statistics are debug kernels                  *
 * (printMatrixStats_T, estimatevec, utility.cu:808-975).                   *
 * ------------------------------------------------------------------------ */

/* Exact number of B2SR tiles (and of non-empty block rows) of a host CSR for
 * T = 4, 8, 16, 32 (arrays of 4, in that order; h_nonempty4 may be NULL). */
int bgb_tile_stats(int nrows, int ncols, const int *h_rowptr, const int *h_colind,
                   long long *h_nblocks4, long long *h_nonempty4);

/* GPU tile statistics of a B2SR, as the reference's printMatrixStats_T
 * (utility.cu:808-930), for every T: for row i of block row I,
 * d_tile_per_row[I*T+i] = tiles in block row I and d_nnz_tile_per_row[I*T+i]
 * = those tiles whose word i is non-zero (arrays of nblockrows*T ints). */
int bgb_tile_row_stats(int T, const void *d_A, const int *d_rowptr, int nblockrows,
                       int *d_tile_per_row, int *d_nnz_tile_per_row, bgb_stream_t stream);

/* The reference's estimatevec (utility.cu:955-975) on a device CSR: *d_out =
 * sum over rows of the runs of equal colind / T. For rows with ascending
 * columns this is the number of non-zero row words of the B2SR (the sum of
 * d_nnz_tile_per_row) and an upper bound on its tiles. *d_out is overwritten
 * (the reference adds into a caller-zeroed int). */
int bgb_estimate_tiles(int T, int nrows, const int *d_rowptr, const int *d_colind,
                       long long *d_out, bgb_stream_t stream);

/* Tile profile of a host CSR for T = 4, 8, 16, 32 (index 0..3): the
 * quantities the reference's printMatrixStats_T and estimatevec report
 * (utility.cu:808-975), summed over the matrix, and the longest block row.
 * These are the autotuner's features. */
typedef struct {
  long long tiles[4];          /* B2SR tiles (nblocks) */
  long long nonempty_rows[4];  /* block rows holding at least one tile */
  long long max_row_tiles[4];  /* most tiles in one block row */
  long long row_words[4];      /* non-zero tile row words: sum of nnz_tile_per_row
                                  = estimatevec's count for sorted rows */
  long long nnz;
  int nrows, ncols;
} bgb_tile_profile;

int bgb_tile_profile_csr(int nrows, int ncols, const int *h_rowptr, const int *h_colind,
                         bgb_tile_profile *out);

/* The kernel configuration the library would run for an op family (0:
 * bin_bin_*, 1: bin_full_full, 2: K=16 MFMA, 3: a pull level of the batched
 * BFS) on a matrix with this profile: T (allowed_mask bit t enables T = 4 << t)
 * with the lowest predicted time, the run length per lane group and the
 * packed-x placement (1 = gathered from memory, 2 = staged in LDS). The model
 * predicts each T from the kernel path it would take -- short-row kernel
 * (T <= 8 and no block row over 32 tiles), LDS-staged or gathered x, FP
 * set-bit walk -- as max(bytes / bandwidth, per-tile + per-row-word +
 * per-block-row issue cost), constants fitted to MI355X runs. cost4_us
 * (may be NULL) receives the predictions. */
int bgb_autotune_profile(int op, const bgb_tile_profile *prof, int allowed_mask, int *T_out,
                         int *run_length_out, int *x_staging_out, double *cost4_us);

/* Measured tuning (no reference counterpart): packs the device CSR at every
 * allowed T (allowed_mask bit t: T = 4 << t) into temporary buffers and times
 * the op's BMV -- op 0: bin_bin_full (int32) with x Bernoulli(0.05), op 1:
 * bin_full_full with x uniform in [0, 1) -- for run lengths 8 and 16 and, for
 * op 0 at T <= 16, x gathered from memory and staged in LDS, with HIP events
 * over `reps` calls after a warm-up call on `stream`. Returns the fastest
 * configuration; h_times_us[16] (may be NULL) receives each time in
 * microseconds at [t*4 + s*2 + x] (t: T index, s: run length 8 / 16, x:
 * gathered / LDS), 0 where not run. Sets the process-wide knobs while it
 * runs and restores them; synchronises. */
int bgb_autotune_measure(int op, int nrows, int ncols, int nnz, const int *d_csrRowPtr,
                         const int *d_csrColInd, int allowed_mask, int reps, int *T_out,
                         int *run_length_out, int *x_staging_out, double *h_times_us,
                         bgb_stream_t stream);

/* Picks T for an op family (0: bin_bin_*, 1: bin_full_full, 2: K=16 MFMA,
 * 3: a level of the batched BFS)
 * from the tile counts with a cost model fitted to MI355X runs of these
 * kernels. allowed_mask bit t enables T = 4 << t. cost4_us (may be NULL)
 * receives the predicted microseconds per T. */
int bgb_autotune_tiledim(int op, int nrows, const long long *h_nblocks4, int allowed_mask,
                         int *T_out, double *cost4_us);

/* The whole per-graph kernel configuration for an op family: T (as
 * bgb_autotune_tiledim), the run length per lane group (bgb_set_run_length;
 * 8 for every measured case on MI355X) and the packed-x placement
 * (bgb_set_x_staging: 1 = gathered from memory, 2 = staged in LDS; bit ops
 * only, when x fits and T <= 16, else 1). Any output pointer may be NULL. */
int bgb_autotune(int op, int nrows, const long long *h_nblocks4, int allowed_mask, int *T_out,
                 int *run_length_out, int *x_staging_out, double *cost4_us);

/* ------------------------------------------------------------------------ *
 * Host graph generators and CSR normalisation (no reference counterpart    *
 * except the normalisation, readMtx.hpp:152-306: symmetrise, drop           *
 * self-loops and duplicates, sort by (row, col), COO -> CSR).                *
 * All random streams are counter-based splitmix64, so results do not depend *
 * on the thread count.                                                      *
 * ------------------------------------------------------------------------ */
typedef struct bgb_csr bgb_csr; /* opaque host CSR (int32 indices) */

int bgb_csr_from_coo(int nrows, int ncols, long long nedges, const int *h_src,
                     const int *h_dst, int undirected, int drop_self_loops,
                     bgb_csr **out);
/* MatrixMarket coordinate file (real / integer / pattern; values unused) ->
 * normalised CSR, as the reference's readMtx + coo2csr (readMtx.hpp:353-422,
 * test_bin_bin_full.cu:139-160): directed 0 = undirected iff the banner says
 * symmetric (the drivers' setting), 1 = directed, 2 = undirected; self-loops
 * and duplicates dropped. BGB_ERR_ARG if the file cannot be opened,
 * BGB_ERR_FORMAT for a bad banner / size line, a short file or an entry out
 * of range (where the reference exits or reads on). */
int bgb_read_mtx(const char *path, int directed, bgb_csr **out);
/* Erdos-Renyi G(n, m): exactly m distinct undirected edges, no self-loops. */
int bgb_gen_er(int n, long long m, uint64_t seed, bgb_csr **out);
/* R-MAT(scale, edge_factor*2^scale edges, a, b, c, d = 1-a-b-c), no noise or
 * permutation, symmetrised, self-loops and duplicates dropped. */
int bgb_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, bgb_csr **out);
/* width x height 4-neighbour grid, vertex id y*width + x. */
int bgb_gen_grid2d(int width, int height, bgb_csr **out);
int bgb_csr_nrows(const bgb_csr *g);
int bgb_csr_ncols(const bgb_csr *g);
long long bgb_csr_nnz(const bgb_csr *g);
const int *bgb_csr_rowptr(const bgb_csr *g);
const int *bgb_csr_colind(const bgb_csr *g);
void bgb_csr_free(bgb_csr *g);

/* Deterministic vectors: out[i] = (u_i < p) ? 1 : 0, or lo + (hi-lo)*u_i,
 * u_i = uniform [0,1) from (seed, i). */
int bgb_gen_bernoulli_f32(long long n, double p, uint64_t seed, float *h_out);
int bgb_gen_uniform_f32(long long n, double lo, double hi, uint64_t seed,
                        float *h_out);

#ifdef __cplusplus
}
#endif

#endif /* BGB_H */
