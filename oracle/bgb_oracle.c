/*
 * bgb_oracle.c -- CPU oracle for the B2SR bit-packed BMV path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline; the product (libbgb) never calls it.
 *
 * Every function restates the reference's semantics for this path:
 *   orc_csr_normalize      readMtx.hpp:152-218 (removeSelfloop), :56-82
 *                          (customSort), :248-306 (coo2csr)
 *   orc_b2sr_nnzb/_fill    cusparseXcsr2bsrNnz + cusparseScsr2bsr (textbook BSR,
 *                          row-major dense tiles) + ToBit{T}Col bit order
 *                          (bsrbmv.cu:33-272; column c at bit T-1-c), as driven
 *                          by csr2bsr_batch_{4,8,16,32,64} (csr2bsr_batch_bsrbmv.cu:
 *                          34-133, 510-628; T = 64 words are ullong)
 *   orc_pack_vec           ToBit{T}Row (bsrbmv.cu:65-272) on x padded to NBR*T
 *   orc_b2sr_bmv_*         bmv{T}_bin_bin_bin / _bin_bin_full / _bin_full_full /
 *                          _bin_bin_bin_masked (bsrbmv.cu:277-1229), accumulate
 *                          semantics of the drivers (test_bin_bin_full.cu:311-353);
 *                          at T = 64 (the reference packs 64-bit tiles but has no
 *                          BMV for them) the same loops over 64-bit words
 *   orc_csr_*              textbook CSR SpMV over the (OR, AND), (+, x) and FP
 *                          semirings -- the "CPU CSR computation" the B2SR path
 *                          must match (the reference itself only compares with
 *                          cuSPARSE csrmv, test_bin_bin_bin.cu:284-323,428-466)
 *   orc_csr2csc            cusparseScsr2csc(..., CUSPARSE_ACTION_NUMERIC,
 *                          CUSPARSE_INDEX_BASE_ZERO) as used by transposeMtx
 *                          (test_bin_bin_bin.cu:110-136): counting sort by
 *                          column, row order kept inside each column
 *   orc_b2sr_transpose     B2SR of A^T from the B2SR of A: block (I, J) becomes
 *                          (J, I) with its T x T bit tile transposed (no
 *                          reference counterpart; equal to packing the CSR of A^T)
 *   orc_tile_row_stats     printMatrixStats_{4,8,16,32} (utility.cu:808-930):
 *                          per row of every block row, the block row's tile count
 *                          and the number of those tiles whose row word is non-zero
 *   orc_estimate_tiles     estimatevec (utility.cu:955-975): per CSR row, the
 *                          number of runs of equal colind/T, summed
 *   orc_bfs, orc_bfs_multi, orc_pagerank_f64, orc_csr_spmm_f64
 *                          build-only rows (multi-source BFS, PageRank, K-RHS)
 *                          with no reference counterpart: plain queue BFS (one
 *                          per source, sources in parallel), FP64 power
 *                          iteration and FP64 CSR SpMM.
 * How each is pinned (golden fixtures, dense numpy, scipy, the reference's own
 * readMtx.hpp built into oracle/_ref) is stated in the tests that use it.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int wbytes(int T) { return T == 64 ? 8 : T == 32 ? 4 : T == 16 ? 2 : 1; }

/* W_T words: u8 (T = 4, 8), u16 (16), u32 (32), u64 (64) */
static uint64_t get_word(const void *p, int T, long long i) {
  if (T == 64) return ((const uint64_t *)p)[i];
  if (T == 32) return ((const uint32_t *)p)[i];
  if (T == 16) return ((const uint16_t *)p)[i];
  return ((const uint8_t *)p)[i];
}
static void set_word(void *p, int T, long long i, uint64_t v) {
  if (T == 64) ((uint64_t *)p)[i] = v;
  else if (T == 32) ((uint32_t *)p)[i] = (uint32_t)v;
  else if (T == 16) ((uint16_t *)p)[i] = (uint16_t)v;
  else ((uint8_t *)p)[i] = (uint8_t)v;
}
static int popc(uint64_t v) { return __builtin_popcountll(v); }

/* ---------------- CSR normalisation (readMtx.hpp:152-306) ---------------- */
typedef struct { int r, c; } rc_t;
static int cmp_rc(const void *a, const void *b) {
  const rc_t *x = (const rc_t *)a, *y = (const rc_t *)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}

/* colind_out needs room for up to 2*nedges entries. Returns 0 or -1 on a bad index. */
int orc_csr_normalize(int nrows, int ncols, long long nedges, const int *src, const int *dst,
                      int undirected, int drop_self, int *rowptr_out, int *colind_out,
                      long long *nnz_out) {
  long long cap = undirected ? 2 * nedges : nedges, m = 0;
  rc_t *t = (rc_t *)malloc(sizeof(rc_t) * (cap > 0 ? cap : 1));
  for (long long e = 0; e < nedges; ++e) {
    if (src[e] < 0 || src[e] >= nrows || dst[e] < 0 || dst[e] >= ncols) {
      free(t);
      return -1;
    }
    t[m].r = src[e];
    t[m].c = dst[e];
    ++m;
    /* the reverse of every non-loop edge (readMtx.hpp:160-168) */
    if (undirected && src[e] != dst[e]) {
      t[m].r = dst[e];
      t[m].c = src[e];
      ++m;
    }
  }
  qsort(t, (size_t)m, sizeof(rc_t), cmp_rc);
  /* drop self-loops and repeated (row, col) pairs (readMtx.hpp:177-194) */
  long long k = 0;
  for (long long e = 0; e < m; ++e) {
    if (drop_self && t[e].r == t[e].c) continue;
    if (k > 0 && t[k - 1].r == t[e].r && t[k - 1].c == t[e].c) continue;
    t[k++] = t[e];
  }
  /* coo2csr (readMtx.hpp:248-306) */
  for (int i = 0; i <= nrows; ++i) rowptr_out[i] = 0;
  for (long long e = 0; e < k; ++e) rowptr_out[t[e].r + 1]++;
  for (int i = 0; i < nrows; ++i) rowptr_out[i + 1] += rowptr_out[i];
  for (long long e = 0; e < k; ++e) colind_out[e] = t[e].c;
  *nnz_out = k;
  free(t);
  return 0;
}

/* ---------------- CSR -> B2SR ---------------- */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* distinct block columns of block row I, ascending; returns count */
static int block_cols(int T, int nrows, const int *rowptr, const int *colind, int I, int *buf) {
  int n = 0;
  int r1 = (I + 1) * T < nrows ? (I + 1) * T : nrows;
  for (int i = I * T; i < r1; ++i)
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) buf[n++] = colind[k] / T;
  qsort(buf, (size_t)n, sizeof(int), cmp_int);
  int u = 0;
  for (int k = 0; k < n; ++k)
    if (u == 0 || buf[u - 1] != buf[k]) buf[u++] = buf[k];
  return u;
}

int orc_b2sr_nnzb(int T, int nrows, int ncols, const int *rowptr, const int *colind,
                  int *bsrRowPtr) {
  (void)ncols;
  int nbr = (nrows + T - 1) / T;
  int *buf = (int *)malloc(sizeof(int) * ((size_t)rowptr[nrows] + 1));
  bsrRowPtr[0] = 0;
  for (int I = 0; I < nbr; ++I)
    bsrRowPtr[I + 1] = bsrRowPtr[I] + block_cols(T, nrows, rowptr, colind, I, buf);
  free(buf);
  return bsrRowPtr[nbr];
}

/* Per tile: a dense row-major T x T FP32 tile (cusparseScsr2bsr,
 * CUSPARSE_DIRECTION_ROW), then ToBit{T}Col: word r has bit T-1-c set iff
 * tile[r*T + c] > 0. */
void orc_b2sr_fill(int T, int nrows, int ncols, const int *rowptr, const int *colind,
                   const int *bsrRowPtr, int *bsrColInd, void *bsrVal) {
  (void)ncols;
  int nbr = (nrows + T - 1) / T;
  int *buf = (int *)malloc(sizeof(int) * ((size_t)rowptr[nrows] + 1));
  float *dense = (float *)malloc(sizeof(float) * T * T);
  for (int I = 0; I < nbr; ++I) {
    int nb = block_cols(T, nrows, rowptr, colind, I, buf);
    for (int b = 0; b < nb; ++b) {
      int J = buf[b], k = bsrRowPtr[I] + b;
      bsrColInd[k] = J;
      memset(dense, 0, sizeof(float) * T * T);
      for (int r = 0; r < T; ++r) {
        int i = I * T + r;
        if (i >= nrows) break;
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
          if (colind[e] / T == J) dense[r * T + colind[e] % T] = 1.0f;
      }
      for (int r = 0; r < T; ++r) {
        uint64_t w = 0;
        for (int c = 0; c < T; ++c)
          if (dense[r * T + c] > 0) w |= 1ull << (T - 1 - c);
        set_word(bsrVal, T, (long long)k * T + r, w);
      }
    }
  }
  free(dense);
  free(buf);
}

/* ToBit{T}Row on x[0..n) zero-padded: nwords = ceil(NBR/T)*T words */
void orc_pack_vec(int T, const float *x, int n, long long nwords, void *xb) {
  for (long long J = 0; J < nwords; ++J) {
    uint64_t w = 0;
    for (int c = 0; c < T; ++c) {
      long long i = J * T + c;
      if (i < n && x[i] > 0) w |= 1ull << (T - 1 - c);
    }
    set_word(xb, T, J, w);
  }
}

/* packs a 0/1 byte per vertex the same way (for bit outputs) */
void orc_pack_u8(int T, const uint8_t *bits, int n, long long nwords, void *out) {
  for (long long J = 0; J < nwords; ++J) {
    uint64_t w = 0;
    for (int c = 0; c < T; ++c) {
      long long i = J * T + c;
      if (i < n && bits[i]) w |= 1ull << (T - 1 - c);
    }
    set_word(out, T, J, w);
  }
}

/* ---------------- B2SR tile loops (bsrbmv.cu:277-1229) ---------------- */
void orc_b2sr_bmv_bin_bin_bin(int T, const void *A, const void *B, void *C, const int *rowptr,
                              const int *colind, int nbr) {
#pragma omp parallel for schedule(dynamic, 64) /* block rows write disjoint outputs */
  for (int I = 0; I < nbr; ++I) {
    uint64_t out = 0;
    for (int r = 0; r < T; ++r) {
      int cnt = 0;
      for (int k = rowptr[I]; k < rowptr[I + 1]; ++k)
        cnt += popc(get_word(A, T, (long long)k * T + r) & get_word(B, T, colind[k]));
      if (cnt > 0) out |= 1ull << (T - 1 - r);
    }
    set_word(C, T, I, get_word(C, T, I) | out); /* C[I] |= (bsrbmv.cu:308,612) */
  }
}

void orc_b2sr_bmv_bin_bin_full(int T, const void *A, const void *B, int *C, const int *rowptr,
                               const int *colind, int nbr) {
#pragma omp parallel for schedule(dynamic, 64) /* block rows write disjoint outputs */
  for (int I = 0; I < nbr; ++I)
    for (int r = 0; r < T; ++r) {
      int cnt = 0;
      for (int k = rowptr[I]; k < rowptr[I + 1]; ++k)
        cnt += popc(get_word(A, T, (long long)k * T + r) & get_word(B, T, colind[k]));
      C[(long long)I * T + r] += cnt; /* C[I*T+r] += (bsrbmv.cu:674,819) */
    }
}

/* FP32, each row summed in ascending (tile, column) order */
void orc_b2sr_bmv_bin_full_full(int T, const void *A, const float *B, float *C, const int *rowptr,
                                const int *colind, int nbr) {
#pragma omp parallel for schedule(dynamic, 64) /* block rows write disjoint outputs */
  for (int I = 0; I < nbr; ++I)
    for (int r = 0; r < T; ++r) {
      float sum = 0.0f;
      for (int k = rowptr[I]; k < rowptr[I + 1]; ++k) {
        uint64_t w = get_word(A, T, (long long)k * T + r);
        for (int c = 0; c < T; ++c)
          if ((w >> (T - 1 - c)) & 1ull) sum += B[(long long)colind[k] * T + c];
      }
      C[(long long)I * T + r] += sum;
    }
}

/* C[I] = bits & ~mask[I] over all T bits (bsrbmv.cu:1086-1229, the T=8/16
 * shift/nibble defects corrected) */
void orc_b2sr_bmv_masked(int T, const void *A, const void *B, void *C, const int *rowptr,
                         const int *colind, int nbr, const void *mask) {
#pragma omp parallel for schedule(dynamic, 64) /* block rows write disjoint outputs */
  for (int I = 0; I < nbr; ++I) {
    uint64_t out = 0;
    for (int r = 0; r < T; ++r) {
      int cnt = 0;
      for (int k = rowptr[I]; k < rowptr[I + 1]; ++k)
        cnt += popc(get_word(A, T, (long long)k * T + r) & get_word(B, T, colind[k]));
      if (cnt > 0) out |= 1ull << (T - 1 - r);
    }
    set_word(C, T, I, out & ~get_word(mask, T, I));
  }
}

/* ---------------- CSR semirings ---------------- */
void orc_csr_spmv_count(int nrows, const int *rowptr, const int *colind, const uint8_t *x, int *y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nrows; ++i) {
    int s = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += x[colind[k]] != 0;
    y[i] = s;
  }
}

void orc_csr_spmv_or(int nrows, const int *rowptr, const int *colind, const uint8_t *x, uint8_t *y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nrows; ++i) {
    uint8_t s = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1] && !s; ++k) s = x[colind[k]] != 0;
    y[i] = s;
  }
}

void orc_csr_spmv_f32(int nrows, const int *rowptr, const int *colind, const float *x, float *y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nrows; ++i) {
    float s = 0.0f;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += x[colind[k]];
    y[i] = s;
  }
}

void orc_csr_spmv_f64(int nrows, const int *rowptr, const int *colind, const double *x, double *y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nrows; ++i) {
    double s = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += x[colind[k]];
    y[i] = s;
  }
}

/* Y[n x k] = A * X[n x k], row-major, FP64 */
void orc_csr_spmm_f64(int nrows, const int *rowptr, const int *colind, int kdim, const double *X,
                      double *Y) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nrows; ++i) {
    for (int j = 0; j < kdim; ++j) Y[(long long)i * kdim + j] = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
      for (int j = 0; j < kdim; ++j)
        Y[(long long)i * kdim + j] += X[(long long)colind[e] * kdim + j];
  }
}

/* ---------------- BFS / PageRank ---------------- */
/* level[v] = hop distance from src, -1 if unreached; returns the number of levels */
int orc_bfs(int n, const int *rowptr, const int *colind, int src, int *level) {
  int *q = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int v = 0; v < n; ++v) level[v] = -1;
  int head = 0, tail = 0, depth = 0;
  level[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    int v = q[head++];
    if (level[v] > depth) depth = level[v];
    for (int k = rowptr[v]; k < rowptr[v + 1]; ++k) {
      int u = colind[k];
      if (level[u] < 0) {
        level[u] = level[v] + 1;
        q[tail++] = u;
      }
    }
  }
  free(q);
  return depth + 1;
}

/* orc_bfs from each of nsrc sources, sources in parallel (OpenMP):
 * level[s*n + v] = hop distance of v from srcs[s]; returns the largest
 * number of levels over the sources. */
int orc_bfs_multi(int n, const int *rowptr, const int *colind, int nsrc, const int *srcs, int *level) {
  int most = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : most)
  for (int s = 0; s < nsrc; ++s) {
    const int nl = orc_bfs(n, rowptr, colind, srcs[s], level + (long long)s * n);
    if (nl > most) most = nl;
  }
  return most;
}

/* pr_{t+1}[v] = (1-d)/n + d * (sum_{u in N(v)} pr_t[u]/deg(u) + dangling_t/n),
 * pr_0 = 1/n; residual[t] = ||pr_{t+1} - pr_t||_1 (may be NULL). */
void orc_pagerank_f64(int n, const int *rowptr, const int *colind, double d, int iters, double *pr,
                      double *residual) {
  double *nx = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  double *y = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int v = 0; v < n; ++v) pr[v] = 1.0 / n;
  for (int it = 0; it < iters; ++it) {
    double dangling = 0.0;
    for (int v = 0; v < n; ++v) {
      int deg = rowptr[v + 1] - rowptr[v];
      if (deg == 0) dangling += pr[v];
      nx[v] = deg ? pr[v] / deg : 0.0;
    }
    orc_csr_spmv_f64(n, rowptr, colind, nx, y);
    double res = 0.0;
    for (int v = 0; v < n; ++v) {
      double p = (1.0 - d) / n + d * (y[v] + dangling / n);
      res += p > pr[v] ? p - pr[v] : pr[v] - p;
      pr[v] = p;
    }
    if (residual) residual[it] = res;
  }
  free(y);
  free(nx);
}

/* ---------------- transposes ---------------- */
/* cusparseScsr2csc semantics (zero base, numeric): cscColPtr[j] = entries with
 * column < j; entries of column j in ascending CSR position (= ascending row).
 * vals / cscVal may be NULL (structure only). */
void orc_csr2csc(int nrows, int ncols, long long nnz, const float *vals, const int *rowptr,
                 const int *colind, float *cscVal, int *cscRowInd, int *cscColPtr) {
  memset(cscColPtr, 0, sizeof(int) * ((size_t)ncols + 1));
  for (long long e = 0; e < nnz; ++e) cscColPtr[colind[e] + 1]++;
  for (int j = 0; j < ncols; ++j) cscColPtr[j + 1] += cscColPtr[j];
  int *next = (int *)malloc(sizeof(int) * ((size_t)ncols + 1));
  memcpy(next, cscColPtr, sizeof(int) * ((size_t)ncols + 1));
  for (int r = 0; r < nrows; ++r)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int p = next[colind[e]]++;
      cscRowInd[p] = r;
      if (vals && cscVal) cscVal[p] = vals[e];
    }
  free(next);
}

/* Transposed B2SR. nbc = block columns of A (= block rows of A^T). Bit c of
 * word r of tile (I, J) is A(I*T + r, J*T + c) (bit T-1-c); the transposed tile
 * (J, I) holds it as bit T-1-r of word c. */
void orc_b2sr_transpose(int T, int nbr, int nbc, const int *rowptr, const int *colind, const void *val,
                        int *t_rowptr, int *t_colind, void *t_val) {
  const long long nblocks = rowptr[nbr];
  memset(t_rowptr, 0, sizeof(int) * ((size_t)nbc + 1));
  for (long long k = 0; k < nblocks; ++k) t_rowptr[colind[k] + 1]++;
  for (int J = 0; J < nbc; ++J) t_rowptr[J + 1] += t_rowptr[J];
  int *next = (int *)malloc(sizeof(int) * ((size_t)nbc + 1));
  memcpy(next, t_rowptr, sizeof(int) * ((size_t)nbc + 1));
  for (int I = 0; I < nbr; ++I)
    for (int k = rowptr[I]; k < rowptr[I + 1]; ++k) {
      const int p = next[colind[k]]++;
      t_colind[p] = I;
      for (int c = 0; c < T; ++c) {
        uint64_t w = 0;
        for (int r = 0; r < T; ++r)
          if ((get_word(val, T, (long long)k * T + r) >> (T - 1 - c)) & 1) w |= 1ull << (T - 1 - r);
        set_word(t_val, T, (long long)p * T + c, w);
      }
    }
  free(next);
}

/* ---------------- tile statistics (utility.cu:808-975) ---------------- */
/* printMatrixStats_T: for row i of block row I, tile_per_row[I*T+i] = tiles in
 * block row I, nnz_tile_per_row[I*T+i] = those tiles whose word i is non-zero. */
void orc_tile_row_stats(int T, int nbr, const int *rowptr, const void *val, int *tile_per_row,
                        int *nnz_tile_per_row) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int I = 0; I < nbr; ++I) {
    const int load = rowptr[I + 1] - rowptr[I];
    for (int i = 0; i < T; ++i) {
      int nz = 0;
      for (int j = rowptr[I]; j < rowptr[I + 1]; ++j) nz += get_word(val, T, (long long)j * T + i) != 0;
      tile_per_row[(long long)I * T + i] = load;
      nnz_tile_per_row[(long long)I * T + i] = nz;
    }
  }
}

/* estimatevec: sum over CSR rows of the number of runs of equal colind / T (an
 * upper bound on the tiles: a block column is counted once per row). */
long long orc_estimate_tiles(int nrows, const int *rowptr, const int *colind, int T) {
  long long total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 1024)
  for (int r = 0; r < nrows; ++r) {
    int count = 0, last = -1;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int J = colind[e] / T;
      if (e == rowptr[r] || J != last) ++count;
      last = J;
    }
    total += count;
  }
  return total;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int t) { omp_set_num_threads(t); }
int orc_word_bytes(int T) { return wbytes(T); }
