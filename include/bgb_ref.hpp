// bgb_ref.hpp -- C++ host mirror of the reference's operator interface.
//
// Same names, argument order and meaning as the reference's kernels and
// packers (bmv_eval/backend/bsrbmv.cu, csr2bsr_batch_bsrbmv.cu), so a driver
// written against the reference (test_bin_bin_full.cu and friends) ports by
// replacing `kernel<<<grid, block>>>(args)` with `kernel(args)`:
//   * launch geometry is chosen inside (the reference's cube grids are not
//     needed);
//   * every call runs on the default stream and accumulates into C exactly as
//     the reference's kernels do (|= for bits, += for counts / sums; the
//     masked variants overwrite);
//   * like the reference the functions return void; a failure is reported by
//     bgb_last_error() and bgb_ref::last_status() instead of being ignored.
// Each BMV reads nblocks = rowptr[nblockrows] back from the device (one small
// synchronous copy), because the reference's signature does not carry it; the
// C ABI in bgb.h takes nblocks directly and avoids that copy.
#pragma once

#include "bgb.h"

namespace bgb_ref {

typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned long long ullong;

// Status of the last call on this thread (BGB_OK on success).
int last_status();

// cusparseXcsr2bsrNnz on a device CSR (test_bin_bin_full.cu:199-200).
void csr2bsr_nnz(const int *csrRowPtr, const int *csrColInd, int nrows, int ncols, int nnz, int blocksize,
                 int *bsrRowPtr, int *nblocks);

// cusparseScsr2csc(handle, m, n, nnz, csrVal, csrRowPtr, csrColInd, cscVal,
// cscRowInd, cscColPtr, CUSPARSE_ACTION_NUMERIC, CUSPARSE_INDEX_BASE_ZERO) as
// transposeMtx calls it (test_bin_bin_bin.cu:110-136); device arrays. csrVal
// and cscVal may both be NULL (structure only).
void csr2csc(int m, int n, int nnz, const float *csrVal, const int *csrRowPtr, const int *csrColInd,
             float *cscVal, int *cscRowInd, int *cscColPtr);

// csr2bsr_batch_{4,8,16,32,64} (csr2bsr_batch_bsrbmv.cu:16,140,263,386,510): host
// CSR in, device B2SR out; bsrRowPtr already filled by the nnz pass.
void csr2bsr_batch_4(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz,
                     int *bsrRowPtr, int *bsrColInd, uchar *bsrVal, int blocksize, int nblockrows, int nblocks);
void csr2bsr_batch_8(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz,
                     int *bsrRowPtr, int *bsrColInd, uchar *bsrVal, int blocksize, int nblockrows, int nblocks);
void csr2bsr_batch_16(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz,
                      int *bsrRowPtr, int *bsrColInd, ushort *bsrVal, int blocksize, int nblockrows, int nblocks);
void csr2bsr_batch_32(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz,
                      int *bsrRowPtr, int *bsrColInd, unsigned *bsrVal, int blocksize, int nblockrows,
                      int nblocks);
void csr2bsr_batch_64(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz,
                      int *bsrRowPtr, int *bsrColInd, ullong *bsrVal, int blocksize, int nblockrows,
                      int nblocks);

// ToBit{T}Row (bsrbmv.cu:65,117,163,206,253): x padded to nblockrows*T floats ->
// ceil(nblockrows/T)*T packed words.
void ToBit4Row(const float *A, uchar *B, int nblockrows);
void ToBit8Row(const float *A, uchar *B, int nblockrows);
void ToBit16Row(const float *A, ushort *B, int nblockrows);
void ToBit32Row(const float *A, unsigned *B, int nblockrows);
void ToBit64Row(const float *A, ullong *B, int nblockrows);

// T = 64: the reference packs 64-bit tiles (ToBit64Col/Row, csr2bsr_batch_64)
// but has no BMV for them; these follow the same contract as the T <= 32 ones.
void bmv64_bin_bin_bin(const ullong *A, const ullong *B, ullong *C, const int *rowptr, const int *colind,
                       int nblockrows);
void bmv64_bin_bin_full(const ullong *A, const ullong *B, float *C, const int *rowptr, const int *colind,
                        int nblockrows);
void bmv64_bin_full_full(const ullong *A, const float *B, float *C, const int *rowptr, const int *colind,
                         int nblockrows);
void bmv64_bin_bin_bin_masked(const ullong *A, const ullong *B, ullong *C, const int *rowptr, const int *colind,
                              int nblockrows, const ullong *mask);

// bmv{T}_bin_bin_bin (bsrbmv.cu:277,437,508,579): C[I] |= bits.
void bmv4_bin_bin_bin(const uchar *A, const uchar *B, uchar *C, const int *rowptr, const int *colind,
                      int nblockrows);
void bmv8_bin_bin_bin(const uchar *A, const uchar *B, uchar *C, const int *rowptr, const int *colind,
                      int nblockrows);
void bmv16_bin_bin_bin(const ushort *A, const ushort *B, ushort *C, const int *rowptr, const int *colind,
                       int nblockrows);
void bmv32_bin_bin_bin(const unsigned *A, const unsigned *B, unsigned *C, const int *rowptr, const int *colind,
                       int nblockrows);

// bmv{T}_bin_bin_full (bsrbmv.cu:618,678,738,789): C[I*T+r] += count (float).
void bmv4_bin_bin_full(const uchar *A, const uchar *B, float *C, const int *rowptr, const int *colind,
                       int nblockrows);
void bmv8_bin_bin_full(const uchar *A, const uchar *B, float *C, const int *rowptr, const int *colind,
                       int nblockrows);
void bmv16_bin_bin_full(const ushort *A, const ushort *B, float *C, const int *rowptr, const int *colind,
                        int nblockrows);
void bmv32_bin_bin_full(const unsigned *A, const unsigned *B, float *C, const int *rowptr, const int *colind,
                        int nblockrows);

// bmv{T}_bin_full_full (bsrbmv.cu:826,899,975,1040): C[I*T+r] += sum.
void bmv4_bin_full_full(const uchar *A, const float *B, float *C, const int *rowptr, const int *colind,
                        int nblockrows);
void bmv8_bin_full_full(const uchar *A, const float *B, float *C, const int *rowptr, const int *colind,
                        int nblockrows);
void bmv16_bin_full_full(const ushort *A, const float *B, float *C, const int *rowptr, const int *colind,
                         int nblockrows);
void bmv32_bin_full_full(const unsigned *A, const float *B, float *C, const int *rowptr, const int *colind,
                         int nblockrows);

// bmv{T}_bin_bin_bin_masked (bsrbmv.cu:1086,1122,1158,1194): C[I] = bits & ~mask[I].
void bmv4_bin_bin_bin_masked(const uchar *A, const uchar *B, uchar *C, const int *rowptr, const int *colind,
                             int nblockrows, const uchar *mask);
void bmv8_bin_bin_bin_masked(const uchar *A, const uchar *B, uchar *C, const int *rowptr, const int *colind,
                             int nblockrows, const uchar *mask);
void bmv16_bin_bin_bin_masked(const ushort *A, const ushort *B, ushort *C, const int *rowptr, const int *colind,
                              int nblockrows, const ushort *mask);
void bmv32_bin_bin_bin_masked(const unsigned *A, const unsigned *B, unsigned *C, const int *rowptr,
                              const int *colind, int nblockrows, const unsigned *mask);

}  // namespace bgb_ref
