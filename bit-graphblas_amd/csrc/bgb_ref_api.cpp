// C++ host mirror of the reference's operator interface over the C ABI.
// See include/bgb_ref.hpp for the conventions.
#include <hip/hip_runtime_api.h>

#include "../../include/bgb_ref.hpp"
#include "bgb_internal.h"

namespace bgb_ref {
namespace {

thread_local int g_status = BGB_OK;

int record(int st) {
  g_status = st;
  return st;
}

int nblocks_of(const int *rowptr, int nblockrows) {
  int nb = 0;
  if (hipMemcpy(&nb, rowptr + nblockrows, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
    record(bgb::set_error(BGB_ERR_HIP, "reading rowptr[nblockrows] failed"));
    return -1;
  }
  return nb;
}

void batch(int T, const int *h_rowptr, const int *h_colind, int nrows, int ncols, int nnz, int *bsrRowPtr,
           int *bsrColInd, void *bsrVal, int blocksize, int nblocks) {
  if (blocksize != T) {
    record(bgb::set_error(BGB_ERR_TILEDIM, "blocksize %d != %d", blocksize, T));
    return;
  }
  int *d_rp = nullptr, *d_ci = nullptr;
  if (hipMalloc(&d_rp, sizeof(int) * (nrows + 1)) != hipSuccess ||
      hipMalloc(&d_ci, sizeof(int) * (nnz > 0 ? nnz : 1)) != hipSuccess) {
    record(bgb::set_error(BGB_ERR_ALLOC, "device CSR allocation failed"));
  } else if (hipMemcpy(d_rp, h_rowptr, sizeof(int) * (nrows + 1), hipMemcpyHostToDevice) != hipSuccess ||
             (nnz > 0 && hipMemcpy(d_ci, h_colind, sizeof(int) * nnz, hipMemcpyHostToDevice) != hipSuccess)) {
    record(bgb::set_error(BGB_ERR_HIP, "CSR upload failed"));
  } else {
    record(bgb_csr2b2sr(T, nrows, ncols, nnz, d_rp, d_ci, bsrRowPtr, nblocks, bsrColInd, bsrVal, nullptr));
  }
  if (d_rp) (void)hipFree(d_rp);
  if (d_ci) (void)hipFree(d_ci);
}

void tobit(int T, const float *A, void *B, int nblockrows) {
  record(bgb_pack_vec(T, A, nblockrows * T, B, nullptr));
}

void bbb(int T, const void *A, const void *B, void *C, const int *rp, const int *ci, int nbr) {
  const int nb = nblocks_of(rp, nbr);
  if (nb >= 0) record(bgb_bmv_bin_bin_bin(T, A, B, C, rp, ci, nbr, nb, 1, nullptr));
}
void bbf(int T, const void *A, const void *B, float *C, const int *rp, const int *ci, int nbr) {
  const int nb = nblocks_of(rp, nbr);
  if (nb >= 0) record(bgb_bmv_bin_bin_full_f32(T, A, B, C, rp, ci, nbr, nb, 1, nullptr));
}
void bff(int T, const void *A, const float *B, float *C, const int *rp, const int *ci, int nbr) {
  const int nb = nblocks_of(rp, nbr);
  if (nb >= 0) record(bgb_bmv_bin_full_full(T, A, B, C, rp, ci, nbr, nb, 1, nullptr));
}
void bbm(int T, const void *A, const void *B, void *C, const int *rp, const int *ci, int nbr, const void *mask) {
  const int nb = nblocks_of(rp, nbr);
  if (nb >= 0) record(bgb_bmv_bin_bin_bin_masked(T, A, B, C, rp, ci, nbr, nb, mask, nullptr));
}

}  // namespace

int last_status() { return g_status; }

void csr2bsr_nnz(const int *csrRowPtr, const int *csrColInd, int nrows, int ncols, int nnz, int blocksize,
                 int *bsrRowPtr, int *nblocks) {
  record(bgb_b2sr_nnzb(blocksize, nrows, ncols, nnz, csrRowPtr, csrColInd, bsrRowPtr, nblocks, nullptr));
}

void csr2csc(int m, int n, int nnz, const float *csrVal, const int *csrRowPtr, const int *csrColInd,
             float *cscVal, int *cscRowInd, int *cscColPtr) {
  record(bgb_csr2csc(m, n, nnz, csrVal, csrRowPtr, csrColInd, cscVal, cscRowInd, cscColPtr, nullptr));
}

void csr2bsr_batch_4(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz, int *bsrRowPtr,
                     int *bsrColInd, uchar *bsrVal, int blocksize, int, int nblocks) {
  batch(4, h_csrRowPtr, h_csrColInd, nrows, ncols, nnz, bsrRowPtr, bsrColInd, bsrVal, blocksize, nblocks);
}
void csr2bsr_batch_8(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz, int *bsrRowPtr,
                     int *bsrColInd, uchar *bsrVal, int blocksize, int, int nblocks) {
  batch(8, h_csrRowPtr, h_csrColInd, nrows, ncols, nnz, bsrRowPtr, bsrColInd, bsrVal, blocksize, nblocks);
}
void csr2bsr_batch_16(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz, int *bsrRowPtr,
                      int *bsrColInd, ushort *bsrVal, int blocksize, int, int nblocks) {
  batch(16, h_csrRowPtr, h_csrColInd, nrows, ncols, nnz, bsrRowPtr, bsrColInd, bsrVal, blocksize, nblocks);
}
void csr2bsr_batch_32(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz, int *bsrRowPtr,
                      int *bsrColInd, unsigned *bsrVal, int blocksize, int, int nblocks) {
  batch(32, h_csrRowPtr, h_csrColInd, nrows, ncols, nnz, bsrRowPtr, bsrColInd, bsrVal, blocksize, nblocks);
}

void csr2bsr_batch_64(const int *h_csrRowPtr, const int *h_csrColInd, int nrows, int ncols, int nnz, int *bsrRowPtr,
                      int *bsrColInd, ullong *bsrVal, int blocksize, int, int nblocks) {
  batch(64, h_csrRowPtr, h_csrColInd, nrows, ncols, nnz, bsrRowPtr, bsrColInd, bsrVal, blocksize, nblocks);
}

void ToBit4Row(const float *A, uchar *B, int nblockrows) { tobit(4, A, B, nblockrows); }
void ToBit8Row(const float *A, uchar *B, int nblockrows) { tobit(8, A, B, nblockrows); }
void ToBit16Row(const float *A, ushort *B, int nblockrows) { tobit(16, A, B, nblockrows); }
void ToBit32Row(const float *A, unsigned *B, int nblockrows) { tobit(32, A, B, nblockrows); }
void ToBit64Row(const float *A, ullong *B, int nblockrows) { tobit(64, A, B, nblockrows); }
void bmv64_bin_bin_bin(const ullong *A, const ullong *B, ullong *C, const int *rp, const int *ci, int nbr) { bbb(64, A, B, C, rp, ci, nbr); }
void bmv64_bin_bin_full(const ullong *A, const ullong *B, float *C, const int *rp, const int *ci, int nbr) { bbf(64, A, B, C, rp, ci, nbr); }
void bmv64_bin_full_full(const ullong *A, const float *B, float *C, const int *rp, const int *ci, int nbr) { bff(64, A, B, C, rp, ci, nbr); }
void bmv64_bin_bin_bin_masked(const ullong *A, const ullong *B, ullong *C, const int *rp, const int *ci, int nbr, const ullong *m) { bbm(64, A, B, C, rp, ci, nbr, m); }

void bmv4_bin_bin_bin(const uchar *A, const uchar *B, uchar *C, const int *rp, const int *ci, int nbr) { bbb(4, A, B, C, rp, ci, nbr); }
void bmv8_bin_bin_bin(const uchar *A, const uchar *B, uchar *C, const int *rp, const int *ci, int nbr) { bbb(8, A, B, C, rp, ci, nbr); }
void bmv16_bin_bin_bin(const ushort *A, const ushort *B, ushort *C, const int *rp, const int *ci, int nbr) { bbb(16, A, B, C, rp, ci, nbr); }
void bmv32_bin_bin_bin(const unsigned *A, const unsigned *B, unsigned *C, const int *rp, const int *ci, int nbr) { bbb(32, A, B, C, rp, ci, nbr); }

void bmv4_bin_bin_full(const uchar *A, const uchar *B, float *C, const int *rp, const int *ci, int nbr) { bbf(4, A, B, C, rp, ci, nbr); }
void bmv8_bin_bin_full(const uchar *A, const uchar *B, float *C, const int *rp, const int *ci, int nbr) { bbf(8, A, B, C, rp, ci, nbr); }
void bmv16_bin_bin_full(const ushort *A, const ushort *B, float *C, const int *rp, const int *ci, int nbr) { bbf(16, A, B, C, rp, ci, nbr); }
void bmv32_bin_bin_full(const unsigned *A, const unsigned *B, float *C, const int *rp, const int *ci, int nbr) { bbf(32, A, B, C, rp, ci, nbr); }

void bmv4_bin_full_full(const uchar *A, const float *B, float *C, const int *rp, const int *ci, int nbr) { bff(4, A, B, C, rp, ci, nbr); }
void bmv8_bin_full_full(const uchar *A, const float *B, float *C, const int *rp, const int *ci, int nbr) { bff(8, A, B, C, rp, ci, nbr); }
void bmv16_bin_full_full(const ushort *A, const float *B, float *C, const int *rp, const int *ci, int nbr) { bff(16, A, B, C, rp, ci, nbr); }
void bmv32_bin_full_full(const unsigned *A, const float *B, float *C, const int *rp, const int *ci, int nbr) { bff(32, A, B, C, rp, ci, nbr); }

void bmv4_bin_bin_bin_masked(const uchar *A, const uchar *B, uchar *C, const int *rp, const int *ci, int nbr, const uchar *m) { bbm(4, A, B, C, rp, ci, nbr, m); }
void bmv8_bin_bin_bin_masked(const uchar *A, const uchar *B, uchar *C, const int *rp, const int *ci, int nbr, const uchar *m) { bbm(8, A, B, C, rp, ci, nbr, m); }
void bmv16_bin_bin_bin_masked(const ushort *A, const ushort *B, ushort *C, const int *rp, const int *ci, int nbr, const ushort *m) { bbm(16, A, B, C, rp, ci, nbr, m); }
void bmv32_bin_bin_bin_masked(const unsigned *A, const unsigned *B, unsigned *C, const int *rp, const int *ci, int nbr, const unsigned *m) { bbm(32, A, B, C, rp, ci, nbr, m); }

}  // namespace bgb_ref
