// Driver in the shape of the reference's test_bin_bin_full.cu /
// test_bin_bin_bin.cu / test_bin_full_full.cu (readMtxCSR -> CSR2HSBSF ->
// packRandVec -> evalHSBSFBmv with TEST_TIMES accumulating launches ->
// verifyResult), written against the bgb_ref mirror of the reference's
// operator names. Graphs come from the library's generators instead of .mtx
// files; the result is checked against the CPU oracle's CSR loop (x
// TEST_TIMES, since the launches accumulate). Test infrastructure.
//   bmv_driver <bin_bin_full|bin_bin_bin|bin_full_full> <T> <rmat:S:EF | er:N:DEG | grid:W:H>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "bgb.h"
#include "bgb_ref.hpp"

#define TEST_TIMES 5

extern "C" {  // oracle/liboracle.so
void orc_csr_spmv_count(int nrows, const int *rowptr, const int *colind, const uint8_t *x, int *y);
void orc_csr_spmv_f64(int nrows, const int *rowptr, const int *colind, const double *x, double *y);
}

static int nrows, ncols, nnz, tiledim, nblockrows, nblocks;
static std::vector<int> h_csrRowPtr, h_csrColInd;
static int *csrRowPtr, *csrColInd, *bsrRowPtr, *bsrColInd;
static void *tA, *tB;
static float *fB, *fC;
static std::vector<float> B;

static int fail(const char *what) {
  fprintf(stderr, "%s: %s\n", what, bgb_last_error());
  return 1;
}

static int readGraphCSR(const char *spec) {
  bgb_csr *g = nullptr;
  int a = 0, b = 0, st;
  if (sscanf(spec, "rmat:%d:%d", &a, &b) == 2) st = bgb_gen_rmat(a, b, 0.57, 0.19, 0.19, 1, &g);
  else if (sscanf(spec, "er:%d:%d", &a, &b) == 2) st = bgb_gen_er(a, (long long)a * b / 2, 1, &g);
  else if (sscanf(spec, "grid:%d:%d", &a, &b) == 2) st = bgb_gen_grid2d(a, b, &g);
  else if (strncmp(spec, "mtx:", 4) == 0) st = bgb_read_mtx(spec + 4, 0, &g);  // readMtx(..., directed = 0)
  else return 2;
  if (st) return fail("generator");
  nrows = bgb_csr_nrows(g);
  ncols = bgb_csr_ncols(g);
  nnz = (int)bgb_csr_nnz(g);
  h_csrRowPtr.assign(bgb_csr_rowptr(g), bgb_csr_rowptr(g) + nrows + 1);
  h_csrColInd.assign(bgb_csr_colind(g), bgb_csr_colind(g) + nnz);
  bgb_csr_free(g);
  hipMalloc(&csrRowPtr, sizeof(int) * (nrows + 1));
  hipMalloc(&csrColInd, sizeof(int) * (nnz > 0 ? nnz : 1));
  hipMemcpy(csrRowPtr, h_csrRowPtr.data(), sizeof(int) * (nrows + 1), hipMemcpyHostToDevice);
  hipMemcpy(csrColInd, h_csrColInd.data(), sizeof(int) * nnz, hipMemcpyHostToDevice);
  return 0;
}

// transposeMtx (test_bin_bin_bin.cu:110-136): csr2csc on the device, then the
// CSC arrays serve as the CSR (of A^T).
static int transposeMtx() {
  int *cscRowInd = nullptr, *cscColPtr = nullptr;
  hipMalloc(&cscRowInd, sizeof(int) * (nnz > 0 ? nnz : 1));
  hipMalloc(&cscColPtr, sizeof(int) * (ncols + 1));
  bgb_ref::csr2csc(nrows, ncols, nnz, nullptr, csrRowPtr, csrColInd, nullptr, cscRowInd, cscColPtr);
  if (bgb_ref::last_status()) return fail("csr2csc");
  std::swap(nrows, ncols);
  hipFree(csrRowPtr);
  hipFree(csrColInd);
  csrRowPtr = cscColPtr;
  csrColInd = cscRowInd;
  h_csrRowPtr.resize(nrows + 1);
  hipMemcpy(h_csrRowPtr.data(), csrRowPtr, sizeof(int) * (nrows + 1), hipMemcpyDeviceToHost);
  hipMemcpy(h_csrColInd.data(), csrColInd, sizeof(int) * nnz, hipMemcpyDeviceToHost);
  return 0;
}

static int CSR2HSBSF() {
  nblockrows = (nrows + tiledim - 1) / tiledim;
  hipMalloc(&bsrRowPtr, sizeof(int) * (nblockrows + 1));
  bgb_ref::csr2bsr_nnz(csrRowPtr, csrColInd, nrows, ncols, nnz, tiledim, bsrRowPtr, &nblocks);
  if (bgb_ref::last_status()) return fail("csr2bsr_nnz");
  const int wb = bgb_word_bytes(tiledim);
  hipMalloc(&bsrColInd, sizeof(int) * (nblocks > 0 ? nblocks : 1));
  hipMalloc(&tA, (size_t)(nblocks > 0 ? nblocks : 1) * tiledim * wb);
  switch (tiledim) {
    case 4: bgb_ref::csr2bsr_batch_4(h_csrRowPtr.data(), h_csrColInd.data(), nrows, ncols, nnz, bsrRowPtr, bsrColInd, (unsigned char *)tA, tiledim, nblockrows, nblocks); break;
    case 8: bgb_ref::csr2bsr_batch_8(h_csrRowPtr.data(), h_csrColInd.data(), nrows, ncols, nnz, bsrRowPtr, bsrColInd, (unsigned char *)tA, tiledim, nblockrows, nblocks); break;
    case 16: bgb_ref::csr2bsr_batch_16(h_csrRowPtr.data(), h_csrColInd.data(), nrows, ncols, nnz, bsrRowPtr, bsrColInd, (unsigned short *)tA, tiledim, nblockrows, nblocks); break;
    case 32: bgb_ref::csr2bsr_batch_32(h_csrRowPtr.data(), h_csrColInd.data(), nrows, ncols, nnz, bsrRowPtr, bsrColInd, (unsigned *)tA, tiledim, nblockrows, nblocks); break;
    default: bgb_ref::csr2bsr_batch_64(h_csrRowPtr.data(), h_csrColInd.data(), nrows, ncols, nnz, bsrRowPtr, bsrColInd, (unsigned long long *)tA, tiledim, nblockrows, nblocks); break;
  }
  return bgb_ref::last_status() ? fail("csr2bsr_batch") : 0;
}

static void genRandVec(bool binary) {
  B.resize(nrows);
  if (binary) bgb_gen_bernoulli_f32(nrows, 0.5, 7, B.data());
  else bgb_gen_uniform_f32(nrows, 0.0, 1.0, 7, B.data());
}

static int packRandVec() {
  hipMalloc(&fB, sizeof(float) * nblockrows * tiledim);
  hipMemset(fB, 0, sizeof(float) * nblockrows * tiledim);
  hipMemcpy(fB, B.data(), sizeof(float) * nrows, hipMemcpyHostToDevice);  // the rest is padding
  const long long words = bgb_packed_vec_words(tiledim, nblockrows);
  hipMalloc(&tB, words * bgb_word_bytes(tiledim));
  hipMemset(tB, 0, words * bgb_word_bytes(tiledim));
  switch (tiledim) {
    case 4: bgb_ref::ToBit4Row(fB, (unsigned char *)tB, nblockrows); break;
    case 8: bgb_ref::ToBit8Row(fB, (unsigned char *)tB, nblockrows); break;
    case 16: bgb_ref::ToBit16Row(fB, (unsigned short *)tB, nblockrows); break;
    case 32: bgb_ref::ToBit32Row(fB, (unsigned *)tB, nblockrows); break;
    default: bgb_ref::ToBit64Row(fB, (unsigned long long *)tB, nblockrows); break;
  }
  return bgb_ref::last_status() ? fail("ToBitRow") : 0;
}

static double evalHSBSFBmv(const std::string &op) {
  const size_t outb = op == "bin_bin_bin" ? (size_t)nblockrows * bgb_word_bytes(tiledim)
                                          : sizeof(float) * nblockrows * tiledim;
  hipMalloc(&fC, outb);
  hipMemset(fC, 0, outb);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, nullptr);
  for (int i = 0; i < TEST_TIMES; i++) {
    if (op == "bin_bin_full") {
      switch (tiledim) {
        case 4: bgb_ref::bmv4_bin_bin_full((unsigned char *)tA, (unsigned char *)tB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 8: bgb_ref::bmv8_bin_bin_full((unsigned char *)tA, (unsigned char *)tB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 16: bgb_ref::bmv16_bin_bin_full((unsigned short *)tA, (unsigned short *)tB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 32: bgb_ref::bmv32_bin_bin_full((unsigned *)tA, (unsigned *)tB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        default: bgb_ref::bmv64_bin_bin_full((unsigned long long *)tA, (unsigned long long *)tB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
      }
    } else if (op == "bin_full_full") {
      switch (tiledim) {
        case 4: bgb_ref::bmv4_bin_full_full((unsigned char *)tA, fB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 8: bgb_ref::bmv8_bin_full_full((unsigned char *)tA, fB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 16: bgb_ref::bmv16_bin_full_full((unsigned short *)tA, fB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 32: bgb_ref::bmv32_bin_full_full((unsigned *)tA, fB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
        default: bgb_ref::bmv64_bin_full_full((unsigned long long *)tA, fB, fC, bsrRowPtr, bsrColInd, nblockrows); break;
      }
    } else {
      switch (tiledim) {
        case 4: bgb_ref::bmv4_bin_bin_bin((unsigned char *)tA, (unsigned char *)tB, (unsigned char *)fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 8: bgb_ref::bmv8_bin_bin_bin((unsigned char *)tA, (unsigned char *)tB, (unsigned char *)fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 16: bgb_ref::bmv16_bin_bin_bin((unsigned short *)tA, (unsigned short *)tB, (unsigned short *)fC, bsrRowPtr, bsrColInd, nblockrows); break;
        case 32: bgb_ref::bmv32_bin_bin_bin((unsigned *)tA, (unsigned *)tB, (unsigned *)fC, bsrRowPtr, bsrColInd, nblockrows); break;
        default: bgb_ref::bmv64_bin_bin_bin((unsigned long long *)tA, (unsigned long long *)tB, (unsigned long long *)fC, bsrRowPtr, bsrColInd, nblockrows); break;
      }
    }
    if (bgb_ref::last_status()) return -1.0;
  }
  hipEventRecord(e1, nullptr);
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  return ms / TEST_TIMES;
}

static bool verifyResult(const std::string &op) {
  std::vector<int> cnt(nrows);
  std::vector<uint8_t> x8(nrows);
  for (int i = 0; i < nrows; ++i) x8[i] = B[i] > 0;
  if (op == "bin_full_full") {
    std::vector<double> xd(B.begin(), B.end()), y(nrows);
    orc_csr_spmv_f64(nrows, h_csrRowPtr.data(), h_csrColInd.data(), xd.data(), y.data());
    std::vector<float> got(nrows);
    hipMemcpy(got.data(), fC, sizeof(float) * nrows, hipMemcpyDeviceToHost);
    for (int i = 0; i < nrows; ++i)
      if (std::fabs(got[i] - TEST_TIMES * y[i]) > 1e-6 * TEST_TIMES * std::fabs(y[i]) + 1e-5) return false;
    return true;
  }
  orc_csr_spmv_count(nrows, h_csrRowPtr.data(), h_csrColInd.data(), x8.data(), cnt.data());
  if (op == "bin_bin_full") {
    std::vector<float> got(nrows);
    hipMemcpy(got.data(), fC, sizeof(float) * nrows, hipMemcpyDeviceToHost);
    for (int i = 0; i < nrows; ++i)
      if (got[i] != (float)(TEST_TIMES * cnt[i])) return false;
    return true;
  }
  const int wb = bgb_word_bytes(tiledim);
  std::vector<unsigned char> got((size_t)nblockrows * wb);
  hipMemcpy(got.data(), fC, got.size(), hipMemcpyDeviceToHost);
  for (int I = 0; I < nblockrows; ++I) {
    uint64_t w = 0;
    memcpy(&w, &got[(size_t)I * wb], wb);
    for (int r = 0; r < tiledim; ++r) {
      const int v = I * tiledim + r;
      const bool want = v < nrows && cnt[v] > 0;
      if ((((w >> (tiledim - 1 - r)) & 1ull) != 0) != want) return false;
    }
  }
  return true;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <op> <T> <graph: rmat:S:EF | er:N:DEG | grid:W:H | mtx:PATH> [transpose]\n",
            argv[0]);
    return 2;
  }
  const std::string op = argv[1];
  tiledim = atoi(argv[2]);
  const int transpose = argc > 4 ? atoi(argv[4]) : 0;  // the reference drivers' argv[2]
  hipSetDevice(0);
  if (readGraphCSR(argv[3])) return 1;
  if (transpose && transposeMtx()) return 1;
  genRandVec(op != "bin_full_full");
  if (CSR2HSBSF() || packRandVec()) return 1;
  const double ms = evalHSBSFBmv(op);
  if (ms < 0) return fail("bmv");
  const bool ok = verifyResult(op);
  printf("%s T=%d n=%d nnz=%d nblocks=%d %f ms\nsuccess: %d\n", op.c_str(), tiledim, nrows, nnz, nblocks, ms, ok);
  return ok ? 0 : 3;
}
