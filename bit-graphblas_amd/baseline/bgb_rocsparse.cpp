// rocSPARSE CSR SpMV timing, the GPU baseline reported beside the B2SR BMV
// (the counterpart of the reference's cuSPARSE baseline,
// test_spmv_baseline_a100.cu:92-146 / test_bin_bin_full.cu:271-309).
// Measurement aid only: built into libbgb_baseline.so, never called by libbgb.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdio>

#define RS_TRY(x)                                                   \
  do {                                                              \
    rocsparse_status _s = (x);                                      \
    if (_s != rocsparse_status_success) {                           \
      fprintf(stderr, "%s failed: %d\n", #x, (int)_s);              \
      rc = 1;                                                       \
      goto done;                                                    \
    }                                                               \
  } while (0)
#define HIP_TRY(x)                                                  \
  do {                                                              \
    hipError_t _e = (x);                                            \
    if (_e != hipSuccess) {                                         \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(_e)); \
      rc = 2;                                                       \
      goto done;                                                    \
    }                                                               \
  } while (0)

extern "C" int bgbx_rocsparse_csrmv(int n, int nnz, const int *d_rowptr, const int *d_colind,
                                    const float *d_val, const float *d_x, float *d_y, int alg,
                                    int reps, double *ms_per_spmv, double *ms_preprocess,
                                    void *stream) {
  int rc = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  rocsparse_handle h = nullptr;
  rocsparse_spmat_descr A = nullptr;
  rocsparse_dnvec_descr X = nullptr, Y = nullptr;
  void *buf = nullptr;
  size_t bytes = 0;
  const float alpha = 1.0f, beta = 0.0f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.0f;
  const rocsparse_spmv_alg a = alg == 1 ? rocsparse_spmv_alg_csr_rowsplit : rocsparse_spmv_alg_csr_adaptive;
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
  RS_TRY(rocsparse_create_handle(&h));
  RS_TRY(rocsparse_set_stream(h, s));
  RS_TRY(rocsparse_create_csr_descr(&A, n, n, nnz, (void *)d_rowptr, (void *)d_colind, (void *)d_val,
                                    rocsparse_indextype_i32, rocsparse_indextype_i32,
                                    rocsparse_index_base_zero, rocsparse_datatype_f32_r));
  RS_TRY(rocsparse_create_dnvec_descr(&X, n, (void *)d_x, rocsparse_datatype_f32_r));
  RS_TRY(rocsparse_create_dnvec_descr(&Y, n, (void *)d_y, rocsparse_datatype_f32_r));
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  RS_TRY(rocsparse_spmv(h, rocsparse_operation_none, &alpha, A, X, &beta, Y, rocsparse_datatype_f32_r, a,
                        rocsparse_spmv_stage_buffer_size, &bytes, nullptr));
  HIP_TRY(hipMalloc(&buf, bytes > 0 ? bytes : 256));
  HIP_TRY(hipEventRecord(e0, s));
  RS_TRY(rocsparse_spmv(h, rocsparse_operation_none, &alpha, A, X, &beta, Y, rocsparse_datatype_f32_r, a,
                        rocsparse_spmv_stage_preprocess, &bytes, buf));
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_preprocess = ms;
  for (int i = 0; i < 3; ++i)
    RS_TRY(rocsparse_spmv(h, rocsparse_operation_none, &alpha, A, X, &beta, Y, rocsparse_datatype_f32_r, a,
                          rocsparse_spmv_stage_compute, &bytes, buf));
  HIP_TRY(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i)
    RS_TRY(rocsparse_spmv(h, rocsparse_operation_none, &alpha, A, X, &beta, Y, rocsparse_datatype_f32_r, a,
                          rocsparse_spmv_stage_compute, &bytes, buf));
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_spmv = ms / (reps > 0 ? reps : 1);
#pragma clang diagnostic pop
done:
  if (buf) (void)hipFree(buf);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (X) rocsparse_destroy_dnvec_descr(X);
  if (Y) rocsparse_destroy_dnvec_descr(Y);
  if (A) rocsparse_destroy_spmat_descr(A);
  if (h) rocsparse_destroy_handle(h);
  return rc;
}
