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

// Read-only HBM stream (16-byte loads, grid-stride, XOR-reduced so the loads
// are kept): the read ceiling the BMV kernels are compared with, beside the
// device-to-device copy ceiling.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__global__ void stream_read_kernel(const u32x4 *__restrict__ p, size_t n16, unsigned *out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  unsigned acc = 0;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const u32x4 a = __builtin_nontemporal_load(p + i);
    const u32x4 b = __builtin_nontemporal_load(p + i + stride);
    const u32x4 c = __builtin_nontemporal_load(p + i + 2 * stride);
    const u32x4 d = __builtin_nontemporal_load(p + i + 3 * stride);
    acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
  }
  for (; i < n16; i += stride) {
    const u32x4 a = p[i];
    acc ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  if (acc == 0x9E3779B9u) atomicXor(out, acc);  // practically never taken; keeps the loads live
}

extern "C" int bgbx_stream_read(const void *d_p, size_t bytes, int blocks, int reps, double *ms_per_pass,
                                void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned *out = nullptr;
  if (hipMalloc(&out, sizeof(unsigned)) != hipSuccess) return 2;
  const size_t n16 = bytes / 16;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  stream_read_kernel<<<blocks, 256, 0, s>>>(static_cast<const u32x4 *>(d_p), n16, out);
  (void)hipEventRecord(e0, s);
  for (int r = 0; r < reps; ++r) stream_read_kernel<<<blocks, 256, 0, s>>>(static_cast<const u32x4 *>(d_p), n16, out);
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *ms_per_pass = ms / (reps > 0 ? reps : 1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}
