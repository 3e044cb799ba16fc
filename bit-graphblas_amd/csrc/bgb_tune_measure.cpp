// Measured tile-dim / kernel-configuration tuning (no reference counterpart;
// the reference fixes T at compile time, Makefile:12-46): packs a device CSR
// at every allowed T, times the BMV for each run length and packed-x
// placement with HIP events, and returns the fastest configuration.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <vector>

#include "bgb_internal.h"

namespace {

#define TRY_HIP(expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return bgb::set_error(BGB_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));  \
  } while (0)

// Device buffers of one candidate tile dim, freed on the stream.
struct Scratch {
  std::vector<void *> ptrs;
  hipStream_t s;
  explicit Scratch(hipStream_t st) : s(st) {}
  hipError_t alloc(void **p, size_t bytes) {
    hipError_t e = hipMallocAsync(p, bytes < 256 ? 256 : bytes, s);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
  ~Scratch() {
    for (void *p : ptrs) (void)hipFreeAsync(p, s);
  }
};

struct Knobs {  // process-wide knobs, restored on exit (without touching the error message)
  int run[5], xs[5];
  Knobs() {
    for (int t = 0; t < 5; ++t) {
      run[t] = bgb::g_run_length[t];
      xs[t] = bgb::g_x_staging[t];
    }
  }
  ~Knobs() {
    for (int t = 0; t < 5; ++t) {
      bgb::g_run_length[t] = run[t];
      bgb::g_x_staging[t] = xs[t];
    }
  }
};

int measure(int op, int nrows, int ncols, int nnz, const int *d_rp, const int *d_ci, int allowed_mask, int reps,
            int *T_out, int *run_out, int *xs_out, double *times, hipStream_t s) {
  const int nx = std::max(nrows, ncols);
  // inputs: x Bernoulli(0.05) for the bit op, uniform [0, 1) for FP32 (the
  // benchmark's vectors), generated on the host once
  std::vector<float> hx((size_t)std::max(nx, 1));
  int st = op == 0 ? bgb_gen_bernoulli_f32(nx, 0.05, 2, hx.data()) : bgb_gen_uniform_f32(nx, 0.0, 1.0, 3, hx.data());
  if (st) return st;
  Knobs saved;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TRY_HIP(hipEventCreate(&e0));
  TRY_HIP(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  double best = 1e300;
  for (int t = 0; t < 4; ++t) {
    if (!((allowed_mask >> t) & 1)) continue;
    const int T = 4 << t;
    const int nbr = (nrows + T - 1) / T;
    const int nbx = (nx + T - 1) / T;
    const long long words = bgb_packed_vec_words(T, nbx);
    const int wb = bgb_word_bytes(T);
    Scratch sc(s);
    void *rowptr = nullptr, *colind = nullptr, *val = nullptr, *x = nullptr, *xb = nullptr, *y = nullptr;
    TRY_HIP(sc.alloc(&rowptr, sizeof(int) * ((size_t)nbr + 1)));
    int nblocks = 0;
    if ((st = bgb_b2sr_nnzb(T, nrows, ncols, nnz, d_rp, d_ci, static_cast<int *>(rowptr), &nblocks, s))) return st;
    TRY_HIP(sc.alloc(&colind, sizeof(int) * ((size_t)nblocks + 4)));
    TRY_HIP(sc.alloc(&val, (size_t)wb * T * ((size_t)nblocks + 16)));
    if ((st = bgb_csr2b2sr(T, nrows, ncols, nnz, d_rp, d_ci, static_cast<const int *>(rowptr), nblocks,
                           static_cast<int *>(colind), val, s)))
      return st;
    TRY_HIP(sc.alloc(&x, sizeof(float) * (size_t)nbx * T));
    TRY_HIP(hipMemsetAsync(x, 0, sizeof(float) * (size_t)nbx * T, s));
    TRY_HIP(hipMemcpyAsync(x, hx.data(), sizeof(float) * (size_t)nx, hipMemcpyHostToDevice, s));
    if (op == 0) {
      TRY_HIP(sc.alloc(&xb, (size_t)wb * (size_t)std::max<long long>(words, 1) + 16));
      if ((st = bgb_pack_vec(T, static_cast<const float *>(x), nbx * T, xb, s))) return st;
    }
    TRY_HIP(sc.alloc(&y, sizeof(float) * (size_t)std::max(nbr, 1) * T));
    auto call = [&]() -> int {
      return op == 0 ? bgb_bmv_bin_bin_full_i32(T, val, xb, static_cast<int *>(y), static_cast<const int *>(rowptr),
                                                static_cast<const int *>(colind), nbr, nblocks, 0, s)
                     : bgb_bmv_bin_full_full(T, val, static_cast<const float *>(x), static_cast<float *>(y),
                                             static_cast<const int *>(rowptr), static_cast<const int *>(colind), nbr,
                                             nblocks, 0, s);
    };
    for (int si = 0; si < 2; ++si) {
      const int S = si ? 16 : 8;
      for (int xi = 0; xi < 2; ++xi) {
        // packed-x placement: a choice of the bit op at T <= 16 only
        if (xi == 1 && !(op == 0 && T <= 16)) continue;
        if ((st = bgb_set_run_length(T, S))) return st;
        if ((st = bgb_set_x_staging(T, op == 0 && T <= 16 ? (xi ? 2 : 1) : 0))) return st;
        if ((st = call())) return st;  // warm-up (and the per-matrix facts)
        TRY_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r)
          if ((st = call())) return st;
        TRY_HIP(hipEventRecord(e1, s));
        TRY_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / reps;
        if (times) times[t * 4 + si * 2 + xi] = us;
        if (us < best) {
          best = us;
          *T_out = T;
          if (run_out) *run_out = S;
          if (xs_out) *xs_out = op == 0 && T <= 16 ? (xi ? 2 : 1) : 1;
        }
      }
    }
    TRY_HIP(hipStreamSynchronize(s));  // the candidate's buffers are freed on the stream
  }
  return BGB_OK;
}

}  // namespace

extern "C" int bgb_autotune_measure(int op, int nrows, int ncols, int nnz, const int *d_csrRowPtr,
                                    const int *d_csrColInd, int allowed_mask, int reps, int *T_out,
                                    int *run_length_out, int *x_staging_out, double *h_times_us,
                                    bgb_stream_t stream) {
  return bgb::host_guard([&]() -> int {
    bgb::clear_error();
    BGB_REQUIRE(op == 0 || op == 1, BGB_ERR_ARG, "measured tuning covers op 0 (bin_bin_*) and 1 (bin_full_full)");
    BGB_REQUIRE(nrows >= 1 && ncols >= 1 && nnz >= 0 && d_csrRowPtr && (nnz == 0 || d_csrColInd) && T_out,
                BGB_ERR_ARG, "bad arguments");
    BGB_REQUIRE((allowed_mask & 0xF) != 0, BGB_ERR_ARG, "no tile dim allowed");
    BGB_REQUIRE(reps >= 1 && reps <= 10000, BGB_ERR_ARG, "reps must be 1..10000");
    if (h_times_us)
      for (int i = 0; i < 16; ++i) h_times_us[i] = 0.0;
    return measure(op, nrows, ncols, nnz, d_csrRowPtr, d_csrColInd, allowed_mask, reps, T_out, run_length_out,
                   x_staging_out, h_times_us, static_cast<hipStream_t>(stream));
  });
}
