// PageRank over a symmetric B2SR matrix using the bin_full_full BMV (gfx950).
//
// No reference counterpart: the reference's bin_full_full BMV
// (bsrbmv.cu:826-1081) is the building block. Iteration (pr_0 = 1/n):
//   x_t[v]   = pr_t[v] / deg(v)            (0 for deg(v) = 0)
//   y        = A x_t                        (bin_full_full, FP32 accumulate)
//   pr_t+1   = (1-d)/n + d * (y + dangling_t / n),   dangling_t = sum_{deg=0} pr_t
//   res_t    = || pr_t+1 - pr_t ||_1
// Precision modes:
//   0  FP32 vector, FP32 BMV, FP32 residual / dangling sums
//   1  x stored as bfloat16 (round to nearest even), FP32 accumulation and
//      sums -- halves the gathered bytes of x
//   2  FP32 vector and BMV, residual and dangling / teleport sums in FP64
// When the BMV takes its short-row path (T <= 8, no block row over 32 tiles:
// one lane per block row, whose sums are final in the lane), the update is
// fused into it: one kernel reads x_t, pr_t and deg and writes pr_t+1 and
// x_t+1 (into a second x buffer, since other lanes still gather x_t), so y
// never goes through memory.
// The update kernel fuses the next x, the dangling sum and the residual;
// reductions are two-level (per-block partials, one finishing block) so the
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "bgb_device.h"
#include "bgb_tile.h"

extern "C" int bgb_get_short_row_path(void);

extern "C" int bgb_bmv_bin_full_full_bf16(int T, const void *d_A, const uint16_t *d_B, float *d_C,
                                          const int *d_rowptr, const int *d_colind, int nblockrows,
                                          int nblocks, int accumulate, bgb_stream_t stream);

namespace bgb {
namespace {

constexpr int kPrBlock = 256;
// Update / init grids: at most this many workgroups (grid-stride loops), so
// the single finishing block folds at most this many partials per sum.
constexpr int kPrMaxBlocks = 1024;

// float -> bfloat16 with round-to-nearest-even on the upper 16 bits.
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Sum over a workgroup of NT threads (the total in thread 0).
template <class Acc, int NT = kPrBlock>
__device__ __forceinline__ Acc block_sum(Acc v, Acc *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  Acc t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

// x from pr, dangling partials; used at start. Like the update kernel, a
// grid-stride loop over a grid of at most kPrMaxBlocks workgroups, so the
// finishing block folds few partials.
template <class Acc, bool BF>
__global__ __launch_bounds__(kPrBlock) void pr_init_kernel(float *pr, const int *deg, int n, long long npad,
                                                           void *x, Acc *part_dang) {
  __shared__ Acc sh[kPrBlock / 64];
  Acc dang = 0;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < npad;
       v += (long long)gridDim.x * blockDim.x) {
    float p = v < n ? 1.0f / (float)n : 0.0f;
    if (v < n) pr[v] = p;
    const int dg = v < n ? deg[v] : 0;
    const float xv = dg > 0 ? p / (float)dg : 0.0f;
    if (BF) static_cast<uint16_t *>(x)[v] = f32_to_bf16(xv);
    else static_cast<float *>(x)[v] = xv;
    if (v < n && dg == 0) dang += (Acc)p;
  }
  const Acc t = block_sum<Acc>(dang, sh);
  if (threadIdx.x == 0) part_dang[blockIdx.x] = t;
}

template <class Acc, bool BF>
__global__ __launch_bounds__(kPrBlock) void pr_update_kernel(const float *y, const float *pr, float *pr_next,
                                                             const int *deg, int n, long long npad, float damping,
                                                             const Acc *dang_total, void *x, Acc *part_res,
                                                             Acc *part_dang) {
  __shared__ Acc sh[kPrBlock / 64];
  const Acc tele = (Acc)(1.0f - damping) / (Acc)n;
  const Acc dn = *dang_total / (Acc)n;
  Acc res = 0, dang = 0;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < npad;
       v += (long long)gridDim.x * blockDim.x) {
    float xv = 0.0f;
    if (v < n) {
      const Acc pn = tele + (Acc)damping * ((Acc)y[v] + dn);
      const Acc po = (Acc)pr[v];
      res += pn > po ? pn - po : po - pn;
      const float p = (float)pn;
      pr_next[v] = p;
      const int dg = deg[v];
      xv = dg > 0 ? p / (float)dg : 0.0f;
      if (dg == 0) dang += (Acc)p;
    }
    if (BF) static_cast<uint16_t *>(x)[v] = f32_to_bf16(xv);
    else static_cast<float *>(x)[v] = xv;
  }
  const Acc r = block_sum<Acc>(res, sh);
  const Acc d = block_sum<Acc>(dang, sh);
  if (threadIdx.x == 0) {
    part_res[blockIdx.x] = r;
    part_dang[blockIdx.x] = d;
  }
}

// Short-row BMV + update, one lane per block row (one partial pair per
// workgroup; a grid-stride loop only past 2^20 workgroups).

template <int T, class Acc, bool BF, int U>
__global__ __launch_bounds__(kPrBlock) void pr_rows_kernel(const void *A, const int *rowptr, const int *colind,
                                                           int nbr, int n, float damping, const Acc *dang_total,
                                                           const void *xin, void *xout, const float *pr,
                                                           float *pr_next, const int *deg, Acc *part_res,
                                                           Acc *part_dang) {
  __shared__ Acc sh[kPrBlock / 64];
  const Acc tele = (Acc)(1.0f - damping) / (Acc)n;
  const Acc dn = *dang_total / (Acc)n;
  Acc res = 0, dang = 0;
  for (long long I = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; I < nbr;
       I += (long long)gridDim.x * blockDim.x) {
    float y[T];
    block_row_sums<T, BF ? kRowSumBF16 : kRowSumF32, U>(A, colind, xin, rowptr[I], rowptr[I + 1], y);
    const long long v0 = I * T;
    if (v0 + T <= n) {  // whole block row: 16-byte accesses (pr has n entries, no padding)
      float po[T], pn_f[T], xv[T];
      int dg[T];
#pragma unroll
      for (int q = 0; q < T / 4; ++q) {
        const float4 a4 = reinterpret_cast<const float4 *>(pr + v0)[q];
        const int4 d4 = reinterpret_cast<const int4 *>(deg + v0)[q];
        po[4 * q] = a4.x, po[4 * q + 1] = a4.y, po[4 * q + 2] = a4.z, po[4 * q + 3] = a4.w;
        dg[4 * q] = d4.x, dg[4 * q + 1] = d4.y, dg[4 * q + 2] = d4.z, dg[4 * q + 3] = d4.w;
      }
#pragma unroll
      for (int r = 0; r < T; ++r) {
        const Acc pn = tele + (Acc)damping * ((Acc)y[r] + dn);
        const Acc pa = (Acc)po[r];
        res += pn > pa ? pn - pa : pa - pn;
        pn_f[r] = (float)pn;
        xv[r] = dg[r] > 0 ? pn_f[r] / (float)dg[r] : 0.0f;
        if (dg[r] == 0) dang += (Acc)pn_f[r];
      }
#pragma unroll
      for (int q = 0; q < T / 4; ++q) {
        reinterpret_cast<float4 *>(pr_next + v0)[q] =
            make_float4(pn_f[4 * q], pn_f[4 * q + 1], pn_f[4 * q + 2], pn_f[4 * q + 3]);
        if (BF)
          reinterpret_cast<uint2 *>(static_cast<uint16_t *>(xout) + v0)[q] =
              make_uint2((uint32_t)f32_to_bf16(xv[4 * q]) | ((uint32_t)f32_to_bf16(xv[4 * q + 1]) << 16),
                         (uint32_t)f32_to_bf16(xv[4 * q + 2]) | ((uint32_t)f32_to_bf16(xv[4 * q + 3]) << 16));
        else
          reinterpret_cast<float4 *>(static_cast<float *>(xout) + v0)[q] =
              make_float4(xv[4 * q], xv[4 * q + 1], xv[4 * q + 2], xv[4 * q + 3]);
      }
    } else {  // the last, partial block row
#pragma unroll
      for (int r = 0; r < T; ++r) {
        const long long v = v0 + r;
        float xs = 0.0f;
        if (v < n) {
          const Acc pn = tele + (Acc)damping * ((Acc)y[r] + dn);
          const Acc pa = (Acc)pr[v];
          res += pn > pa ? pn - pa : pa - pn;
          const float p = (float)pn;
          pr_next[v] = p;
          const int d = deg[v];
          xs = d > 0 ? p / (float)d : 0.0f;
          if (d == 0) dang += (Acc)p;
        }
        if (BF) static_cast<uint16_t *>(xout)[v] = f32_to_bf16(xs);
        else static_cast<float *>(xout)[v] = xs;
      }
    }
  }
  const Acc r = block_sum<Acc>(res, sh);
  const Acc d = block_sum<Acc>(dang, sh);
  if (threadIdx.x == 0) {
    part_res[blockIdx.x] = r;
    part_dang[blockIdx.x] = d;
  }
}

template <int T, class Acc, bool BF>
void launch_pr_rows(int U, unsigned grid, hipStream_t s, const void *A, const int *rowptr, const int *colind, int nbr,
                    int n, float damping, const Acc *dtot, const void *xin, void *xout, const float *pr,
                    float *pr_next, const int *deg, Acc *pres, Acc *pdang) {
#define BGB_PR_ROWS(UU)                                                                                        \
  pr_rows_kernel<T, Acc, BF, UU><<<grid, kPrBlock, 0, s>>>(A, rowptr, colind, nbr, n, damping, dtot, xin, xout, \
                                                           pr, pr_next, deg, pres, pdang)
  switch (U) {
    case 4: BGB_PR_ROWS(4); break;
    case 5: BGB_PR_ROWS(5); break;
    case 6: BGB_PR_ROWS(6); break;
    default: BGB_PR_ROWS(8); break;
  }
#undef BGB_PR_ROWS
}

// One workgroup of kFinishBlock threads folds the partials (in a fixed
// order): residual -> res_out[it] (as double), dangling -> dang_total.
constexpr int kFinishBlock = 1024;

template <class Acc>
__global__ __launch_bounds__(kFinishBlock) void pr_finish_kernel(const Acc *part_res, const Acc *part_dang,
                                                                 int nparts, double *res_out, Acc *dang_total) {
  __shared__ Acc sh[kFinishBlock / 64];
  Acc r = 0, d = 0;
#pragma unroll 8
  for (int i = threadIdx.x; i < nparts; i += kFinishBlock) {
    if (part_res) r += part_res[i];
    d += part_dang[i];
  }
  r = block_sum<Acc, kFinishBlock>(r, sh);
  d = block_sum<Acc, kFinishBlock>(d, sh);
  if (threadIdx.x == 0) {
    if (res_out) *res_out = (double)r;
    *dang_total = d;
  }
}

// deg(v) = popcount of row v over its block row's tiles.
template <int T>
__device__ __forceinline__ int row_degree(const void *A, const int *rowptr, int I, int r) {
  int d = 0;
  for (int k = rowptr[I]; k < rowptr[I + 1]; ++k) {
    if constexpr (T == 64) {
      d += __popcll(static_cast<const uint64_t *>(A)[(long long)k * T + r]);
    } else {
      uint32_t w;
      if (T == 32) w = static_cast<const uint32_t *>(A)[(long long)k * T + r];
      else if (T == 16) w = static_cast<const uint16_t *>(A)[(long long)k * T + r];
      else w = static_cast<const uint8_t *>(A)[(long long)k * T + r];
      d += __popc(w);
    }
  }
  return d;
}

template <int T>
__global__ void row_degree_kernel(const void *A, const int *rowptr, int nbr, int *deg) {
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < (long long)nbr * T;
       v += (long long)gridDim.x * blockDim.x)
    deg[v] = row_degree<T>(A, rowptr, (int)(v / T), (int)(v % T));
}

template <class Acc, bool BF>
int pagerank_impl(int T, const void *A, const int *rowptr, const int *colind, int nbr, int nblocks, int n,
                  float damping, int iters, float *pr_out, double *h_res, hipStream_t s) {
  const long long npad = (long long)nbr * T;
  const unsigned blocks = (unsigned)(ceil_div(npad, kPrBlock) < kPrMaxBlocks ? ceil_div(npad, kPrBlock) : kPrMaxBlocks);
  const unsigned deg_blocks = capped_blocks(npad, kPrBlock);
  // fused short-row iteration when the BMV would take its short-row path
  int mr = -1;
  const int path = bgb_get_short_row_path();
  if (T <= 8 && path != 1) {
    if (path == 2) mr = 0;
    else if (int st = max_row_tiles(s, rowptr, nbr, nblocks, &mr)) return st;
  }
  // (its 16-byte accesses to pr need the caller's array 16-byte aligned)
  const bool fused = mr >= 0 && mr <= kShortRowMax && (reinterpret_cast<uintptr_t>(pr_out) & 15) == 0;
  const unsigned rows_blocks = capped_blocks(nbr, kPrBlock);
  const unsigned nparts = std::max(blocks, rows_blocks);
  DeviceBuffer deg, x, x2, y, pr2, pres, pdang, dtot, res;
  BGB_HIP_TRY(deg.alloc(sizeof(int) * npad, s));
  BGB_HIP_TRY(x.alloc((BF ? 2 : 4) * npad + 16, s));
  if (fused) BGB_HIP_TRY(x2.alloc((BF ? 2 : 4) * npad + 16, s));
  else BGB_HIP_TRY(y.alloc(sizeof(float) * npad, s));
  BGB_HIP_TRY(pr2.alloc(sizeof(float) * n, s));
  BGB_HIP_TRY(pres.alloc(sizeof(Acc) * nparts, s));
  BGB_HIP_TRY(pdang.alloc(sizeof(Acc) * nparts, s));
  BGB_HIP_TRY(dtot.alloc(sizeof(Acc), s));
  BGB_HIP_TRY(res.alloc(sizeof(double) * (iters > 0 ? iters : 1), s));
  switch (T) {
    case 4: row_degree_kernel<4><<<deg_blocks, kPrBlock, 0, s>>>(A, rowptr, nbr, deg.as<int>()); break;
    case 8: row_degree_kernel<8><<<deg_blocks, kPrBlock, 0, s>>>(A, rowptr, nbr, deg.as<int>()); break;
    case 16: row_degree_kernel<16><<<deg_blocks, kPrBlock, 0, s>>>(A, rowptr, nbr, deg.as<int>()); break;
    case 32: row_degree_kernel<32><<<deg_blocks, kPrBlock, 0, s>>>(A, rowptr, nbr, deg.as<int>()); break;
    default: row_degree_kernel<64><<<deg_blocks, kPrBlock, 0, s>>>(A, rowptr, nbr, deg.as<int>()); break;
  }
  pr_init_kernel<Acc, BF><<<blocks, kPrBlock, 0, s>>>(pr_out, deg.as<int>(), n, npad, x.ptr, pdang.as<Acc>());
  pr_finish_kernel<Acc><<<1, kFinishBlock, 0, s>>>(nullptr, pdang.as<Acc>(), (int)blocks, nullptr, dtot.as<Acc>());
  BGB_HIP_TRY(hipGetLastError());
  float *cur = pr_out, *nxt = pr2.as<float>();
  void *xa = x.ptr, *xb = x2.ptr;
  for (int it = 0; it < iters && fused; ++it) {
    const int U = short_row_batch(mr);
    if (T == 4)
      launch_pr_rows<4, Acc, BF>(U, rows_blocks, s, A, rowptr, colind, nbr, n, damping, dtot.as<Acc>(), xa, xb, cur,
                                 nxt, deg.as<int>(), pres.as<Acc>(), pdang.as<Acc>());
    else
      launch_pr_rows<8, Acc, BF>(U, rows_blocks, s, A, rowptr, colind, nbr, n, damping, dtot.as<Acc>(), xa, xb, cur,
                                 nxt, deg.as<int>(), pres.as<Acc>(), pdang.as<Acc>());
    pr_finish_kernel<Acc><<<1, kFinishBlock, 0, s>>>(pres.as<Acc>(), pdang.as<Acc>(), (int)rows_blocks,
                                                 res.as<double>() + it, dtot.as<Acc>());
    BGB_HIP_TRY(hipGetLastError());
    std::swap(xa, xb);
    std::swap(cur, nxt);
  }
  for (int it = 0; it < iters && !fused; ++it) {
    int st = BF ? bgb_bmv_bin_full_full_bf16(T, A, x.as<uint16_t>(), y.as<float>(), rowptr, colind, nbr,
                                             nblocks, 0, s)
                : bgb_bmv_bin_full_full(T, A, x.as<float>(), y.as<float>(), rowptr, colind, nbr, nblocks, 0, s);
    if (st) return st;
    pr_update_kernel<Acc, BF><<<blocks, kPrBlock, 0, s>>>(y.as<float>(), cur, nxt, deg.as<int>(), n, npad, damping,
                                                         dtot.as<Acc>(), x.ptr, pres.as<Acc>(), pdang.as<Acc>());
    pr_finish_kernel<Acc><<<1, kFinishBlock, 0, s>>>(pres.as<Acc>(), pdang.as<Acc>(), (int)blocks,
                                                 res.as<double>() + it, dtot.as<Acc>());
    BGB_HIP_TRY(hipGetLastError());
    float *t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != pr_out)
    BGB_HIP_TRY(hipMemcpyAsync(pr_out, cur, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
  if (h_res && iters > 0)
    BGB_HIP_TRY(hipMemcpyAsync(h_res, res.ptr, sizeof(double) * iters, hipMemcpyDeviceToHost, s));
  BGB_HIP_TRY(hipStreamSynchronize(s));
  return BGB_OK;
}

}  // namespace
}  // namespace bgb

using namespace bgb;

extern "C" int bgb_pagerank(int T, const void *d_A, const int *d_rowptr, const int *d_colind, int nblockrows,
                            int nblocks, int n, float damping, int iters, int mode, float *d_pr,
                            double *h_residuals, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
    BGB_REQUIRE(d_rowptr && d_pr && (nblocks == 0 || (d_A && d_colind)), BGB_ERR_ARG, "null argument");
    BGB_REQUIRE(n >= 1 && nblockrows == (int)ceil_div(n, T), BGB_ERR_ARG, "nblockrows must be ceil(n/T)");
    BGB_REQUIRE(iters >= 0 && damping >= 0.0f && damping <= 1.0f, BGB_ERR_ARG, "bad iters / damping");
    BGB_REQUIRE(mode >= 0 && mode <= 2, BGB_ERR_ARG, "mode must be 0 (fp32), 1 (bf16 x) or 2 (fp64 sums)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (mode) {
      case 0: return pagerank_impl<float, false>(T, d_A, d_rowptr, d_colind, nblockrows, nblocks, n, damping, iters, d_pr, h_residuals, s);
      case 1: return pagerank_impl<float, true>(T, d_A, d_rowptr, d_colind, nblockrows, nblocks, n, damping, iters, d_pr, h_residuals, s);
      default: return pagerank_impl<double, false>(T, d_A, d_rowptr, d_colind, nblockrows, nblocks, n, damping, iters, d_pr, h_residuals, s);
    }
  });
}
