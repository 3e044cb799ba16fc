// Tile statistics on the GPU: the counterparts of the reference's debug
// kernels printMatrixStats_{4,8,16,32} and estimatevec (utility.cu:808-975),
// here for every tile dim including 64, and without their one-thread-per-row
// loops (a hub block row is walked by a whole wave).
//
//   bgb_tile_row_stats: for row i of block row I, tile_per_row[I*T+i] = tiles in
//     block row I and nnz_tile_per_row[I*T+i] = those tiles whose word i is
//     non-zero. A wave walks the block row's words in order; since 64 is a
//     multiple of T, lane l always reads word l % T of some tile, so the counts
//     are per-lane sums folded over the 64/T lanes of equal l % T.
//   bgb_estimate_tiles: sum over CSR rows of the runs of equal colind / T
//     (estimatevec). For sorted rows it is the number of non-zero row words of
//     the B2SR, i.e. the sum of nnz_tile_per_row.
#include <hip/hip_runtime.h>

#include "bgb_device.h"
#include "bgb_tile.h"

namespace bgb {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

template <int T>
__global__ void tile_row_stats_kernel(const void *__restrict__ A, const int *__restrict__ rowptr, int nbr,
                                      int *__restrict__ tile_per_row, int *__restrict__ nnz_tile_per_row) {
  using W = typename Tile<T>::W;
  const W *a = static_cast<const W *>(A);
  const int lane = threadIdx.x & 63;
  const long long waves = (long long)gridDim.x * kWavesPerBlock;
  for (long long I = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); I < nbr; I += waves) {
    const int s = rowptr[I], e = rowptr[I + 1];
    int cnt = 0;
    for (long long q = (long long)s * T + lane; q < (long long)e * T; q += 64) cnt += a[q] != 0;
#pragma unroll
    for (int m = T; m < 64; m <<= 1) cnt += __shfl_xor(cnt, m);
    if (lane < T) {
      tile_per_row[I * T + lane] = e - s;
      nnz_tile_per_row[I * T + lane] = cnt;
    }
  }
}

__global__ void zero_i64_kernel(long long *p) {
  if (threadIdx.x == 0) p[0] = 0;
}

// One wave per CSR row (grid-stride); each wave keeps its own total and adds
// it to *out once.
__global__ void estimate_tiles_kernel(const int *__restrict__ rowptr, const int *__restrict__ colind, int nrows,
                                      int lg, unsigned long long *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long waves = (long long)gridDim.x * kWavesPerBlock;
  unsigned long long total = 0;
  for (long long row = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); row < nrows; row += waves) {
    const int s = rowptr[row], e = rowptr[row + 1];
    for (int k = s + lane; k < e; k += 64)
      total += (k == s) || ((colind[k] >> lg) != (colind[k - 1] >> lg));
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) total += __shfl_xor(total, m);
  if (lane == 0 && total) atomicAdd(out, total);
}

}  // namespace
}  // namespace bgb

using namespace bgb;

extern "C" {

int bgb_tile_row_stats(int T, const void *d_A, const int *d_rowptr, int nblockrows, int *d_tile_per_row,
                       int *d_nnz_tile_per_row, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
    BGB_REQUIRE(nblockrows >= 0 && d_rowptr && (nblockrows == 0 || (d_tile_per_row && d_nnz_tile_per_row)),
                BGB_ERR_ARG, "bad arguments");
    if (nblockrows == 0) return BGB_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned grid = capped_blocks((long long)nblockrows * 64, kBlock);
    switch (T) {
      case 4: tile_row_stats_kernel<4><<<grid, kBlock, 0, s>>>(d_A, d_rowptr, nblockrows, d_tile_per_row, d_nnz_tile_per_row); break;
      case 8: tile_row_stats_kernel<8><<<grid, kBlock, 0, s>>>(d_A, d_rowptr, nblockrows, d_tile_per_row, d_nnz_tile_per_row); break;
      case 16: tile_row_stats_kernel<16><<<grid, kBlock, 0, s>>>(d_A, d_rowptr, nblockrows, d_tile_per_row, d_nnz_tile_per_row); break;
      case 32: tile_row_stats_kernel<32><<<grid, kBlock, 0, s>>>(d_A, d_rowptr, nblockrows, d_tile_per_row, d_nnz_tile_per_row); break;
      default: tile_row_stats_kernel<64><<<grid, kBlock, 0, s>>>(d_A, d_rowptr, nblockrows, d_tile_per_row, d_nnz_tile_per_row); break;
    }
    BGB_HIP_TRY(hipGetLastError());
    return BGB_OK;
  });
}

int bgb_estimate_tiles(int T, int nrows, const int *d_rowptr, const int *d_colind, long long *d_out,
                       bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
    BGB_REQUIRE(nrows >= 0 && d_rowptr && d_out, BGB_ERR_ARG, "bad arguments");  // colind unread when every row is empty
    hipStream_t s = static_cast<hipStream_t>(stream);
    zero_i64_kernel<<<1, 64, 0, s>>>(d_out);
    BGB_HIP_TRY(hipGetLastError());
    if (nrows == 0) return BGB_OK;
    estimate_tiles_kernel<<<capped_blocks((long long)nrows * 64, kBlock), kBlock, 0, s>>>(
        d_rowptr, d_colind, nrows, ilog2_tiledim(T), reinterpret_cast<unsigned long long *>(d_out));
    BGB_HIP_TRY(hipGetLastError());
    return BGB_OK;
  });
}

}  // extern "C"
