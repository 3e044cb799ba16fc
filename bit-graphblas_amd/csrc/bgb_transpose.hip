// Transposes on the GPU (gfx950, wave64).
//
// The reference transposes a directed graph before packing it
// (transposeMtx, test_bin_bin_bin.cu:110-136): cusparseScsr2csc, then the CSC
// arrays are used as the CSR of A^T. bgb_csr2csc does that conversion here;
// bgb_b2sr_transpose goes one step further and builds the B2SR of A^T straight
// from the B2SR of A (no reference counterpart), tile by tile, so a packed
// graph can be walked in both directions without going back to the CSR.
//
// Both are the same three passes over "entries" (nonzeros / tiles):
//   1. one 64-bit key per entry, (column << pos_bits) | position -- one wave per
//      row walks the row, so hub rows do not serialise, and records the row of
//      every position;
//   2. a stable radix sort on the column bits only: inside a column the
//      positions, and therefore the rows, stay ascending (what csr2csc gives);
//   3. per output position: the source position from the key, the row from the
//      recorded rows, the value moved (csr2csc) or the T x T bit tile
//      transposed with one 64-lane ballot per tile column (B2SR); the output
//      row pointer by a binary search per column over the sorted keys.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bgb_device.h"
#include "bgb_tile.h"

namespace bgb {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

inline int bits_for(long long count) {  // bits needed to hold values in [0, count)
  int b = 1;
  while (b < 62 && (1LL << b) < count) ++b;
  return b;
}

// rowptr monotone with rowptr[0] = 0 and rowptr[nrows] = nnz; every index in
// [0, ncols). Bits of *bad: 1 first entry, 2 last entry, 4 non-monotone, 8 index.
__global__ void check_rows_kernel(const int *__restrict__ rowptr, int nrows, long long nnz, int *__restrict__ bad) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= nrows; i += stride) {
    const int s = rowptr[i];
    if (i == 0 && s != 0) atomicOr(bad, 1);
    if (i == nrows) {
      if (s != nnz) atomicOr(bad, 2);
    } else {
      const int e = rowptr[i + 1];
      if (e < s || s < 0 || e > nnz) atomicOr(bad, 4);
    }
  }
}

__global__ void check_index_kernel(const int *__restrict__ idx, long long n, int bound, int *__restrict__ bad) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    if ((unsigned)idx[k] >= (unsigned)bound) atomicOr(bad, 8);
}

// Step 1: one wave per row, grid-stride over rows.
__global__ void entry_keys_kernel(const int *__restrict__ rowptr, const int *__restrict__ colind, int nrows,
                                  int pos_bits, uint64_t *__restrict__ keys, int *__restrict__ row_of) {
  const int lane = threadIdx.x & 63;
  const long long waves = (long long)gridDim.x * kWavesPerBlock;
  for (long long row = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); row < nrows; row += waves) {
    const int s = rowptr[row], e = rowptr[row + 1];
    for (int k = s + lane; k < e; k += 64) {
      keys[k] = ((uint64_t)(uint32_t)colind[k] << pos_bits) | (uint64_t)k;
      row_of[k] = (int)row;
    }
  }
}

// Output row pointer: ptr[j] = first sorted position whose column is >= j.
__global__ void column_bounds_kernel(const uint64_t *__restrict__ keys, long long n, int pos_bits, int ncols,
                                     int *__restrict__ ptr) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j <= ncols; j += stride) {
    long long lo = 0, hi = n;
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if ((long long)(keys[mid] >> pos_bits) < j) lo = mid + 1;
      else hi = mid;
    }
    ptr[j] = (int)lo;
  }
}

// Step 3 of csr2csc: row index and (optionally) value of every output position.
__global__ void csc_scatter_kernel(const uint64_t *__restrict__ keys, long long nnz, int pos_bits,
                                   const int *__restrict__ row_of, const float *__restrict__ val,
                                   int *__restrict__ rowind, float *__restrict__ cval) {
  const uint64_t pmask = (1ull << pos_bits) - 1;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
    const long long k = (long long)(keys[p] & pmask);
    rowind[p] = row_of[k];
    if (val) cval[p] = val[k];
  }
}

// Step 3 of the B2SR transpose. A lane group of T lanes owns one output tile
// (64/T tiles per wave); lane r holds row word r of the source tile. For tile
// column c, one ballot of bit (T-1-c) gives every group the column as a T-bit
// field with row r at bit r; the transposed tile's word c wants row r at bit
// T-1-r, so lane c of the group keeps the field and reverses it.
template <int T>
__global__ void tile_transpose_kernel(const uint64_t *__restrict__ keys, long long nblocks, int pos_bits,
                                      const int *__restrict__ row_of, const void *__restrict__ val,
                                      int *__restrict__ t_colind, void *__restrict__ t_val) {
  using W = typename Tile<T>::W;
  constexpr int G = 64 / T;
  const int lane = threadIdx.x & 63;
  const int g = lane / T, r = lane % T;
  const uint64_t pmask = (1ull << pos_bits) - 1;
  const uint64_t fmask = T == 64 ? ~0ull : ((1ull << (T % 64)) - 1);
  const long long waves = (long long)gridDim.x * kWavesPerBlock;
  const W *src = static_cast<const W *>(val);
  W *dst = static_cast<W *>(t_val);
  for (long long base = ((long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) * G; base < nblocks;
       base += waves * G) {
    const long long p = base + g;
    const bool ok = p < nblocks;
    const long long q = ok ? p : nblocks - 1;  // keeps the loads in bounds, results unused
    const long long k = (long long)(keys[q] & pmask);
    const uint64_t w = (uint64_t)src[k * T + r];
    uint64_t mine = 0;
#pragma unroll 8
    for (int c = 0; c < T; ++c) {
      const uint64_t b = __ballot((w >> (T - 1 - c)) & 1);
      mine = c == r ? b : mine;
    }
    mine = (mine >> (g * T)) & fmask;
    const uint64_t out = __builtin_bitreverse64(mine) >> (64 - T);
    if (ok) {
      dst[p * T + r] = (W)out;
      if (r == 0) t_colind[p] = row_of[k];
    }
  }
}

// Validates rowptr / index ranges of an (nrows x ncols) structure with nnz
// entries (synchronises the stream).
int check_structure(const int *rowptr, const int *colind, int nrows, int ncols, long long nnz, hipStream_t s,
                    const char *what) {
  DeviceBuffer bad;
  BGB_HIP_TRY(bad.alloc(sizeof(int), s));
  BGB_HIP_TRY(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
  check_rows_kernel<<<capped_blocks((long long)nrows + 1, kBlock), kBlock, 0, s>>>(rowptr, nrows, nnz,
                                                                                  bad.as<int>());
  BGB_HIP_TRY(hipGetLastError());
  if (nnz > 0) {
    check_index_kernel<<<capped_blocks(nnz, kBlock), kBlock, 0, s>>>(colind, nnz, ncols, bad.as<int>());
    BGB_HIP_TRY(hipGetLastError());
  }
  int h_bad = 0;
  BGB_HIP_TRY(hipMemcpyAsync(&h_bad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  BGB_HIP_TRY(hipStreamSynchronize(s));
  BGB_REQUIRE(h_bad == 0, BGB_ERR_FORMAT,
              "invalid %s (code %d: 1/2 row pointer ends, 4 non-monotone row pointer, 8 index out of range)",
              what, h_bad);
  return BGB_OK;
}

// Steps 1-2 plus the output row pointer; leaves sorted keys and row_of.
int sorted_entries(const int *rowptr, const int *colind, int nrows, int ncols, long long nnz, hipStream_t s,
                   DeviceBuffer &keys, DeviceBuffer &row_of, int *out_ptr, int *pos_bits) {
  const int pb = bits_for(nnz);
  const int cb = bits_for((long long)ncols);
  BGB_REQUIRE(pb + cb <= 64, BGB_ERR_SIZE, "too many entries for 64-bit sort keys");
  *pos_bits = pb;
  DeviceBuffer raw;
  BGB_HIP_TRY(raw.alloc(sizeof(uint64_t) * (size_t)nnz, s));
  BGB_HIP_TRY(keys.alloc(sizeof(uint64_t) * (size_t)nnz, s));
  BGB_HIP_TRY(row_of.alloc(sizeof(int) * (size_t)nnz, s));
  if (nnz > 0) {
    entry_keys_kernel<<<capped_blocks((long long)nrows * 64, kBlock), kBlock, 0, s>>>(
        rowptr, colind, nrows, pb, raw.as<uint64_t>(), row_of.as<int>());
    BGB_HIP_TRY(hipGetLastError());
    BGB_HIP_TRY(radix_sort_u64_bits(raw.as<uint64_t>(), keys.as<uint64_t>(), (size_t)nnz, pb, pb + cb, s));
  }
  column_bounds_kernel<<<capped_blocks((long long)ncols + 1, kBlock), kBlock, 0, s>>>(keys.as<uint64_t>(), nnz, pb,
                                                                                     ncols, out_ptr);
  BGB_HIP_TRY(hipGetLastError());
  return BGB_OK;
}

template <int T>
int launch_tile_transpose(const uint64_t *keys, long long nblocks, int pb, const int *row_of, const void *val,
                          int *t_colind, void *t_val, hipStream_t s) {
  const long long waves = (nblocks + 64 / T - 1) / (64 / T);
  tile_transpose_kernel<T><<<capped_blocks(waves * 64, kBlock), kBlock, 0, s>>>(keys, nblocks, pb, row_of, val,
                                                                                t_colind, t_val);
  BGB_HIP_TRY(hipGetLastError());
  return BGB_OK;
}

}  // namespace
}  // namespace bgb

using namespace bgb;

extern "C" {

int bgb_csr2csc(int m, int n, int nnz, const float *d_csrVal, const int *d_csrRowPtr, const int *d_csrColInd,
                float *d_cscVal, int *d_cscRowInd, int *d_cscColPtr, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, BGB_ERR_ARG, "negative size");
    BGB_REQUIRE(d_csrRowPtr && d_cscColPtr && (nnz == 0 || (d_csrColInd && d_cscRowInd)), BGB_ERR_ARG,
                "null CSR / CSC array");
    BGB_REQUIRE((d_csrVal == nullptr) == (d_cscVal == nullptr), BGB_ERR_ARG,
                "values: pass both csrVal and cscVal, or neither (structure only)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st = check_structure(d_csrRowPtr, d_csrColInd, m, n, nnz, s, "CSR");
    if (st) return st;
    DeviceBuffer keys, row_of;
    int pb = 0;
    if ((st = sorted_entries(d_csrRowPtr, d_csrColInd, m, n, nnz, s, keys, row_of, d_cscColPtr, &pb))) return st;
    if (nnz > 0) {
      csc_scatter_kernel<<<capped_blocks(nnz, kBlock), kBlock, 0, s>>>(keys.as<uint64_t>(), nnz, pb, row_of.as<int>(),
                                                                       d_csrVal, d_cscRowInd, d_cscVal);
      BGB_HIP_TRY(hipGetLastError());
    }
    return BGB_OK;
  });
}

int bgb_b2sr_transpose(int T, int nblockrows, int nblockcols, int nblocks, const int *d_rowptr,
                       const int *d_colind, const void *d_val, int *d_t_rowptr, int *d_t_colind, void *d_t_val,
                       bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
    BGB_REQUIRE(nblockrows >= 0 && nblockcols >= 0 && nblocks >= 0, BGB_ERR_ARG, "negative size");
    BGB_REQUIRE(d_rowptr && d_t_rowptr && (nblocks == 0 || (d_colind && d_val && d_t_colind && d_t_val)),
                BGB_ERR_ARG, "null B2SR array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st = check_structure(d_rowptr, d_colind, nblockrows, nblockcols, nblocks, s, "B2SR");
    if (st) return st;
    DeviceBuffer keys, row_of;
    int pb = 0;
    if ((st = sorted_entries(d_rowptr, d_colind, nblockrows, nblockcols, nblocks, s, keys, row_of, d_t_rowptr, &pb)))
      return st;
    if (nblocks == 0) return BGB_OK;
    const uint64_t *k = keys.as<uint64_t>();
    const int *ro = row_of.as<int>();
    switch (T) {
      case 4: return launch_tile_transpose<4>(k, nblocks, pb, ro, d_val, d_t_colind, d_t_val, s);
      case 8: return launch_tile_transpose<8>(k, nblocks, pb, ro, d_val, d_t_colind, d_t_val, s);
      case 16: return launch_tile_transpose<16>(k, nblocks, pb, ro, d_val, d_t_colind, d_t_val, s);
      case 32: return launch_tile_transpose<32>(k, nblocks, pb, ro, d_val, d_t_colind, d_t_val, s);
      default: return launch_tile_transpose<64>(k, nblocks, pb, ro, d_val, d_t_colind, d_t_val, s);
    }
  });
}

}  // extern "C"
