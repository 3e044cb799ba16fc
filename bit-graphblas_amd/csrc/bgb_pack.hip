// CSR -> B2SR packing and bit-vector packing on the GPU (gfx950, wave64).
//
// The reference converts one block row at a time on the host: it copies the
// row's CSR slice to the device, expands it to dense FP32 tiles with
// cusparseScsr2bsr and packs those with a 32-lane ballot
// (csr2bsr_batch_bsrbmv.cu:34-133, bsrbmv.cu:33-222). Here the whole matrix is
// packed in a handful of launches with no dense expansion:
//   1. one 64-bit key (I, J, r, c) per nonzero (tile row/col, row/col in tile);
//   2. radix sort of the keys -- CSR rows are already ordered by I, so the sort
//      only reorders inside block rows;
//   3. tile-head flags + exclusive scan: the scan at a block row's first CSR
//      entry is that row's bsrRowPtr, the scan total is nblocks;
//   4. one thread per tile ORs its nonzeros into T row words (column c at bit
//      T-1-c) and writes them whole, so bsrVal needs no memset or atomics.
// Vector packing uses the 64-bit ballot: one wave covers 64 consecutive x
// entries, i.e. 16/8/4/2 packed words for T = 4/8/16/32.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bgb_device.h"

namespace bgb {
namespace {

constexpr int kBlock = 256;

inline int bits_for(long long count) {  // bits needed to hold values in [0, count)
  int b = 1;
  while (b < 62 && (1LL << b) < count) ++b;
  return b;
}

struct KeyLayout {
  int lg, bitsJ, bitsI;
  int end_bit() const { return bitsI + bitsJ + 2 * lg; }
};

// CSR checks, parallel over rows and over nonzeros (a per-row loop would
// serialise on hub rows): rowptr bounds/monotonicity per row, row-start marks,
// then per nonzero the column range and strict increase inside its row.
__global__ void validate_rows_kernel(const int *__restrict__ rowptr, int nrows, int nnz,
                                     uint8_t *__restrict__ row_start, int *__restrict__ bad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nrows) return;
  const int s = rowptr[i];
  if (i == 0 && s != 0) atomicOr(bad, 1);
  if (i == nrows) {
    if (s != nnz) atomicOr(bad, 2);
    return;
  }
  const int e = rowptr[i + 1];
  if (e < s || e > nnz || s < 0) {
    atomicOr(bad, 4);
    return;
  }
  if (s < e) row_start[s] = 1;
}

__global__ void validate_cols_kernel(const int *__restrict__ colind, long long nnz, int ncols,
                                     const uint8_t *__restrict__ row_start, int *__restrict__ bad) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int j = colind[k];
  if (j < 0 || j >= ncols || (k > 0 && !row_start[k] && colind[k - 1] >= j)) atomicOr(bad, 8);
}

// One wave per CSR row (grid-stride over rows: 64 work-items per row would
// pass the dispatch's 32-bit grid size beyond 2^26 rows); lanes stride over
// the row's nonzeros.
__global__ void make_keys_kernel(const int *__restrict__ rowptr, const int *__restrict__ colind,
                                 int nrows, KeyLayout kl, uint64_t *__restrict__ keys) {
  const int lane = threadIdx.x & 63;
  const int mask = (1 << kl.lg) - 1;
  const long long waves = (long long)gridDim.x * (kBlock / 64);
  for (long long row = (long long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); row < nrows; row += waves) {
    const uint64_t I = (uint64_t)(row >> kl.lg);
    const uint64_t r = (uint64_t)(row & mask);
    const uint64_t hi = (I << (kl.bitsJ + 2 * kl.lg)) | (r << kl.lg);
    const int s = rowptr[row], e = rowptr[row + 1];
    for (int k = s + lane; k < e; k += 64) {
      const uint32_t j = (uint32_t)colind[k];
      keys[k] = hi | ((uint64_t)(j >> kl.lg) << (2 * kl.lg)) | (uint64_t)(j & mask);
    }
  }
}

__global__ void tile_flags_kernel(const uint64_t *__restrict__ keys, long long nnz, int lg,
                                  uint8_t *__restrict__ flags) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e > nnz) return;
  if (e == nnz) {
    flags[e] = 0;
    return;
  }
  flags[e] = (e == 0) || ((keys[e] >> (2 * lg)) != (keys[e - 1] >> (2 * lg)));
}

__global__ void block_rowptr_kernel(const int *__restrict__ csr_rowptr, int nrows, int T,
                                    int nblockrows, const long long *__restrict__ excl,
                                    int *__restrict__ bsr_rowptr) {
  long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I > nblockrows) return;
  long long i = I * T < nrows ? I * T : nrows;
  bsr_rowptr[I] = (int)excl[csr_rowptr[i]];
}

__global__ void tile_heads_kernel(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ flags,
                                  const long long *__restrict__ excl, long long nnz, KeyLayout kl,
                                  int *__restrict__ colind_out, int *__restrict__ head_pos) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) {
    if (e == nnz) head_pos[excl[nnz]] = (int)nnz;
    return;
  }
  if (!flags[e]) return;
  const long long t = excl[e];
  head_pos[t] = (int)e;
  colind_out[t] = (int)((keys[e] >> (2 * kl.lg)) & ((1ull << kl.bitsJ) - 1));
}

template <int T>
__global__ void tile_fill_kernel(const uint64_t *__restrict__ keys, const int *__restrict__ head_pos,
                                 int nblocks, void *__restrict__ val) {
  constexpr int lg = T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : 5;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblocks) return;
  uint32_t w[T];
#pragma unroll
  for (int r = 0; r < T; ++r) w[r] = 0;
  const int s = head_pos[t], e = head_pos[t + 1];
  for (int k = s; k < e; ++k) {
    const uint32_t rc = (uint32_t)(keys[k] & ((1u << (2 * lg)) - 1));
    const int r = rc >> lg, c = rc & (T - 1);
    // rows are few and compile-time indexed; select the word without a
    // dynamically indexed register array
#pragma unroll
    for (int rr = 0; rr < T; ++rr)
      if (rr == r) w[rr] |= 1u << (T - 1 - c);
  }
  if constexpr (T == 4) {
    static_cast<uint32_t *>(val)[t] = w[0] | (w[1] << 8) | (w[2] << 16) | (w[3] << 24);
  } else if constexpr (T == 8) {
    uint2 v;
    v.x = w[0] | (w[1] << 8) | (w[2] << 16) | (w[3] << 24);
    v.y = w[4] | (w[5] << 8) | (w[6] << 16) | (w[7] << 24);
    static_cast<uint2 *>(val)[t] = v;
  } else if constexpr (T == 16) {
    uint4 *o = static_cast<uint4 *>(val) + t * 2;
    o[0] = make_uint4(w[0] | (w[1] << 16), w[2] | (w[3] << 16), w[4] | (w[5] << 16), w[6] | (w[7] << 16));
    o[1] = make_uint4(w[8] | (w[9] << 16), w[10] | (w[11] << 16), w[12] | (w[13] << 16),
                      w[14] | (w[15] << 16));
  } else {
    uint4 *o = static_cast<uint4 *>(val) + t * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
  }
}

// One wave packs 64 consecutive entries of x with a 64-bit ballot.
// T = 64: one thread per (tile, row) word. The tile's nonzeros are sorted by
// (row, column), so the thread ORs the bits of its row's run; 64 words of 8 B
// per tile would not fit one thread's registers.
// Grid-stride: 64 words per tile pass the 32-bit grid size beyond 2^26 tiles.
__global__ void tile_fill64_kernel(const uint64_t *__restrict__ keys, const int *__restrict__ head_pos,
                                   long long nwords, uint64_t *__restrict__ val) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nwords; t += stride) {
    const long long tile = t >> 6;
    const int r = (int)(t & 63);
    const int s = head_pos[tile], e = head_pos[tile + 1];
    uint64_t w = 0;
    for (int k = s; k < e; ++k) {
      const uint32_t rc = (uint32_t)(keys[k] & 0xFFFu);  // 6 row bits, 6 column bits
      if ((int)(rc >> 6) == r) w |= 1ull << (63 - (rc & 63));
    }
    val[t] = w;
  }
}

template <int T>
__global__ void pack_vec_kernel(const float *__restrict__ x, long long n, long long nwords,
                                void *__restrict__ xb) {
  constexpr int PER = 64 / T;  // packed words per wave
  const long long base = ((long long)blockIdx.x * blockDim.x + threadIdx.x) & ~63LL;
  const int lane = threadIdx.x & 63;
  const long long i = base + lane;
  const bool bit = (i < n) && (x[i] > 0.0f);
  const uint64_t m = __ballot(bit);
  const uint64_t rb = __builtin_bitreverse64(m);  // x[base + l] now at bit 63 - l
  const long long word = base / T + lane;
  if (lane < PER && word < nwords) {
    if constexpr (T == 64) {
      static_cast<uint64_t *>(xb)[word] = rb;
      return;
    }
    const uint64_t v = (rb >> (64 - T * (lane + 1))) & ((T == 32) ? 0xffffffffull : ((1ull << T) - 1));
    if constexpr (T == 32) static_cast<uint32_t *>(xb)[word] = (uint32_t)v;
    else if constexpr (T == 16) static_cast<uint16_t *>(xb)[word] = (uint16_t)v;
    else static_cast<uint8_t *>(xb)[word] = (uint8_t)v;
  }
}

unsigned grid_for(long long n, int block = kBlock) { return (unsigned)ceil_div(std::max(n, 1LL), block); }

int check_common(int T, int nrows, int ncols, int nnz, const int *rowptr, const int *colind) {
  BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
  BGB_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, BGB_ERR_ARG, "negative size");
  BGB_REQUIRE(rowptr && (nnz == 0 || colind), BGB_ERR_ARG, "null CSR array");
  return BGB_OK;
}

// Checks the CSR structure on the device (synchronises the stream).
int validate_csr(const int *rowptr, const int *colind, int nrows, int ncols, int nnz, hipStream_t s) {
  DeviceBuffer bad, starts;
  BGB_HIP_TRY(bad.alloc(sizeof(int), s));
  BGB_HIP_TRY(starts.alloc((size_t)nnz + 1, s));
  BGB_HIP_TRY(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
  BGB_HIP_TRY(hipMemsetAsync(starts.ptr, 0, (size_t)nnz + 1, s));
  validate_rows_kernel<<<grid_for((long long)nrows + 1), kBlock, 0, s>>>(rowptr, nrows, nnz, starts.as<uint8_t>(),
                                                                        bad.as<int>());
  BGB_HIP_TRY(hipGetLastError());
  if (nnz > 0) {
    validate_cols_kernel<<<grid_for(nnz), kBlock, 0, s>>>(colind, nnz, ncols, starts.as<uint8_t>(), bad.as<int>());
    BGB_HIP_TRY(hipGetLastError());
  }
  int h_bad = 0;
  BGB_HIP_TRY(hipMemcpyAsync(&h_bad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  BGB_HIP_TRY(hipStreamSynchronize(s));
  BGB_REQUIRE(h_bad == 0, BGB_ERR_FORMAT,
              "invalid CSR (code %d: 1/2 rowptr ends, 4 non-monotone rowptr, 8 colind unsorted, "
              "duplicated or out of range)",
              h_bad);
  return BGB_OK;
}

// Steps 1-3 shared by both packing calls: sorted keys, flags and their scan.
int sorted_tiles(int T, int nrows, int ncols, int nnz, const int *rowptr, const int *colind,
                 hipStream_t s, DeviceBuffer &keys, DeviceBuffer &flags, DeviceBuffer &excl,
                 KeyLayout &kl) {
  const int lg = ilog2_tiledim(T);
  const long long nbr = ceil_div(nrows, T), nbc = ceil_div(ncols, T);
  kl = KeyLayout{lg, bits_for(nbc), bits_for(nbr)};
  BGB_REQUIRE(kl.end_bit() <= 64, BGB_ERR_SIZE, "matrix too large for 64-bit tile keys");
  DeviceBuffer raw;
  BGB_HIP_TRY(raw.alloc(sizeof(uint64_t) * (size_t)nnz, s));
  BGB_HIP_TRY(keys.alloc(sizeof(uint64_t) * (size_t)nnz, s));
  BGB_HIP_TRY(flags.alloc((size_t)nnz + 1, s));
  BGB_HIP_TRY(excl.alloc(sizeof(long long) * ((size_t)nnz + 1), s));
  if (nrows > 0)
    make_keys_kernel<<<capped_blocks((long long)nrows * 64, kBlock), kBlock, 0, s>>>(rowptr, colind, nrows, kl,
                                                                     raw.as<uint64_t>());
  BGB_HIP_TRY(hipGetLastError());
  if (nnz > 0) BGB_HIP_TRY(radix_sort_u64(raw.as<uint64_t>(), keys.as<uint64_t>(), nnz, kl.end_bit(), s));
  tile_flags_kernel<<<grid_for((long long)nnz + 1), kBlock, 0, s>>>(keys.as<uint64_t>(), nnz, lg,
                                                                    flags.as<uint8_t>());
  BGB_HIP_TRY(hipGetLastError());
  BGB_HIP_TRY(exclusive_scan_u8_to_i64(flags.as<uint8_t>(), excl.as<long long>(), (size_t)nnz + 1, s));
  return BGB_OK;
}

}  // namespace
}  // namespace bgb

using namespace bgb;

extern "C" {

int bgb_b2sr_nnzb(int T, int nrows, int ncols, int nnz, const int *d_csrRowPtr,
                  const int *d_csrColInd, int *d_bsrRowPtr, int *h_nblocks, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    int st = check_common(T, nrows, ncols, nnz, d_csrRowPtr, d_csrColInd);
    if (st) return st;
    BGB_REQUIRE(d_bsrRowPtr && h_nblocks, BGB_ERR_ARG, "null output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = validate_csr(d_csrRowPtr, d_csrColInd, nrows, ncols, nnz, s))) return st;
    DeviceBuffer keys, flags, excl;
    KeyLayout kl;
    if ((st = sorted_tiles(T, nrows, ncols, nnz, d_csrRowPtr, d_csrColInd, s, keys, flags, excl, kl)))
      return st;
    const int nbr = (int)ceil_div(nrows, T);
    block_rowptr_kernel<<<grid_for((long long)nbr + 1), kBlock, 0, s>>>(d_csrRowPtr, nrows, T, nbr,
                                                                       excl.as<long long>(), d_bsrRowPtr);
    BGB_HIP_TRY(hipGetLastError());
    long long total = 0;
    BGB_HIP_TRY(hipMemcpyAsync(&total, excl.as<long long>() + nnz, sizeof(long long),
                               hipMemcpyDeviceToHost, s));
    BGB_HIP_TRY(hipStreamSynchronize(s));
    *h_nblocks = (int)total;
    return BGB_OK;
  });
}

int bgb_csr2b2sr(int T, int nrows, int ncols, int nnz, const int *d_csrRowPtr,
                 const int *d_csrColInd, const int *d_bsrRowPtr, int nblocks, int *d_bsrColInd,
                 void *d_bsrVal, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    int st = check_common(T, nrows, ncols, nnz, d_csrRowPtr, d_csrColInd);
    if (st) return st;
    BGB_REQUIRE(nblocks >= 0 && d_bsrRowPtr, BGB_ERR_ARG, "bad block row pointer / nblocks");
    BGB_REQUIRE(nblocks == 0 || (d_bsrColInd && d_bsrVal), BGB_ERR_ARG, "null output");
    BGB_REQUIRE(nblocks <= nnz, BGB_ERR_ARG, "nblocks %d exceeds nnz %d", nblocks, nnz);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = validate_csr(d_csrRowPtr, d_csrColInd, nrows, ncols, nnz, s))) return st;
    if (nblocks == 0 && nnz == 0) return BGB_OK;
    DeviceBuffer keys, flags, excl, heads;
    KeyLayout kl;
    if ((st = sorted_tiles(T, nrows, ncols, nnz, d_csrRowPtr, d_csrColInd, s, keys, flags, excl, kl)))
      return st;
    long long total = 0;
    BGB_HIP_TRY(hipMemcpyAsync(&total, excl.as<long long>() + nnz, sizeof(long long),
                               hipMemcpyDeviceToHost, s));
    BGB_HIP_TRY(hipStreamSynchronize(s));
    BGB_REQUIRE(total == nblocks, BGB_ERR_ARG, "nblocks %d does not match the matrix (%lld tiles)",
                nblocks, total);
    BGB_HIP_TRY(heads.alloc(sizeof(int) * ((size_t)nblocks + 1), s));
    tile_heads_kernel<<<grid_for((long long)nnz + 1), kBlock, 0, s>>>(
        keys.as<uint64_t>(), flags.as<uint8_t>(), excl.as<long long>(), nnz, kl, d_bsrColInd,
        heads.as<int>());
    BGB_HIP_TRY(hipGetLastError());
    const unsigned g = grid_for(nblocks);
    switch (T) {
      case 4: tile_fill_kernel<4><<<g, kBlock, 0, s>>>(keys.as<uint64_t>(), heads.as<int>(), nblocks, d_bsrVal); break;
      case 8: tile_fill_kernel<8><<<g, kBlock, 0, s>>>(keys.as<uint64_t>(), heads.as<int>(), nblocks, d_bsrVal); break;
      case 16: tile_fill_kernel<16><<<g, kBlock, 0, s>>>(keys.as<uint64_t>(), heads.as<int>(), nblocks, d_bsrVal); break;
      case 32: tile_fill_kernel<32><<<g, kBlock, 0, s>>>(keys.as<uint64_t>(), heads.as<int>(), nblocks, d_bsrVal); break;
      default:
        tile_fill64_kernel<<<capped_blocks(64LL * nblocks, kBlock), kBlock, 0, s>>>(keys.as<uint64_t>(), heads.as<int>(),
                                                                       64LL * nblocks, static_cast<uint64_t *>(d_bsrVal));
        break;
    }
    BGB_HIP_TRY(hipGetLastError());
    return BGB_OK;
  });
}

int bgb_pack_vec(int T, const float *d_x, int n, void *d_xb, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
    BGB_REQUIRE(n >= 0 && d_xb && (n == 0 || d_x), BGB_ERR_ARG, "bad vector arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long nbr = ceil_div(n, T);
    const long long nwords = ceil_div(nbr, T) * T;
    if (nwords == 0) return BGB_OK;
    const long long threads = ceil_div(nwords * T, 64) * 64;
    const unsigned g = grid_for(threads);
    switch (T) {
      case 4: pack_vec_kernel<4><<<g, kBlock, 0, s>>>(d_x, n, nwords, d_xb); break;
      case 8: pack_vec_kernel<8><<<g, kBlock, 0, s>>>(d_x, n, nwords, d_xb); break;
      case 16: pack_vec_kernel<16><<<g, kBlock, 0, s>>>(d_x, n, nwords, d_xb); break;
      case 32: pack_vec_kernel<32><<<g, kBlock, 0, s>>>(d_x, n, nwords, d_xb); break;
      default: pack_vec_kernel<64><<<g, kBlock, 0, s>>>(d_x, n, nwords, d_xb); break;
    }
    BGB_HIP_TRY(hipGetLastError());
    return BGB_OK;
  });
}

}  // extern "C"
