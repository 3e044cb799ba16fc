// Device-side helpers shared by the HIP translation units of libbgb.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bgb_internal.h"

#define BGB_HIP_TRY(expr)                                                         \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess)                                                         \
      return ::bgb::set_error(BGB_ERR_HIP, "%s failed: %s (%s:%d)", #expr,        \
                              hipGetErrorString(_e), __FILE__, __LINE__);         \
  } while (0)

namespace bgb {

// Stream-ordered temporary device allocation, released on the same stream.
struct DeviceBuffer {
  void *ptr = nullptr;
  hipStream_t stream = nullptr;
  hipError_t alloc(size_t bytes, hipStream_t s) {
    stream = s;
    if (bytes == 0) bytes = 256;
    return hipMallocAsync(&ptr, bytes, s);
  }
  ~DeviceBuffer() {
    if (ptr) (void)hipFreeAsync(ptr, stream);
  }
  template <class P>
  P *as() const { return static_cast<P *>(ptr); }
};

hipError_t radix_sort_u64(const uint64_t *in, uint64_t *out, size_t n, int end_bit, hipStream_t s);
hipError_t radix_sort_u64_bits(const uint64_t *in, uint64_t *out, size_t n, int begin_bit, int end_bit,
                               hipStream_t s);
hipError_t exclusive_scan_u8_to_i64(const uint8_t *in, long long *out, size_t n, hipStream_t s);
hipError_t exclusive_scan_i32(const int *in, int *out, size_t n, hipStream_t s);

// Per-(device, stream) grow-only scratch used by the BMV carry records. The
// pointer stays valid until a later call on the same stream asks for more.
int stream_workspace(hipStream_t s, size_t bytes, void **out);
// Chunk -> first block row for chunks of CH tiles, cached per (device, stream,
// rowptr address, sizes); the kernels validate and repair entries.
int chunk_partition(hipStream_t s, const int *rowptr, int nbr, int nblocks, int CH, int **out);
int max_row_tiles(hipStream_t s, const int *rowptr, int nbr, int nblocks, int *out);
// Largest block column of colind[0..nblocks), cached per (device, colind
// address, nblocks); -1 when unknown (stream being captured).
int max_block_col(hipStream_t s, const int *colind, int nblocks, int *out);

// Per-device launch facts, queried once per device: compute units and the
// LDS bytes a workgroup / a CU may use.
struct DeviceFacts {
  int cus = 0;
  int lds_per_block = 0;
  int lds_per_cu = 0;
};
int device_facts(DeviceFacts *out);

// Workgroups for n items of `block` threads, capped: the dispatch packet
// carries the grid size in work-items as a 32-bit number, so launches whose
// item count can pass 2^32 (per-carry-value fix-ups, fills of 16 words per
// vertex) use at most 2^20 workgroups and walk their range with a
// grid-stride loop.
inline unsigned capped_blocks(long long n, int block) {
  const long long b = (n + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : (b > (1LL << 20) ? (1LL << 20) : b));
}

inline int ilog2_tiledim(int T) { return T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 6; }

}  // namespace bgb
