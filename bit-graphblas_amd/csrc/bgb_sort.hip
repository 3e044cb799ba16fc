// Library sort / scan primitives used by the CSR -> B2SR packer, kept in their
// own translation unit because the rocPRIM templates dominate compile time.
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "bgb_device.h"

namespace bgb {

hipError_t radix_sort_u64(const uint64_t *in, uint64_t *out, size_t n, int end_bit,
                          hipStream_t s) {
  size_t bytes = 0;
  hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, in, out, n, 0, end_bit, s);
  if (e != hipSuccess) return e;
  DeviceBuffer tmp;
  if ((e = tmp.alloc(bytes, s)) != hipSuccess) return e;
  return rocprim::radix_sort_keys(tmp.ptr, bytes, in, out, n, 0, end_bit, s);
}

// Stable sort on key bits [begin_bit, end_bit): keys equal in those bits keep
// their input order (rocPRIM's radix sort is an LSD sort of stable passes).
hipError_t radix_sort_u64_bits(const uint64_t *in, uint64_t *out, size_t n, int begin_bit, int end_bit,
                               hipStream_t s) {
  size_t bytes = 0;
  hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, in, out, n, begin_bit, end_bit, s);
  if (e != hipSuccess) return e;
  DeviceBuffer tmp;
  if ((e = tmp.alloc(bytes, s)) != hipSuccess) return e;
  return rocprim::radix_sort_keys(tmp.ptr, bytes, in, out, n, begin_bit, end_bit, s);
}

hipError_t exclusive_scan_u8_to_i64(const uint8_t *in, long long *out, size_t n,
                                    hipStream_t s) {
  size_t bytes = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, 0LL, n,
                                         rocprim::plus<long long>(), s);
  if (e != hipSuccess) return e;
  DeviceBuffer tmp;
  if ((e = tmp.alloc(bytes, s)) != hipSuccess) return e;
  return rocprim::exclusive_scan(tmp.ptr, bytes, in, out, 0LL, n, rocprim::plus<long long>(), s);
}

hipError_t exclusive_scan_i32(const int *in, int *out, size_t n, hipStream_t s) {
  size_t bytes = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, 0, n, rocprim::plus<int>(), s);
  if (e != hipSuccess) return e;
  DeviceBuffer tmp;
  if ((e = tmp.alloc(bytes, s)) != hipSuccess) return e;
  return rocprim::exclusive_scan(tmp.ptr, bytes, in, out, 0, n, rocprim::plus<int>(), s);
}

}  // namespace bgb
