// Internal helpers shared by the host (.cpp) and device (.hip) translation
// units of libbgb: status reporting, the counter-based random stream and
// tile-dimension traits on the host side.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <new>
#include <stdexcept>

#include "../../include/bgb.h"

namespace bgb {

// Records a message for bgb_last_error() on this host thread and returns status.
int set_error(int status, const char *fmt, ...);
void clear_error();

// Runs an entry point's body and turns a C++ exception from host code (a
// host allocation that fails, a container length beyond its limit) into a
// status: no exception crosses the C ABI.
template <class F>
int host_guard(F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return set_error(BGB_ERR_ALLOC, "host allocation failed");
  } catch (const std::length_error &e) {
    return set_error(BGB_ERR_SIZE, "size exceeds a host container limit (%s)", e.what());
  } catch (const std::exception &e) {
    return set_error(BGB_ERR_ARG, "host exception: %s", e.what());
  } catch (...) {
    return set_error(BGB_ERR_ARG, "unknown host exception");
  }
}

// BMV tuning knobs per tile dim (index log2(T) - 2), set by
// bgb_set_run_length / bgb_set_x_staging.
extern int g_run_length[5];
extern int g_x_staging[5];

inline bool valid_tiledim(int T) { return T == 4 || T == 8 || T == 16 || T == 32 || T == 64; }
inline int word_bytes(int T) {
  return T == 64 ? 8 : T == 32 ? 4 : (T == 16 ? 2 : (T == 4 || T == 8 ? 1 : 0));
}
inline long long ceil_div(long long a, long long b) { return (a + b - 1) / b; }

// splitmix64 finaliser (Steele, Lea, Flood 2014). The random value for index i
// of stream (seed, sid) is mix64(mix64(seed + sid * golden) + i * golden), so
// every element is computable independently of the others.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t sid) {
  return mix64(seed + sid * 0x9E3779B97F4A7C15ull);
}
inline uint64_t rand_at(uint64_t key, uint64_t i) {
  return mix64(key ^ (i * 0xD1B54A32D192ED03ull));
}
// Uniform double in [0, 1) with 53 random bits.
inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace bgb

#define BGB_REQUIRE(cond, status, ...)                   \
  do {                                                    \
    if (!(cond)) return ::bgb::set_error(status, __VA_ARGS__); \
  } while (0)
