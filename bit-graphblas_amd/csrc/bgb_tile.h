// B2SR tile traits and per-lane fragment access for gfx950 (wave64).
//
// A lane group of L lanes owns one tile at a time; each lane holds R of its T
// rows in NW 32-bit registers (its "fragment"):
//   T=4 : 1 lane,  4 rows, 1 dword  -- byte r = row r, low nibble, col c at bit 3-c
//   T=8 : 1 lane,  8 rows, 2 dwords -- byte r = row r, col c at bit 7-c
//   T=16: 2 lanes, 8 rows, 4 dwords -- half-word i = row 8q+i, col c at bit 15-c
//   T=32: 8 lanes, 4 rows, 4 dwords -- dword i = row 4q+i, col c at bit 31-c
//   T=64: 16 lanes, 4 rows, 8 dwords -- qword i = row 4q+i, col c at bit 63-c
// Row words are 32-bit (Word) up to T = 32 and 64-bit at T = 64.
// Every lane therefore loads 4, 8 or 16 contiguous bytes per tile and a wave
// streams 64/L tiles per instruction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace bgb {

template <int T>
struct Tile;
template <>
struct Tile<4> {
  using W = uint8_t;
  static constexpr int L = 1, R = 4, NW = 1, LG = 2;
};
template <>
struct Tile<8> {
  using W = uint8_t;
  static constexpr int L = 1, R = 8, NW = 2, LG = 3;
};
template <>
struct Tile<16> {
  using W = uint16_t;
  static constexpr int L = 2, R = 8, NW = 4, LG = 4;
};
template <>
struct Tile<32> {
  using W = uint32_t;
  static constexpr int L = 8, R = 4, NW = 4, LG = 5;
};
template <>
struct Tile<64> {
  using W = uint64_t;
  static constexpr int L = 16, R = 4, NW = 8, LG = 6;
};

// Register type of one row word (and of a packed-x word).
template <int T>
using Word = typename std::conditional<T == 64, uint64_t, uint32_t>::type;

__device__ __forceinline__ int bit_count(uint32_t w) { return __popc(w); }
__device__ __forceinline__ int bit_count(uint64_t w) { return __popcll(w); }
// Index of the highest set bit of a non-zero word.
__device__ __forceinline__ int top_bit(uint32_t w) { return 31 - __clz((int)w); }
__device__ __forceinline__ int top_bit(uint64_t w) { return 63 - __clzll((long long)w); }

// x where bit b of w is set, else +0.0 (whose bits do not depend on x, so a
// non-finite x in a clear column never reaches a sum): the sign-extended
// 1-bit field of w is an all-ones or all-zero mask -- an extract and an AND,
// no compare, no condition-register hazard.
__device__ __forceinline__ float bit_select(uint32_t w, int b, float x) {
  return __uint_as_float(__float_as_uint(x) & (uint32_t)__builtin_amdgcn_sbfe((int)w, b, 1));
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// 16-byte non-temporal load (streamed data that is read once per pass).
__device__ __forceinline__ uint4 load_nt16(const void *p) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 v = __builtin_nontemporal_load(static_cast<const u32x4 *>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Row i (0 <= i < R) of a lane fragment, as a T-bit word.
template <int T>
__device__ __forceinline__ Word<T> row_word(const uint32_t (&F)[Tile<T>::NW], int i) {
  if constexpr (T == 4) return (F[0] >> (8 * i)) & 0xFu;
  else if constexpr (T == 8) return (F[i >> 2] >> (8 * (i & 3))) & 0xFFu;
  else if constexpr (T == 16) return (F[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
  else if constexpr (T == 32) return F[i];
  else return (uint64_t)F[2 * i] | ((uint64_t)F[2 * i + 1] << 32);
}

// Fragment of tile p for lane q of its group.
template <int T>
__device__ __forceinline__ void load_frag(const void *__restrict__ A, long long p, int q,
                                          uint32_t (&F)[Tile<T>::NW]) {
  if constexpr (T == 4) {
    F[0] = static_cast<const uint32_t *>(A)[p];
  } else if constexpr (T == 8) {
    const uint2 v = static_cast<const uint2 *>(A)[p];
    F[0] = v.x;
    F[1] = v.y;
  } else if constexpr (T <= 32) {
    constexpr int L = Tile<T>::L;
    // T = 32 streams 128 B per tile: non-temporal loads keep the tile stream
    // from displacing x, rowptr and the outputs in L2 (measured faster at
    // T = 32 only; the smaller tiles lose with them)
    const uint4 v = T == 32 ? load_nt16(static_cast<const uint4 *>(A) + p * L + q)
                            : static_cast<const uint4 *>(A)[p * L + q];
    F[0] = v.x;
    F[1] = v.y;
    F[2] = v.z;
    F[3] = v.w;
  } else {
    // T = 64: 512 B per tile, 32 B (rows 4q..4q+3) per lane, non-temporal
    const uint4 *src = static_cast<const uint4 *>(A) + (p * 16 + q) * 2;
    const uint4 v0 = load_nt16(src), v1 = load_nt16(src + 1);
    F[0] = v0.x;
    F[1] = v0.y;
    F[2] = v0.z;
    F[3] = v0.w;
    F[4] = v1.x;
    F[5] = v1.y;
    F[6] = v1.z;
    F[7] = v1.w;
  }
}

// Loads the fragments and block columns of tiles [pa, pa+n), 1 <= n <= S.
// With n == S and pa a multiple of S (S a multiple of 4) the L=1 tile dims
// use 16-byte loads over the lane's contiguous run.
// kFull: the caller guarantees n == S (no short-run code is emitted, which
// keeps the compiler from merging the two paths into per-dword loads).
template <int T, int S, bool kFull = false>
__device__ __forceinline__ void load_run(const void *__restrict__ A, const int *__restrict__ colind,
                                         long long pa, int n, int q,
                                         uint32_t (&F)[S][Tile<T>::NW], int (&J)[S]) {
  static_assert(S % 4 == 0, "run length must be a multiple of 4");
  if (kFull || n == S) {
    const int4 *ci = reinterpret_cast<const int4 *>(colind + pa);
#pragma unroll
    for (int s = 0; s < S / 4; ++s) {
      int4 c;
      if constexpr (T >= 32) {
        const uint4 u = load_nt16(ci + s);
        c = make_int4((int)u.x, (int)u.y, (int)u.z, (int)u.w);
      } else {
        c = ci[s];
      }
      J[4 * s] = c.x;
      J[4 * s + 1] = c.y;
      J[4 * s + 2] = c.z;
      J[4 * s + 3] = c.w;
    }
    if constexpr (T == 4) {
      const uint4 *a = static_cast<const uint4 *>(A) + pa / 4;
#pragma unroll
      for (int s = 0; s < S / 4; ++s) {
        const uint4 v = a[s];
        F[4 * s][0] = v.x;
        F[4 * s + 1][0] = v.y;
        F[4 * s + 2][0] = v.z;
        F[4 * s + 3][0] = v.w;
      }
    } else if constexpr (T == 8) {
      const uint4 *a = static_cast<const uint4 *>(A) + pa / 2;
#pragma unroll
      for (int s = 0; s < S / 2; ++s) {
        const uint4 v = a[s];
        F[2 * s][0] = v.x;
        F[2 * s][1] = v.y;
        F[2 * s + 1][0] = v.z;
        F[2 * s + 1][1] = v.w;
      }
    } else {
#pragma unroll
      for (int s = 0; s < S; ++s) load_frag<T>(A, pa + s, q, F[s]);
    }
  } else if constexpr (kFull) {
  } else if constexpr (T <= 8) {
    // A short run (the matrix's last). Entries s >= n repeat tile pa+n-1, so
    // every element is written unconditionally (callers only use s < n);
    // conditional writes here leave a pointer select over two stack slots
    // (scratch) in the T = 4 kernels.
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const long long p = pa + (s < n ? s : n - 1);
      J[s] = colind[p];
      load_frag<T>(A, p, q, F[s]);
    }
  } else {
    // T >= 16: zero-filled; the unconditional form above would be merged
    // with the full-run path (the same load_frag calls) into per-dword loads.
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (s < n) {
        J[s] = colind[pa + s];
        load_frag<T>(A, pa + s, q, F[s]);
      } else {
        J[s] = 0;
#pragma unroll
        for (int w = 0; w < Tile<T>::NW; ++w) F[s][w] = 0;
      }
    }
  }
}

// Largest I in [lo, hi] with rowptr[I] <= p, given rowptr[lo] <= p, searched
// by the whole wave: each step probes 64 evenly spaced entries and keeps the
// interval between the last probe <= p and the next one (log64 steps). Must
// be called by all 64 lanes.
__device__ __forceinline__ int wave_find_row(const int *__restrict__ rowptr, int lo, int hi, long long p,
                                             int lane) {
  while (hi > lo) {
    const long long span = hi - lo;
    const int idx = lo + (int)((span * (lane + 1)) >> 6);
    const int cnt = __popcll(__ballot(rowptr[idx] <= p));
    const int nlo = cnt > 0 ? lo + (int)((span * cnt) >> 6) : lo;
    const int nhi = cnt < 64 ? lo + (int)((span * (cnt + 1)) >> 6) - 1 : hi;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// The rows holding tiles p0 <= p1, found together: lanes 0-31 search for p0,
// lanes 32-63 for p1, each half probing 32 entries per step. Must be called
// by all 64 lanes; returns the pair in every lane.
__device__ __forceinline__ int2 wave_find_rows2(const int *__restrict__ rowptr, int hi0, long long p0, long long p1,
                                                int lane) {
  const int half = lane >> 5, hl = lane & 31;
  const long long p = half ? p1 : p0;
  int lo = 0, hi = hi0;
  for (;;) {
    const bool open = hi > lo;
    if (!__any(open)) break;
    const long long span = hi - lo;
    const int idx = lo + (int)((span * (hl + 1)) >> 5);
    const unsigned long long m = __ballot(rowptr[idx] <= p);
    const int cnt = __popc(half ? (unsigned)(m >> 32) : (unsigned)m);
    if (open) {
      const int nlo = cnt > 0 ? lo + (int)((span * cnt) >> 5) : lo;
      const int nhi = cnt < 32 ? lo + (int)((span * (cnt + 1)) >> 5) - 1 : hi;
      lo = nlo;
      hi = nhi;
    }
  }
  return make_int2(__shfl(lo, 0), __shfl(lo, 32));
}

// Row of tile pa for one lane group, searched inside the wave's row range [lo, hi].
__device__ __forceinline__ int group_find_row(const int *__restrict__ rowptr, int lo, int hi, long long pa) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rowptr[mid] <= pa) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Row of each lane group's first tile pa (rowptr[I] <= pa < rowptr[I+1]),
// given the rows [Ia, Ib] holding the wave's first and last tiles. When the
// wave spans fewer than 64 rows (the common case) one coalesced load brings
// the row ends rowptr[Ia+1 .. Ib] into lanes and each lane binary-searches
// that register window with shuffles; otherwise it binary-searches memory.
// Must be called by all 64 lanes.
__device__ __forceinline__ int groups_find_row(const int *__restrict__ rowptr, int Ia, int Ib, long long pa,
                                               bool active, int lane) {
  const int m = Ib - Ia;  // wave-uniform
  if (m < 64) {
    const int w = lane < m ? rowptr[Ia + 1 + lane] : 0x7fffffff;
    int lo = 0, hi = m;  // number of row ends <= pa, in [lo, hi]
#pragma unroll
    for (int step = 0; step < 6; ++step) {
      const int mid = (lo + hi + 1) >> 1;
      const int wm = __shfl(w, mid > 0 ? mid - 1 : 0);
      if (mid > lo) {
        if (wm <= pa) lo = mid;
        else hi = mid - 1;
      }
    }
    return Ia + lo;
  }
  return active ? group_find_row(rowptr, Ia, Ib, pa) : Ia;
}

// part[c] = block row holding tile c*CH, for every chunk c < nchunks (each
// chunk's first tile lies in exactly one non-empty row). One lane per block
// row writes the chunks starting in its row; rows where more than 8 chunks
// start (hubs) are written by the whole wave together.
template <int Unused = 0>  // a template: every translation unit that launches it gets the kernel
__global__ __launch_bounds__(256) void chunk_rows_kernel(const int *__restrict__ rowptr, int nbr, int CH,
                                                         int nchunks, int *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long c0 = 0, c1 = 0;  // chunks [c0, c1) start in row I
  if (I < nbr) {
    const long long r0 = rowptr[I], r1 = rowptr[I + 1];
    c0 = (r0 + CH - 1) / CH;
    c1 = (r1 + CH - 1) / CH;
    if (c1 > nchunks) c1 = nchunks;
    if (c1 < c0) c1 = c0;
  }
  const bool big = c1 - c0 > 8;
  if (!big)
    for (long long c = c0; c < c1; ++c) part[c] = (int)I;
  unsigned long long m = __ballot(big);
  while (m) {  // wave-uniform
    const int src = __ffsll((long long)m) - 1;
    m &= m - 1;
    const long long b0 = __shfl(c0, src), b1 = __shfl(c1, src);
    const int row = __shfl((int)I, src);
    for (long long c = b0 + lane; c < b1; c += 64) part[c] = row;
  }
}

// The row ends a wave needs, staged in LDS: rowptr[Ia .. Ib+1], where Ia is
// the row of the wave's first tile and Ib bounds the row of the first tile
// after its chunk. When they do not fit (a chunk spanning more than kWin-2
// block rows, e.g. long runs of empty rows) lookups go to memory.
constexpr int kWin = 256;
// LDS ints per wave: the window plus 64 for the group-row pass
// (window_group_rows).
constexpr int kWinAlloc = kWin + 64;
struct RowWindow {
  const int *__restrict__ rowptr;
  int *win;
  int Ia, Ib;
  bool staged;  // wave-uniform

  // Entries rowptr[ia .. ib+1] into registers, kWin/64 per lane (lane + 64j);
  // issued early by the pipelined kernels. `ia`, `ib` are wave-uniform.
  // The loads are unconditional (indices clamped into rowptr[ia .. ib+1],
  // entries past the window unused): a load inside a branch leaves the
  // compiler unsure how many loads are in flight, and its next wait then
  // drains all of them, prefetches included.
  __device__ __forceinline__ static void load(const int *__restrict__ rp, int ia, int ib, int lane,
                                              int (&wv)[kWin / 64]) {
    const int m = ib + 2 - ia;
    const int mm = m < kWin ? m : kWin;
#pragma unroll
    for (int j = 0; j < kWin / 64; ++j) {
      const int i = lane + 64 * j;
      wv[j] = rp[ia + (i < mm ? i : mm - 1)];
    }
  }
  // Makes the window of rows [ia, ib] current, from registers filled by load().
  // Must be called by all 64 lanes.
  __device__ __forceinline__ void put(const int *__restrict__ rp, int *w, int ia, int ib, int lane,
                                      const int (&wv)[kWin / 64]) {
    rowptr = rp;
    win = w;
    Ia = ia;
    Ib = ib;
    const int m = ib + 2 - ia;
    staged = m <= kWin;
    // Every slot is written, staged or not (entries past the window are not
    // read): a load whose value is consumed only on some paths stays pending
    // at the join, and the first reuse of its register then waits for every
    // load in flight.
#pragma unroll
    for (int j = 0; j < kWin / 64; ++j) win[lane + 64 * j] = wv[j];
    win[kWin + lane] = 0;  // group-row slots (window_group_rows), cleared under the same wait
    // the wave reads entries other lanes wrote (wave-private LDS, in order)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // Must be called by all 64 lanes.
  __device__ __forceinline__ void stage(const int *__restrict__ rp, int *w, int ia, int ib, int lane) {
    int wv[kWin / 64];
    load(rp, ia, ib, lane, wv);
    put(rp, w, ia, ib, lane, wv);
  }
  // rowptr[i] for Ia <= i <= Ib + 1
  __device__ __forceinline__ int at(int i) const { return staged ? win[i - Ia] : rowptr[i]; }
  // The same with the placement known at compile time (ST = staged): the
  // runtime select above makes the compiler use flat loads, whose waits also
  // drain every global load in flight; callers branch once, wave-uniformly,
  // and use these in the two copies of their walk.
  template <bool ST>
  __device__ __forceinline__ int at_t(int i) const {
    if constexpr (ST) return win[i - Ia];
    else return rowptr[i];
  }
  template <bool ST>
  __device__ __forceinline__ int find_t(long long p) const {
    int lo = Ia, hi = Ib;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (at_t<ST>(mid) <= p) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
  // Calls f(std::true_type) when the window is staged, f(std::false_type)
  // otherwise (a wave-uniform branch).
  template <class F>
  __device__ __forceinline__ void dispatch(F &&f) const {
    if (__builtin_amdgcn_readfirstlane((int)staged)) f(std::true_type{});
    else f(std::false_type{});
  }
  // Row of tile p (rowptr[I] <= p < rowptr[I+1]) for p inside the chunk.
  __device__ __forceinline__ int find(long long p) const {
    int lo = Ia, hi = Ib;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (at(mid) <= p) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
};

// Rows [Ia, Ib] of a chunk: Ia holds tile k0 (the chunk's first), Ib holds
// tile kend (the next chunk's first; nbr-1 for the last chunk). They come
// from a partition (chunk_rows_kernel) that may be cached from an earlier
// call on the same rowptr address, so the raw entries (ia = part[chunk],
// ib = part[chunk+1]) are clamped here and checked by put_checked.
__device__ __forceinline__ int2 clamp_rows(int ia, int ib, bool last, int nbr) {
  int Ia = ia < 0 ? 0 : (ia > nbr - 1 ? nbr - 1 : ia);
  int Ib = last ? nbr - 1 : (ib < Ia ? Ia : (ib > nbr - 1 ? nbr - 1 : ib));
  return make_int2(Ia, Ib);
}

// Puts the window of a chunk's rows from registers (RowWindow::load on the
// clamped partition entries) and checks them against rowptr: rowptr[Ia] <=
// k0 < rowptr[Ia+1] and, unless last, rowptr[Ib] <= kend < rowptr[Ib+1]. On
// a mismatch (rowptr rewritten since the partition was made) the rows are
// searched, the entry is rewritten (vector stores of the same value by every
// lane) and the window restaged from memory. Must be called by all lanes.
__device__ __forceinline__ bool put_checked(RowWindow &rw, const int *__restrict__ rowptr, int *part, int *win,
                                            int chunk, int nchunks, int nbr, long long k0, long long kend,
                                            int2 rows, int lane, const int (&wv)[kWin / 64]) {
  const bool last = chunk + 1 >= nchunks;
  rw.put(rowptr, win, rows.x, rows.y, lane, wv);
  bool ok = false;
  rw.dispatch([&](auto st) {
    constexpr bool ST = decltype(st)::value;
    if constexpr (ST) {
      // the four entries are in the lanes' registers (entry i in lane i % 64,
      // slot i / 64): each lane checks the ones it holds, one ballot joins them
      const int mlast = rows.y + 1 - rows.x;  // entry of rowptr[Ib + 1]
      bool bad = false;
#pragma unroll
      for (int j = 0; j < kWin / 64; ++j) {
        const int i = lane + 64 * j;
        const long long e = wv[j];
        bad |= (i == 0 && e > k0) || (i == 1 && e <= k0);
        if (!last) bad |= (i == mlast - 1 && e > kend) || (i == mlast && e <= kend);
      }
      ok = __ballot(bad) == 0;
    } else {
      ok = rw.at_t<ST>(rows.x) <= k0 && k0 < rw.at_t<ST>(rows.x + 1);
      if (!last) ok = ok && rw.at_t<ST>(rows.y) <= kend && kend < rw.at_t<ST>(rows.y + 1);
    }
  });
  if (!__builtin_amdgcn_readfirstlane((int)ok)) {
    const int2 ab = wave_find_rows2(rowptr, nbr - 1, k0, last ? k0 : kend, lane);
    const int Ia = ab.x, Ib = last ? nbr - 1 : ab.y;
    part[chunk] = Ia;
    rw.stage(rowptr, win, Ia, Ib, lane);
    return true;
  }
  return false;
}

// Value of lane (lane - D) inside the lane's 16-lane row (DPP row_shr:D, a
// VALU operand modifier, no LDS round trip); lanes with (lane % 16) < D get
// `fill`.
template <int D>
__device__ __forceinline__ uint32_t dpp_shr(uint32_t v, uint32_t fill) {
  static_assert(D >= 1 && D < 16, "row_shr distance");
  return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x110 + D, 0xF, 0xF, false);
}

// OR of v over the lane's 16-lane row (DPP row_ror 1, 2, 4, 8: every lane of
// the row ends with the whole row's OR).
__device__ __forceinline__ uint32_t row_or16(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
  return v;
}

// Value of lane (lane - D) across the wave (lanes < D keep their own value),
// DPP for in-row distances, ds_bpermute otherwise.
template <int D>
__device__ __forceinline__ uint32_t lane_up(uint32_t v, uint32_t fill) {
  if constexpr (D < 16) return dpp_shr<D>(v, fill);
  else return (uint32_t)__shfl_up((int)v, D);
}
template <int D>
__device__ __forceinline__ float lane_up(float v, float fill) {
  return __uint_as_float(lane_up<D>(__float_as_uint(v), __float_as_uint(fill)));
}
template <int D>
__device__ __forceinline__ uint64_t lane_up(uint64_t v, uint64_t fill) {
  const uint32_t lo = lane_up<D>((uint32_t)v, (uint32_t)fill);
  const uint32_t hi = lane_up<D>((uint32_t)(v >> 32), (uint32_t)(fill >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Row broadcasts (GFX9 DPP row_bcast): Bcast15 gives rows 1 and 3 the value of
// lanes 15 and 47, Bcast31 gives rows 2 and 3 the value of lane 31; other
// lanes get `old`. Together with in-row row_shr steps they finish a
// 64-lane scan without LDS traffic. WaveShr1: the value of lane - 1 (lane 0
// gets `old`).
enum DppWide { kBcast15 = 0x142, kBcast31 = 0x143, kWaveShr1 = 0x138 };
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_wide(uint32_t v, uint32_t old) {
  constexpr int kRowMask = CTRL == kBcast15 ? 0xA : CTRL == kBcast31 ? 0xC : 0xF;
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, kRowMask, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_wide(float v, float old) {
  return __uint_as_float(dpp_wide<CTRL>(__float_as_uint(v), __float_as_uint(old)));
}
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_wide(uint64_t v, uint64_t old) {
  const uint32_t lo = dpp_wide<CTRL>((uint32_t)v, (uint32_t)old);
  const uint32_t hi = dpp_wide<CTRL>((uint32_t)(v >> 32), (uint32_t)(old >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Sum of v over the wave, returned in every lane: DPP row_shr steps give each
// lane its 16-lane row's inclusive prefix, the row broadcasts carry the row
// totals up to lane 63, which is read out -- VALU only (a shuffle reduction
// is six LDS round trips per value).
__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
  v += dpp_shr<1>(v, 0u);
  v += dpp_shr<2>(v, 0u);
  v += dpp_shr<4>(v, 0u);
  v += dpp_shr<8>(v, 0u);
  v += dpp_wide<kBcast15>(v, 0u);
  v += dpp_wide<kBcast31>(v, 0u);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Value of lane (lane - L) for lanes >= L (others get `old`): the previous
// lane group's lane in the same position. DPP wave_shr:1 for L = 1,
// ds_bpermute otherwise.
template <int L>
__device__ __forceinline__ int group_up(int v) {
  if constexpr (L == 1) return (int)dpp_wide<kWaveShr1>((uint32_t)v, 0u);
  else return __shfl_up(v, L);
}
template <int L, class V>
__device__ __forceinline__ V group_up(V v) {
  if constexpr (L == 1) return dpp_wide<kWaveShr1>(v, (V)0);
  else return __shfl_up(v, L);
}

// Row of the first tile of every lane group of a staged window, all groups at
// once (group g of L lanes starts at tile k0 + g*S): with e_i = rowptr[Ia+i]
// and h_i = the first group whose start tile is >= e_i, the row of group g is
// Ia + max{i <= Ib - Ia : h_i <= g}. Each window entry (wv, as passed to
// put) does one LDS max into a per-group slot, then a prefix max runs over
// the lanes -- two LDS round trips instead of a binary search of ~log2(kWin)
// dependent ones per lane. Must be called by all 64 lanes.
template <int L, int S>
__device__ __forceinline__ int window_group_rows(const RowWindow &rw, long long k0, int lane,
                                                 const int (&wv)[kWin / 64]) {
  static_assert((S & (S - 1)) == 0, "run length must be a power of two");
  constexpr int G = 64 / L;
  // put() cleared the slots; h_0 = 0 (rowptr[Ia] <= k0), so every group's
  // answer is >= 0
  int *tmp = rw.win + kWin;
  const int last = rw.Ib - rw.Ia;  // entries 0 .. last start rows Ia .. Ib
#pragma unroll
  for (int j = 0; j < kWin / 64; ++j) {
    const int i = lane + 64 * j;
    const long long d = (long long)wv[j] - k0;
    const long long h = d <= 0 ? 0 : (d + S - 1) / S;
    if (i <= last && h < G) atomicMax(&tmp[(int)h], i);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int v = tmp[lane / L];
  // inclusive prefix max over the lanes (groups are in lane order)
  v = max(v, (int)dpp_shr<1>((uint32_t)v, 0u));
  v = max(v, (int)dpp_shr<2>((uint32_t)v, 0u));
  v = max(v, (int)dpp_shr<4>((uint32_t)v, 0u));
  v = max(v, (int)dpp_shr<8>((uint32_t)v, 0u));
  v = max(v, (int)dpp_wide<kBcast15>((uint32_t)v, 0u));
  v = max(v, (int)dpp_wide<kBcast31>((uint32_t)v, 0u));
  return rw.Ia + v;
}

// Inclusive segmented scan over lane groups of L lanes (L divides 16; keys
// non-decreasing over groups, so equal keys are contiguous): afterwards sc[i]
// of group g holds the combination of groups g' <= g with key(g') == key(g).
// Hillis-Steele inside each 16-lane row with DPP row_shr (lanes whose source
// lies in an earlier row read the sentinel key and keep their value), then two
// steps across rows: row r takes the last group of row r-1, then row r >= 2
// the last group of row r-2 (covering rows r-3..r-2 by then). Keys are never INT_MIN.
// Must be called by all 64 lanes.
template <int L, int R, class Acc, class Op>
__device__ __forceinline__ void group_scan(int key, Acc (&sc)[R], Op op) {
  static_assert(16 % L == 0, "groups must tile the 16-lane rows");
  constexpr uint32_t kNone = 0x80000000u;
#define BGB_ROW_STEP(DL)                                                       \
  if constexpr ((DL) < 16) {                                                   \
    const int ku = (int)dpp_shr<(DL)>((uint32_t)key, kNone);                    \
    _Pragma("unroll") for (int i = 0; i < R; ++i) {                            \
      const Acc up = lane_up<(DL)>(sc[i], (Acc)0);                             \
      if (ku == key) sc[i] = op(sc[i], up);                                    \
    }                                                                          \
  }
  BGB_ROW_STEP(L)
  BGB_ROW_STEP(2 * L)
  BGB_ROW_STEP(4 * L)
  BGB_ROW_STEP(8 * L)
#undef BGB_ROW_STEP
  if constexpr (L == 1) {
    // one lane per group: the last group of rows r-1 / r-2 is lane 15, 31 or
    // 47, which the row broadcasts deliver (rows 1, 3 then rows 2, 3)
    {
      const int ku = (int)dpp_wide<kBcast15>((uint32_t)key, kNone);
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const Acc up = dpp_wide<kBcast15>(sc[i], (Acc)0);
        if (ku == key) sc[i] = op(sc[i], up);
      }
    }
    {
      const int ku = (int)dpp_wide<kBcast31>((uint32_t)key, kNone);
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const Acc up = dpp_wide<kBcast31>(sc[i], (Acc)0);
        if (ku == key) sc[i] = op(sc[i], up);
      }
    }
    return;
  }
  const int lane = threadIdx.x & 63, row = lane >> 4;
  // last group of row r-1, same lane position inside the group
  const int src1 = row * 16 - L + (lane & (L - 1));
  {
    const int ku = __shfl(key, src1 & 63);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const Acc up = __shfl(sc[i], src1 & 63);
      if (row >= 1 && ku == key) sc[i] = op(sc[i], up);
    }
  }
  const int src2 = row * 16 - 16 - L + (lane & (L - 1));  // last group of row r-2
  {
    const int ku = __shfl(key, src2 & 63);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const Acc up = __shfl(sc[i], src2 & 63);
      if (row >= 2 && ku == key) sc[i] = op(sc[i], up);
    }
  }
}

// Workgroup index remapped so that consecutive index ranges stay on one XCD
// (workgroups are dispatched round-robin over the 8 XCDs of an MI355X).
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// One block row's results for the short-row path (T <= 8, one lane per block
// row): the row's tiles [s, e) in batches of U, every load of a batch issued
// unconditionally (indices clamped into the row, the padding slots' words
// zeroed so they add nothing). kMode: popcount(a_r & x_J) counts over packed
// x (x: W_T words), or FP sums over an FP32 / bfloat16 x (T values per block
// column) in the CSR order -- tiles in order, ascending columns, set bits only.
constexpr int kRowCount = 0, kRowSumF32 = 1, kRowSumBF16 = 2;
// The short-row path applies when no block row holds more than this many
// tiles; its batch U follows the longest row (4 up to 4 tiles, then 5, 6, 8).
constexpr int kShortRowMax = 32;
inline int short_row_batch(int max_row) { return max_row <= 4 ? 4 : max_row <= 6 ? max_row : 8; }

template <int T, int kMode, int U, class Acc>
__device__ __forceinline__ void block_row_sums(const void *__restrict__ A, const int *__restrict__ colind,
                                               const void *__restrict__ x, int s, int e,
                                               Acc (&acc)[Tile<T>::R]) {
  static_assert(Tile<T>::L == 1, "one lane per block row needs T <= 8");
  constexpr int R = Tile<T>::R, NW = Tile<T>::NW;
  using W = typename Tile<T>::W;
  // FP sums are folded in three levels -- a batch of U tiles, up to 16
  // batches, the row -- so that a long row (the path forced on hub rows) is
  // not one FP32 chain of thousands of terms; the terms keep the CSR order.
  Acc mid[R];
  int nbatch = 0;
#pragma unroll
  for (int i = 0; i < R; ++i) acc[i] = mid[i] = 0;
  for (int k = s; k < e; k += U) {
    uint32_t F[U][NW];
    int J[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long p = k + u < e ? k + u : e - 1;
      J[u] = colind[p];
      load_frag<T>(A, p, 0, F[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k + u >= e) {
#pragma unroll
        for (int w = 0; w < NW; ++w) F[u][w] = 0u;  // padding: no set bits
      }
    if constexpr (kMode != kRowCount) {
      float XS[U][T];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if constexpr (kMode == kRowSumF32) {
          const float4 *xp = reinterpret_cast<const float4 *>(static_cast<const float *>(x) + (long long)J[u] * T);
#pragma unroll
          for (int v = 0; v < T / 4; ++v) {
            const float4 f = xp[v];
            XS[u][4 * v] = f.x;
            XS[u][4 * v + 1] = f.y;
            XS[u][4 * v + 2] = f.z;
            XS[u][4 * v + 3] = f.w;
          }
        } else {
          const uint2 *xp = reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(x) + (long long)J[u] * T);
#pragma unroll
          for (int v = 0; v < T / 4; ++v) {
            const uint2 h = xp[v];
            XS[u][4 * v] = bf16_to_f32((uint16_t)(h.x & 0xFFFFu));
            XS[u][4 * v + 1] = bf16_to_f32((uint16_t)(h.x >> 16));
            XS[u][4 * v + 2] = bf16_to_f32((uint16_t)(h.y & 0xFFFFu));
            XS[u][4 * v + 3] = bf16_to_f32((uint16_t)(h.y >> 16));
          }
        }
      }
      // (Skipping slots that are padding in every lane behind a wave-uniform
      // branch let the compiler sink each slot's loads into its branch, one
      // round trip per tile: slower.)
      Acc bs[R];
#pragma unroll
      for (int i = 0; i < R; ++i) bs[i] = 0;
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const Word<T> w = row_word<T>(F[u], i);
#pragma unroll
          for (int c = 0; c < T; ++c) bs[i] += bit_select(w, T - 1 - c, XS[u][c]);
        }
#pragma unroll
      for (int i = 0; i < R; ++i) mid[i] += bs[i];
      if (++nbatch == 16) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          acc[i] += mid[i];
          mid[i] = 0;
        }
        nbatch = 0;
      }
    } else {
      Word<T> X[U];
#pragma unroll
      for (int u = 0; u < U; ++u) X[u] = (Word<T>) static_cast<const W *>(x)[J[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i] += bit_count(row_word<T>(F[u], i) & X[u]);
    }
  }
  if constexpr (kMode != kRowCount) {
#pragma unroll
    for (int i = 0; i < R; ++i) acc[i] += mid[i];
  }
}

}  // namespace bgb
