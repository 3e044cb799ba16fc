// B2SR binary SpMV (BMV) kernels for gfx950: bin_bin_bin, bin_bin_full (int32
// and float counts), bin_full_full (FP32) and masked bin_bin_bin.
//
// Semantics follow bsrbmv.cu:274-1229: for block row I, row r of the result
// combines row r of every tile (I, J) with the x segment of block column J
// (a packed word for the bin_bin ops, T floats for bin_full_full).
//
// Work decomposition (not the reference's one-warp-per-block-row loop,
// bsrbmv.cu:1046, which serialises R-MAT hub rows):
//   * the tile array is cut into fixed chunks of CH = (64/L)*S tiles, one
//     chunk per wave, so every wave streams the same number of bytes;
//   * inside a chunk each lane group walks S consecutive tiles, accumulating
//     per-row partials and flushing a row to C whenever the row ends (rows
//     that end at the same position -- empty rows -- are flushed as zeros);
//   * the first row a group flushes may have started in an earlier group: a
//     wave-level segmented scan over the groups' open rows supplies that
//     carry, so rows that start and end inside one wave are written once;
//   * a row still open at the end of the chunk leaves a carry record; a
//     fix-up pass adds those records to C (integer atomics are exact, bit
//     results use atomic OR, FP32 uses float atomics as the reference does);
//     an all-zero record is not published. (Merging those rows with atomics
//     inside the main pass instead was measured slower on MI355X.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>
#include <climits>
#include <cmath>
#include <type_traits>
#include <map>
#include <mutex>

#include "bgb_device.h"
#include "bgb_tile.h"

namespace bgb {

__device__ __forceinline__ long long lmin(long long a, long long b) { return a < b ? a : b; }

enum BmvOp : int {
  OP_BITS = 0,
  OP_MASKED = 1,
  OP_COUNT_I32 = 2,
  OP_COUNT_F32 = 3,
  OP_FULL_F32 = 4,
  OP_FULL_BF16 = 5  // x in bfloat16, FP32 accumulation
};

struct BmvArgs {
  const void *A;
  const void *B;
  void *C;
  const int *rowptr;
  const int *colind;
  const void *mask;
  int nbr;
  int nblocks;
  int accumulate;
  int nchunks;
  int *carry_key;
  void *carry_val;
  int *part;  // part[c]: block row of chunk c's first tile (cached per stream and rowptr)
  int fixup_blocks;  // workgroups of the fix-up pass proper; any after them zero the empty block rows
  int xl_words;      // LDS kernels: packed-x words [0, xl_words) staged, the rest read from B
  int x_words;       // packed-x words in B (the range of its buffer resource)
  // Piece kernel with a non-zero bitmap (XL 3): bit j of xbm = word xl_words + j
  // of x is non-zero; bm_dwords dwords (a multiple of 4), staged at LDS byte bm_off.
  const uint32_t *xbm;
  int bm_dwords;
  int bm_off;
};

constexpr int kWavesPerBlock = 4;

// Tiles per lane-group run (S), per tile dim (index log2(T)-2): 8 or 16.
// Longer runs put more bytes in flight per wave, which matters for the small
// tiles; set with bgb_set_run_length.
int g_run_length[5] = {8, 8, 8, 8, 8};
// Packed-x staging per tile dim: 0 automatic, 1 never, 2 always (all of x
// when it fits in LDS, else its first words, the rest read from memory).
int g_x_staging[5] = {0, 0, 0, 0, 0};
// Short-row path (T <= 8): 0 automatic (every block row at most kShortRowMax
// tiles), 1 never, 2 always.
int g_short_rows = 0;
// Row pieces for the count ops at T <= 8 (bgb_set_row_pieces): 0 automatic
// (when the packed x fits in LDS and the pieces are mostly full), 1 never
// (the chunk walk), 2 always.
int g_row_pieces = 0;

template <int OP>
struct AccOf {
  using type = uint32_t;
};
template <>
struct AccOf<OP_FULL_F32> {
  using type = float;
};
template <>
struct AccOf<OP_FULL_BF16> {
  using type = float;
};
constexpr bool is_full(int op) { return op == OP_FULL_F32 || op == OP_FULL_BF16; }

template <int T, int OP>
__device__ __forceinline__ void flush_row(const BmvArgs &a, int I, int q,
                                          const typename AccOf<OP>::type (&v)[Tile<T>::R]) {
  constexpr int R = Tile<T>::R, L = Tile<T>::L;
  using W = typename Tile<T>::W;
  if constexpr (OP == OP_BITS || OP == OP_MASKED) {
    using Wd = Word<T>;
    Wd bits = 0;
#pragma unroll
    for (int i = 0; i < R; ++i) bits |= (Wd)(v[i] != 0u ? 1u : 0u) << (T - 1 - (q * R + i));
#pragma unroll
    for (int m = 1; m < L; m <<= 1) bits |= __shfl_xor(bits, m);
    if (q == 0) {
      W *C = static_cast<W *>(a.C);
      if constexpr (OP == OP_MASKED) bits &= ~(Wd) static_cast<const W *>(a.mask)[I];
      else if (a.accumulate) bits |= (Wd)C[I];
      C[I] = (W)bits;
    }
  } else {
    using V = typename std::conditional<OP == OP_COUNT_I32, int, float>::type;
    V out[R];
#pragma unroll
    for (int i = 0; i < R; ++i) out[i] = (V)v[i];
    V *C = static_cast<V *>(a.C) + (long long)I * T + q * R;
    static_assert(R % 4 == 0, "rows per lane must fill 16-byte stores");
#pragma unroll
    for (int j = 0; j < R / 4; ++j) {
      typedef V V4 __attribute__((ext_vector_type(4)));
      V4 *p = reinterpret_cast<V4 *>(C) + j;
      V4 w = {out[4 * j], out[4 * j + 1], out[4 * j + 2], out[4 * j + 3]};
      if (a.accumulate) w += *p;
      *p = w;
    }
  }
}

// Packed-x word J of a copy staged in LDS: T = 4 keeps two nibbles per byte
// there (word J in byte J/2, high nibble for odd J), T >= 8 the words as
// they are.
template <int T>
__device__ __forceinline__ Word<T> lds_xword(const uint8_t *xl, int J) {
  if constexpr (T == 4) return (uint32_t)(xl[J >> 1] >> ((J & 1) * 4)) & 0xFu;
  else return (Word<T>) reinterpret_cast<const typename Tile<T>::W *>(xl)[J];
}

// What one chunk needs before its walk: each lane group's tile run, the
// raw-clamped row range and the row-window entries in registers. Issued ahead
// of the walk (a chunk early in the persistent kernel), so that its latency
// overlaps other work.
template <int T, int S>
struct ChunkIn {
  uint32_t F[S][Tile<T>::NW];
  int J[S];
  int wv[kWin / 64];
  int2 rows;
};

// Issues the loads of chunk `chunk` with row range `rows` (wave-uniform).
// kFull: the chunk's CH tiles all exist (every lane group has a whole run).
template <int T, int S, bool kFull = false>
__device__ __forceinline__ void load_chunk(const BmvArgs &a, int chunk, int lane, int2 rows, ChunkIn<T, S> &in) {
  constexpr int L = Tile<T>::L, G = 64 / L, CH = G * S;
  const int g = lane / L, q = lane % L;
  const long long k0 = (long long)chunk * CH;
  const long long pa = lmin(k0 + (long long)g * S, a.nblocks);
  const long long pb = lmin(pa + S, a.nblocks);
  if (kFull) load_run<T, S, true>(a.A, a.colind, pa, S, q, in.F, in.J);
  else if (pa < pb) load_run<T, S>(a.A, a.colind, pa, (int)(pb - pa), q, in.F, in.J);
  RowWindow::load(a.rowptr, rows.x, rows.y, lane, in.wv);
  in.rows = rows;
}

// One chunk of CH tiles for one wave (all 64 lanes) from its loaded inputs:
// `win` is the wave's LDS row window, `xl` the packed x staged in LDS
// (XL = 1: all of it; XL = 2: words [0, a.xl_words)) or unused (XL = 0: x
// gathered from memory).
template <int T, int OP, int S, int XL>
__device__ __forceinline__ void bmv_chunk(const BmvArgs &a, int chunk, int lane, int *win, const uint8_t *xl,
                                          const ChunkIn<T, S> &in) {
  constexpr int L = Tile<T>::L, R = Tile<T>::R, NW = Tile<T>::NW, G = 64 / L, CH = G * S;
  using Acc = typename AccOf<OP>::type;
  using W = typename Tile<T>::W;
  const int g = lane / L, q = lane % L;
  const long long k0 = (long long)chunk * CH;
  const long long pa = lmin(k0 + (long long)g * S, a.nblocks);
  const long long pb = lmin(pa + S, a.nblocks);
  const bool active = pa < pb;
  const int n = (int)(pb - pa);
  const uint32_t(&F)[S][NW] = in.F;
  const int(&J)[S] = in.J;

  // x words of the run's tiles. The gathers are unconditional (block column
  // 0 stands in for missing tiles, whose word is then zeroed) so that they
  // issue back to back instead of one predicated region and wait per tile.
  Word<T> X[S];
  if constexpr (!is_full(OP)) {
    // XL = 2: every lane reads LDS (its index clamped into the staged words)
    // and memory through a buffer resource over x, with an out-of-range
    // offset where the word is staged -- no memory access for those lanes,
    // and both loads unconditional
    __amdgpu_buffer_rsrc_t xr;
    if constexpr (XL == 2)
      xr = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.B), (short)0, a.x_words * (int)sizeof(W),
                                             0x00020000);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const bool v = active && s < n;
      const int j = v ? J[s] : 0;
      if constexpr (XL == 1) {
        X[s] = lds_xword<T>(xl, j);
      } else if constexpr (XL == 2) {
        const bool staged = j < a.xl_words;
        const Word<T> xs = lds_xword<T>(xl, staged ? j : 0);
        const int off = staged ? a.x_words * (int)sizeof(W) : j * (int)sizeof(W);
        Word<T> xm;
        if constexpr (sizeof(W) == 1) xm = __builtin_amdgcn_raw_buffer_load_b8(xr, off, 0, 0);
        else if constexpr (sizeof(W) == 2) xm = __builtin_amdgcn_raw_buffer_load_b16(xr, off, 0, 0);
        else xm = __builtin_amdgcn_raw_buffer_load_b32(xr, off, 0, 0);
        X[s] = staged ? xs : xm;
      } else {
        X[s] = (Word<T>) static_cast<const W *>(a.B)[j];
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) X[s] = (active && s < n) ? X[s] : (Word<T>)0;
  }

  // FP ops with T <= 8: the whole x segment of each tile (T values) is loaded
  // up front with the tile, so the sums need no dependent gathers.
  constexpr bool kSeg = is_full(OP) && T <= 8;
  float XS[kSeg ? S : 1][kSeg ? T : 1];
  if constexpr (kSeg) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (active && s < n) {
        if constexpr (OP == OP_FULL_F32) {
          const float4 *xp = reinterpret_cast<const float4 *>(static_cast<const float *>(a.B) + (long long)J[s] * T);
#pragma unroll
          for (int v = 0; v < T / 4; ++v) {
            const float4 f = xp[v];
            XS[s][4 * v] = f.x;
            XS[s][4 * v + 1] = f.y;
            XS[s][4 * v + 2] = f.z;
            XS[s][4 * v + 3] = f.w;
          }
        } else {
          const uint2 *xp = reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(a.B) + (long long)J[s] * T);
#pragma unroll
          for (int v = 0; v < T / 4; ++v) {
            const uint2 h = xp[v];
            XS[s][4 * v] = bf16_to_f32((uint16_t)(h.x & 0xFFFFu));
            XS[s][4 * v + 1] = bf16_to_f32((uint16_t)(h.x >> 16));
            XS[s][4 * v + 2] = bf16_to_f32((uint16_t)(h.y & 0xFFFFu));
            XS[s][4 * v + 3] = bf16_to_f32((uint16_t)(h.y >> 16));
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < T; ++c) XS[s][c] = 0.0f;
      }
    }
  }

  // Block row holding tile pa: rowptr[I] <= pa < rowptr[I+1], searched in
  // the wave's staged row ends (checked against rowptr first).
  RowWindow rw;
  const bool restaged = __builtin_amdgcn_readfirstlane(
      (int)put_checked(rw, a.rowptr, a.part, win, chunk, a.nchunks, a.nbr, k0, k0 + CH, in.rows, lane, in.wv));
  // The walk, compiled twice: row ends read from the LDS window (ST, the
  // common case) or from rowptr in memory (chunks spanning > kWin rows).
  rw.dispatch([&](auto st) {
  constexpr bool ST = decltype(st)::value;
  // A chunk over one or two block rows (not the first or last chunk; 46 % and
  // 19 % of the chunks at T = 4 on R-MAT 20): rows Ia = Ib (the row neither
  // starts nor ends here, the next chunk starts in it) or Ib = Ia + 1 (row Ia
  // ends at E inside the chunk, row Ib runs on into the next one). Each lane
  // splits its run's counts at E, two wave sums give the rows' totals, row Ia
  // is flushed and row Ib (Ia for one row) becomes the chunk's carry record
  // -- no group-row pass, row-end walk or segmented scan. Count and bit ops
  // (integer sums, exact in any order).
  // Two rows only where x is gathered from memory: the LDS-staged kernels run
  // 16 waves of at most 128 VGPRs, and the second row's sums made T = 8 spill
  // there (R-MAT 20 T = 8: 99 -> 113 us; from memory, R-MAT 22 T = 8 510 -> 482).
  constexpr int kMaxRows = XL ? 1 : 2;
  if constexpr (ST && !is_full(OP)) {
    if (k0 > 0 && chunk + 1 < a.nchunks && rw.Ib - rw.Ia < kMaxRows) {  // wave-uniform
      const bool two = kMaxRows == 2 && rw.Ib != rw.Ia;
      const long long E = two ? (long long)rw.at_t<true>(rw.Ia + 1) : LLONG_MAX;
      Acc ta[R], tb[R];  // row Ia (tiles before E), row Ib (tiles from E)
      if constexpr (T == 4) {
        uint32_t pka = 0, pkb = 0;  // four byte-wise row counts, at most 8 x 4 each
#pragma unroll
        for (int s = 0; s < S; ++s) {
          uint32_t m = F[s][0] & ((uint32_t)X[s] * 0x01010101u);
          m = m - ((m >> 1) & 0x55555555u);
          m = (m & 0x33333333u) + ((m >> 2) & 0x33333333u);
          const bool before = pa + s < E;
          pka += before ? m : 0u;
          pkb += before ? 0u : m;
        }
        // rows 0, 2 and 1, 3 in 16-bit fields (a wave sum is at most 2048)
        auto wave_rows = [&](uint32_t pk, Acc (&t)[R]) {
          const uint32_t lo = wave_total(pk & 0x00FF00FFu), hi = wave_total((pk >> 8) & 0x00FF00FFu);
          t[0] = lo & 0xFFFFu, t[1] = hi & 0xFFFFu, t[2] = lo >> 16, t[3] = hi >> 16;
        };
        wave_rows(pka, ta);
        if (two) wave_rows(pkb, tb);
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) ta[i] = tb[i] = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const bool before = pa + s < E;
#pragma unroll
          for (int i = 0; i < R; ++i) {
            const Acc c = (Acc)bit_count(row_word<T>(F[s], i) & X[s]);
            ta[i] += before ? c : (Acc)0;
            tb[i] += before ? (Acc)0 : c;
          }
        }
        auto wave_rows = [&](Acc (&t)[R]) {
          if constexpr (L == 1 && !is_full(OP)) {
            // two rows per register in 16-bit fields (a row's chunk total is
            // at most CH x T <= 8192), summed over the wave with DPP
#pragma unroll
            for (int j = 0; j < R / 2; ++j) {
              const uint32_t w = wave_total(t[2 * j] | (t[2 * j + 1] << 16));
              t[2 * j] = w & 0xFFFFu;
              t[2 * j + 1] = w >> 16;
            }
          } else {
#pragma unroll
            for (int m = L; m < 64; m <<= 1)
#pragma unroll
              for (int i = 0; i < R; ++i) t[i] += (Acc)__shfl_xor((int)t[i], m);
          }
        };
        wave_rows(ta);
        if (two) wave_rows(tb);
      }
      if (two && g == 0) flush_row<T, OP>(a, rw.Ia, q, ta);
      Acc *open_t = two ? tb : ta;
      if (g == G - 1) {
        bool open = true;
        if constexpr (OP == OP_BITS || OP == OP_MASKED) {
          bool nz = false;
#pragma unroll
          for (int i = 0; i < R; ++i) nz |= open_t[i] != (Acc)0;
#pragma unroll
          for (int m = 1; m < L; m <<= 1) nz |= __shfl_xor((int)nz, m) != 0;
          open = nz;
        }
        if (q == 0) a.carry_key[chunk] = open ? rw.Ib : -2 - rw.Ib;
        if (open) {
          Acc *cv = static_cast<Acc *>(a.carry_val) + (long long)chunk * T + q * R;
#pragma unroll
          for (int i = 0; i < R; ++i) cv[i] = open_t[i];
        }
      }
      return;
    }
  }
  int I;
  if constexpr (ST) {
    // the registers hold the window unless put_checked restaged it
    if (!restaged) I = window_group_rows<L, S>(rw, k0, lane, in.wv);
    else I = rw.find_t<true>(pa);
  } else {
    I = rw.find_t<false>(pa);
  }
  I = active ? I : rw.Ia;
  const int I0 = I;

  Acc acc[R];
#pragma unroll
  for (int i = 0; i < R; ++i) acc[i] = (Acc)0;

  int held = -1;
  Acc hv[R];
#pragma unroll
  for (int i = 0; i < R; ++i) hv[i] = (Acc)0;
  // e = end of row I, e2 = end of row I+1, read one row ahead so that a row
  // end does not wait on the LDS (or memory) read of the next one. Indices
  // past the window / rowptr are clamped: such rows lie beyond the chunk and
  // their ends are never compared.
  const int hi = ST ? rw.Ib + 1 : a.nbr;
  auto row_end = [&](int r) { return r < a.nbr ? rw.at_t<ST>(r + 1 < hi ? r + 1 : hi) : INT_MAX; };
  int e = active ? row_end(I) : INT_MAX;
  int e2 = active ? row_end(I + 1) : INT_MAX;

  // T = 4 bit ops, branch-free: a tile's four row counts are byte-wise
  // popcounts of its word (row r in byte r), accumulated packed (a run adds
  // at most 8 x 4 per byte); a run with at most one row end keeps a snapshot
  // of the packed counts at that end instead of branching per tile. Waves in
  // which some lane's run holds a second row end (or an empty row) take the
  // general walk below.
  bool walked = false;
  if constexpr (T == 4 && !is_full(OP) && ST) {
    const long long lastpos = pa + n;
    const bool multi = active && (long long)e2 <= lastpos;
    if (__ballot(multi) == 0) {
      uint32_t pk = 0, snap = 0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        // X[s] is 0 past the run, so those tiles add nothing
        uint32_t m = F[s][0] & ((uint32_t)X[s] * 0x01010101u);
        m = m - ((m >> 1) & 0x55555555u);
        m = (m & 0x33333333u) + ((m >> 2) & 0x33333333u);
        pk += m;
        snap = (int)(pa + s) + 1 == e ? pk : snap;
      }
      const bool ends = active && (long long)e <= lastpos;
      const uint32_t open = ends ? pk - snap : pk;  // bytes never borrow: snap <= pk per byte
#pragma unroll
      for (int i = 0; i < R; ++i) acc[i] = (Acc)((open >> (8 * i)) & 0xFFu);
      if (ends) {
        held = I;
#pragma unroll
        for (int i = 0; i < R; ++i) hv[i] = (Acc)((snap >> (8 * i)) & 0xFFu);
        ++I;
      }
      walked = true;
    }
  }

#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (!walked && s < n) {
      if constexpr (kSeg) {
        // ascending columns, adding 0 for clear bits: the CSR summation order
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const Word<T> w = row_word<T>(F[s], i);
#pragma unroll
          for (int c = 0; c < T; ++c) acc[i] += bit_select(w, T - 1 - c, XS[s][c]);
        }
      } else if constexpr (OP == OP_FULL_F32) {
        const float *__restrict__ x = static_cast<const float *>(a.B) + (long long)J[s] * T;
#pragma unroll
        for (int i = 0; i < R; ++i) {
          Word<T> w = row_word<T>(F[s], i);
          while (w) {
            const int h = top_bit(w);
            acc[i] += x[T - 1 - h];
            w &= ~((Word<T>)1 << h);
          }
        }
      } else if constexpr (OP == OP_FULL_BF16) {
        const uint16_t *__restrict__ x = static_cast<const uint16_t *>(a.B) + (long long)J[s] * T;
#pragma unroll
        for (int i = 0; i < R; ++i) {
          Word<T> w = row_word<T>(F[s], i);
          while (w) {
            const int h = top_bit(w);
            acc[i] += bf16_to_f32(x[T - 1 - h]);
            w &= ~((Word<T>)1 << h);
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i] += bit_count(row_word<T>(F[s], i) & X[s]);
      }
      const int pos = (int)(pa + s) + 1;
      if (pos == e) {
        if (held < 0) {
          held = I;
#pragma unroll
          for (int i = 0; i < R; ++i) hv[i] = acc[i];
        } else {
          flush_row<T, OP>(a, I, q, acc);
        }
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i] = (Acc)0;
        ++I;
        e = e2;
        e2 = row_end(I + 1);
        if (e == pos) {
          // Empty block rows (no tiles) follow: the walk steps over them (the
          // fix-up pass zeroes them when C is overwritten) to the first row
          // ending after pos. In the LDS window (at most kWin rows) one row
          // end per step; in memory, where a chunk can span any number of
          // empty rows, a binary search over the row ends.
          if constexpr (ST) {
            do {
              ++I;
              e = e2;
              e2 = row_end(I + 1);
            } while (e == pos);
          } else {
            int lo = I + 1, up = a.nbr;
            while (lo < up) {
              const int mid = (lo + up) >> 1;
              if (row_end(mid) > pos) up = mid;
              else lo = mid + 1;
            }
            I = lo;
            e = row_end(I);
            e2 = row_end(I + 1);
          }
        }
      }
    }
  }

  // Segmented inclusive scan of the open-row partials across groups (keys
  // are non-decreasing over groups, so equal keys are contiguous).
  // Count and bit ops scan two rows per register, in 16-bit fields: a row's
  // count within one chunk is at most (tiles of the chunk) x T <= 8192, so
  // the fields never carry into each other -- half the scan's DPP steps.
  const int key = active ? I : INT_MAX;
  Acc sc[R];
  Acc cin[R];
  const int kin = group_up<L>(key);
  const bool cin_on = g > 0 && kin == I0;
  if constexpr (!is_full(OP)) {
    static_assert(R % 2 == 0 && CH * T < 65536, "two 16-bit row counts per register");
    uint32_t pk[R / 2];
#pragma unroll
    for (int j = 0; j < R / 2; ++j) pk[j] = acc[2 * j] | (acc[2 * j + 1] << 16);
    group_scan<L, R / 2, uint32_t>(key, pk, [](uint32_t x, uint32_t y) { return x + y; });
    // Carry into this group's first flushed row from the groups before it.
#pragma unroll
    for (int j = 0; j < R / 2; ++j) {
      const uint32_t upv = group_up<L, uint32_t>(pk[j]);  // every lane takes part
      const uint32_t up = cin_on ? upv : 0u;
      sc[2 * j] = pk[j] & 0xFFFFu;
      sc[2 * j + 1] = pk[j] >> 16;
      cin[2 * j] = up & 0xFFFFu;
      cin[2 * j + 1] = up >> 16;
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; ++i) sc[i] = acc[i];
    group_scan<L, R, Acc>(key, sc, [](Acc x, Acc y) { return x + y; });
    // Carry into this group's first flushed row from the groups before it.
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const Acc up = group_up<L, Acc>(sc[i]);
      cin[i] = cin_on ? up : (Acc)0;
    }
  }
  if (held >= 0) {
#pragma unroll
    for (int i = 0; i < R; ++i) hv[i] += cin[i];
    flush_row<T, OP>(a, held, q, hv);
  }
  // The last active group publishes the chunk's open row I (nbr: none). Bit
  // carries that are all zero publish no record (key -2 - I, so the row stays
  // known to the fix-up's tail wave); count / FP carries are always published
  // so that the chunks carrying into one row stay contiguous for the fix-up.
  const int last = (int)lmin(G - 1, (a.nblocks - 1 - k0) / S);
  if (g == last) {
    bool open = key < a.nbr;
    if constexpr (OP == OP_BITS || OP == OP_MASKED) {
      bool nz = false;
#pragma unroll
      for (int i = 0; i < R; ++i) nz |= sc[i] != (Acc)0;
#pragma unroll
      for (int m = 1; m < L; m <<= 1) nz |= __shfl_xor((int)nz, m) != 0;
      open = open && nz;
    }
    if (q == 0) a.carry_key[chunk] = open ? key : -2 - key;
    if (open) {
      Acc *cv = static_cast<Acc *>(a.carry_val) + (long long)chunk * T + q * R;
#pragma unroll
      for (int i = 0; i < R; ++i) cv[i] = sc[i];
    }
  }
  });
}

template <int T, int OP, int S>
__global__ __launch_bounds__(64 * kWavesPerBlock) void bmv_kernel(BmvArgs a) {
  constexpr int CH = (64 / Tile<T>::L) * S;
  __shared__ int s_win[kWavesPerBlock][kWinAlloc];
  const int lane = threadIdx.x & 63;
  const int chunk = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock +
                                                   (threadIdx.x >> 6));
  if (chunk >= a.nchunks) return;  // uniform per wave
  const bool last = chunk + 1 >= a.nchunks;
  const int2 rows = clamp_rows(a.part[chunk], last ? 0 : a.part[chunk + 1], last, a.nbr);
  ChunkIn<T, S> in;
  load_chunk<T, S>(a, chunk, lane, rows, in);
  (void)CH;
  bmv_chunk<T, OP, S, 0>(a, chunk, lane, s_win[threadIdx.x >> 6], nullptr, in);
}

// Raw partition entries of chunk c (part[c], part[c+1]) as vector loads, so
// that they can be in flight across other work (vector loads complete in
// order, scalar ones would hold up every LDS wait).
__device__ __forceinline__ int2 load_part_vec(const BmvArgs &a, int c) {
  int cv = c;
  asm volatile("" : "+v"(cv));  // a VGPR address: vector loads
  const int ia = a.part[cv];
  const int ib = a.part[cv + 1 < a.nchunks ? cv + 1 : cv];  // unconditional (see RowWindow::load)
  return make_int2(ia, cv + 1 < a.nchunks ? ib : 0);
}

// Copies the packed x (x_bytes bytes of a.B, 16-byte aligned) into LDS at xl,
// all threads of the workgroup; the caller synchronises. T = 4 keeps two
// words (low nibbles) per LDS byte.
template <int T>
__device__ __forceinline__ void stage_x(const BmvArgs &a, int x_bytes, uint8_t *xl) {
  // stage the packed x (x_bytes bytes, 16-byte aligned): whole 16-byte
  // units, then the tail byte by byte
  const uint4 *src = static_cast<const uint4 *>(a.B);
  const uint8_t *srcb = static_cast<const uint8_t *>(a.B);
  const int units = x_bytes / 16;
  if constexpr (T == 4) {
    // two words (low nibbles) per LDS byte: 16 source bytes -> 8 LDS bytes
    auto pack = [](uint32_t w) {  // bytes b0..b3 -> (b0 | b1 << 4) | (b2 | b3 << 4) << 8
      return (w & 0xFu) | ((w >> 4) & 0xF0u) | ((w >> 8) & 0xF00u) | ((w >> 12) & 0xF000u);
    };
    uint2 *dst = reinterpret_cast<uint2 *>(xl);
    for (int i = threadIdx.x; i < units; i += blockDim.x) {
      const uint4 v = src[i];
      dst[i] = make_uint2(pack(v.x) | (pack(v.y) << 16), pack(v.z) | (pack(v.w) << 16));
    }
    for (int k = units * 8 + threadIdx.x; 2 * k < x_bytes; k += blockDim.x) {
      const uint32_t lo = srcb[2 * k] & 0xFu;
      const uint32_t hi = 2 * k + 1 < x_bytes ? (srcb[2 * k + 1] & 0xFu) : 0u;
      xl[k] = (uint8_t)(lo | (hi << 4));
    }
  } else {
    uint4 *dst = reinterpret_cast<uint4 *>(xl);
    for (int i = threadIdx.x; i < units; i += blockDim.x) dst[i] = src[i];
    for (int k = units * 16 + threadIdx.x; k < x_bytes; k += blockDim.x) xl[k] = srcb[k];
  }
}

// Bit-vector ops with the whole packed x in LDS (it fits for n up to ~1.1M):
// one workgroup of lds_waves<T, S>() waves per CU (or two, if they fit) stages x
// once, then each wave walks chunks gw, gw + W, ... (W waves in the grid).
// The per-tile x gathers become LDS reads instead of one L2 request each.
// 16 waves (128 VGPRs each) hold two chunks' inputs for T <= 8 and run
// length 8; the larger fragments need the 256 VGPRs of 8-wave workgroups.
template <int T, int S>
constexpr int lds_waves() { return (T >= 16 || S > 8) ? 8 : 16; }
// XL = 1: all of x staged (x_bytes of it); XL = 2: its first a.xl_words words
// (x_bytes bytes), the other words read from memory.
template <int T, int OP, int S, int XL>
__global__ __launch_bounds__((64 * lds_waves<T, S>())) void bmv_lds_kernel(BmvArgs a, int x_bytes) {
  constexpr int kLdsWaves = lds_waves<T, S>();
  extern __shared__ __align__(16) uint8_t s_dyn[];
  int *win_all = reinterpret_cast<int *>(s_dyn);
  uint8_t *xl = s_dyn + sizeof(int) * kLdsWaves * kWinAlloc;
  stage_x<T>(a, x_bytes, xl);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = gridDim.x * kLdsWaves;
  int c = (int)blockIdx.x * kLdsWaves + wave;
  if (c >= a.nchunks) return;  // uniform per wave; no barrier follows
  int *win = win_all + wave * kWinAlloc;
  // Software pipeline over this wave's chunks c, c + nw, ...: while chunk c is
  // walked, the tiles and row window of c + nw and the partition entries of
  // c + 2nw are in flight.
  auto rows_of = [&](int cc, int2 raw) {
    const bool last = cc + 1 >= a.nchunks;
    return clamp_rows(__builtin_amdgcn_readfirstlane(raw.x), __builtin_amdgcn_readfirstlane(raw.y), last, a.nbr);
  };
  // One pipeline step over whole chunks (c < nfull, every run has S tiles):
  // walk chunk c from `cur` after issuing the loads of c + nw into `nxt`;
  // returns false after the wave's last whole chunk. Two buffers alternate
  // roles, so no chunk inputs are copied.
  constexpr int CH = (64 / Tile<T>::L) * S;
  const int nfull = a.nblocks / CH;  // the last chunk may be short
  // Every prefetch is issued unconditionally -- past the wave's last whole
  // chunk it re-reads an existing one (clamped index, result unused) -- so
  // that the walk's waits count exactly the loads it needs.
  int2 pnext = load_part_vec(a, c + nw < nfull ? c + nw : c);
  auto step = [&](ChunkIn<T, S> &cur, ChunkIn<T, S> &nxt) {
    const int chunk = __builtin_amdgcn_readfirstlane(c);
    const int cn = c + nw;
    const bool more = cn < nfull;  // wave-uniform
    const int cl = more ? cn : chunk;
    load_chunk<T, S, true>(a, cl, lane, rows_of(cl, pnext), nxt);
    pnext = load_part_vec(a, cn + nw < nfull ? cn + nw : chunk);
    bmv_chunk<T, OP, S, XL>(a, chunk, lane, win, xl, cur);
    c = cn;
    return more;
  };
  if (c < nfull) {
    ChunkIn<T, S> b0, b1;
    load_chunk<T, S, true>(a, c, lane, rows_of(c, load_part_vec(a, c)), b0);
    while (step(b0, b1) && step(b1, b0)) {
    }
  }
  // the short last chunk, if any, by the wave whose sequence reaches it
  if (c == nfull && nfull < a.nchunks) {
    const int chunk = __builtin_amdgcn_readfirstlane(c);
    ChunkIn<T, S> in;
    load_chunk<T, S>(a, chunk, lane, rows_of(chunk, load_part_vec(a, chunk)), in);
    bmv_chunk<T, OP, S, XL>(a, chunk, lane, win, xl, in);
  }
}

// Bit carries of the output words in the last, partial dword of C (words
// tail..nbr-1, at most three): the chunks carrying into them are the last
// chunks (open rows never decrease over chunks), so one wave walks the carry
// keys back from the last chunk, 64 at a time, until a chunk's open row lies
// before `tail`, ORs each word's carries over the wave and merges them into C
// with plain stores of those words alone -- no byte past C[nbr-1] is read or
// written, so C may be exactly nbr words (the reference caller's allocation,
// test_bin_bin_bin.cu:329). The other fix-up threads skip these rows.
template <int T, int OP>
__device__ __forceinline__ void bits_tail(const BmvArgs &a, int tail, int lane) {
  using W = typename Tile<T>::W;
  static_assert(sizeof(W) < 4, "only sub-dword words have a partial last dword");
  uint32_t w[3] = {0u, 0u, 0u};
  for (long long base = (long long)a.nchunks - 1; base >= 0; base -= 64) {
    const long long c = base - lane;
    const int k = c >= 0 ? a.carry_key[c] : -2;
    const int row = k >= 0 ? k : -2 - k;
    uint32_t bits = 0;
    if (k >= tail) {  // a non-zero record of a tail row (k < nbr)
      const uint32_t *cv = static_cast<const uint32_t *>(a.carry_val) + c * T;
#pragma unroll
      for (int r = 0; r < T; ++r) bits |= (cv[r] != 0u ? 1u : 0u) << (T - 1 - r);
      if constexpr (OP == OP_MASKED) bits &= ~(uint32_t) static_cast<const W *>(a.mask)[k];
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) w[t] |= k - tail == t ? bits : 0u;
    if (__ballot(c < 0 || row < tail)) break;  // earlier chunks end before the tail rows
  }
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) w[t] |= (uint32_t)__shfl_xor((int)w[t], m);
  if (lane == 0) {
    W *C = static_cast<W *>(a.C);
#pragma unroll
    for (int t = 0; t < 3; ++t)
      if (tail + t < a.nbr && w[t]) C[tail + t] = (W)((uint32_t)C[tail + t] | w[t]);
  }
}

// Adds the chunk carries of rows that cross chunk boundaries after the main
// pass.
//
// Bit carries (no record when zero, key < 0): one thread per chunk, one
// 32-bit OR into the aligned dword holding the row's word (a 64-bit OR at
// T = 64); the words of a partial last dword are merged by bits_tail in the
// grid's last workgroup.
//
// Count / FP carries: one wave per 64 consecutive chunks, lane j owning chunk
// c0 + j. Carry keys are non-decreasing over chunks (negative only past the
// last block row), so the chunks carrying into one row are contiguous; a segmented
// scan over the lanes sums each run and the run's last lane adds the sum to
// the output -- a plain read-modify-write when the run lies inside this
// wave's 64 chunks (it is then the row's only writer in this pass), atomics
// when the run continues into a neighbouring wave's chunks.
template <int T, int OP>
__global__ __launch_bounds__(256) void bmv_fixup_kernel(BmvArgs a) {
  using Acc = typename AccOf<OP>::type;
  using W = typename Tile<T>::W;
  if ((int)blockIdx.x >= a.fixup_blocks) {
    // Workgroups past the fix-up proper (launched when C is overwritten)
    // zero the empty block rows, which the main pass steps over. No carry
    // lands in an empty row, so these stores meet no other writer.
    const long long stride = (long long)(gridDim.x - a.fixup_blocks) * blockDim.x;
    for (long long I = (long long)(blockIdx.x - a.fixup_blocks) * blockDim.x + threadIdx.x; I < a.nbr;
         I += stride) {
      if (a.rowptr[I] != a.rowptr[I + 1]) continue;
      if constexpr (OP == OP_BITS || OP == OP_MASKED) {
        static_cast<W *>(a.C)[I] = (W)0;
      } else {
        typedef uint32_t U4 __attribute__((ext_vector_type(4)));
        U4 *y = reinterpret_cast<U4 *>(static_cast<uint32_t *>(a.C) + I * T);
#pragma unroll
        for (int r = 0; r < T / 4; ++r) y[r] = U4{0u, 0u, 0u, 0u};
      }
    }
    return;
  }
  if constexpr (OP == OP_BITS || OP == OP_MASKED) {
    // first word whose aligned dword reaches past C[nbr-1] (nbr: none)
    const int tail = sizeof(W) < 4 ? (int)((((long long)a.nbr * sizeof(W)) & ~3ll) / sizeof(W)) : a.nbr;
    if ((int)blockIdx.x == a.fixup_blocks - 1) {  // the tail workgroup
      if constexpr (sizeof(W) < 4)
        if (threadIdx.x < 64 && tail < a.nbr) bits_tail<T, OP>(a, tail, threadIdx.x);
      return;
    }
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.nchunks) return;
    const int key = a.carry_key[tid];
    if (key < 0 || key >= tail) return;  // no record, or a tail word
    const Acc *cv = static_cast<const Acc *>(a.carry_val) + tid * T;
    Word<T> bits = 0;
#pragma unroll
    for (int r = 0; r < T; ++r) bits |= (Word<T>)(cv[r] != 0u ? 1u : 0u) << (T - 1 - r);
    if constexpr (OP == OP_MASKED) bits &= ~(Word<T>) static_cast<const W *>(a.mask)[key];
    if (bits == 0) return;
    if constexpr (T == 64) {
      atomicOr(reinterpret_cast<unsigned long long *>(static_cast<W *>(a.C) + key), (unsigned long long)bits);
    } else {
      // OR into the aligned dword holding word `key` (W is 1, 2 or 4 bytes).
      const uintptr_t addr = reinterpret_cast<uintptr_t>(static_cast<W *>(a.C) + key);
      const int shift = (int)(addr & 3) * 8;
      atomicOr(reinterpret_cast<unsigned int *>(addr & ~(uintptr_t)3), (uint32_t)bits << shift);
    }
  } else {
    typedef Acc Acc4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const long long c0 = ((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
    if (c0 >= a.nchunks) return;  // uniform over the wave
    const long long c = c0 + lane;
    const bool in = c < a.nchunks;
    // every load is issued up front, unconditionally (indices clamped, values
    // of chunks without a record zeroed after), so they share one round trip:
    // the keys, the records and the neighbouring waves' boundary keys
    const long long cc = in ? c : a.nchunks - 1;
    int key = a.carry_key[cc];
    const int kprev = a.carry_key[c0 > 0 ? c0 - 1 : 0];
    const int knext_wave = a.carry_key[c0 + 64 < a.nchunks ? c0 + 64 : a.nchunks - 1];
    Acc v[T];
    const Acc4 *cv = reinterpret_cast<const Acc4 *>(static_cast<const Acc *>(a.carry_val) + cc * T);
#pragma unroll
    for (int r = 0; r < T / 4; ++r) {
      const Acc4 f = cv[r];
      v[4 * r] = f.x;
      v[4 * r + 1] = f.y;
      v[4 * r + 2] = f.z;
      v[4 * r + 3] = f.w;
    }
    if (!in || key < 0) key = INT_MAX;  // no open row: never written
#pragma unroll
    for (int r = 0; r < T; ++r) v[r] = key != INT_MAX ? v[r] : (Acc)0;
    group_scan<1, T, Acc>(key, v, [](Acc x, Acc y) { return x + y; });
    const int key0 = __shfl(key, 0), kend = __shfl(key, 63);
    const int knext = __shfl_down(key, 1);
    const bool run_end = lane == 63 || knext != key;
    if (!run_end || key == INT_MAX) return;
    // does the run continue from the previous / into the next wave's chunks?
    const bool from_prev = key == key0 && c0 > 0 && kprev == key;
    const bool into_next = lane == 63 && c0 + 64 < a.nchunks && knext_wave == kend;
    // output element type: int32 counts, else float (counts converted)
    using Out = typename std::conditional<OP == OP_COUNT_I32, int, float>::type;
    typedef Out Out4 __attribute__((ext_vector_type(4)));
    Out *y = static_cast<Out *>(a.C) + (long long)key * T;
    if (from_prev || into_next) {
#pragma unroll
      for (int r = 0; r < T; ++r) {
        if (v[r] == (Acc)0) continue;
        if constexpr (OP == OP_COUNT_I32) atomicAdd(y + r, (int)v[r]);
        else unsafeAtomicAdd(y + r, (float)v[r]);
      }
    } else {
      bool nz = false;
#pragma unroll
      for (int r = 0; r < T; ++r) nz |= v[r] != (Acc)0;
      if (!nz) return;
      Out4 *y4 = reinterpret_cast<Out4 *>(y);
#pragma unroll
      for (int r = 0; r < T / 4; ++r) {
        Out4 t = y4[r];
        t.x += (Out)v[4 * r];
        t.y += (Out)v[4 * r + 1];
        t.z += (Out)v[4 * r + 2];
        t.w += (Out)v[4 * r + 3];
        y4[r] = t;
      }
    }
  }
}

// Short-row path (T <= 8, every block row at most kShortRowMax tiles): one
// lane per block row walks the row's tiles in batches of U tiles with
// unconditional loads (indices clamped into the row, contributions of the
// padding masked), then writes the row once -- no chunk partition, row
// window, segmented scan or carries. For regular matrices (a 2-D grid has 2-3
// tiles per block row) the chunk walk would otherwise spend most of its time
// on row ends: a run of 8 tiles crosses ~3 of them.
// U: tiles per batch of loads, picked from the longest block row (a row of
// at most 8 tiles is one batch with the fewest padding slots: on the 4096^2
// grid, 5 tiles per block row, batches of 5 run T = 8 in 67 us, of 6 in 72,
// of 4 -- two batches -- in 83).
template <int T, int OP, int U>
__global__ __launch_bounds__(256) void bmv_rows_kernel(BmvArgs a) {
  static_assert(Tile<T>::L == 1, "one lane per block row needs T <= 8");
  constexpr int kMode = OP == OP_FULL_F32 ? kRowSumF32 : OP == OP_FULL_BF16 ? kRowSumBF16 : kRowCount;
  using Acc = typename AccOf<OP>::type;
  // consecutive block rows on one XCD: the x segments a row gathers (for a
  // grid, block columns I and I +- one grid row) stay in that XCD's L2
  for (long long I = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; I < a.nbr;
       I += (long long)gridDim.x * blockDim.x) {
    Acc acc[Tile<T>::R];
    block_row_sums<T, kMode, U>(a.A, a.colind, a.B, a.rowptr[I], a.rowptr[I + 1], acc);
    flush_row<T, OP>(a, (int)I, 0, acc);
  }
}

// ---------------------------------------------------------------------------
// Row pieces (count ops, T <= 8). The chunk walk cuts the tile array into
// fixed chunks regardless of rows, and pays for it per chunk: the group-row
// pass, a row-end test per tile, the segmented scan and carries. Here the
// work is cut along rows instead, classified once per matrix:
//   * short block rows (at most kShortRowMax tiles): one lane per row, by a
//     few waves of each piece-kernel workgroup (x read as the pieces read it);
//   * longer rows: pieces of piece_tiles<T>() tiles on a grid aligned to the row's
//     first 16-byte unit of tile words, one wave per piece (the workgroup's
//     other waves): lanes stream whole 16-byte units (4 tiles at T = 4, 2 at
//     T = 8), count with byte permutes and v_bcnt, and one DPP wave sum per
//     two rows gives the piece's row totals -- no row ends inside a piece. A
//     row's last piece writes the row (plain store or accumulate); its other
//     pieces leave carry records, which the chunk path's fix-up adds (records
//     of one row are consecutive).
// Slots outside the piece (the aligned unit's head, the tail of the row's
// last unit) count nothing (their x words are zeroed); units past the
// matrix read zeros through the buffer resources' range checks.
// Tiles per piece: 256 at T = 4 (one 16-byte unit per lane), 256 at T = 8
// (two units per lane).
template <int T>
constexpr int piece_tiles() { return 256; }

// Pieces of block row I (0: a short row, taken by the one-lane-per-row pass).
template <int T>
__device__ __forceinline__ int row_pieces(int s, int e) {
  constexpr int A = T == 4 ? 4 : 2;  // tiles per 16-byte unit
  return e - s > kShortRowMax ? (e - (s & ~(A - 1)) + piece_tiles<T>() - 1) / piece_tiles<T>() : 0;
}

template <int T>
__global__ __launch_bounds__(256) void piece_count_kernel(const int *__restrict__ rowptr, int nbr,
                                                          int *__restrict__ np,
                                                          unsigned long long *__restrict__ long_tiles) {
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long lt = 0;  // tiles of the rows cut into pieces
  if (I < nbr) {
    const int s = rowptr[I], e = rowptr[I + 1], n = row_pieces<T>(s, e);
    np[I] = n;
    lt = n > 0 ? (unsigned long long)(e - s) : 0ull;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lt += __shfl_xor(lt, o);
  if ((threadIdx.x & 63) == 0 && lt != 0) atomicAdd(long_tiles, lt);
}

// Piece descriptors: {row, 2 * (piece of the row) + (last piece of the row),
// the row's first and end tile when the plan was made}.
__global__ void piece_fill_kernel(const int *__restrict__ rowptr, const int *__restrict__ np,
                                  const int *__restrict__ off, int nbr, int4 *__restrict__ pieces) {
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const int n = np[I], o = off[I], s = rowptr[I], e = rowptr[I + 1];
  for (int j = 0; j < n; ++j) pieces[o + j] = make_int4((int)I, 2 * j + (j == n - 1 ? 1 : 0), s, e);
}

// One wave per piece, persistent over the piece list (waves take pieces
// w, w + nw, ...). XL: packed x in LDS -- 1 all of it, 2 its first
// a.xl_words words (the rest from memory), 0 none (gathered from memory).
// A piece's inputs: its tile words and block columns (one 16-byte unit of
// tile words per lane and step; units outside the piece read nothing, their
// offsets lying past the buffers), the row's extent now, the packed-x read
// of each tile, and the piece's tile range.
template <int T>
struct PieceIn {
  static constexpr int A = T == 4 ? 4 : 2;          // tiles per 16-byte unit
  static constexpr int STEPS = piece_tiles<T>() / (64 * A);  // units per lane
  uint4 wv[STEPS];
  uint32_t jv[STEPS][A];
  uint32_t xw[STEPS][A];
  int s_now, e_now;
  int base, lo, hi;  // the piece's aligned start and tiles [lo, hi) (wave-uniform)
};

template <int T>
__device__ __forceinline__ int piece_base(int s, int jl) {
  return (s & ~(PieceIn<T>::A - 1)) + (jl >> 1) * piece_tiles<T>();
}
// End of the tiles piece d covers (a row's last piece: the row's end).
template <int T>
__device__ __forceinline__ int piece_hi(int4 d, int base) {
  return (d.y & 1) ? d.w : min(base + piece_tiles<T>(), d.w);
}

// Issues the tile-word, block-column and row-extent loads of piece d.
template <int T>
__device__ __forceinline__ void piece_load(const BmvArgs &a, __amdgpu_buffer_rsrc_t vr, __amdgpu_buffer_rsrc_t cr,
                                           int4 d, int lane, PieceIn<T> &in) {
  constexpr int A = PieceIn<T>::A, TB = T == 4 ? 4 : 8;
  const int base = piece_base<T>(d.z, d.y);
  const int hi = piece_hi<T>(d, base);
#pragma unroll
  for (int st = 0; st < PieceIn<T>::STEPS; ++st) {
    const int t0 = base + (lane + 64 * st) * A;
    const bool need = t0 < hi;
    in.wv[st] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                              vr, need ? t0 * TB : a.nblocks * TB, 0, 0));
    if constexpr (A == 4) {
      const uint4 c = __builtin_bit_cast(
          uint4, __builtin_amdgcn_raw_buffer_load_b128(cr, need ? t0 * 4 : a.nblocks * 4, 0, 0));
      in.jv[st][0] = c.x, in.jv[st][1] = c.y, in.jv[st][2] = c.z, in.jv[st][3] = c.w;
    } else {
      const uint2 c = __builtin_bit_cast(
          uint2, __builtin_amdgcn_raw_buffer_load_b64(cr, need ? t0 * 4 : a.nblocks * 4, 0, 0));
      in.jv[st][0] = c.x, in.jv[st][1] = c.y;
    }
  }
  // vector loads (a scalar load's wait would hold up every LDS read)
  int iv = d.x;
  asm volatile("" : "+v"(iv));
  in.s_now = a.rowptr[iv];
  in.e_now = a.rowptr[iv + 1];
}

// Packed-x word J as the piece walk reads it: from LDS (XL 1); from LDS when
// staged and else from memory (XL 2: both loads issued, the memory one with
// an offset past the buffer when the word is staged); from memory (XL 0);
// XL 3: from LDS when staged, else from memory only when the LDS bitmap of
// the unstaged words says it is non-zero (zero otherwise, with no memory
// access).
template <int T, int XL>
__device__ __forceinline__ uint32_t piece_xword(const BmvArgs &a, const uint8_t *xl, __amdgpu_buffer_rsrc_t xr,
                                                int J) {
  using W = typename Tile<T>::W;
  if constexpr (XL == 1) {
    return lds_xword<T>(xl, J);
  } else if constexpr (XL == 2 || XL == 3) {
    const bool staged = J < a.xl_words;
    const uint32_t xs = lds_xword<T>(xl, staged ? J : 0);
    bool mem = !staged;
    if constexpr (XL == 3) {
      const int jb = staged ? 0 : J - a.xl_words;
      const uint32_t bw = reinterpret_cast<const uint32_t *>(xl + a.bm_off)[jb >> 5];
      mem = mem && ((bw >> (jb & 31)) & 1u) != 0;
    }
    const uint32_t xm =
        __builtin_amdgcn_raw_buffer_load_b8(xr, mem ? J * (int)sizeof(W) : a.x_words * (int)sizeof(W), 0, 0);
    return staged ? xs : xm;
  } else {
    return __builtin_amdgcn_raw_buffer_load_b8(xr, J * (int)sizeof(W), 0, 0);
  }
}

// Issues the packed-x reads of piece d's tiles (its block columns loaded):
// XL 1 all of x staged in LDS, 2 its first a.xl_words words (the others
// from memory), 3 those and the non-zero bitmap of the others, 0 none. Every slot reads the word of its block column (a
// slot outside the piece holds a tile of a neighbouring row, or reads zeros
// past the matrix: either way a valid column); piece_count zeroes the slots
// outside [lo, hi). With x all in LDS only the loads are issued here (the
// raw LDS byte is kept; piece_count extracts the word), so their latency
// overlaps the count of the piece before.
template <int T, int XL>
__device__ __forceinline__ void piece_xload(const BmvArgs &a, const uint8_t *xl, __amdgpu_buffer_rsrc_t xr, int4 d,
                                            int lane, PieceIn<T> &in) {
  using W = typename Tile<T>::W;
  constexpr int A = PieceIn<T>::A;
  const int base = piece_base<T>(d.z, d.y);
  in.base = base;
  in.lo = base > d.z ? base : d.z;
  in.hi = piece_hi<T>(d, base);
#pragma unroll
  for (int st = 0; st < PieceIn<T>::STEPS; ++st)
#pragma unroll
    for (int u = 0; u < A; ++u) {
      const int J = (int)in.jv[st][u];
      uint32_t x;
      if constexpr (XL == 1) {
        x = xl[T == 4 ? (J >> 1) : J];
      } else if constexpr (XL == 0) {
        x = __builtin_amdgcn_raw_buffer_load_b8(xr, J * (int)sizeof(W), 0, 0);
      } else {
        x = piece_xword<T, XL>(a, xl, xr, J);
      }
      in.xw[st][u] = x;
    }
}

// Row totals t[] += the counts of a loaded piece. Per lane: the x word of
// each of its 4 tiles (0 outside the piece), then the rows' bits of all 4
// tiles gathered into one register per row with byte permutes, one v_bcnt
// per row; then one DPP wave sum per two rows (16-bit fields: a piece adds
// at most 256 * T per row).
//   T = 4: tile words hold rows 0..3 as the low nibbles of bytes 0..3; tiles
//          0 and 1 (2 and 3) share a register, tile 1 (3) in the high
//          nibbles, ANDed with their x words likewise placed; row r's 16 bits
//          are byte r of both registers.
//   T = 8: tile k's rows 0..3 and 4..7 are bytes of two registers; a 4x4
//          byte transpose of the four tiles' low (high) registers gives row
//          r in one register, byte k from tile k, ANDed with the four x
//          bytes.
template <int T, int XL>
__device__ __forceinline__ void piece_count(const PieceIn<T> &in, int lane, uint32_t (&t)[T]) {
  constexpr int A = PieceIn<T>::A, STEPS = PieceIn<T>::STEPS;
  constexpr int G = STEPS * A / 4;  // groups of 4 tiles per lane
  static_assert(STEPS * A % 4 == 0, "whole groups of four tiles per lane");
  const bool edge = !(in.lo == in.base && in.hi == in.base + piece_tiles<T>());  // wave-uniform
  uint32_t lt[T];
#pragma unroll
  for (int r = 0; r < T; ++r) lt[r] = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    uint32_t xv[4], w[4 * (T / 4)];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int st = (4 * g + k) / A, u = (4 * g + k) % A;
      uint32_t x = in.xw[st][u];
      if constexpr (XL == 1 && T == 4) x = __builtin_amdgcn_ubfe(x, (in.jv[st][u] & 1u) << 2, 4);
      else if constexpr (T == 4) x &= 0xFu;
      else x &= 0xFFu;
      if (edge) {  // slots outside the piece count nothing
        const int tt = in.base + (lane + 64 * st) * A + u;
        x = tt >= in.lo && tt < in.hi ? x : 0u;
      }
      xv[k] = x;
    }
#pragma unroll
    for (int q = 0; q < T / 4; ++q) {
      const uint4 v = in.wv[g * (T / 4) + q];
      w[4 * q] = v.x, w[4 * q + 1] = v.y, w[4 * q + 2] = v.z, w[4 * q + 3] = v.w;
    }
    if constexpr (T == 4) {
      const uint32_t c0 = (w[0] & 0x0F0F0F0Fu) | ((w[1] << 4) & 0xF0F0F0F0u);
      const uint32_t c1 = (w[2] & 0x0F0F0F0Fu) | ((w[3] << 4) & 0xF0F0F0F0u);
      const uint32_t m0 = c0 & __builtin_amdgcn_perm(0u, xv[0] | (xv[1] << 4), 0u);  // byte 0 to every byte
      const uint32_t m1 = c1 & __builtin_amdgcn_perm(0u, xv[2] | (xv[3] << 4), 0u);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lt[r] += __builtin_popcount(__builtin_amdgcn_perm(m1, m0, 0x0C0C0000u | ((4u + r) << 8) | (uint32_t)r));
    } else {
      // w: tile 0 rows 0-3, rows 4-7, tile 1 rows 0-3, rows 4-7, tile 2 ..., tile 3 ...
      const uint32_t X4 = xv[0] | (xv[1] << 8) | (xv[2] << 16) | (xv[3] << 24);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t r0 = w[h], r1 = w[2 + h], r2 = w[4 + h], r3 = w[6 + h];  // tiles 0..3
        const uint32_t p0 = __builtin_amdgcn_perm(r1, r0, 0x05010400u);  // t0.b0 t1.b0 t0.b1 t1.b1
        const uint32_t p1 = __builtin_amdgcn_perm(r1, r0, 0x07030602u);  // t0.b2 t1.b2 t0.b3 t1.b3
        const uint32_t p2 = __builtin_amdgcn_perm(r3, r2, 0x05010400u);
        const uint32_t p3 = __builtin_amdgcn_perm(r3, r2, 0x07030602u);
        lt[4 * h + 0] += __builtin_popcount(__builtin_amdgcn_perm(p2, p0, 0x05040100u) & X4);
        lt[4 * h + 1] += __builtin_popcount(__builtin_amdgcn_perm(p2, p0, 0x07060302u) & X4);
        lt[4 * h + 2] += __builtin_popcount(__builtin_amdgcn_perm(p3, p1, 0x05040100u) & X4);
        lt[4 * h + 3] += __builtin_popcount(__builtin_amdgcn_perm(p3, p1, 0x07060302u) & X4);
      }
    }
  }
  // wave totals, rows r and r + T/2 in one register (a lane adds at most
  // piece_tiles / 64 * T per row, a wave piece_tiles * T)
#pragma unroll
  for (int r = 0; r < T / 2; ++r) {
    const uint32_t v = wave_total(lt[r] | (lt[r + T / 2] << 16));
    t[r] += v & 0xFFFFu;
    t[r + T / 2] += v >> 16;
  }
}

// Counts of a short block row (tiles [s, e), s < e) for the piece kernel:
// batches of 8 tiles, every load of a batch issued unconditionally (indices
// clamped into the row, the padding slots' words zeroed), x words read as
// the pieces read them.
template <int T, int XL>
__device__ __forceinline__ void short_row_counts(const BmvArgs &a, const uint8_t *xl, __amdgpu_buffer_rsrc_t xr,
                                                 int s, int e, uint32_t (&acc)[T]) {
  constexpr int U = 8, NW = Tile<T>::NW;
#pragma unroll
  for (int i = 0; i < T; ++i) acc[i] = 0;
  for (int k = s; k < e; k += U) {
    uint32_t F[U][NW];
    int J[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = k + u < e ? k + u : e - 1;
      J[u] = a.colind[p];
      load_frag<T>(a.A, p, 0, F[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k + u >= e) {
#pragma unroll
        for (int w = 0; w < NW; ++w) F[u][w] = 0u;
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t x = piece_xword<T, XL>(a, xl, xr, J[u]);
#pragma unroll
      for (int i = 0; i < T; ++i) acc[i] += bit_count(row_word<T>(F[u], i) & x);
    }
  }
}

// One wave per piece, persistent over the piece list (waves take pieces
// w, w + nw, ...), software-pipelined four deep: while piece p is counted,
// the x words of the wave's next piece and the tile words, block columns
// and row extents of the two after it are in flight. Each step issues the
// descriptor load first and the tile loads last, so that waiting for a
// descriptor or for x never waits for the tiles issued after it (the load
// counter counts in issue order). XL: packed x in LDS -- 1
// all of it, 2 its first a.xl_words words (the rest from memory), 3 those
// and the non-zero bitmap of the rest (only non-zero words read from
// memory), 0 none.
template <int T, int OP, int XL>
__global__ __launch_bounds__(1024) void bmv_piece_kernel(BmvArgs a, int x_bytes, const int4 *__restrict__ pieces,
                                                         int npieces, const int *__restrict__ np, int ns) {
  static_assert(T == 4 || T == 8, "one lane per tile");
  using W = typename Tile<T>::W;
  using Acc = typename AccOf<OP>::type;
  static_assert(std::is_same<Acc, uint32_t>::value, "count ops");
  constexpr int TB = T == 4 ? 4 : 8;  // bytes of tile words per tile
  extern __shared__ __align__(16) uint8_t s_dyn[];
  uint8_t *xl = s_dyn;
  if constexpr (XL != 0) {
    stage_x<T>(a, x_bytes, xl);
    if constexpr (XL == 3) {
      const uint4 *src = reinterpret_cast<const uint4 *>(a.xbm);
      uint4 *dst = reinterpret_cast<uint4 *>(xl + a.bm_off);
      for (int i = threadIdx.x; i < a.bm_dwords / 4; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  // The first ns waves of each workgroup count the short rows, the others
  // walk the pieces: the short rows' dependent loads (row extents, then
  // tiles) run beside the piece pipeline instead of delaying it.
  const int wv = (int)threadIdx.x >> 6, wpg = (int)blockDim.x >> 6;
  const bool srw = wv < ns;
  const int nw = (int)gridDim.x * (wpg - ns);
  const int nsw = (int)gridDim.x * ns;
  int p = __builtin_amdgcn_readfirstlane(srw ? (int)blockIdx.x * ns + wv : (int)blockIdx.x * (wpg - ns) + wv - ns);
  const int p0 = p;
  const __amdgpu_buffer_rsrc_t cr =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(a.colind), (short)0, a.nblocks * 4, 0x00020000);
  const __amdgpu_buffer_rsrc_t vr =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.A), (short)0, a.nblocks * TB, 0x00020000);
  const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.B), (short)0,
                                                                      a.x_words * (int)sizeof(W), 0x00020000);
  // The short rows (np[I] == 0), one lane each, in blocks of 64 consecutive
  // block rows, short-row wave w taking blocks w, w + nsw, ..., with x read
  // as the pieces read it (from LDS where staged). Empty rows get zeros when
  // C is overwritten (the fix-up zeroes none on this path).
  auto short_rows = [&]() {
    for (long long blk = p0; blk * 64 < a.nbr; blk += nsw) {
      const long long I = blk * 64 + lane;
      if (I < a.nbr && np[I] == 0) {
        const int s = a.rowptr[I], e = a.rowptr[I + 1];
        uint32_t acc[T];
        if (s < e) short_row_counts<T, XL>(a, xl, xr, s, e, acc);
        else {
#pragma unroll
          for (int r = 0; r < T; ++r) acc[r] = 0;
        }
        if (s < e || !a.accumulate) flush_row<T, OP>(a, (int)I, 0, acc);
      }
    }
  };
  if (srw) {  // uniform per wave; no barrier follows
    short_rows();
    return;
  }
  if (p >= npieces) return;
  // descriptors as vector loads; past the wave's last piece every load
  // (descriptor, tiles, x) repeats its first piece, results unused, so that
  // all loads stay unconditional and the waits count them exactly
  auto load_desc = [&](int pp) {
    int pv = pp < npieces ? pp : p0;
    asm volatile("" : "+v"(pv));
    return pieces[pv];
  };
  auto uni = [](int4 d) {
    return make_int4(__builtin_amdgcn_readfirstlane(d.x), __builtin_amdgcn_readfirstlane(d.y),
                     __builtin_amdgcn_readfirstlane(d.z), __builtin_amdgcn_readfirstlane(d.w));
  };
  // d0: piece p (x in flight or arrived), d1: p + nw (tiles in flight),
  // dn: the descriptor of p + 2nw (in flight)
  int4 d0 = uni(load_desc(p));
  int4 d1 = uni(load_desc(p + nw));
  int4 d2 = uni(load_desc(p + 2 * nw));
  int4 dn = load_desc(p + 3 * nw);
  auto step = [&](PieceIn<T> &cur, PieceIn<T> &nxt, PieceIn<T> &mid, PieceIn<T> &far) {
    const int pc = p;
    const bool more = p + nw < npieces;  // wave-uniform
    const int4 d3 = uni(dn);
    dn = load_desc(p + 4 * nw);
    piece_xload<T, XL>(a, xl, xr, d1, lane, nxt);  // x of p + nw
    piece_load<T>(a, vr, cr, d3, lane, far);   // tiles of p + 3nw
    const int I = d0.x, jl = d0.y;
    const bool last = (jl & 1) != 0;
    uint32_t t[T];
#pragma unroll
    for (int r = 0; r < T; ++r) t[r] = 0;
    const int s_now = __builtin_amdgcn_readfirstlane(cur.s_now), e_now = __builtin_amdgcn_readfirstlane(cur.e_now);
    if (s_now == d0.z && e_now == d0.w) {
      piece_count<T, XL>(cur, lane, t);
    } else {
      // rowptr was rewritten since the plan: count the row as it is now, a
      // window at a time (the row's last piece runs on to the row's end)
      const int base0 = piece_base<T>(s_now, jl);
      constexpr int P = piece_tiles<T>();
      const int kend = last ? e_now : min(base0 + P, e_now);
      for (int kb = base0; kb < kend; kb += P) {
        PieceIn<T> w;
        const int4 dw = make_int4(I, 1, kb, min(kb + P, kend));  // one window, from kb
        piece_load<T>(a, vr, cr, dw, lane, w);
        piece_xload<T, XL>(a, xl, xr, make_int4(I, 1, kb > s_now ? kb : s_now, dw.w), lane, w);
        piece_count<T, XL>(w, lane, t);
      }
      // drain everything here, so that after the join the compiler still
      // counts the prefetches of the common path exactly (rather than
      // waiting for all of them before the next piece)
      __builtin_amdgcn_s_waitcnt(0);
    }
    if (lane == 0) {
      if (last) {
        flush_row<T, OP>(a, I, 0, t);
        a.carry_key[pc] = -1;
      } else {
        a.carry_key[pc] = I;
        typedef uint32_t U4 __attribute__((ext_vector_type(4)));
        U4 *cv = reinterpret_cast<U4 *>(static_cast<uint32_t *>(a.carry_val) + (long long)pc * T);
#pragma unroll
        for (int r = 0; r < T / 4; ++r) cv[r] = U4{t[4 * r], t[4 * r + 1], t[4 * r + 2], t[4 * r + 3]};
      }
    }
    d0 = d1;
    d1 = d2;
    d2 = d3;
    p += nw;
    return more;
  };
  PieceIn<T> b0, b1, b2, b3;
  piece_load<T>(a, vr, cr, d0, lane, b0);
  piece_load<T>(a, vr, cr, d1, lane, b1);
  piece_load<T>(a, vr, cr, d2, lane, b2);
  piece_xload<T, XL>(a, xl, xr, d0, lane, b0);
  while (step(b0, b1, b2, b3) && step(b1, b2, b3, b0) && step(b2, b3, b0, b1) && step(b3, b0, b1, b2)) {
  }
}

// Most tiles in one block row (device-wide max of rowptr[I+1] - rowptr[I]):
// a grid-stride loop over at most kMaxRowBlocks workgroups of 256, a
// workgroup max, one atomic per workgroup (one per wave on one address
// serialised: 385 us for 4 M block rows).
constexpr int kMaxRowBlocks = 1024;
__global__ __launch_bounds__(256) void max_row_tiles_kernel(const int *rowptr, int nbr, int *out) {
  __shared__ int s_m[4];
  int m = 0;
  for (long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x; I < nbr;
       I += (long long)gridDim.x * blockDim.x) {
    const int d = rowptr[I + 1] - rowptr[I];
    m = d > m ? d : m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int t = __shfl_xor(m, o);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int b = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
    if (b > 0) atomicMax(out, b);
  }
}

// Largest block column (max of colind[0..n)) into *out (pre-zeroed), same
// shape as max_row_tiles_kernel.
__global__ __launch_bounds__(256) void max_col_kernel(const int *colind, long long n, int *out) {
  __shared__ int s_m[4];
  int m = 0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    m = colind[k] > m ? colind[k] : m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int t = __shfl_xor(m, o);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int b = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
    if (b > 0) atomicMax(out, b);
  }
}

template <class V>
__global__ void fill_kernel(V *p, long long n, V v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Scratch of the allocation-free entry point (bgb_bmv_bin_bin_full_i32_noalloc):
// carry keys, chunk partition and carry records of one call, 64 MiB per
// device, held from the library's load until the process ends. Every call on
// a device shares that device's copy, so such calls must not overlap on one
// device (they would race on it); nothing in it needs clearing between calls
// (every chunk writes its key and partition entry, records are read only
// under a key that says they were written). Host code reaches it only through
// the address hipGetSymbolAddress returns for the current device: the name
// itself is a host-side stand-in, not a device address.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices];  // per device, set on its first call
}  // namespace

// Device address of the current device's copy of `scratch`, looked up once
// per device (with that device current) and kept.
int static_scratch(uint8_t **out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  BGB_REQUIRE(dev >= 0 && dev < kMaxDevices, BGB_ERR_ARG, "device %d beyond the %d the scratch table holds", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    BGB_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BGB_REQUIRE(p, BGB_ERR_HIP, "hipGetSymbolAddress returned no address for the static scratch");
    g_scratch_ptr[dev] = static_cast<uint8_t *>(p);
  }
  *out = g_scratch_ptr[dev];
  return BGB_OK;
}

// Piece plans of the row-piece path, one per (device, rowptr address, nbr,
// nblocks, T): per block row its number of pieces (0: short), and the piece
// descriptors. Made on the first call for a matrix (one read-back of the
// piece count, which synchronises the stream) and kept for the next. The
// kernels read the row extents from rowptr itself, and a row's last piece
// runs on to the row's end, so a rowptr rewritten in place (same nbr) still
// gives exact results from a stale plan.
namespace {
struct PieceEntry {
  int dev;
  const int *rowptr;
  int nbr, nblocks, T;
  int *np;        // nbr + 1 ints
  int4 *pieces;   // npieces
  int npieces;
  long long long_tiles;  // tiles of the rows cut into pieces (the others: short rows)
};
constexpr int kPieceCache = 8;
std::mutex g_piece_mu;
std::vector<PieceEntry> g_piece;  // most recently used last

void free_piece(const PieceEntry &e) {
  (void)hipSetDevice(e.dev);
  (void)hipDeviceSynchronize();  // queued work on any stream may still read it
  (void)hipFree(e.np);
  (void)hipFree(e.pieces);
}
}  // namespace

template <int T>
int piece_plan(hipStream_t s, const int *rowptr, int nbr, int nblocks, PieceEntry *out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_piece_mu);
    for (size_t i = 0; i < g_piece.size(); ++i) {
      const PieceEntry e = g_piece[i];
      if (e.dev == dev && e.rowptr == rowptr && e.nbr == nbr && e.nblocks == nblocks && e.T == T) {
        g_piece.erase(g_piece.begin() + (long)i);
        g_piece.push_back(e);
        *out = e;
        return BGB_OK;
      }
    }
  }
  PieceEntry ne{dev, rowptr, nbr, nblocks, T, nullptr, nullptr, 0, 0};
  DeviceBuffer off;
  hipError_t he = hipMalloc(&ne.np, sizeof(int) * ((size_t)nbr + 1));
  if (he != hipSuccess) return set_error(BGB_ERR_ALLOC, "piece plan hipMalloc failed: %s", hipGetErrorString(he));
  auto fail = [&](int code, const char *what, hipError_t err) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(ne.np);
    if (ne.pieces) (void)hipFree(ne.pieces);
    return set_error(code, "piece plan: %s: %s", what, hipGetErrorString(err));
  };
  // offsets (nbr + 1 ints), then the long-row tile counter (8 bytes, aligned)
  const size_t lt_off = (sizeof(int) * ((size_t)nbr + 1) + 7) & ~(size_t)7;
  if ((he = off.alloc(lt_off + sizeof(unsigned long long), s)) != hipSuccess)
    return fail(BGB_ERR_ALLOC, "alloc", he);
  unsigned long long *d_lt = reinterpret_cast<unsigned long long *>(static_cast<char *>(off.ptr) + lt_off);
  fill_kernel<int><<<1, 64, 0, s>>>(ne.np + nbr, 1, 0);
  fill_kernel<unsigned long long><<<1, 64, 0, s>>>(d_lt, 1, 0ull);
  piece_count_kernel<T><<<(unsigned)ceil_div(nbr, 256), 256, 0, s>>>(rowptr, nbr, ne.np, d_lt);
  if ((he = hipGetLastError()) != hipSuccess) return fail(BGB_ERR_HIP, "count", he);
  if ((he = exclusive_scan_i32(ne.np, off.as<int>(), (size_t)nbr + 1, s)) != hipSuccess)
    return fail(BGB_ERR_HIP, "scan", he);
  int total = 0;
  unsigned long long lt = 0;
  if ((he = hipMemcpyAsync(&total, off.as<int>() + nbr, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (he = hipMemcpyAsync(&lt, d_lt, sizeof(lt), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (he = hipStreamSynchronize(s)) != hipSuccess)
    return fail(BGB_ERR_HIP, "read-back", he);
  ne.npieces = total;
  ne.long_tiles = (long long)lt;
  if ((he = hipMalloc(&ne.pieces, sizeof(int4) * (size_t)std::max(total, 1))) != hipSuccess)
    return fail(BGB_ERR_ALLOC, "hipMalloc", he);
  if (total > 0) {
    piece_fill_kernel<<<(unsigned)ceil_div(nbr, 256), 256, 0, s>>>(rowptr, ne.np, off.as<int>(), nbr, ne.pieces);
    if ((he = hipGetLastError()) != hipSuccess) return fail(BGB_ERR_HIP, "fill", he);
  }
  std::lock_guard<std::mutex> lk(g_piece_mu);
  if ((int)g_piece.size() >= kPieceCache) {
    free_piece(g_piece.front());
    g_piece.erase(g_piece.begin());
    (void)hipSetDevice(dev);
  }
  g_piece.push_back(ne);
  *out = ne;
  return BGB_OK;
}

// Non-zero bitmap of the packed-x words past the first `from`: bit j of
// bm[j / 32] = (word from + j != 0) (T = 4: its low nibble), bm_dwords
// dwords (those past the words written as 0). One thread per 16 words (one
// 16-byte load; x and `from` are 16-byte aligned), two threads per dword.
template <int T>
__global__ __launch_bounds__(256) void x_nonzero_bitmap_kernel(const uint8_t *__restrict__ x, int from, int words,
                                                               uint32_t *__restrict__ bm, int bm_dwords) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long j0 = from + 16 * t;
  uint32_t m = 0;  // bit k: word j0 + k non-zero
  if (j0 + 16 <= words) {
    const uint4 v = *reinterpret_cast<const uint4 *>(x + j0);
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t b = (w4[k >> 2] >> (8 * (k & 3))) & (T == 4 ? 0xFu : 0xFFu);
      m |= (b != 0u ? 1u : 0u) << k;
    }
  } else {
    for (int k = 0; k < 16 && j0 + k < words; ++k) {
      const uint32_t b = x[j0 + k] & (T == 4 ? 0xFu : 0xFFu);
      m |= (b != 0u ? 1u : 0u) << k;
    }
  }
  const uint32_t hi = __shfl_xor(m, 1);  // the odd thread's 16 words
  if ((t & 1) == 0 && t / 2 < bm_dwords) bm[t / 2] = m | (hi << 16);
}

// Where the piece kernel reads the packed x from (XL of bmv_piece_kernel):
// 1 all of it in LDS; 3 its first `words` words in LDS with a bitmap of the
// others' non-zero words (read from memory only where non-zero), when that
// fits and x is larger than LDS; 2 its first words in LDS, the rest from
// memory; 0 all from memory. lds: dynamic LDS bytes; bm_bytes: global
// bitmap bytes the launch needs (XL 3).
struct PieceX {
  int xl = 0;
  long long words = 0, x_bytes = 0, stage_words = 0, lds = 0;
  int bm_dwords = 0, bm_off = 0;
  size_t bm_bytes = 0;
};
template <int T>
int piece_x_plan(const BmvArgs &a, hipStream_t s, PieceX *px) {
  using W = typename Tile<T>::W;
  DeviceFacts f;
  int st = device_facts(&f);
  if (st) return st;
  // the largest block column sizes x (the range of its buffer resource and of
  // the staged words): a wide matrix reads past ceil(NBR/T)*T words
  int maxcol = -1;
  const int mode = g_x_staging[ilog2_tiledim(T) - 2];
  if ((st = max_block_col(s, a.colind, a.nblocks, &maxcol))) return st;
  BGB_REQUIRE(maxcol >= 0 || a.nblocks == 0, BGB_ERR_ARG, "largest block column unknown (stream capture?)");
  const long long words = std::max<long long>(ceil_div(a.nbr, T) * T, (long long)maxcol + 1);
  const long long x_bytes = words * (long long)sizeof(W);
  BGB_REQUIRE(x_bytes < (1ll << 31), BGB_ERR_SIZE, "packed x of %lld bytes past a buffer resource's range",
              x_bytes);
  const long long xl = ((T == 4 ? (x_bytes + 1) / 2 : x_bytes) + 15) & ~15ll;  // LDS bytes of all of x
  const long long lds_max = f.lds_per_block & ~15ll;
  const bool aligned = (reinterpret_cast<uintptr_t>(a.B) & 15) == 0;
  PieceX p;
  p.words = words;
  p.x_bytes = x_bytes;
  if (mode != 1 && aligned && xl <= lds_max) {
    p.xl = 1;
    p.stage_words = words;
    p.lds = xl;
  } else if (mode != 1 && aligned) {
    // bitmap: prefix S words (S/2 LDS bytes at T = 4) + (words - S)/8 bytes
    // of bitmap within lds_max; S a multiple of 128 (16-byte prefix)
    const long long num = 8 * lds_max - words;
    long long S = num <= 0 ? -1 : ((T == 4 ? num / 3 : num / 7) & ~127ll);
    auto pre_bytes = [&](long long sw) { return ((T == 4 ? (sw + 1) / 2 : sw) + 15) & ~15ll; };
    auto bm_dw = [&](long long sw) { return (ceil_div(words - sw, 32) + 3) & ~3ll; };
    while (S >= 0 && pre_bytes(S) + 4 * bm_dw(S) > lds_max) S -= 128;  // the bitmap's padding
    if (S >= 0) {
      p.xl = 3;
      p.stage_words = S;
      p.bm_dwords = (int)bm_dw(S);
      p.bm_off = (int)pre_bytes(S);
      p.lds = p.bm_off + 4ll * p.bm_dwords;
      p.bm_bytes = (size_t)(4ll * p.bm_dwords);
    }
    if (p.xl == 0) {
      const long long part_words = T == 4 ? 2 * lds_max : lds_max / (long long)sizeof(W);
      if (mode == 2 || 8 * part_words >= words) {
        p.xl = 2;
        p.stage_words = part_words;
        p.lds = lds_max;
      }
    }
  }
  if (p.xl == 0) p.lds = 0;
  *px = p;
  return BGB_OK;
}

// The row-piece path's main pass: one persistent launch doing the short rows
// (one lane per block row) and then the pieces, the packed x placed as
// piece_x_plan decided (XL 3: the bitmap made first, into `bm`).
template <int T, int OP>
int launch_pieces(BmvArgs a, hipStream_t s, const PieceEntry &pe, const PieceX &px, uint32_t *bm) {
  using W = typename Tile<T>::W;
  DeviceFacts f;
  int st = device_facts(&f);
  if (st) return st;
  a.x_words = (int)px.words;
  a.xl_words = px.xl == 2 || px.xl == 3 ? (int)px.stage_words : 0;
  a.xbm = bm;
  a.bm_dwords = px.bm_dwords;
  a.bm_off = px.bm_off;
  if (px.xl == 3) {
    x_nonzero_bitmap_kernel<T><<<(unsigned)ceil_div((long long)px.bm_dwords * 2, 256), 256, 0, s>>>(
        static_cast<const uint8_t *>(a.B), a.xl_words, a.x_words, bm, px.bm_dwords);
    BGB_HIP_TRY(hipGetLastError());
  }
  void (*k)(BmvArgs, int, const int4 *, int, const int *, int) = px.xl == 1 ? &bmv_piece_kernel<T, OP, 1>
                                                          : px.xl == 2 ? &bmv_piece_kernel<T, OP, 2>
                                                          : px.xl == 3 ? &bmv_piece_kernel<T, OP, 3>
                                                                       : &bmv_piece_kernel<T, OP, 0>;
  if (px.lds > 0)
    BGB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)px.lds));
  int per_cu = 0;
  BGB_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k), 1024,
                                                           (size_t)px.lds));
  // short-row waves per workgroup of 16: with the short rows' share of the
  // tiles (a short-row tile costs a lane's dependent loads, about twice a
  // piece slot), at least 2 -- R-MAT 20 and 22 at T = 4: 2 of 16 (4: 56.9 /
  // 332 us, 2: 56.9 / 326 us, 6: 61 / 350 us)
  const double short_frac = a.nblocks ? (double)(a.nblocks - pe.long_tiles) / (double)a.nblocks : 1.0;
  const int ns = std::max(2, std::min(8, (int)std::ceil(32.0 * short_frac)));
  // workgroups: one per 16 - ns pieces or per ns blocks of 64 short-row
  // candidates, at most the occupancy
  const unsigned grid = (unsigned)std::min<long long>(
      (long long)f.cus * std::max(per_cu, 1),
      std::max<long long>(1, std::max(ceil_div(pe.npieces, 16 - ns), ceil_div(ceil_div(a.nbr, 64), ns))));
  const int stage_bytes = px.xl == 0 ? 0 : (int)(px.stage_words * (long long)sizeof(W));
  k<<<grid, 1024, (size_t)px.lds, s>>>(a, stage_bytes, pe.pieces, pe.npieces, pe.np, ns);
  BGB_HIP_TRY(hipGetLastError());
  return BGB_OK;
}

// scr: the static scratch of the allocation-free entry point (carry records
// and a partition made on every call, x gathered from memory), or null for
// the per-stream workspace, the cached partition and the automatic paths.
template <int T, int OP, int S>
int launch_bmv_s(BmvArgs a, hipStream_t s, uint8_t *scr) {
  constexpr int CH = (64 / Tile<T>::L) * S;
  using Acc = typename AccOf<OP>::type;
  int st = 0;
  // Count ops at T <= 8 take the row-piece path (not with the static scratch:
  // its plan is cached and made with a read-back).
  constexpr bool kPieceOp = (T == 4 || T == 8) && (OP == OP_COUNT_I32 || OP == OP_COUNT_F32);
  PieceEntry pe{};
  PieceX px;
  bool pieces = false;
  if constexpr (kPieceOp) {
    if (!scr && g_row_pieces != 1 && (long long)a.nblocks * 8 < (1ll << 31)) {
      // automatic: when the whole packed x fits in LDS, or its first words and
      // the non-zero bitmap of the others do (XL 3; with x partly in memory
      // and no bitmap the pieces' per-slot x reads cost more than the walk
      // saves: R-MAT 22, T = 4, 396 vs 385 us), and the pieces are well
      // filled (R-MAT 20, T = 8: 71 % full, pieces 74 us, chunk walk 83 us)
      if ((st = piece_x_plan<T>(a, s, &px))) return st;
      bool want = g_row_pieces == 2 || (g_x_staging[ilog2_tiledim(T) - 2] != 1 && (px.xl == 1 || px.xl == 3));
      if (want) {
        if ((st = piece_plan<T>(s, a.rowptr, a.nbr, a.nblocks, &pe))) return st;
        // tiles of the long rows per piece slot, and the short rows' share
        const double fill = (double)pe.long_tiles / ((double)std::max(pe.npieces, 1) * piece_tiles<T>());
        const double short_frac = a.nblocks ? 1.0 - (double)pe.long_tiles / (double)a.nblocks : 1.0;
        pieces = g_row_pieces == 2 || (fill >= 0.6 && short_frac <= 0.25);
      }
    }
  }
  const long long nchunks = pieces ? (long long)pe.npieces : ceil_div(a.nblocks, CH);
  BGB_REQUIRE(nchunks < INT_MAX / 2, BGB_ERR_SIZE, "too many chunks");
  a.nchunks = (int)nchunks;
  void *ws = nullptr;
  const size_t key_bytes = (sizeof(int) * (size_t)std::max(nchunks, 1ll) + 255) & ~(size_t)255;
  const size_t val_bytes = (sizeof(Acc) * (size_t)std::max(nchunks, 1ll) * T + 255) & ~(size_t)255;
  if (pieces) {
    const size_t bm_bytes = (px.bm_bytes + 255) & ~(size_t)255;
    if ((st = stream_workspace(s, key_bytes + val_bytes + bm_bytes, &ws))) return st;
    a.carry_key = static_cast<int *>(ws);
    a.carry_val = static_cast<char *>(ws) + key_bytes;
    uint32_t *bm = px.bm_bytes ? reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + key_bytes + val_bytes)
                               : nullptr;
    if constexpr (kPieceOp)
      if ((st = launch_pieces<T, OP>(a, s, pe, px, bm))) return st;
  } else if (scr) {
    const size_t need = 2 * key_bytes + val_bytes;  // keys, partition, records
    BGB_REQUIRE(need <= kScratchBytes, BGB_ERR_SIZE,
                "%lld tiles at T=%d need %zu bytes of scratch; the static scratch holds %zu", (long long)a.nblocks,
                T, need, kScratchBytes);
    a.carry_key = reinterpret_cast<int *>(scr);
    a.part = reinterpret_cast<int *>(scr + key_bytes);
    a.carry_val = scr + 2 * key_bytes;
    chunk_rows_kernel<<<(unsigned)ceil_div(a.nbr, 256), 256, 0, s>>>(a.rowptr, a.nbr, CH, (int)nchunks, a.part);
    BGB_HIP_TRY(hipGetLastError());
  } else {
    if ((st = stream_workspace(s, key_bytes + val_bytes, &ws))) return st;
    a.carry_key = static_cast<int *>(ws);
    a.carry_val = static_cast<char *>(ws) + key_bytes;
    if ((st = chunk_partition(s, a.rowptr, a.nbr, a.nblocks, CH, &a.part))) return st;
  }
  bool lds_path = pieces;
  if constexpr (!is_full(OP)) {
    if (!pieces) {
    // Packed x in LDS when the tile stream is large next to staging it once
    // per workgroup: all of it when it fits, else its first words. The words
    // cover every block column the matrix references (a wide matrix reads
    // past ceil(NBR/T)*T).
    DeviceFacts f;
    if ((st = device_facts(&f))) return st;
    int maxcol = -1;
    const int mode = g_x_staging[ilog2_tiledim(T) - 2];
    if (!scr && mode != 1 && (mode == 2 || T <= 16))
      if ((st = max_block_col(s, a.colind, a.nblocks, &maxcol))) return st;
    const long long words = std::max<long long>(ceil_div(a.nbr, T) * T, (long long)maxcol + 1);
    using W = typename Tile<T>::W;
    const long long x_bytes = words * (long long)sizeof(W);
    const long long xl = ((T == 4 ? (x_bytes + 1) / 2 : x_bytes) + 15) & ~15ll;  // LDS bytes of all of x
    constexpr int kLdsWaves = lds_waves<T, S>();
    const long long win_bytes = (long long)sizeof(int) * kLdsWaves * kWinAlloc;
    const bool fits = win_bytes + xl <= f.lds_per_block;
    // x larger than LDS: its first words are staged (as many as the LDS left
    // after the row windows holds, in whole 16-byte units), the others are
    // read from memory. R-MAT graphs reference their low block columns most
    // (the first quarter of the columns of R-MAT 22 holds about 55 % of the
    // tiles).
    const long long part_lds = (f.lds_per_block - win_bytes) & ~15ll;
    const long long part_words = T == 4 ? 2 * part_lds : part_lds / (long long)sizeof(W);
    // automatic: only while the staged words are at least an eighth of x
    // (uniformly spread columns gain little from a small prefix)
    const bool partial_ok =
        !fits && part_lds >= 16384 && x_bytes < (1ll << 31) && (mode == 2 || 8 * part_words >= words);
    const long long stage_lds = fits ? xl : part_lds;
    const long long tile_bytes = (long long)a.nblocks * (4 + T * (long long)sizeof(W));
    const bool aligned = (reinterpret_cast<uintptr_t>(a.B) & 15) == 0;
    // automatic: T <= 16 only (T = 32 streams 132 B per tile and is
    // bandwidth-bound; the persistent grid's lower occupancy costs more there
    // than the x gathers it saves, measured on MI355X)
    const bool worth = mode == 2 || (mode == 0 && T <= 16 && tile_bytes >= 2 * stage_lds * f.cus);
    if (maxcol >= 0 && aligned && (fits || partial_ok) && worth) {
      static bool attr_ok[2] = {
          hipFuncSetAttribute(reinterpret_cast<const void *>(&bmv_lds_kernel<T, OP, S, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)f.lds_per_block) == hipSuccess,
          hipFuncSetAttribute(reinterpret_cast<const void *>(&bmv_lds_kernel<T, OP, S, 2>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)f.lds_per_block) == hipSuccess};
      (void)hipGetLastError();
      if (attr_ok[fits ? 0 : 1]) {
        const long long lds = win_bytes + stage_lds;
        const int per_cu = std::max(1, std::min(2, f.lds_per_cu / (int)lds));
        const long long want = ceil_div(nchunks, kLdsWaves);
        const unsigned grid = (unsigned)std::min<long long>((long long)f.cus * per_cu, want);
        if (fits) {
          bmv_lds_kernel<T, OP, S, 1><<<grid, 64 * kLdsWaves, (size_t)lds, s>>>(a, (int)x_bytes);
        } else {
          a.xl_words = (int)part_words;
          a.x_words = (int)words;
          bmv_lds_kernel<T, OP, S, 2><<<grid, 64 * kLdsWaves, (size_t)lds, s>>>(
              a, (int)(part_words * (long long)sizeof(W)));
        }
        BGB_HIP_TRY(hipGetLastError());
        lds_path = true;
      }
    }
    }
  }
  if (!lds_path) {
    const unsigned blocks = (unsigned)ceil_div(nchunks, kWavesPerBlock);
    bmv_kernel<T, OP, S><<<blocks, 64 * kWavesPerBlock, 0, s>>>(a);
    BGB_HIP_TRY(hipGetLastError());
  }
  // bits: one thread per chunk, plus the tail workgroup; count / FP: one wave
  // per 64 chunks
  constexpr bool kBits = OP == OP_BITS || OP == OP_MASKED;
  const long long fix_blocks = kBits ? ceil_div(nchunks, 256) + 1 : ceil_div(ceil_div(nchunks, 64) * 64, 256);
  a.fixup_blocks = (int)fix_blocks;
  // overwriting calls also zero the empty block rows (accumulate leaves
  // them); the piece kernel zeroes its own
  const long long zero_blocks = ((a.accumulate && OP != OP_MASKED) || pieces)
                                    ? 0
                                    : std::min<long long>(ceil_div(a.nbr, 256), 1024);
  if (fix_blocks + zero_blocks == 0) return BGB_OK;  // no records (all rows short) and nothing to zero
  bmv_fixup_kernel<T, OP><<<(unsigned)(fix_blocks + zero_blocks), 256, 0, s>>>(a);
  BGB_HIP_TRY(hipGetLastError());
  return BGB_OK;
}

template <int T, int OP>
int launch_bmv(BmvArgs a, hipStream_t s, uint8_t *scr) {
  if constexpr (Tile<T>::L == 1) {
    int mr = -1;  // most tiles in a block row; -1: unknown (always with scr)
    if (scr) {
    } else if (g_short_rows == 2) {
      mr = 0;
    } else if (g_short_rows == 0) {
      const int st = max_row_tiles(s, a.rowptr, a.nbr, a.nblocks, &mr);
      if (st) return st;
    }
    if (mr >= 0 && mr <= kShortRowMax) {
      const unsigned grid = capped_blocks(a.nbr, 256);
      switch (short_row_batch(mr)) {
        case 4: bmv_rows_kernel<T, OP, 4><<<grid, 256, 0, s>>>(a); break;
        case 5: bmv_rows_kernel<T, OP, 5><<<grid, 256, 0, s>>>(a); break;
        case 6: bmv_rows_kernel<T, OP, 6><<<grid, 256, 0, s>>>(a); break;
        default: bmv_rows_kernel<T, OP, 8><<<grid, 256, 0, s>>>(a); break;
      }
      BGB_HIP_TRY(hipGetLastError());
      return BGB_OK;
    }
  }
  return g_run_length[ilog2_tiledim(T) - 2] == 16 ? launch_bmv_s<T, OP, 16>(a, s, scr)
                                                  : launch_bmv_s<T, OP, 8>(a, s, scr);
}

// noalloc: the allocation-free variant (static scratch, no caches, nothing
// allocated or synchronised).
template <int OP>
int bmv_entry(int T, const void *A, const void *B, void *C, const int *rowptr, const int *colind,
              int nbr, int nblocks, int accumulate, const void *mask, bgb_stream_t stream, bool noalloc = false) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
    BGB_REQUIRE(nbr >= 0 && nblocks >= 0, BGB_ERR_ARG, "negative size");
    BGB_REQUIRE(C && rowptr, BGB_ERR_ARG, "null output or rowptr");
    BGB_REQUIRE(nblocks == 0 || (A && B && colind), BGB_ERR_ARG, "null matrix / vector");
    BGB_REQUIRE(OP != OP_MASKED || mask, BGB_ERR_ARG, "null mask");
    BGB_REQUIRE(accumulate == 0 || accumulate == 1, BGB_ERR_ARG, "accumulate must be 0 or 1");
    BGB_REQUIRE((long long)nbr * T < INT_MAX, BGB_ERR_SIZE, "NBR*T exceeds the int32 range");
    auto aligned16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    BGB_REQUIRE(aligned16(A) && aligned16(colind), BGB_ERR_ARG,
                "bsrVal and bsrColInd must be 16-byte aligned");
    BGB_REQUIRE(OP == OP_BITS || OP == OP_MASKED || aligned16(C), BGB_ERR_ARG,
                "count / float outputs must be 16-byte aligned");
    BGB_REQUIRE((reinterpret_cast<uintptr_t>(C) & 3) == 0, BGB_ERR_ARG, "outputs must be 4-byte aligned");
    BGB_REQUIRE(T != 64 || !(OP == OP_BITS || OP == OP_MASKED) ||
                    ((reinterpret_cast<uintptr_t>(C) | reinterpret_cast<uintptr_t>(mask)) & 7) == 0,
                BGB_ERR_ARG, "T=64 bit outputs and masks must be 8-byte aligned");
    if (nbr == 0) return BGB_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t *scr = nullptr;
    if (noalloc) {
      int st = static_scratch(&scr);
      if (st) return st;
    } else {
      // a captured call would keep pointers to the per-stream workspace and
      // the cached partition, which later calls may free or reuse
      hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
      BGB_HIP_TRY(hipStreamIsCapturing(s, &cap));
      BGB_REQUIRE(cap == hipStreamCaptureStatusNone, BGB_ERR_ARG,
                  "this BMV entry point cannot run while its stream is being captured (its carry workspace and "
                  "chunk partition are library-cached and may be freed or reused by later calls); "
                  "bgb_bmv_bin_bin_full_i32_noalloc can");
    }
    if (nblocks == 0) {  // every block row is empty
      if (accumulate && OP != OP_MASKED) return BGB_OK;
      const long long n = (OP == OP_BITS || OP == OP_MASKED) ? (long long)nbr * word_bytes(T)
                                                             : (long long)nbr * T * 4;
      fill_kernel<uint8_t><<<capped_blocks(n, 256), 256, 0, s>>>(static_cast<uint8_t *>(C), n, 0);
      BGB_HIP_TRY(hipGetLastError());
      return BGB_OK;
    }
    BmvArgs a{A, B, C, rowptr, colind, mask, nbr, nblocks, accumulate, 0, nullptr, nullptr, nullptr, 0};
    switch (T) {
      case 4: return launch_bmv<4, OP>(a, s, scr);
      case 8: return launch_bmv<8, OP>(a, s, scr);
      case 16: return launch_bmv<16, OP>(a, s, scr);
      case 32: return launch_bmv<32, OP>(a, s, scr);
      default: return launch_bmv<64, OP>(a, s, scr);
    }
  });
}

// ---------------------------------------------------------------------------
// Per-(device, stream) scratch.
namespace {
struct Ws {
  void *ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, Ws> g_ws;
}  // namespace

int stream_workspace(hipStream_t s, size_t bytes, void **out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Ws &w = g_ws[{dev, s}];
  if (w.bytes < bytes) {
    if (w.ptr) {
      // the previous buffer may still be read by queued work on this stream
      BGB_HIP_TRY(hipStreamSynchronize(s));
      BGB_HIP_TRY(hipFree(w.ptr));
      w.ptr = nullptr;
      w.bytes = 0;
    }
    const size_t want = std::max(bytes + bytes / 4, (size_t)1 << 20);
    hipError_t e = hipMalloc(&w.ptr, want);
    if (e != hipSuccess) {
      w.ptr = nullptr;
      return set_error(BGB_ERR_ALLOC, "workspace hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    w.bytes = want;
  }
  *out = w.ptr;
  return BGB_OK;
}

namespace {
std::mutex g_facts_mu;
std::map<int, DeviceFacts> g_facts;
}  // namespace

int device_facts(DeviceFacts *out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_facts_mu);
  auto it = g_facts.find(dev);
  if (it == g_facts.end()) {
    DeviceFacts f;
    BGB_HIP_TRY(hipDeviceGetAttribute(&f.cus, hipDeviceAttributeMultiprocessorCount, dev));
    BGB_HIP_TRY(hipDeviceGetAttribute(&f.lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    BGB_HIP_TRY(hipDeviceGetAttribute(&f.lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    if (f.lds_per_cu < f.lds_per_block) f.lds_per_cu = f.lds_per_block;
    it = g_facts.emplace(dev, f).first;
  }
  *out = it->second;
  return BGB_OK;
}

// Most tiles in one block row, per (device, rowptr address, nbr, nblocks):
// one reduction pass and a read-back on the first call for a matrix (so the
// first call synchronises the stream), kept for the next. Unknown (-1) while
// the stream is being captured into a graph. A caller that rewrites rowptr in
// place keeps the old figure: it only selects a kernel, both are exact.
namespace {
struct RowStat {
  int dev;
  const int *rowptr;
  int nbr, nblocks, max_row;
};
constexpr int kRowStatCache = 16;
std::mutex g_rowstat_mu;
std::vector<RowStat> g_rowstat;  // most recently used last
}  // namespace

int max_row_tiles(hipStream_t s, const int *rowptr, int nbr, int nblocks, int *out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_rowstat_mu);
    for (size_t i = 0; i < g_rowstat.size(); ++i) {
      const RowStat e = g_rowstat[i];
      if (e.dev == dev && e.rowptr == rowptr && e.nbr == nbr && e.nblocks == nblocks) {
        g_rowstat.erase(g_rowstat.begin() + (long)i);
        g_rowstat.push_back(e);
        *out = e.max_row;
        return BGB_OK;
      }
    }
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  BGB_HIP_TRY(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) {
    *out = -1;
    return BGB_OK;
  }
  void *ws = nullptr;
  int st = stream_workspace(s, 256, &ws);
  if (st) return st;
  BGB_HIP_TRY(hipMemsetAsync(ws, 0, sizeof(int), s));
  max_row_tiles_kernel<<<std::min<unsigned>(capped_blocks(nbr, 256), kMaxRowBlocks), 256, 0, s>>>(
      rowptr, nbr, static_cast<int *>(ws));
  BGB_HIP_TRY(hipGetLastError());
  int h = 0;
  BGB_HIP_TRY(hipMemcpyAsync(&h, ws, sizeof(int), hipMemcpyDeviceToHost, s));
  BGB_HIP_TRY(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lk(g_rowstat_mu);
  if ((int)g_rowstat.size() >= kRowStatCache) g_rowstat.erase(g_rowstat.begin());
  g_rowstat.push_back(RowStat{dev, rowptr, nbr, nblocks, h});
  *out = h;
  return BGB_OK;
}

// Largest block column of a matrix, per (device, colind address, nblocks):
// one reduction and a read-back on the first call (the stream is
// synchronised), kept for the next; -1 while the stream is being captured.
// It sizes the packed x the LDS-staged kernels copy, so a caller that
// rewrites colind in place with larger block columns must gather x from
// memory (bgb_set_x_staging(T, 1)).
namespace {
struct ColStat {
  int dev;
  const int *colind;
  int nblocks, max_col;
};
std::mutex g_colstat_mu;
std::vector<ColStat> g_colstat;  // most recently used last
}  // namespace

int max_block_col(hipStream_t s, const int *colind, int nblocks, int *out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_colstat_mu);
    for (size_t i = 0; i < g_colstat.size(); ++i) {
      const ColStat e = g_colstat[i];
      if (e.dev == dev && e.colind == colind && e.nblocks == nblocks) {
        g_colstat.erase(g_colstat.begin() + (long)i);
        g_colstat.push_back(e);
        *out = e.max_col;
        return BGB_OK;
      }
    }
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  BGB_HIP_TRY(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone || nblocks == 0) {
    *out = nblocks == 0 ? 0 : -1;
    return BGB_OK;
  }
  // the stream's workspace holds the result: work queued on the stream
  // before this call is done with it by the time the kernel runs
  void *ws = nullptr;
  int st = stream_workspace(s, 256, &ws);
  if (st) return st;
  fill_kernel<int><<<1, 64, 0, s>>>(static_cast<int *>(ws), 1, 0);
  max_col_kernel<<<std::min<unsigned>(capped_blocks(nblocks, 256), kMaxRowBlocks), 256, 0, s>>>(
      colind, nblocks, static_cast<int *>(ws));
  BGB_HIP_TRY(hipGetLastError());
  int h = 0;
  BGB_HIP_TRY(hipMemcpyAsync(&h, ws, sizeof(int), hipMemcpyDeviceToHost, s));
  BGB_HIP_TRY(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lk(g_colstat_mu);
  if ((int)g_colstat.size() >= kRowStatCache) g_colstat.erase(g_colstat.begin());
  g_colstat.push_back(ColStat{dev, colind, nblocks, h});
  *out = h;
  return BGB_OK;
}

// Chunk partitions (chunk -> block row of its first tile), one per (device,
// stream, rowptr address, nbr, nblocks, chunk size), kept for the next call on
// the same matrix. The kernels check every entry against rowptr and repair
// it, so a caller that rewrites rowptr in place only costs a search.
namespace {
struct PartEntry {
  int dev;
  hipStream_t s;
  const int *rowptr;
  int nbr, nblocks, CH;
  int *part;
  size_t cap;  // ints
};
constexpr int kPartCache = 8;
std::mutex g_part_mu;
std::vector<PartEntry> g_part;  // most recently used last
}  // namespace

int chunk_partition(hipStream_t s, const int *rowptr, int nbr, int nblocks, int CH, int **out) {
  int dev = 0;
  BGB_HIP_TRY(hipGetDevice(&dev));
  const long long nchunks = ceil_div(nblocks, CH);
  std::lock_guard<std::mutex> lk(g_part_mu);
  for (size_t i = 0; i < g_part.size(); ++i) {
    PartEntry &e = g_part[i];
    if (e.dev == dev && e.s == s && e.rowptr == rowptr && e.nbr == nbr && e.nblocks == nblocks && e.CH == CH) {
      PartEntry hit = e;
      g_part.erase(g_part.begin() + (long)i);
      g_part.push_back(hit);
      *out = hit.part;
      return BGB_OK;
    }
  }
  PartEntry ne{dev, s, rowptr, nbr, nblocks, CH, nullptr, 0};
  // cache full: evict the least recently used entry. Queued work may still
  // read it; work on this stream is ordered before the new writes, any other
  // stream (possibly destroyed by now) is waited for through the device.
  if ((int)g_part.size() >= kPartCache) {
    PartEntry old = g_part.front();
    g_part.erase(g_part.begin());
    if (old.dev == dev && old.s == s && old.cap >= (size_t)nchunks) {
      ne.part = old.part;
      ne.cap = old.cap;
    } else {
      BGB_HIP_TRY(hipSetDevice(old.dev));
      BGB_HIP_TRY(hipDeviceSynchronize());
      BGB_HIP_TRY(hipFree(old.part));
      BGB_HIP_TRY(hipSetDevice(dev));
    }
  }
  if (!ne.part) {
    ne.cap = (size_t)std::max<long long>(nchunks, 1);
    hipError_t e = hipMalloc(&ne.part, sizeof(int) * ne.cap);
    if (e != hipSuccess)
      return set_error(BGB_ERR_ALLOC, "partition hipMalloc(%zu) failed: %s", sizeof(int) * ne.cap,
                       hipGetErrorString(e));
  }
  chunk_rows_kernel<<<(unsigned)ceil_div(nbr, 256), 256, 0, s>>>(rowptr, nbr, CH, (int)nchunks, ne.part);
  BGB_HIP_TRY(hipGetLastError());
  g_part.push_back(ne);
  *out = ne.part;
  return BGB_OK;
}

}  // namespace bgb

using namespace bgb;

extern "C" {

int bgb_set_run_length(int T, int S) {
  clear_error();
  BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
  BGB_REQUIRE(S == 8 || S == 16, BGB_ERR_ARG, "run length must be 8 or 16 (got %d)", S);
  g_run_length[ilog2_tiledim(T) - 2] = S;
  return BGB_OK;
}

int bgb_get_run_length(int T) { return valid_tiledim(T) ? g_run_length[ilog2_tiledim(T) - 2] : -1; }

int bgb_set_row_pieces(int mode) {
  clear_error();
  BGB_REQUIRE(mode >= 0 && mode <= 2, BGB_ERR_ARG, "row-piece mode must be 0, 1 or 2 (got %d)", mode);
  g_row_pieces = mode;
  return BGB_OK;
}

int bgb_get_row_pieces(void) { return g_row_pieces; }

int bgb_set_x_staging(int T, int mode) {
  clear_error();
  BGB_REQUIRE(valid_tiledim(T), BGB_ERR_TILEDIM, "tile dim %d not in {4,8,16,32,64}", T);
  BGB_REQUIRE(mode >= 0 && mode <= 2, BGB_ERR_ARG, "x staging mode must be 0, 1 or 2 (got %d)", mode);
  g_x_staging[ilog2_tiledim(T) - 2] = mode;
  return BGB_OK;
}

int bgb_get_x_staging(int T) { return valid_tiledim(T) ? g_x_staging[ilog2_tiledim(T) - 2] : -1; }

int bgb_set_short_row_path(int mode) {
  clear_error();
  BGB_REQUIRE(mode >= 0 && mode <= 2, BGB_ERR_ARG, "short-row path mode must be 0, 1 or 2 (got %d)", mode);
  g_short_rows = mode;
  return BGB_OK;
}

int bgb_get_short_row_path(void) { return g_short_rows; }

int bgb_bmv_bin_bin_bin(int T, const void *d_A, const void *d_B, void *d_C, const int *d_rowptr,
                        const int *d_colind, int nblockrows, int nblocks, int accumulate,
                        bgb_stream_t stream) {
  return bmv_entry<OP_BITS>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks, accumulate,
                            nullptr, stream);
}

int bgb_bmv_bin_bin_full_i32(int T, const void *d_A, const void *d_B, int *d_C, const int *d_rowptr,
                             const int *d_colind, int nblockrows, int nblocks, int accumulate,
                             bgb_stream_t stream) {
  return bmv_entry<OP_COUNT_I32>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks,
                                 accumulate, nullptr, stream);
}

int bgb_bmv_bin_bin_full_f32(int T, const void *d_A, const void *d_B, float *d_C,
                             const int *d_rowptr, const int *d_colind, int nblockrows, int nblocks,
                             int accumulate, bgb_stream_t stream) {
  return bmv_entry<OP_COUNT_F32>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks,
                                 accumulate, nullptr, stream);
}

int bgb_bmv_bin_bin_full_i32_noalloc(int T, const void *d_A, const void *d_B, int *d_C, const int *d_rowptr,
                                     const int *d_colind, int nblockrows, int nblocks, int accumulate,
                                     bgb_stream_t stream) {
  return bmv_entry<OP_COUNT_I32>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks, accumulate,
                                 nullptr, stream, /*noalloc=*/true);
}

int bgb_bmv_bin_full_full(int T, const void *d_A, const float *d_B, float *d_C, const int *d_rowptr,
                          const int *d_colind, int nblockrows, int nblocks, int accumulate,
                          bgb_stream_t stream) {
  return bmv_entry<OP_FULL_F32>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks,
                                accumulate, nullptr, stream);
}

int bgb_bmv_bin_full_full_bf16(int T, const void *d_A, const uint16_t *d_B, float *d_C,
                               const int *d_rowptr, const int *d_colind, int nblockrows,
                               int nblocks, int accumulate, bgb_stream_t stream) {
  return bmv_entry<OP_FULL_BF16>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks,
                                 accumulate, nullptr, stream);
}

int bgb_bmv_bin_bin_bin_masked(int T, const void *d_A, const void *d_B, void *d_C,
                               const int *d_rowptr, const int *d_colind, int nblockrows,
                               int nblocks, const void *d_mask, bgb_stream_t stream) {
  return bmv_entry<OP_MASKED>(T, d_A, d_B, d_C, d_rowptr, d_colind, nblockrows, nblocks, 0, d_mask,
                              stream);
}

}  // extern "C"
