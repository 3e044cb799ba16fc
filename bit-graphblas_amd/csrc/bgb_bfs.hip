// Breadth-first search on a symmetric B2SR matrix (gfx950).
//
// The reference has no BFS; its building block is the masked bin_bin_bin BMV
// C[I] = (A x != 0) & ~mask[I] (bsrbmv.cu:1086-1229), i.e. one BFS level of
// one source. Two drivers are built here:
//
// * bgb_bfs_plain -- the per-source loop: one masked BMV per level, then an
//   update kernel (visited |= next, level[v] = d) and a host check of the
//   "any new vertex" flag.
// * bgb_bfs_multi -- up to 64 sources at once: vertex v keeps one uint64
//   frontier word F[v] (bit s = v is on source s's frontier) and one visited
//   word V[v]. A level is one pull pass over the tiles:
//       N[v] = OR over neighbours u with F[u] != 0 of F[u]
//       new  = N & ~V[v];  V[v] |= new;  Fnext[v] = new
//   Only frontier neighbours are gathered: the tile row word is first ANDed
//   with a packed "frontier non-empty" bit vector (the B2SR vector layout),
//   and rows whose 64 sources are all visited skip their gathers. The level
//   of (s, v) is kept as bit-planes P[k][v] (bit s of plane k = bit k of the
//   level), expanded to level[s*n + v] at the end. The level loop (two levels
//   per graph, so the double buffers alternate) is captured in a hipGraph
//   and replayed until a level finds nothing; device-side flags turn the
//   levels after the last one into no-ops.
//   Direction: a pull pass streams every tile, whatever the frontier size.
//   Given the B2SR of A^T (A's own arrays when A is symmetric), a level whose
//   frontier is cheap to push is a push instead: one wave per block row of
//   A^T holding frontier vertices walks its tiles once per frontier row --
//   the vertices v with A[v][u] = 1, which pull from u -- and ORs F[u] into
//   the unvisited ones with atomics. The push cost of a frontier is the tiles
//   of the A^T block rows of its vertices (one walk per vertex); each level
//   sums it for the vertices it adds, and the end-of-level kernel pushes the
//   next level when that sum is below nblocks / kPushDiv (a pull streams all
//   nblocks tiles; a pushed tile costs far more: a random atomic per
//   neighbour). The first level (the sources) is a push.
// The tile walk, group scan and chunk carries are those of the BMV kernels,
// with OR as the combine operator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "bgb_device.h"
#include "bgb_tile.h"

namespace bgb {
namespace {

constexpr int kBfsWaves = 4;
constexpr int kBfsRun = 8;
constexpr int kPlanes = 16;  // level < 2^16
constexpr int kPushDiv = 16;    // push while the next frontier's push cost < nblocks / kPushDiv tiles
constexpr int kCounters = 256;  // push-cost counters, spread to avoid one hot address
constexpr int kPushSeg = 2048;  // push: frontier rows of longer block rows are split into segments
constexpr int kBigCap = 1 << 16;  // push: (block row, row) entries of the long-row list

__device__ __forceinline__ long long lmin2(long long a, long long b) { return a < b ? a : b; }

struct BfsState {
  const void *A;
  const int *rowptr;
  const int *colind;
  int nbr, nblocks, nchunks;
  const uint64_t *F;     // frontier of this level
  uint64_t *Fn;          // frontier of the next level
  const void *fnz;       // packed F != 0 (W_T words, >= ceil(NBR/T)*T)
  void *fnzn;            // packed Fn != 0
  uint64_t *V;           // visited
  uint64_t *P;           // kPlanes bit-planes, stride npad
  long long npad;        // NBR * T
  uint64_t full;         // all source bits
  int *ctl;              // [0] depth of this level, [1] any-new flag, [2] done, [3] push level,
                         // [5] some chunk of this pull level is skipped
  unsigned long long *fcnt;  // kCounters counters of the next frontier's push cost (tiles)
  int *carry_key;
  uint64_t *carry_val;
  const int *part;       // part[c]: block row of chunk c's first tile
  const void *At;        // B2SR of A^T for push levels (null: pull only)
  const int *rowptrT;
  const int *colindT;
  int *big_n;            // push: entries in big_list
  int2 *big_list;        // push: frontier rows (I, r) of block rows longer than kPushSeg tiles
  uint64_t *live;        // sources whose frontier F is non-empty (this level's input)
  unsigned long long *live_acc;  // kCounters words: OR of this level's new bits, spread
  uint8_t *sat;          // pull levels: block row I saturated (every vertex reached by every live source)
  uint8_t *skip;         // pull levels: every block row of chunk c saturated
};

// ORs a wave's new bits into one of the kCounters spread words (no returned
// value, so the wave does not wait for it); the end-of-level kernel folds
// them into the next level's live mask.
__device__ __forceinline__ void publish_live(const BfsState &b, uint64_t ln, int lane, long long slot) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) ln |= __shfl_xor(ln, m);
  if (lane == 0 && ln) atomicOr(b.live_acc + (slot & (kCounters - 1)), (unsigned long long)ln);
}

// Records the vertices first reached at this level (new bits) for one row.
// The level planes are OR-ed with atomics whose result is not used: a plain
// read-modify-write would make the wave wait for the read.
__device__ __forceinline__ bool record_new(const BfsState &b, long long v, uint64_t vold, uint64_t nb,
                                           int depth) {
  // Fn starts the level cleared (bfs_clear_kernel): only new bits are stored
  if (nb) {
    b.Fn[v] = nb;
    b.V[v] = vold | nb;
    // one atomic per set bit of the depth (a loop, not 16 unrolled sites:
    // this runs at every row end of the unrolled walk, and the unrolled form
    // made the level kernel larger than the instruction cache)
#pragma unroll 1
    for (int m = depth; m; m &= m - 1)
      atomicOr(reinterpret_cast<unsigned long long *>(b.P + (__ffs(m) - 1) * b.npad + v), nb);
    return true;
  }
  return false;
}

// New bits of block row I (lanes q of its tile group): acc & ~visited.
// `row_len(I)` is the row's tile count in A^T (its push cost per vertex).
template <int T, class RL>
__device__ __forceinline__ void bfs_flush(const BfsState &b, int I, int q, const uint64_t (&acc)[Tile<T>::R],
                                          const uint64_t (&vr)[Tile<T>::R], int depth, bool &any,
                                          unsigned long long &cost, uint64_t &lnew, RL &&row_len) {
  constexpr int R = Tile<T>::R, L = Tile<T>::L;
  using W = typename Tile<T>::W;
  uint32_t bits = 0;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const long long v = (long long)I * T + q * R + i;
    const uint64_t nb = acc[i] & ~vr[i];
    lnew |= nb;
    if (record_new(b, v, vr[i], nb, depth)) {
      any = true;
      bits |= 1u << (T - 1 - (q * R + i));
    }
  }
  // push cost of the rows added: one walk of A^T's block row I each
  if (b.rowptrT && bits) cost += (unsigned long long)__popc(bits) * row_len(I);
#pragma unroll
  for (int m = 1; m < L; m <<= 1) bits |= (uint32_t)__shfl_xor((int)bits, m);
  if (q == 0 && bits) static_cast<W *>(b.fnzn)[I] = (W)bits;
}

// Row words of a tile restricted to the columns whose frontier word is
// non-empty (x: the packed frontier-non-empty word of the tile's block column,
// bit T-1-c for column c), in place: every row slot of the tile word ANDed
// with x.
template <int T>
__device__ __forceinline__ void mask_cols(uint32_t (&F)[Tile<T>::NW], uint32_t x) {
  constexpr int NW = Tile<T>::NW;
  const uint32_t rep = T == 4 || T == 8 ? x * 0x01010101u : (T == 16 ? x * 0x00010001u : x);
#pragma unroll
  for (int j = 0; j < NW; ++j) F[j] &= rep;
}

// Visited words of a chunk's vertices staged in LDS, per wave: the walk reads
// one row's words at every row change, and from memory each read would be a
// round trip the next tile waits for.
constexpr int kVisStage = 512;

// One chunk of the pull pass for one wave: `win` and `sv` are the wave's LDS
// row window and visited-word stage.
// kG: 0 gathers the frontier word of each set bit of each row in turn, the
// loop stopping as soon as the row is saturated; 1 (T <= 8, one lane per
// tile) gathers once per set column of the tile and ORs the word into every
// row holding that column -- one gather instruction per column instead of
// one per row, with more lanes active in each (the gathers are bound by the
// address units, not by bytes).
template <int T, int S>
struct BfsIn {
  uint32_t F[S][Tile<T>::NW];
  int J[S];
  int wv[kWin / 64];
  int Ia, Ib;  // block rows of the chunk's first tile and of the next chunk's
};

// Issues the loads of chunk `chunk` (rows Ia .. Ib, wave-uniform): its tile
// runs, block columns and row-window entries. kFull: every lane group has a
// whole run of S tiles (the loads are then unconditional, which keeps the
// compiler's count of loads in flight exact across the pipelined walk).
template <int T, int S, bool kFull>
__device__ __forceinline__ void bfs_load(const BfsState &b, int chunk, int lane, int Ia, int Ib,
                                         BfsIn<T, S> &in) {
  constexpr int L = Tile<T>::L, G = 64 / L, CH = G * S;
  const int g = lane / L, q = lane % L;
  const long long k0 = (long long)chunk * CH;
  const long long pa = lmin2(k0 + (long long)g * S, b.nblocks);
  const long long pb = lmin2(pa + S, b.nblocks);
  if (kFull) load_run<T, S, true>(b.A, b.colind, pa, S, q, in.F, in.J);
  else if (pa < pb) load_run<T, S>(b.A, b.colind, pa, (int)(pb - pa), q, in.F, in.J);
  RowWindow::load(b.rowptr, Ia, Ib, lane, in.wv);
  in.Ia = Ia;
  in.Ib = Ib;
}

template <int T, int S, int kG>
__device__ __forceinline__ void bfs_chunk(const BfsState &b, int chunk, int lane, int depth, int *win,
                                          uint64_t *sv, const BfsIn<T, S> &in) {
  constexpr int L = Tile<T>::L, R = Tile<T>::R, NW = Tile<T>::NW, G = 64 / L, CH = G * S;
  using W = typename Tile<T>::W;
  const int g = lane / L, q = lane % L;
  const long long k0 = (long long)chunk * CH;
  const long long pa = lmin2(k0 + (long long)g * S, b.nblocks);
  const long long pb = lmin2(pa + S, b.nblocks);
  const bool active = pa < pb;
  const int *__restrict__ rowptr = b.rowptr;
  uint32_t F[S][NW];
  int J[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    J[s] = in.J[s];
#pragma unroll
    for (int w = 0; w < NW; ++w) F[s][w] = in.F[s][w];
  }
  const int n = (int)(pb - pa);
  const int Ia = in.Ia, Ib = in.Ib;
  const uint64_t dead = ~*b.live;  // sources whose BFS has no frontier left
  RowWindow rw;
  // Only columns with a non-empty frontier word are gathered: the tile words
  // are masked by the packed frontier-non-empty words of their block columns
  // (gathered unconditionally -- block column 0 for missing tiles, whose
  // words are then cleared -- so that the loads issue together).
  {
    uint32_t X[S];
#pragma unroll
    for (int s = 0; s < S; ++s) X[s] = (uint32_t) static_cast<const W *>(b.fnz)[(active && s < n) ? J[s] : 0];
#pragma unroll
    for (int s = 0; s < S; ++s) mask_cols<T>(F[s], (active && s < n) ? X[s] : 0u);
  }
  // visited words of vertices [Ia*T, (Ib+1)*T), when they fit the wave's
  // stage (and the row window is staged): unconditional loads, indices
  // clamped into the range
  const long long vbase = (long long)Ia * T;
  const int vcount = (Ib - Ia + 1) * T;
  const bool vst = vcount <= kVisStage && Ib + 2 - Ia <= kWin;  // wave-uniform
  if (vst) {
    // as many 64-vertex loads as the chunk's vertices need (most chunks lie
    // in one or a few block rows: one load instead of kVisStage / 64)
    auto stage = [&](auto nl) {
      constexpr int NL = decltype(nl)::value;
      uint64_t vt[NL];
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        const int i = lane + 64 * j;
        vt[j] = b.V[vbase + (i < vcount ? i : vcount - 1)];
      }
#pragma unroll
      for (int j = 0; j < NL; ++j) sv[lane + 64 * j] = vt[j];
    };
    if (vcount <= 64) stage(std::integral_constant<int, 1>{});
    else stage(std::integral_constant<int, kVisStage / 64>{});
  }
  const int(&wv)[kWin / 64] = in.wv;
  rw.put(rowptr, win, Ia, Ib, lane, wv);  // (its wave barrier also covers the stage)
  const bool sym = b.rowptrT == rowptr;  // A^T's block rows are A's: lengths from the window
  // The walk, compiled for: row ends and visited words in LDS (the common
  // case), row ends in LDS and visited words in memory, both in memory.
  auto walk = [&](auto st, auto vs) {
  constexpr bool ST = decltype(st)::value, VS = decltype(vs)::value;
  auto load_vis = [&](int I, uint64_t(&vr)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const long long v = (long long)I * T + q * R + i;
      if constexpr (VS) vr[i] = sv[v - vbase];
      else vr[i] = b.V[v];
    }
  };
  auto row_len = [&](int I) -> int {
    if constexpr (ST) {
      if (sym) return rw.at_t<true>(I + 1) - rw.at_t<true>(I);
    }
    return b.rowptrT[I + 1] - b.rowptrT[I];
  };
  int I;
  if constexpr (ST) I = window_group_rows<L, S>(rw, k0, lane, wv);
  else I = rw.find_t<false>(pa);
  I = active ? I : Ia;
  const int I0 = I;

  bool any = false;
  unsigned long long nnew = 0;  // push cost of the rows this wave adds
  uint64_t lnew = 0;            // sources with a new bit in this wave's rows
  uint64_t acc[R], vr[R], hv[R];
#pragma unroll
  for (int i = 0; i < R; ++i) acc[i] = hv[i] = 0;
  if (active) load_vis(I, vr);
  // kG = 1 (one lane per tile): byte r of om[j] is 0xFF while row 4j+r of the
  // open block row is not saturated -- recomputed only when the partial or
  // the visited words change (a gather, a new row), so a tile's column set is
  // a few bitwise ops
  constexpr int kOM = (kG == 1 && L == 1) ? NW : 1;
  uint32_t om[kOM];
  auto update_open = [&]() {
    if constexpr (kG == 1 && L == 1) {
#pragma unroll
      for (int j = 0; j < NW; ++j) om[j] = 0;
#pragma unroll
      for (int i = 0; i < R; ++i)
        if ((acc[i] | vr[i] | dead) != ~0ull) om[i >> 2] |= 0xFFu << (8 * (i & 3));
    }
  };
  update_open();
  int held = -1;
  // end of row I and, read one row ahead, of row I+1 (indices clamped into the
  // window / rowptr; clamped ends belong to rows past the chunk)
  const int hi = ST ? rw.Ib + 1 : b.nbr;
  auto row_end = [&](int r) { return r < b.nbr ? rw.at_t<ST>(r + 1 < hi ? r + 1 : hi) : INT_MAX; };
  int e = active ? row_end(I) : INT_MAX;
  int e2 = active ? row_end(I + 1) : INT_MAX;
  // kG = 1: the frontier word of each tile's highest set column is gathered
  // for a group of kPre tiles at once, before the group is walked (through a
  // buffer resource over F, with an offset past its end -- no memory access
  // -- for tiles without a frontier column): one round trip per group
  // instead of one per tile position. Further columns of a tile (few) are
  // gathered in the walk.
  constexpr int kPre = (kG == 1 && L == 1) ? 4 : 1;
  __amdgpu_buffer_rsrc_t frs;
  int f_end = 0;
  if constexpr (kG == 1 && L == 1) {
    f_end = (int)(b.npad * 8);
    frs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t *>(b.F), (short)0, f_end, 0x00020000);
  }
  // column set of tile s: its frontier columns in rows not saturated yet
  // (om describes the open row; a tile past its end starts a new row)
  auto tile_cols = [&](const uint32_t (&Fw)[NW], bool in_open_row) -> uint32_t {
    uint32_t cw = 0;
#pragma unroll
    for (int j = 0; j < (kG == 1 && L == 1 ? NW : 1); ++j) cw |= Fw[j] & (in_open_row ? om[j] : ~0u);
    cw |= cw >> 16;
    cw |= cw >> 8;
    return cw & ((1u << (T < 32 ? T : 31)) - 1u);
  };
  // The tiles in a loop over groups of kPre, not unrolled: the fully
  // unrolled walk, with a row flush at every tile, was several times the
  // instruction cache.
  static_assert(S % kPre == 0, "groups must tile the run");
#pragma unroll 1
  for (int s0 = 0; s0 < S; s0 += kPre) {
    // this group's tile words and block columns, picked with static indices
    // (a dynamic index into the run's arrays sends them to scratch)
    uint32_t Fq[kPre][NW];
    int Jq[kPre];
#pragma unroll
    for (int g2 = 0; g2 < kPre; ++g2) {
      if constexpr (kPre == 1) {
        Jq[g2] = J[s0];
#pragma unroll
        for (int w = 0; w < NW; ++w) Fq[g2][w] = F[s0][w];
      } else {
        Jq[g2] = J[g2];
#pragma unroll
        for (int w = 0; w < NW; ++w) Fq[g2][w] = F[g2][w];
#pragma unroll
        for (int k = 1; k < S / kPre; ++k)
          if (s0 == k * kPre) {
            Jq[g2] = J[k * kPre + g2];
#pragma unroll
            for (int w = 0; w < NW; ++w) Fq[g2][w] = F[k * kPre + g2][w];
          }
      }
    }
    uint64_t fg[kPre];
    int fh[kPre];
    if constexpr (kG == 1 && L == 1) {
#pragma unroll
      for (int g2 = 0; g2 < kPre; ++g2) {
        const int t = s0 + g2;
        const uint32_t cols = t < n ? tile_cols(Fq[g2], pa + t < (long long)e) : 0u;
        const int h = cols ? 31 - __clz(cols) : -1;
        const int off = cols ? (Jq[g2] * T + (T - 1 - h)) * 8 : f_end;
        fg[g2] = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(frs, off, 0, 0));
        fh[g2] = h;
      }
    }
#pragma unroll
  for (int g2 = 0; g2 < kPre; ++g2) {
    const int s = s0 + g2;
    if (s < n) {
      const uint64_t *__restrict__ fr = b.F + (long long)Jq[g2] * T;
      if constexpr (kG == 1 && L == 1) {
        // One gather per set column of the tile (the union of its rows that
        // are not saturated yet), OR-ed into every row holding the column;
        // the highest column's word was gathered with the group.
        uint32_t cols = tile_cols(Fq[g2], true);
        if (cols) {
          do {
            const int h = 31 - __clz(cols);
            cols &= ~(1u << h);
            const uint64_t f = h == fh[g2] ? fg[g2] : fr[T - 1 - h];
#pragma unroll
            for (int i = 0; i < R; ++i) acc[i] |= ((row_word<T>(Fq[g2], i) >> h) & 1u) ? f : 0ull;
          } while (cols);
          update_open();
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          uint32_t w = row_word<T>(Fq[g2], i);
          // Stop once the partial covers every live source that has not
          // reached the vertex yet: frontier words only hold live sources, so
          // more ORs cannot change acc & ~visited (this includes vertices every
          // live source has reached).
          while (w && (acc[i] | vr[i] | dead) != ~0ull) {
            const int h = 31 - __clz(w);
            acc[i] |= fr[T - 1 - h];
            w &= ~(1u << h);
          }
        }
      }
      const int pos = (int)(pa + s) + 1;
      if (pos == e) {
        if (held < 0) {
          held = I;  // its visited words are read again at the end
#pragma unroll
          for (int i = 0; i < R; ++i) hv[i] = acc[i];
        } else {
          bfs_flush<T>(b, I, q, acc, vr, depth, any, nnew, lnew, row_len);
        }
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i] = 0;
        ++I;
        e = e2;
        e2 = row_end(I + 1);
        if (e == pos) {
          // block rows without tiles follow (their words were cleared by the
          // level's first pass): step to the first row ending after pos, one
          // row end per step in the LDS window (at most kWin rows), by a
          // binary search over the row ends in memory
          if constexpr (ST) {
            do {
              ++I;
              e = e2;
              e2 = row_end(I + 1);
            } while (e == pos);
          } else {
            int lo = I + 1, up = b.nbr;
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
        if (I < b.nbr) load_vis(I, vr);
        update_open();
      }
    }
  }
  }
  const int key = active ? I : INT_MAX;
  uint64_t sc[R];
#pragma unroll
  for (int i = 0; i < R; ++i) sc[i] = acc[i];
  group_scan<L, R, uint64_t>(key, sc, [](uint64_t x, uint64_t y) { return x | y; });
  const int kin = group_up<L>(key);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint64_t up = group_up<L, uint64_t>(sc[i]);
    if (g > 0 && kin == I0) hv[i] |= up;
  }
  if (held >= 0) {
    uint64_t hvr[R];
    load_vis(held, hvr);
    bfs_flush<T>(b, held, q, hv, hvr, depth, any, nnew, lnew, row_len);
  }
  const int last = (int)lmin2(G - 1, (b.nblocks - 1 - k0) / S);
  if (g == last) {
    const bool open = key < b.nbr;
    if (q == 0) b.carry_key[chunk] = open ? key : -1;
    if (open) {
#pragma unroll
      for (int i = 0; i < R; ++i) b.carry_val[(long long)chunk * T + q * R + i] = sc[i];
    }
  }
  if (__ballot(any) != 0 && lane == 0) b.ctl[1] = 1;
  publish_live(b, lnew, lane, chunk);
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) nnew += __shfl_xor(nnew, m);
  if (lane == 0 && nnew) atomicAdd(b.fcnt + (chunk & (kCounters - 1)), nnew);
  };
  if (!__builtin_amdgcn_readfirstlane((int)rw.staged)) walk(std::false_type{}, std::false_type{});
  else if (__builtin_amdgcn_readfirstlane((int)vst)) walk(std::true_type{}, std::true_type{});
  else walk(std::true_type{}, std::false_type{});
}

// The pull pass: a persistent grid whose waves walk chunks w, w + nw, ...
// (nw waves in the grid). A one-chunk-per-wave launch left most of the
// machine idle: a workgroup's slot stays taken until its slowest wave ends,
// and chunk times range from under a microsecond (skipped) to tens of
// microseconds (hundreds of row changes). Chunks whose block rows are all
// saturated (bfs_chunk_skip_kernel, levels where ctl[5] is set) are neither
// loaded nor walked. Push levels and the levels after the last one return at
// once. (Loading the wave's next chunk while walking this one, as the BMV's
// persistent kernels do, measured no faster here: the extra registers cost a
// wave per SIMD.)
// T = 4 is held to 96 VGPRs (5 waves per SIMD).
template <int T, int S, int kG>
__global__ __launch_bounds__(64 * kBfsWaves) __attribute__((amdgpu_waves_per_eu(T == 4 ? (kG == 1 ? 4 : 5) : 1)))
void bfs_level_kernel(BfsState b) {
  __shared__ int s_win[kBfsWaves][kWinAlloc];
  __shared__ uint64_t s_vis[kBfsWaves][kVisStage];
  if (b.ctl[2] || b.ctl[3]) return;  // BFS finished, or a push level: no-op
  const int depth = b.ctl[0];
  const bool skipping = b.ctl[5] != 0;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = __builtin_amdgcn_readfirstlane((int)gridDim.x * kBfsWaves);
  for (int chunk = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kBfsWaves + wid); chunk < b.nchunks;
       chunk += nw) {
    // every block row of the chunk saturated: nothing can be new here and the
    // rows' next-frontier words are already cleared; tested before the tiles
    // are loaded, so a skipped chunk reads none of them
    if (skipping && b.skip[chunk]) {
      if (lane == 0) b.carry_key[chunk] = -1;
      continue;
    }
    const int Ia = b.part[chunk];
    const int Ib = chunk + 1 < b.nchunks ? b.part[chunk + 1] : b.nbr - 1;
    BfsIn<T, S> in;
    bfs_load<T, S, false>(b, chunk, lane, Ia, Ib, in);
    bfs_chunk<T, S, kG>(b, chunk, lane, depth, s_win[wid], s_vis[wid], in);
  }
}

// Folds the carries of rows that cross chunks into this level's results. One
// wave per 64 consecutive chunks, lane j owning chunk 64w + j: carry keys are
// non-decreasing over chunks (none only past the last block row), so the
// chunks carrying into one row are contiguous; a segmented OR-scan over the
// lanes joins each run and only the run's last lane records the new bits
// (atomics, since a run can continue into a neighbouring wave). One thread
// per record instead sent a hub row's hundreds of records to the same
// addresses: 190-240 us on the busiest levels of R-MAT 22.
template <int T>
__global__ __launch_bounds__(256) void bfs_fixup_kernel(BfsState b) {
  using W = typename Tile<T>::W;
  if (b.ctl[2] || b.ctl[3]) return;
  const int lane = threadIdx.x & 63;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int key = c < b.nchunks ? b.carry_key[c] : -1;
  key = key < 0 ? INT_MAX : key;  // no record: past every real key
  const int next = __shfl(key, lane + 1 < 64 ? lane + 1 : 63);
  const bool last_of_run = lane == 63 || next != key;
  const int depth = b.ctl[0];
  bool any = false;
  uint64_t lnew = 0;
  for (int r = 0; r < T; ++r) {
    uint64_t sc[1] = {key != INT_MAX ? b.carry_val[c * T + r] : 0ull};
    group_scan<1, 1, uint64_t>(key, sc, [](uint64_t x, uint64_t y) { return x | y; });
    if (!last_of_run || key == INT_MAX || !sc[0]) continue;
    const long long v = (long long)key * T + r;
    const uint64_t nb = sc[0] & ~b.V[v];
    if (!nb) continue;
    lnew |= nb;
    const unsigned long long fold = atomicOr(reinterpret_cast<unsigned long long *>(b.Fn + v), (unsigned long long)nb);
    if (!fold && b.rowptrT)
      atomicAdd(b.fcnt + (c & (kCounters - 1)), (unsigned long long)(b.rowptrT[key + 1] - b.rowptrT[key]));
    atomicOr(reinterpret_cast<unsigned long long *>(b.V + v), (unsigned long long)nb);
    for (int k = 0; k < kPlanes; ++k)
      if ((depth >> k) & 1) atomicOr(reinterpret_cast<unsigned long long *>(b.P + k * b.npad + v), (unsigned long long)nb);
    const uintptr_t addr = reinterpret_cast<uintptr_t>(static_cast<W *>(b.fnzn) + key);
    atomicOr(reinterpret_cast<unsigned int *>(addr & ~(uintptr_t)3), (1u << (T - 1 - r)) << ((int)(addr & 3) * 8));
    any = true;
  }
  if (any) b.ctl[1] = 1;
  publish_live(b, lnew, lane, c >> 6);
}

// End of a level: stop when nothing new was found, else advance the depth and
// pick the next level's direction from the push cost of the vertices this
// level added (push below push_limit). One block of kCounters threads.
// The sources whose BFS found new vertices (the OR of the spread live words)
// are the next level's live mask.
__global__ void bfs_advance_kernel(int *ctl, unsigned long long *fcnt, unsigned long long push_limit,
                                   unsigned long long *live_acc, uint64_t *live) {
  __shared__ unsigned long long part[kCounters], lor[kCounters];
  const int t = threadIdx.x;
  part[t] = fcnt[t];
  fcnt[t] = 0;
  lor[t] = live_acc[t];
  live_acc[t] = 0;
  __syncthreads();
  for (int m = kCounters / 2; m > 0; m >>= 1) {
    if (t < m) {
      part[t] += part[t + m];
      lor[t] |= lor[t + m];
    }
    __syncthreads();
  }
  if (t != 0 || ctl[2]) return;
  *live = lor[0];
  if (ctl[1] == 0) ctl[2] = 1;
  else ctl[0] += 1;
  ctl[1] = 0;
  ctl[3] = part[0] < push_limit ? 1 : 0;
}

// Start of a level. Push levels write the next
// frontier and its packed non-empty bits with atomics: both start cleared.
// A pull level stores only the new bits of the rows it walks: here every
// next-frontier word and packed word is cleared, and block rows every live
// source has reached (saturated: nothing can be new there) are marked, so the
// level kernel may skip chunks made of such rows.
template <int T>
__global__ void bfs_clear_kernel(BfsState b, long long zbytes) {
  using W = typename Tile<T>::W;
  if (b.ctl[2]) return;
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (b.ctl[3]) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.big_n = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < b.npad; i += stride) b.Fn[i] = 0;
    uint8_t *z = static_cast<uint8_t *>(b.fnzn);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < zbytes; i += stride) z[i] = 0;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) b.ctl[5] = 0;  // raised by bfs_chunk_skip_kernel
  const uint64_t dead = ~*b.live;
  const int lane = threadIdx.x & 63;
  // whole waves over 64 consecutive vertices (T divides 64); padding rows
  // (v >= n) hold every source bit in V
  const long long first = ((long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63));
  for (long long base = first; base < b.npad; base += stride) {
    const long long v = base + lane;
    const bool in = v < b.npad;
    const bool full = !in || (b.V[v] | dead) == ~0ull;
    const unsigned long long m = __ballot(full);
    const unsigned long long grp = (T == 64 ? ~0ull : ((1ull << T) - 1)) << (lane & ~(T - 1));
    const bool row_sat = (m & grp) == grp;
    // every next-frontier word and packed word starts the level cleared (a
    // source's seed bit included): the level kernel stores only new bits,
    // and steps over block rows without tiles and skips saturated chunks
    const long long I = in ? v / T : 0;
    if (in) b.Fn[v] = 0;
    if (in && (lane & (T - 1)) == 0) {
      b.sat[I] = row_sat ? 1 : 0;
      static_cast<W *>(b.fnzn)[I] = (W)0;
    }
  }
}

// Pull levels, after the saturation pass: chunk c may skip its tiles when
// every block row it touches (part[c] .. part[c+1]) is saturated; chunks over
// more than 32 block rows are not checked. One thread per chunk.
__global__ void bfs_chunk_skip_kernel(BfsState b) {
  if (b.ctl[2] || b.ctl[3]) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool all = false;
  if (c < b.nchunks) {
    const int Ia = b.part[c];
    const int Ib = c + 1 < b.nchunks ? b.part[c + 1] : b.nbr - 1;
    all = Ib - Ia <= 32;
    for (int I = Ia; all && I <= Ib; ++I) all = b.sat[I] != 0;
    b.skip[c] = all ? 1 : 0;
  }
  if (__ballot(all) != 0 && (threadIdx.x & 63) == 0) b.ctl[5] = 1;
}

// Pushes frontier row r of A^T's block row I over its tiles [k0, k1) (lanes
// stride over them): F[u] & ~V[v] is OR-ed into Fn[v] for every neighbour v
// -- one atomic per pushed edge. V still holds the previous levels only, so
// Fn collects exactly this level's new bits; bfs_push_finish_kernel then
// folds them into V, the level planes and the packed non-empty words.
template <int T>
__device__ __forceinline__ void push_row_segment(const BfsState &b, long long I, int r, int k0, int k1,
                                                 int lane) {
  using W = typename Tile<T>::W;
  const uint64_t Fu = b.F[I * T + r];
  for (int k = k0 + lane; k < k1; k += 64) {
    uint32_t w = (uint32_t) static_cast<const W *>(b.At)[(long long)k * T + r];
    const long long J = b.colindT[k];
    while (w) {
      const int hh = 31 - __clz(w);
      w &= ~(1u << hh);
      const long long v = J * T + (T - 1 - hh);
      const uint64_t nb = Fu & ~b.V[v];
      if (nb) atomicOr(reinterpret_cast<unsigned long long *>(b.Fn + v), (unsigned long long)nb);
    }
  }
}

// End of a push pass: the any-new flag and the wave's push-cost sum.
__device__ __forceinline__ void push_finish(const BfsState &b, bool any, unsigned long long cost, int lane,
                                            int slot) {
  if (__ballot(any) != 0 && lane == 0) b.ctl[1] = 1;
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) cost += __shfl_xor(cost, m);
  if (lane == 0 && cost) atomicAdd(b.fcnt + (slot & (kCounters - 1)), cost);
}

// Push level, part 1: one wave per block row of A^T (grid-stride) whose
// packed frontier word is non-zero. Rows of block rows up to kPushSeg tiles
// are pushed here; those of longer block rows go to big_list (pushed whole
// here only if the list is full).
template <int T>
__global__ __launch_bounds__(256) void bfs_push_kernel(BfsState b) {
  using W = typename Tile<T>::W;
  if (b.ctl[2] || !b.ctl[3]) return;
  const int lane = threadIdx.x & 63;
  const long long waves = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long I = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); I < b.nbr; I += waves) {
    uint32_t fw = (uint32_t) static_cast<const W *>(b.fnz)[I];
    if (!fw) continue;
    const int ks = b.rowptrT[I], ke = b.rowptrT[I + 1];
    while (fw) {
      const int h = 31 - __clz(fw);
      fw &= ~(1u << h);
      const int r = T - 1 - h;
      if (ke - ks > kPushSeg) {
        int slot = 0;
        if (lane == 0) slot = atomicAdd(b.big_n, 1);
        slot = __shfl(slot, 0);
        if (slot < kBigCap) {
          if (lane == 0) b.big_list[slot] = make_int2((int)I, r);
          continue;
        }
      }
      push_row_segment<T>(b, I, r, ks, ke, lane);
    }
  }
}

// Push level, part 2: the long rows, in kPushSeg-tile segments spread over
// all waves (segment s of an entry goes to wave s mod waves).
template <int T>
__global__ __launch_bounds__(256) void bfs_push_long_kernel(BfsState b) {
  if (b.ctl[2] || !b.ctl[3]) return;
  const int lane = threadIdx.x & 63;
  const int nbig = min(*b.big_n, kBigCap);
  const long long waves = (long long)gridDim.x * (blockDim.x >> 6);
  const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int i = 0; i < nbig; ++i) {
    const int2 e = b.big_list[i];
    const int ks = b.rowptrT[e.x], ke = b.rowptrT[e.x + 1];
    const long long nseg = (ke - ks + kPushSeg - 1) / kPushSeg;
    for (long long sg = wave; sg < nseg; sg += waves) {
      const int k0 = ks + (int)sg * kPushSeg;
      push_row_segment<T>(b, e.x, e.y, k0, min(ke, k0 + kPushSeg), lane);
    }
  }
}

// Push level, part 3: one thread per block row folds the new bits Fn into V
// and the level planes, writes the packed non-empty word, and sums the push
// cost of the vertices added (for the next level's direction).
template <int T>
__global__ void bfs_push_finish_kernel(BfsState b) {
  using W = typename Tile<T>::W;
  if (b.ctl[2] || !b.ctl[3]) return;
  const int depth = b.ctl[0];
  const int lane = threadIdx.x & 63;
  bool any = false;
  unsigned long long cost = 0;
  uint64_t lnew = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x; I < b.nbr; I += stride) {
    uint32_t bits = 0;
#pragma unroll 4
    for (int r = 0; r < T; ++r) {
      const long long v = I * T + r;
      const uint64_t nb = b.Fn[v];
      if (!nb) continue;
      lnew |= nb;
      b.V[v] |= nb;
      for (int k = 0; k < kPlanes; ++k)
        if ((depth >> k) & 1) b.P[k * b.npad + v] |= nb;
      bits |= 1u << (T - 1 - r);
    }
    static_cast<W *>(b.fnzn)[I] = (W)bits;
    if (bits) {
      any = true;
      cost += (unsigned long long)__popc(bits) * (b.rowptrT[I + 1] - b.rowptrT[I]);
    }
  }
  push_finish(b, any, cost, lane, blockIdx.x);
  publish_live(b, lnew, lane, blockIdx.x);
}

template <class V>
__global__ void bfs_fill_kernel(V *p, long long n, V v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Sources: V[src] and F[src] get bit s; the packed frontier word gets the bit.
template <int T>
__global__ void bfs_seed_kernel(const int *src, int nsrc, uint64_t *F, uint64_t *V, void *fnz) {
  using W = typename Tile<T>::W;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int s = 0; s < nsrc; ++s) {
    const int v = src[s];
    F[v] |= 1ull << s;
    V[v] |= 1ull << s;
    static_cast<W *>(fnz)[v / T] |= (W)(1u << (T - 1 - v % T));
  }
}

__global__ void bfs_expand_kernel(const uint64_t *V, const uint64_t *P, long long npad, int n, int nsrc,
                                  int nplanes, int *level) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint64_t vis = V[v];
  // plane-outer over the planes in use (ceil(log2(levels)), not all kPlanes):
  // 64 level registers, one bit extract and OR per (plane, source)
  int lev[64];
#pragma unroll
  for (int s = 0; s < 64; ++s) lev[s] = 0;
  for (int k = 0; k < nplanes; ++k) {
    const uint64_t pk = P[k * npad + v];
    const uint32_t lo = (uint32_t)pk, hi = (uint32_t)(pk >> 32);
#pragma unroll
    for (int s = 0; s < 64; ++s)
      lev[s] |= (int)(((s < 32 ? lo >> s : hi >> (s - 32)) & 1u) << k);
  }
#pragma unroll
  for (int s = 0; s < 64; ++s)
    if (s < nsrc) level[(long long)s * n + v] = ((vis >> s) & 1) ? lev[s] : -1;
}

// ---- plain per-source loop ----
template <int T>
__global__ void plain_update_kernel(const void *next, void *visited, int nbr, int n, int *level, int depth,
                                    int *any) {
  using W = typename Tile<T>::W;
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const uint32_t c = static_cast<const W *>(next)[I];
  if (!c) return;
  W *vis = static_cast<W *>(visited);
  vis[I] = (W)(vis[I] | c);
  for (int r = 0; r < T; ++r)
    if ((c >> (T - 1 - r)) & 1u) {
      const long long v = I * T + r;
      if (v < n) level[v] = depth;
    }
  *any = 1;
}

template <int T>
__global__ void plain_seed_kernel(int src, void *front, void *visited, int *level) {
  using W = typename Tile<T>::W;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  static_cast<W *>(front)[src / T] = (W)(1u << (T - 1 - src % T));
  static_cast<W *>(visited)[src / T] = (W)(1u << (T - 1 - src % T));
  level[src] = 0;
}

unsigned nblk(long long n, int b = 256) { return (unsigned)ceil_div(n > 0 ? n : 1, b); }

// Frontier gathers of the pull levels at T <= 8: 0 one per set bit of each
// tile row, 1 one per set column of each tile (bgb_set_bfs_gather_mode).
int g_bfs_gather = 1;

// Per-level timing of the eager level loop (bgb_set_bfs_level_timing): the
// GPU time of each level of the last timed call and whether it pulled.
bool g_level_timing = false;
struct LevelTimes {
  std::vector<float> ms;
  std::vector<int> pulled;
};
LevelTimes &level_times() {
  static thread_local LevelTimes t;
  return t;
}
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t create() {
    hipError_t e = hipEventCreate(&a);
    return e != hipSuccess ? e : hipEventCreate(&b);
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};
// grids of the grid-stride fills (up to 16 words per vertex)
unsigned nfill(long long n) { return capped_blocks(n, 256); }

template <int T>
int bfs_multi_impl(const void *A, const int *rowptr, const int *colind, int nbr, int nblocks, const void *At,
                   const int *rowptrT, const int *colindT, int n, const int *h_src, int nsrc, int *level,
                   int use_graph, int *h_nlevels, hipStream_t s) {
  constexpr int S = kBfsRun, CH = (64 / Tile<T>::L) * S;
  const long long npad = (long long)nbr * T;
  const long long nwords = ceil_div(nbr, T) * T;
  const int wb = word_bytes(T);
  const long long nchunks = ceil_div(nblocks, CH);
  DeviceBuffer F0, F1, Vb, Pb, z0, z1, ctl, src, ck, cv, pt, fc, bign, bigl, lv, lacc, sat, skp;
  BGB_HIP_TRY(F0.alloc(8 * npad, s));
  BGB_HIP_TRY(F1.alloc(8 * npad, s));
  BGB_HIP_TRY(Vb.alloc(8 * npad, s));
  BGB_HIP_TRY(Pb.alloc(8 * npad * kPlanes, s));
  BGB_HIP_TRY(z0.alloc(wb * nwords + 16, s));
  BGB_HIP_TRY(z1.alloc(wb * nwords + 16, s));
  BGB_HIP_TRY(ctl.alloc(8 * sizeof(int), s));
  BGB_HIP_TRY(src.alloc(sizeof(int) * nsrc, s));
  BGB_HIP_TRY(ck.alloc(sizeof(int) * (nchunks + 1), s));
  BGB_HIP_TRY(cv.alloc(8 * (nchunks + 1) * T, s));
  BGB_HIP_TRY(pt.alloc(sizeof(int) * (nchunks + 1), s));
  BGB_HIP_TRY(fc.alloc(sizeof(unsigned long long) * kCounters, s));
  BGB_HIP_TRY(bign.alloc(sizeof(int), s));
  BGB_HIP_TRY(bigl.alloc(sizeof(int2) * (At ? kBigCap : 1), s));
  BGB_HIP_TRY(lv.alloc(sizeof(uint64_t), s));
  BGB_HIP_TRY(lacc.alloc(sizeof(unsigned long long) * kCounters, s));
  BGB_HIP_TRY(sat.alloc(nbr + 64, s));
  BGB_HIP_TRY(skp.alloc(nchunks + 64, s));
  if (nchunks > 0)  // chunk -> first block row, once for every level
    chunk_rows_kernel<<<nblk(nbr), 256, 0, s>>>(rowptr, nbr, CH, (int)nchunks, pt.as<int>());
  const uint64_t full = nsrc == 64 ? ~0ull : ((1ull << nsrc) - 1);
  bfs_fill_kernel<uint64_t><<<nfill(npad), 256, 0, s>>>(F0.as<uint64_t>(), npad, 0ull);
  bfs_fill_kernel<uint64_t><<<nfill(npad), 256, 0, s>>>(F1.as<uint64_t>(), npad, 0ull);
  bfs_fill_kernel<uint64_t><<<nfill(npad * kPlanes), 256, 0, s>>>(Pb.as<uint64_t>(), npad * kPlanes, 0ull);
  // padding rows count as visited by every source
  bfs_fill_kernel<uint64_t><<<nfill(npad), 256, 0, s>>>(Vb.as<uint64_t>(), npad, 0ull);
  if (npad > n)
    bfs_fill_kernel<uint64_t><<<nfill(npad - n), 256, 0, s>>>(Vb.as<uint64_t>() + n, npad - n, full);
  bfs_fill_kernel<uint8_t><<<nfill(wb * nwords + 16), 256, 0, s>>>(z0.as<uint8_t>(), wb * nwords + 16, 0);
  bfs_fill_kernel<uint8_t><<<nfill(wb * nwords + 16), 256, 0, s>>>(z1.as<uint8_t>(), wb * nwords + 16, 0);
  bfs_fill_kernel<unsigned long long><<<1, kCounters, 0, s>>>(fc.as<unsigned long long>(), kCounters, 0ull);
  // every source starts live
  bfs_fill_kernel<uint64_t><<<1, 64, 0, s>>>(lv.as<uint64_t>(), 1, full);
  bfs_fill_kernel<unsigned long long><<<1, kCounters, 0, s>>>(lacc.as<unsigned long long>(), kCounters, 0ull);
  const bool push = At != nullptr;
  const int init_ctl[8] = {1, 0, 0, push ? 1 : 0, 0, 0, 0, 0};  // depth 1: a push level (the sources)
  BGB_HIP_TRY(hipMemcpyAsync(ctl.ptr, init_ctl, sizeof(init_ctl), hipMemcpyHostToDevice, s));
  BGB_HIP_TRY(hipMemcpyAsync(src.ptr, h_src, sizeof(int) * nsrc, hipMemcpyHostToDevice, s));
  bfs_seed_kernel<T><<<1, 64, 0, s>>>(src.as<int>(), nsrc, F0.as<uint64_t>(), Vb.as<uint64_t>(), z0.ptr);
  BGB_HIP_TRY(hipGetLastError());

  BfsState st[2];
  for (int p = 0; p < 2; ++p) {
    BfsState &b = st[p];
    b.A = A;
    b.rowptr = rowptr;
    b.colind = colind;
    b.nbr = nbr;
    b.nblocks = nblocks;
    b.nchunks = (int)nchunks;
    b.F = p == 0 ? F0.as<uint64_t>() : F1.as<uint64_t>();
    b.Fn = p == 0 ? F1.as<uint64_t>() : F0.as<uint64_t>();
    b.fnz = p == 0 ? z0.ptr : z1.ptr;
    b.fnzn = p == 0 ? z1.ptr : z0.ptr;
    b.V = Vb.as<uint64_t>();
    b.P = Pb.as<uint64_t>();
    b.npad = npad;
    b.full = full;
    b.ctl = ctl.as<int>();
    b.carry_key = ck.as<int>();
    b.carry_val = cv.as<uint64_t>();
    b.part = pt.as<int>();
    b.fcnt = fc.as<unsigned long long>();
    b.At = At;
    b.rowptrT = rowptrT;
    b.colindT = colindT;
    b.big_n = bign.as<int>();
    b.big_list = bigl.as<int2>();
    b.live = lv.as<uint64_t>();
    b.live_acc = lacc.as<unsigned long long>();
    b.sat = sat.as<uint8_t>();
    b.skip = skp.as<uint8_t>();
  }
  // pull pass: a persistent grid of the workgroups the GPU holds at once
  // (every wave gets the same share of chunks, so none may wait for a slot)
  DeviceFacts f;
  if (int e = device_facts(&f)) return e;
  // frontier gathers per set column of a tile (T <= 8; F within a buffer
  // resource's 32-bit byte range) or per set bit of a row
  void (*level_kernel)(BfsState) = (T <= 8 && g_bfs_gather == 1 && 8 * npad < (1ll << 31))
                                       ? &bfs_level_kernel<T, S, (T <= 8 ? 1 : 0)>
                                       : &bfs_level_kernel<T, S, 0>;
  int per_cu = 0;
  BGB_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(level_kernel), 64 * kBfsWaves, 0));
  const unsigned blocks = (unsigned)std::max<long long>(
      1, std::min<long long>(ceil_div(nchunks > 0 ? nchunks : 1, kBfsWaves), (long long)f.cus * std::max(per_cu, 1)));
  const unsigned fblocks = nblk(ceil_div(nchunks, 64) * 64);  // one lane per chunk, whole waves
  const unsigned pblocks = (unsigned)std::min<long long>(ceil_div(nbr, 4), 8192);
  const unsigned long long push_limit = push ? (unsigned long long)nblocks / kPushDiv : 0ull;
  const long long zbytes = wb * nwords + 16;
  auto one_level = [&](int p) {
    bfs_clear_kernel<T><<<nfill(npad), 256, 0, s>>>(st[p], zbytes);
    if (nchunks > 0) {
      bfs_chunk_skip_kernel<<<nblk(nchunks), 256, 0, s>>>(st[p]);
      level_kernel<<<blocks, 64 * kBfsWaves, 0, s>>>(st[p]);
      bfs_fixup_kernel<T><<<fblocks, 256, 0, s>>>(st[p]);
      if (push) {
        bfs_push_kernel<T><<<pblocks, 256, 0, s>>>(st[p]);
        bfs_push_long_kernel<T><<<1024, 256, 0, s>>>(st[p]);
        bfs_push_finish_kernel<T><<<nblk(nbr), 256, 0, s>>>(st[p]);
      }
    }
    bfs_advance_kernel<<<1, kCounters, 0, s>>>(ctl.as<int>(), fc.as<unsigned long long>(), push_limit,
                                                lacc.as<unsigned long long>(), lv.as<uint64_t>());
  };
  auto level_pair = [&]() {
    for (int p = 0; p < 2; ++p) one_level(p);
  };
  int h_ctl[4] = {0, 0, 0, 0};
  const int max_pairs = (1 << kPlanes) / 2 - 1;
  if (use_graph) {
    // Capture is not allowed on the legacy default stream: capture and replay
    // on a private non-blocking stream (default priority), ordered after the
    // setup work queued on s.
    BGB_HIP_TRY(hipStreamSynchronize(s));
    hipStream_t cs = nullptr;
    BGB_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int st_code = BGB_OK;
    const char *what = "";
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      st_code = BGB_ERR_HIP;
      what = "hipStreamBeginCapture";
    } else {
      const hipStream_t saved = s;
      s = cs;
      level_pair();
      s = saved;
      if (hipStreamEndCapture(cs, &graph) != hipSuccess) {
        st_code = BGB_ERR_HIP;
        what = "hipStreamEndCapture";
      } else if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        st_code = BGB_ERR_HIP;
        what = "hipGraphInstantiate";
      }
    }
    int pairs = 0;
    while (st_code == BGB_OK && pairs < max_pairs) {
      // replay a few pairs between host checks; levels after the last are no-ops
      for (int r = 0; r < 4 && st_code == BGB_OK; ++r, ++pairs)
        if (hipGraphLaunch(exec, cs) != hipSuccess) {
          st_code = BGB_ERR_HIP;
          what = "hipGraphLaunch";
        }
      if (st_code == BGB_OK &&
          (hipMemcpyAsync(h_ctl, ctl.ptr, sizeof(h_ctl), hipMemcpyDeviceToHost, cs) != hipSuccess ||
           hipStreamSynchronize(cs) != hipSuccess)) {
        st_code = BGB_ERR_HIP;
        what = "level flag read-back";
      }
      if (h_ctl[2]) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamSynchronize(cs);
    (void)hipStreamDestroy(cs);
    if (st_code) return set_error(st_code, "BFS level graph: %s failed", what);
  } else if (g_level_timing) {
    // one level at a time, each between two events; the flags read before a
    // level say whether it pushes
    LevelTimes &lt = level_times();
    lt.ms.clear();
    lt.pulled.clear();
    EventPair ev;
    BGB_HIP_TRY(ev.create());
    for (int lv = 0; lv < 2 * max_pairs; ++lv) {
      BGB_HIP_TRY(hipMemcpyAsync(h_ctl, ctl.ptr, sizeof(h_ctl), hipMemcpyDeviceToHost, s));
      BGB_HIP_TRY(hipStreamSynchronize(s));
      if (h_ctl[2]) break;
      BGB_HIP_TRY(hipEventRecord(ev.a, s));
      one_level(lv & 1);
      BGB_HIP_TRY(hipGetLastError());
      BGB_HIP_TRY(hipEventRecord(ev.b, s));
      BGB_HIP_TRY(hipEventSynchronize(ev.b));
      float ms = 0.f;
      BGB_HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
      lt.ms.push_back(ms);
      lt.pulled.push_back(h_ctl[3] ? 0 : 1);
    }
    BGB_HIP_TRY(hipMemcpyAsync(h_ctl, ctl.ptr, sizeof(h_ctl), hipMemcpyDeviceToHost, s));
    BGB_HIP_TRY(hipStreamSynchronize(s));
  } else {
    for (int pairs = 0; pairs < max_pairs; ++pairs) {
      level_pair();
      BGB_HIP_TRY(hipGetLastError());
      BGB_HIP_TRY(hipMemcpyAsync(h_ctl, ctl.ptr, sizeof(h_ctl), hipMemcpyDeviceToHost, s));
      BGB_HIP_TRY(hipStreamSynchronize(s));
      if (h_ctl[2]) break;
    }
  }
  BGB_REQUIRE(h_ctl[2], BGB_ERR_SIZE, "BFS deeper than %d levels", 1 << kPlanes);
  const int nlevels = h_ctl[0];  // depth of the first empty level == number of levels
  int nplanes = 0;
  while (nplanes < kPlanes && (1 << nplanes) <= nlevels) ++nplanes;
  bfs_expand_kernel<<<nblk(n), 256, 0, s>>>(Vb.as<uint64_t>(), Pb.as<uint64_t>(), npad, n, nsrc, nplanes,
                                            level);
  BGB_HIP_TRY(hipGetLastError());
  *h_nlevels = nlevels;
  return BGB_OK;
}

template <int T>
int bfs_plain_impl(const void *A, const int *rowptr, const int *colind, int nbr, int nblocks, int n,
                   int source, int *level, int *h_nlevels, bgb_stream_t stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long nwords = ceil_div(nbr, T) * T;
  const int wb = word_bytes(T);
  DeviceBuffer f0, f1, vis, any;
  BGB_HIP_TRY(f0.alloc(wb * nwords + 16, s));
  BGB_HIP_TRY(f1.alloc(wb * nwords + 16, s));
  BGB_HIP_TRY(vis.alloc(wb * nwords + 16, s));
  BGB_HIP_TRY(any.alloc(sizeof(int), s));
  for (DeviceBuffer *d : {&f0, &f1, &vis})
    bfs_fill_kernel<uint8_t><<<nfill(wb * nwords + 16), 256, 0, s>>>(d->as<uint8_t>(), wb * nwords + 16, 0);
  bfs_fill_kernel<int><<<nfill(n), 256, 0, s>>>(level, n, -1);
  plain_seed_kernel<T><<<1, 64, 0, s>>>(source, f0.ptr, vis.ptr, level);
  BGB_HIP_TRY(hipGetLastError());
  void *cur = f0.ptr, *nxt = f1.ptr;
  int depth = 1;
  for (;; ++depth) {
    BGB_REQUIRE(depth < INT_MAX, BGB_ERR_SIZE, "BFS too deep");
    int st = bgb_bmv_bin_bin_bin_masked(T, A, cur, nxt, rowptr, colind, nbr, nblocks, vis.ptr, stream);
    if (st) return st;
    bfs_fill_kernel<int><<<1, 1, 0, s>>>(any.as<int>(), 1, 0);
    plain_update_kernel<T><<<nblk(nbr), 256, 0, s>>>(nxt, vis.ptr, nbr, n, level, depth, any.as<int>());
    BGB_HIP_TRY(hipGetLastError());
    int h_any = 0;
    BGB_HIP_TRY(hipMemcpyAsync(&h_any, any.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    BGB_HIP_TRY(hipStreamSynchronize(s));
    if (!h_any) break;
    void *t = cur;
    cur = nxt;
    nxt = t;
  }
  *h_nlevels = depth;
  return BGB_OK;
}

int bfs_check(int T, const void *A, const int *rowptr, const int *colind, int nbr, int nblocks, int n,
              int *level, int *h_nlevels) {
  BGB_REQUIRE(valid_tiledim(T) && T <= 32, BGB_ERR_TILEDIM, "the BFS drivers take tile dims 4, 8, 16, 32 (got %d)",
              T);
  BGB_REQUIRE(rowptr && level && h_nlevels, BGB_ERR_ARG, "null argument");
  BGB_REQUIRE(nblocks == 0 || (A && colind), BGB_ERR_ARG, "null matrix");
  BGB_REQUIRE(n >= 1 && nbr == (int)ceil_div(n, T), BGB_ERR_ARG, "nblockrows must be ceil(n/T)");
  auto aligned16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  BGB_REQUIRE(aligned16(A) && aligned16(colind), BGB_ERR_ARG, "bsrVal and bsrColInd must be 16-byte aligned");
  return BGB_OK;
}

}  // namespace
}  // namespace bgb

using namespace bgb;

extern "C" {

int bgb_bfs_multi(int T, const void *d_A, const int *d_rowptr, const int *d_colind, int nblockrows,
                  int nblocks, int n, const int *h_sources, int nsrc, int *d_levels, int use_graph,
                  int *h_nlevels, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    return bgb_bfs_multi_dir(T, d_A, d_rowptr, d_colind, nblockrows, nblocks, nullptr, nullptr, nullptr, n,
                             h_sources, nsrc, d_levels, use_graph, h_nlevels, stream);
  });
}

int bgb_bfs_multi_dir(int T, const void *d_A, const int *d_rowptr, const int *d_colind, int nblockrows,
                      int nblocks, const void *d_At, const int *d_rowptrT, const int *d_colindT, int n,
                      const int *h_sources, int nsrc, int *d_levels, int use_graph, int *h_nlevels,
                      bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    int st = bfs_check(T, d_A, d_rowptr, d_colind, nblockrows, nblocks, n, d_levels, h_nlevels);
    if (st) return st;
    BGB_REQUIRE(h_sources && nsrc >= 1 && nsrc <= 64, BGB_ERR_ARG, "need 1..64 sources");
    for (int i = 0; i < nsrc; ++i)
      BGB_REQUIRE(h_sources[i] >= 0 && h_sources[i] < n, BGB_ERR_ARG, "source %d out of range", h_sources[i]);
    BGB_REQUIRE((d_At == nullptr) == (d_rowptrT == nullptr) && (d_At == nullptr) == (d_colindT == nullptr),
                BGB_ERR_ARG, "pass all of At, rowptrT, colindT, or none");
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (T) {
      case 4: return bfs_multi_impl<4>(d_A, d_rowptr, d_colind, nblockrows, nblocks, d_At, d_rowptrT, d_colindT, n, h_sources, nsrc, d_levels, use_graph, h_nlevels, s);
      case 8: return bfs_multi_impl<8>(d_A, d_rowptr, d_colind, nblockrows, nblocks, d_At, d_rowptrT, d_colindT, n, h_sources, nsrc, d_levels, use_graph, h_nlevels, s);
      case 16: return bfs_multi_impl<16>(d_A, d_rowptr, d_colind, nblockrows, nblocks, d_At, d_rowptrT, d_colindT, n, h_sources, nsrc, d_levels, use_graph, h_nlevels, s);
      default: return bfs_multi_impl<32>(d_A, d_rowptr, d_colind, nblockrows, nblocks, d_At, d_rowptrT, d_colindT, n, h_sources, nsrc, d_levels, use_graph, h_nlevels, s);
    }
  });
}

int bgb_set_bfs_gather_mode(int mode) {
  clear_error();
  BGB_REQUIRE(mode == 0 || mode == 1, BGB_ERR_ARG, "gather mode must be 0 or 1 (got %d)", mode);
  g_bfs_gather = mode;
  return BGB_OK;
}

int bgb_get_bfs_gather_mode(void) { return g_bfs_gather; }

int bgb_set_bfs_level_timing(int on) {
  clear_error();
  BGB_REQUIRE(on == 0 || on == 1, BGB_ERR_ARG, "on must be 0 or 1 (got %d)", on);
  g_level_timing = on != 0;
  return BGB_OK;
}

int bgb_bfs_level_times(int cap, float *h_ms, int *h_pulled, int *h_n) {
  clear_error();
  BGB_REQUIRE(cap >= 0 && h_n && (cap == 0 || (h_ms && h_pulled)), BGB_ERR_ARG, "bad arguments");
  const LevelTimes &lt = level_times();
  const int n = (int)std::min<size_t>(lt.ms.size(), (size_t)cap);
  for (int i = 0; i < n; ++i) {
    h_ms[i] = lt.ms[i];
    h_pulled[i] = lt.pulled[i];
  }
  *h_n = (int)lt.ms.size();
  return BGB_OK;
}

int bgb_bfs_plain(int T, const void *d_A, const int *d_rowptr, const int *d_colind, int nblockrows,
                  int nblocks, int n, int source, int *d_level, int *h_nlevels, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    int st = bfs_check(T, d_A, d_rowptr, d_colind, nblockrows, nblocks, n, d_level, h_nlevels);
    if (st) return st;
    BGB_REQUIRE(source >= 0 && source < n, BGB_ERR_ARG, "source out of range");
    switch (T) {
      case 4: return bfs_plain_impl<4>(d_A, d_rowptr, d_colind, nblockrows, nblocks, n, source, d_level, h_nlevels, stream);
      case 8: return bfs_plain_impl<8>(d_A, d_rowptr, d_colind, nblockrows, nblocks, n, source, d_level, h_nlevels, stream);
      case 16: return bfs_plain_impl<16>(d_A, d_rowptr, d_colind, nblockrows, nblocks, n, source, d_level, h_nlevels, stream);
      default: return bfs_plain_impl<32>(d_A, d_rowptr, d_colind, nblockrows, nblocks, n, source, d_level, h_nlevels, stream);
    }
  });
}

}  // extern "C"
