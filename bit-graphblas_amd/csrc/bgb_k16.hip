// bin_full_full with K = 16 right-hand sides on the MFMA units (gfx950).
//
// No reference counterpart (the reference's bin_full_full takes one FP32
// vector, bsrbmv.cu:826-1081). Y[v][0..16) (+)= sum_u A[v][u] X[u][0..16):
// the 0/1 tile is exact in FP16, X is FP16, accumulation is FP32.
//
// One v_mfma_f32_16x16x16_f16 per 16x16 sub-tile:
//   A operand  lane l: row l%16, columns 4*(l/16)..+3 of the tile -> 4 bits of
//              the row word expanded to FP16 0/1
//   B operand  lane l: RHS column l%16, tile columns 4*(l/16)..+3 -> 4 FP16
//              values of X; X is stored pre-permuted in exactly this order
//              (bgb_k16_pack_rhs), so each 16-row block of X is one contiguous
//              512-byte read per wave
//   D          lane l: rows 4*(l/16)..+3, RHS column l%16 (FP32)
// T = 16 is one MFMA per tile; T = 32 is four (two row halves x two column
// halves). Work decomposition follows the BMV kernels: fixed chunks of tiles
// per wave, rows flushed when they end, a carry record for the row open at a
// chunk's end, and a fix-up pass that adds the records (summed per row run;
// atomics only where a run crosses the fix-up's wave boundaries).
// The batched entry point runs many graphs in one persistent launch whose
// waves take chunks from a global queue until it is drained.
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>
#include <vector>

#include "bgb_device.h"
#include "bgb_tile.h"

namespace bgb {
namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kK = 16;        // right-hand sides
constexpr int kChunk = 256;   // tiles per wave
constexpr int kWaves = 4;     // waves per workgroup
// tiles per pipelined group (X gathers in flight per wave)
template <int T>
constexpr int kGroup = 8;

// 4 bits (bit 3 = first column) -> four FP16 values 0.0 / 1.0.
__device__ __forceinline__ half4_t nib_to_h4(uint32_t nib) {
  const uint32_t one = 0x3C00u;
  uint2 u;
  u.x = ((nib >> 3) & 1u) * one | (((nib >> 2) & 1u) * one) << 16;
  u.y = ((nib >> 1) & 1u) * one | ((nib & 1u) * one) << 16;
  return __builtin_bit_cast(half4_t, u);
}

struct K16Graph {  // device-side view of one graph
  const int *rowptr;
  const int *colind;
  const void *val;
  const uint2 *xb;  // 4 FP16 per lane per 16-row block
  float *y;
  int nbr, nblocks, accumulate;
};

template <int T>
__device__ __forceinline__ void k16_flush(const K16Graph &G, int I, int lane, const float4_t (&D)[T / 16]) {
  constexpr int MH = T / 16;
  const int n = lane & 15, rq = lane >> 4;
#pragma unroll
  for (int mh = 0; mh < MH; ++mh)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float *p = G.y + ((long long)I * T + mh * 16 + 4 * rq + j) * kK + n;
      *p = G.accumulate ? *p + D[mh][j] : D[mh][j];
    }
}

// Wave-uniform index load through the scalar data cache (s_load): block
// columns and row ends are the same for every lane, and SMEM loads are counted
// apart from the vector loads, so waiting for one does not drain the other.
typedef __attribute__((address_space(4))) const int const_int_t;
__device__ __forceinline__ int sload(const int *p, long long i) {
  return ((const const_int_t *)p)[i];  // address-space cast: C-style only
}

// Lane's 8 bytes of the X block at block column J (MH 512-byte halves), or 0
// when `use` is false. The buffer resource spans exactly the block, and a
// masked lane's offset lies past its end: the hardware's range check returns
// zero for it with no memory access, so the gather needs no branch.
template <int MH>
__device__ __forceinline__ uint2 x_gather(const uint2 *xb, int J, int h, int lane, bool use) {
  constexpr int kBytes = MH * 512;
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint2 *>(xb + (long long)J * MH * 64), (short)0, kBytes, 0x00020000);
  return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, use ? h * 512 + lane * 8 : kBytes, 0, 0));
}

// Processes tiles [k0, k1) of one graph; returns the row left open (or nbr).
//
// Groups of kGroup<T> tiles: the X gathers of group g are issued, then the
// tile words of group g+1 (block columns through SMEM), then the MFMAs of
// group g wait for the gathers alone. Every load is issued unconditionally
// (indices past the chunk are clamped to its last tile and their words
// zeroed), so the compiler can count the loads in flight instead of draining
// them.
//
// A group's tile words (U*T row words, 256 B at T = 16, 1 KB at T = 32) come
// in with one vector load per lane and go through a per-wave LDS buffer to
// the lanes that use them (lane m takes row m of each tile): one VMEM
// instruction per group instead of one per tile and row half. The vector
// memory pipeline (address unit) is what bounds this kernel: every X gather
// is one VMEM instruction per tile already.
template <int T>
__device__ int k16_chunk(const K16Graph &G, long long k0, long long k1, int lane, float *carry) {
  constexpr int MH = T / 16, U = kGroup<T>;
  using W = typename Tile<T>::W;
  constexpr int kTB = U * T * (int)sizeof(W);  // bytes of one group's tiles
  constexpr int kLB = kTB / 64;                // of which each lane loads kLB
  static_assert(kLB == 4 || kLB == 16, "group tile bytes per lane");
  using GV = typename std::conditional<kLB == 4, uint32_t, uint4>::type;
  __shared__ __align__(16) uint8_t s_tw[kWaves][kTB];
  uint8_t *tw = s_tw[threadIdx.x >> 6];
  const uint8_t *Ab = static_cast<const uint8_t *>(G.val);
  const int m = lane & 15, kq = lane >> 4;
  int I = __builtin_amdgcn_readfirstlane(wave_find_row(G.rowptr, 0, G.nbr - 1, k0, lane));
  float4_t D[MH], Z[MH];
#pragma unroll
  for (int mh = 0; mh < MH; ++mh) D[mh] = Z[mh] = float4_t{0.f, 0.f, 0.f, 0.f};
  // (block rows without tiles, before row I or anywhere in the walk, are left
  // to k16_zero_empty_kernel)
  // e: end of row I; e2: end of row I + 1 (prefetched, read at the next row end)
  int e = sload(G.rowptr, I + 1);
  int e2 = I + 1 < G.nbr ? sload(G.rowptr, I + 2) : INT_MAX;
  const long long last = k1 - 1;
  uint32_t w[U][MH];
  GV gv, gvn;
  int J[U], Jn[U];
  // Tiles p0 .. p0+U-1 (p0 a multiple of U past k0, so the group is kTB-byte
  // aligned): the buffer resource ends at the tile array's end, so bytes past
  // it -- a last partial group, or the prefetch after the chunk's last group
  // -- read as zero without a memory access.
  auto load_group = [&](long long p0, GV &g, int (&jj)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = sload(G.colind, p0 + u < last ? p0 + u : last);
    const long long rem = ((long long)G.nblocks - p0) * T * (long long)sizeof(W);
    const int nrec = rem <= 0 ? 0 : (rem < kTB ? (int)rem : kTB);
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t *>(Ab + p0 * T * (long long)sizeof(W)), (short)0, nrec, 0x00020000);
    if constexpr (kLB == 4) g = __builtin_amdgcn_raw_buffer_load_b32(r, lane * kLB, 0, 0);
    else g = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, lane * kLB, 0, 0));
  };
  // the group's row words from LDS: lane m, row h*16 + m of tile u
  auto stage_group = [&](const GV &g) {
    *reinterpret_cast<GV *>(tw + lane * kLB) = g;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const W *tws = reinterpret_cast<const W *>(tw);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int h = 0; h < MH; ++h) w[u][h] = tws[u * T + h * 16 + m];
  };
  load_group(k0, gv, J);
  for (long long p0 = k0; p0 < k1; p0 += U) {
    stage_group(gv);
    // X rows of a 4-column group only matter where the tile has a set bit in
    // those columns (the A operand is 0 elsewhere): the tile's column mask is
    // the OR of its row words, which sit in one 16-lane DPP row (lane m holds
    // row m), so lanes of groups with no set bit read nothing. Tiles past the
    // chunk get an empty mask, hence zero X and zero A.
    uint2 xv[U][MH];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = p0 + u < k1;
#pragma unroll
      for (int h = 0; h < MH; ++h) w[u][h] = in ? w[u][h] : 0u;
      uint32_t cols = 0;
#pragma unroll
      for (int h = 0; h < MH; ++h) cols |= w[u][h];
      cols = row_or16(cols);
#pragma unroll
      for (int h = 0; h < MH; ++h) {
        const uint32_t grp = (cols >> (T - 4 - (h * 16 + 4 * kq))) & 0xFu;
        xv[u][h] = x_gather<MH>(G.xb, J[u], h, lane, grp != 0);
      }
    }
    load_group(p0 + U, gvn, Jn);
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int kh = 0; kh < MH; ++kh) {
        const half4_t b = __builtin_bit_cast(half4_t, xv[u][kh]);
#pragma unroll
        for (int mh = 0; mh < MH; ++mh) {
          const uint32_t nib = (w[u][mh] >> (T - 4 - (kh * 16 + 4 * kq))) & 0xFu;
          D[mh] = __builtin_amdgcn_mfma_f32_16x16x16f16(nib_to_h4(nib), b, D[mh], 0, 0, 0);
        }
      }
      const long long pos = p0 + u + 1;
      if (pos <= k1 && pos == e) {  // wave-uniform (scalar) test
        k16_flush<T>(G, I, lane, D);
#pragma unroll
        for (int mh = 0; mh < MH; ++mh) D[mh] = Z[mh];
        ++I;
        e = e2;
        e2 = I + 1 < G.nbr ? sload(G.rowptr, I + 2) : INT_MAX;
        if (e == pos) {
          // block rows without tiles follow: gallop over the row ends (1, 2,
          // 4, ... rows ahead), then bisect, to the first row ending after pos
          auto row_end = [&](int r) { return r < G.nbr ? sload(G.rowptr, r + 1) : INT_MAX; };
          int lo = I + 1, step = 1, hi = I + 1;
          while (hi < G.nbr && row_end(hi) <= pos) {
            lo = hi + 1;
            step *= 2;
            hi = I + step;
          }
          if (hi > G.nbr) hi = G.nbr;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (row_end(mid) > pos) hi = mid;
            else lo = mid + 1;
          }
          I = __builtin_amdgcn_readfirstlane(lo);
          e = row_end(I);
          e2 = I + 1 < G.nbr ? sload(G.rowptr, I + 2) : INT_MAX;
        }
      }
    }
    gv = gvn;
#pragma unroll
    for (int u = 0; u < U; ++u) J[u] = Jn[u];
    // the next stage_group overwrites the buffer: every lane's reads of this
    // group are done (LDS operations of a wave complete in order)
  }
  if (I < G.nbr) {
    const int rq = lane >> 4;
#pragma unroll
    for (int mh = 0; mh < MH; ++mh)
#pragma unroll
      for (int j = 0; j < 4; ++j) carry[(mh * 16 + 4 * rq + j) * kK + m] = D[mh][j];
  }
  return I;
}

struct K16Args {
  K16Graph g;
  int nchunks;
  int *carry_key;
  float *carry_val;  // [nchunks][32 * 16]
};

template <int T>
__global__ __launch_bounds__(64 * kWaves) void k16_kernel(K16Args a) {
  const int lane = threadIdx.x & 63;
  const int chunk = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * kWaves + (threadIdx.x >> 6));
  if (chunk >= a.nchunks) return;
  const long long k0 = (long long)chunk * kChunk;
  const long long k1 = k0 + kChunk < a.g.nblocks ? k0 + kChunk : a.g.nblocks;
  const int open = k16_chunk<T>(a.g, k0, k1, lane, a.carry_val + (long long)chunk * 32 * kK);
  if (lane == 0) a.carry_key[chunk] = open < a.g.nbr ? open : -1;
}

// carry records -> Y; `graph_of` maps a chunk to its graph (nullptr for a
// single graph). One wave per kFixChunks consecutive chunks; the lanes take
// 64 consecutive values of a record (coalesced), load them for all the
// wave's chunks at once, then walk the chunks in registers summing each run
// of records of one row (keys are non-decreasing within a graph). A run
// inside the wave's chunks is added with a plain read-modify-write (the
// row's only contribution in this pass); the first and last runs, which may
// continue into the neighbouring waves' chunks, with atomics. (One FP32
// atomic per value -- 19.5 M at R-MAT 20, T = 16 -- took 137 us per call;
// a lane per chunk with a segmented scan read each 2 KB record with strided
// per-lane loads and was slower.)
constexpr int kFixChunks = 16;

__global__ __launch_bounds__(256) void k16_fixup_kernel(const int *carry_key, const float *carry_val,
                                                        long long nchunks, const K16Graph *graphs,
                                                        const int *graph_of, const int *tdim, K16Graph single,
                                                        int single_T) {
  const int lane = threadIdx.x & 63;
  auto gidx = [&](long long c) { return graph_of ? graph_of[c] : 0; };
  const int emax = (graph_of ? 32 : single_T) * kK;  // record values in use
  const long long waves = (long long)gridDim.x * (blockDim.x / 64);
  for (long long w = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
       w * kFixChunks < nchunks; w += waves) {
    const long long c0 = w * kFixChunks;
    const int n = (int)(c0 + kFixChunks < nchunks ? kFixChunks : nchunks - c0);
    // keys, graphs and tile dims of the wave's chunks, lane i holding chunk c0 + i
    const long long cl = c0 + (lane < n ? lane : n - 1);
    const int kl = carry_key[cl], gl = gidx(cl);
    const int tl = graph_of ? tdim[gl] : single_T;
    const int prev_g = c0 > 0 ? gidx(c0 - 1) : -1, prev_k = c0 > 0 ? carry_key[c0 - 1] : -1;
    const int next_g = c0 + n < nchunks ? gidx(c0 + n) : -1, next_k = c0 + n < nchunks ? carry_key[c0 + n] : -1;
    for (int e0 = 0; e0 < emax; e0 += 64) {
      const int e = e0 + lane, row = e / kK;  // the slice's 4 rows: `row < T` is wave-uniform
      float v[kFixChunks];
#pragma unroll
      for (int i = 0; i < kFixChunks; ++i) {
        const long long c = c0 + (i < n ? i : n - 1);
        v[i] = carry_val[c * 32 * kK + e];
      }
      int cg = __shfl(gl, 0), ck = __shfl(kl, 0), ct = __shfl(tl, 0);
      bool first = true;
      float acc = 0.0f;
      auto flush = [&](bool atomic) {
        if (ck < 0 || row >= ct || acc == 0.0f) return;
        float *y = (graph_of ? graphs[cg].y : single.y) + (long long)ck * ct * kK + e;
        if (atomic) unsafeAtomicAdd(y, acc);
        else *y += acc;
      };
#pragma unroll
      for (int i = 0; i < kFixChunks; ++i) {
        if (i >= n) break;  // wave-uniform
        const int g = __shfl(gl, i), k = __shfl(kl, i), t = __shfl(tl, i);
        if (g != cg || k != ck) {  // wave-uniform
          flush(first && cg == prev_g && ck == prev_k);
          first = false;
          cg = g;
          ck = k;
          ct = t;
          acc = 0.0f;
        }
        if (k >= 0 && row < t) acc += v[i];
      }
      flush((first && cg == prev_g && ck == prev_k) || (cg == next_g && ck == next_k));
    }
  }
}

// Zeroes the output rows of block rows without tiles (the walk steps over
// them): one thread per 16 consecutive outputs of a block row, over the
// block rows of every graph (row_start: prefix of the graphs' block-row
// counts; graphs == nullptr for one graph, `single`).
__global__ __launch_bounds__(256) void k16_zero_empty_kernel(const K16Graph *graphs, const long long *row_start,
                                                             const int *tdim, int ngraphs, K16Graph single,
                                                             int single_T) {
  const long long total = graphs ? row_start[ngraphs] * 32 : (long long)single.nbr * single_T;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    K16Graph G = single;
    int T = single_T;
    long long r = t;  // (block row, output row) index within the graph
    if (graphs) {
      // 32 slots per block row whatever T (T = 16 uses the first 16)
      const long long br = t / 32;
      int lo = 0, hi = ngraphs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_start[mid] <= br) lo = mid;
        else hi = mid - 1;
      }
      G = graphs[lo];
      T = tdim[lo];
      const int j = (int)(t % 32);
      if (j >= T) continue;
      r = (br - row_start[lo]) * T + j;
    }
    const long long I = r / T;
    if (G.rowptr[I] != G.rowptr[I + 1]) continue;
    float4_t *y = reinterpret_cast<float4_t *>(G.y + r * kK);
#pragma unroll
    for (int k = 0; k < kK / 4; ++k) y[k] = float4_t{0.f, 0.f, 0.f, 0.f};
  }
}

// Persistent launch over many graphs: waves pull chunk indices from a queue.
struct K16BatchArgs {
  const K16Graph *graphs;
  const int *tdim;
  const long long *chunk_start;  // [ngraphs + 1] prefix of chunk counts
  int ngraphs;
  long long cbegin, cend;      // this launch's chunks: all of one tile dim
  unsigned long long *queue;  // next chunk index, relative to cbegin
  int *carry_key;
  float *carry_val;
  int *graph_of;
};

// Persistent launch over many graphs. Thread 0 of a workgroup takes one
// ticket of kWaves chunks from the global queue; the ticket reaches the
// workgroup's waves through LDS between two barriers, so every wave runs the
// loop in uniform control flow (no lane-0-only region feeds a wave-wide
// shuffle) and the whole workgroup leaves the loop together once the queue
// is drained. Graphs are ordered by tile dim and each tile dim gets its own
// launch (its own register budget: the T = 16 walk needs far fewer VGPRs
// than the T = 32 one, so a shared kernel ran T = 16 graphs at T = 32's
// occupancy).
template <int T>
__global__ __launch_bounds__(64 * kWaves) void k16_persistent_kernel(K16BatchArgs a) {
  __shared__ unsigned long long s_ticket;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) s_ticket = atomicAdd(a.queue, (unsigned long long)kWaves);
    __syncthreads();
    const unsigned long long base = s_ticket + (unsigned long long)a.cbegin;
    __syncthreads();  // every thread has read the ticket before the next one is taken
    if (base >= (unsigned long long)a.cend) break;  // uniform per workgroup
    // the chunk and its graph are wave-uniform: kept in scalar registers, so the
    // graph's pointers feed the buffer resources of the X gathers directly
    const unsigned long long c =
        ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32 |
         (unsigned)__builtin_amdgcn_readfirstlane((int)base)) +
        (unsigned)__builtin_amdgcn_readfirstlane(wave);
    if (c < (unsigned long long)a.cend) {  // uniform per wave
      int lo = 0, hi = a.ngraphs - 1;  // graph with chunk_start[g] <= c < chunk_start[g+1]
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.chunk_start[mid] <= (long long)c) lo = mid;
        else hi = mid - 1;
      }
      lo = __builtin_amdgcn_readfirstlane(lo);
      const K16Graph G = a.graphs[lo];
      const long long k0 = ((long long)c - a.chunk_start[lo]) * kChunk;
      const long long k1 = k0 + kChunk < G.nblocks ? k0 + kChunk : G.nblocks;
      float *carry = a.carry_val + (long long)c * 32 * kK;
      const int open = k16_chunk<T>(G, k0, k1, lane, carry);
      // written by every lane (same value): no lane-0-only region in the loop
      a.carry_key[c] = open < G.nbr ? open : -1;
      a.graph_of[c] = lo;
    }
  }
}

// Xb[(b*64 + l)*4 + j] = X[(16b + 4(l/16) + j)*16 + l%16], zero beyond n.
__global__ void k16_pack_rhs_kernel(const uint16_t *X, long long n, long long nblk16, uint16_t *Xb) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nblk16 * 256;
       t += (long long)gridDim.x * blockDim.x) {
    const long long b = t / 256;
    const int r = (int)(t % 256), l = r / 4, j = r % 4;
    const long long v = 16 * b + 4 * (l / 16) + j;
    Xb[t] = v < n ? X[v * kK + (l % 16)] : (uint16_t)0;
  }
}

template <class V>
__global__ void k16_fill_kernel(V *p, long long n, V v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// grids of the grid-stride kernels above
unsigned nb256(long long n) { return capped_blocks(n, 256); }

int k16_check(int T, const void *A, const void *Xb, const float *Y, const int *rowptr, const int *colind, int nbr,
              int nblocks) {
  BGB_REQUIRE(T == 16 || T == 32, BGB_ERR_TILEDIM, "K=16 MFMA path supports tile dims 16 and 32 (got %d)", T);
  BGB_REQUIRE(Y && rowptr && nbr >= 0 && nblocks >= 0, BGB_ERR_ARG, "bad output / sizes");
  BGB_REQUIRE(nblocks == 0 || (A && Xb && colind), BGB_ERR_ARG, "null matrix / RHS");
  BGB_REQUIRE(((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(Xb)) & 15) == 0, BGB_ERR_ARG,
              "bsrVal and the packed RHS must be 16-byte aligned");
  return BGB_OK;
}

}  // namespace
}  // namespace bgb

using namespace bgb;

extern "C" {

int bgb_k16_pack_rhs(const uint16_t *d_X, int n, void *d_Xb, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(n >= 0 && d_Xb && (n == 0 || d_X), BGB_ERR_ARG, "bad RHS arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long nb16 = ceil_div(n, 16);
    k16_pack_rhs_kernel<<<nb256(nb16 * 256), 256, 0, s>>>(d_X, n, nb16, static_cast<uint16_t *>(d_Xb));
    BGB_HIP_TRY(hipGetLastError());
    return BGB_OK;
  });
}

int bgb_bmv_bin_full_full_k16(int T, const void *d_A, const void *d_Xb, float *d_Y, const int *d_rowptr,
                              const int *d_colind, int nblockrows, int nblocks, int accumulate,
                              bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    int st = k16_check(T, d_A, d_Xb, d_Y, d_rowptr, d_colind, nblockrows, nblocks);
    if (st) return st;
    BGB_REQUIRE(accumulate == 0 || accumulate == 1, BGB_ERR_ARG, "accumulate must be 0 or 1");
    if (nblockrows == 0) return BGB_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nblocks == 0) {
      if (!accumulate)
        k16_fill_kernel<float><<<nb256((long long)nblockrows * T * kK), 256, 0, s>>>(d_Y, (long long)nblockrows * T * kK, 0.f);
      BGB_HIP_TRY(hipGetLastError());
      return BGB_OK;
    }
    const long long nchunks = ceil_div(nblocks, kChunk);
    const size_t key_bytes = (sizeof(int) * (size_t)nchunks + 255) & ~(size_t)255;
    void *ws = nullptr;
    if ((st = stream_workspace(s, key_bytes + sizeof(float) * 32 * kK * (size_t)nchunks, &ws))) return st;
    K16Args a;
    a.g = K16Graph{d_rowptr, d_colind, d_A, static_cast<const uint2 *>(d_Xb), d_Y, nblockrows, nblocks, accumulate};
    a.nchunks = (int)nchunks;
    a.carry_key = static_cast<int *>(ws);
    a.carry_val = reinterpret_cast<float *>(static_cast<char *>(ws) + key_bytes);
    const unsigned blocks = (unsigned)ceil_div(nchunks, kWaves);
    if (T == 16) k16_kernel<16><<<blocks, 64 * kWaves, 0, s>>>(a);
    else k16_kernel<32><<<blocks, 64 * kWaves, 0, s>>>(a);
    BGB_HIP_TRY(hipGetLastError());
    if (!accumulate)
      k16_zero_empty_kernel<<<nb256((long long)nblockrows * T), 256, 0, s>>>(nullptr, nullptr, nullptr, 0, a.g, T);
    k16_fixup_kernel<<<nb256(ceil_div(nchunks, kFixChunks) * 64), 256, 0, s>>>(a.carry_key, a.carry_val, nchunks, nullptr, nullptr,
                                                             nullptr, a.g, T);
    BGB_HIP_TRY(hipGetLastError());
    return BGB_OK;
  });
}

int bgb_k16_batched(int ngraphs, const bgb_k16_graph *h_graphs, const int *d_rowptr, const int *d_colind,
                    const void *d_val, const void *d_Xb, float *d_Y, int nworkgroups, bgb_stream_t stream) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(ngraphs >= 1 && h_graphs && d_rowptr && d_Y, BGB_ERR_ARG, "bad batch arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<K16Graph> gs(ngraphs);
    std::vector<int> td(ngraphs);
    std::vector<long long> cs(ngraphs + 1, 0), rs(ngraphs + 1, 0);
    // T = 16 graphs first, then T = 32: one contiguous chunk range per tile dim
    std::vector<int> order;
    order.reserve(ngraphs);
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < ngraphs; ++i)
        if ((h_graphs[i].T == 32) == (pass == 1)) order.push_back(i);
    int n16 = 0;
    for (int i = 0; i < ngraphs; ++i) n16 += h_graphs[i].T != 32;
    for (int slot = 0; slot < ngraphs; ++slot) {
      const int i = order[slot];
      const bgb_k16_graph &h = h_graphs[i];
      BGB_REQUIRE(h.T == 16 || h.T == 32, BGB_ERR_TILEDIM, "graph %d: tile dim %d not in {16, 32}", i, h.T);
      BGB_REQUIRE(h.n >= 1 && h.nblocks >= 0 && h.val_byte_off % 16 == 0 && h.x_off % 256 == 0, BGB_ERR_ARG,
                  "graph %d: bad sizes or misaligned offsets", i);
      const int nbr = (int)ceil_div(h.n, h.T);
      gs[slot] = K16Graph{d_rowptr + h.rowptr_off, d_colind + h.tile_off,
                          static_cast<const char *>(d_val) + h.val_byte_off,
                          reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(d_Xb) + h.x_off),
                          d_Y + h.y_off, nbr, h.nblocks, 0};
      td[slot] = h.T;
      cs[slot + 1] = cs[slot] + (h.nblocks > 0 ? ceil_div(h.nblocks, kChunk) : 0);
      rs[slot + 1] = rs[slot] + nbr;
      if (h.nblocks == 0)
        k16_fill_kernel<float><<<nb256((long long)nbr * h.T * kK), 256, 0, s>>>(d_Y + h.y_off, (long long)nbr * h.T * kK, 0.f);
    }
    const long long total = cs[ngraphs];
    DeviceBuffer dg, dt, dc, dr, q, ck, cv, go;
    BGB_HIP_TRY(dg.alloc(sizeof(K16Graph) * ngraphs, s));
    BGB_HIP_TRY(dt.alloc(sizeof(int) * ngraphs, s));
    BGB_HIP_TRY(dc.alloc(sizeof(long long) * (ngraphs + 1), s));
    BGB_HIP_TRY(dr.alloc(sizeof(long long) * (ngraphs + 1), s));
    BGB_HIP_TRY(q.alloc(2 * sizeof(unsigned long long), s));  // one queue per tile dim
    BGB_HIP_TRY(ck.alloc(sizeof(int) * (total + 1), s));
    BGB_HIP_TRY(cv.alloc(sizeof(float) * 32 * kK * (total + 1), s));
    BGB_HIP_TRY(go.alloc(sizeof(int) * (total + 1), s));
    BGB_HIP_TRY(hipMemcpyAsync(dg.ptr, gs.data(), sizeof(K16Graph) * ngraphs, hipMemcpyHostToDevice, s));
    BGB_HIP_TRY(hipMemcpyAsync(dt.ptr, td.data(), sizeof(int) * ngraphs, hipMemcpyHostToDevice, s));
    BGB_HIP_TRY(hipMemcpyAsync(dc.ptr, cs.data(), sizeof(long long) * (ngraphs + 1), hipMemcpyHostToDevice, s));
    BGB_HIP_TRY(hipMemcpyAsync(dr.ptr, rs.data(), sizeof(long long) * (ngraphs + 1), hipMemcpyHostToDevice, s));
    // the walk steps over block rows without tiles: zero their outputs
    k16_zero_empty_kernel<<<nb256(rs[ngraphs] * 32), 256, 0, s>>>(dg.as<K16Graph>(), dr.as<long long>(), dt.as<int>(),
                                                                 ngraphs, K16Graph{}, 0);
    k16_fill_kernel<unsigned long long><<<1, 64, 0, s>>>(q.as<unsigned long long>(), 2, 0ull);
    BGB_HIP_TRY(hipGetLastError());
    if (total > 0) {
      const long long split = cs[n16];  // chunks of the T = 16 graphs
      for (int t = 0; t < 2; ++t) {
        const long long cb = t == 0 ? 0 : split, ce = t == 0 ? split : total;
        if (ce <= cb) continue;
        K16BatchArgs a{dg.as<K16Graph>(), dt.as<int>(), dc.as<long long>(), ngraphs, cb, ce,
                       q.as<unsigned long long>() + t, ck.as<int>(), cv.as<float>(), go.as<int>()};
        int wgs = nworkgroups > 0 ? nworkgroups : 2048;
        if ((long long)wgs * kWaves > ce - cb) wgs = (int)ceil_div(ce - cb, kWaves);
        if (t == 0) k16_persistent_kernel<16><<<wgs, 64 * kWaves, 0, s>>>(a);
        else k16_persistent_kernel<32><<<wgs, 64 * kWaves, 0, s>>>(a);
        BGB_HIP_TRY(hipGetLastError());
      }
      K16Graph none{};
      k16_fixup_kernel<<<nb256(ceil_div(total, kFixChunks) * 64), 256, 0, s>>>(ck.as<int>(), cv.as<float>(), total, dg.as<K16Graph>(),
                                                             go.as<int>(), dt.as<int>(), none, 0);
      BGB_HIP_TRY(hipGetLastError());
    }
    BGB_HIP_TRY(hipStreamSynchronize(s));
    return BGB_OK;
  });
}

}  // extern "C"
