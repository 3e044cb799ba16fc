// Per-graph tile-dim autotuner (host side).
//
// The reference leaves T as a compile-time choice (-DTILEDIM, Makefile:12-46)
// and has tile statistics only as debug kernels (printMatrixStats_T,
// estimatevec, utility.cu:808-975, the latter an upper bound: distinct block
// columns per CSR row). Here the statistics are exact -- tiles per T counted
// as distinct (I, J) pairs per block row -- and feed a cost model of the BMV
// kernels: predicted time = max(HBM bytes / bandwidth, tiles * issue cost)
// + a per-block-row term. The constants are fitted to MI355X measurements of
// the chunked kernels (per-tile issue cost per T and op).
#include <omp.h>

#include <algorithm>
#include <vector>

#include "bgb_internal.h"

namespace {

const int kTs[4] = {4, 8, 16, 32};

// Bytes and issue constants per T (index 0..3 = T 4/8/16/32), per op family.
// tile_ns: whole-chip time per tile of the chunked kernels (ns);
// bw_gbs: sustained HBM bandwidth of the streaming loop.
struct Model {
  double tile_ns[4];
  double row_ns[4];
  double bw_gbs;
};
// op 0: bin_bin_full / bin_bin_bin with x gathered from memory, op 1:
// bin_full_full (FP32), op 2: K=16 MFMA, op 3: one level of the batched
// 64-source BFS. Fitted to MI355X runs on R-MAT scale 20 / 22 (bit ops) and the
// 4096^2 grid (FP32): T <= 16 is bound by per-tile issue and latency, T = 32 by
// HBM at ~5.5 TB/s.
const Model kModels[4] = {
    {{0.0045, 0.0048, 0.0091, 0.0200}, {0.0010, 0.0010, 0.0010, 0.0010}, 5500.0},
    {{0.0055, 0.0095, 0.0450, 0.0580}, {0.0010, 0.0010, 0.0010, 0.0010}, 4000.0},
    {{1e9, 1e9, 0.0600, 0.2000}, {1e9, 1e9, 0.0300, 0.0600}, 4000.0},
    {{0.0117, 0.0160, 0.0300, 0.0700}, {0.0050, 0.0080, 0.0150, 0.0300}, 4000.0},
};
// op 0 when the packed x is staged in LDS (persistent workgroups, T <= 16):
// per-tile costs measured on R-MAT scale 20.
const double kTileNsLds[4] = {0.0035, 0.0044, 0.0092, 0.0200};
// LDS a workgroup may use on gfx950, minus the per-wave row windows and
// group-row slots (16 waves x 1280 B).
constexpr long long kLdsForX = 160 * 1024 - 16 * 1280;

// Packed-x bytes the staging kernel keeps in LDS for n rows at tile dim T
// (T = 4: two nibbles per byte).
long long lds_x_bytes(int T, long long nrows) {
  const long long nbr = (nrows + T - 1) / T;
  const long long wb = T == 32 ? 4 : T == 16 ? 2 : 1;
  const long long x = (nbr + T - 1) / T * T * wb;
  return T == 4 ? (x + 1) / 2 : x;
}

double op_bytes(int op, int T, long long nrows, long long nblocks) {
  const double wb = T == 32 ? 4 : T == 16 ? 2 : 1;
  const double nbr = (double)((nrows + T - 1) / T);
  const double tiles = (double)nblocks * (4.0 + T * wb);
  const double rp = 4.0 * (nbr + 1);
  if (op == 0) return tiles + rp + nbr * T * wb / T + 4.0 * nbr * T;  // packed x + int32 y
  if (op == 1) return tiles + rp + 8.0 * nbr * T;                     // FP32 x + y
  if (op == 3) return tiles + rp + 24.0 * nbr * T;                    // frontier, visited, next
  return tiles + rp + nbr * T * 16 * (2.0 + 4.0);                     // FP16 X + FP32 Y
}


// Tile profile of a host CSR at every T: tiles and non-empty block rows
// (distinct block columns per block row), the longest block row, and the
// non-zero tile row words (distinct block columns per CSR row: the sum of
// printMatrixStats_T's nnz_tile_per_row, utility.cu:808-930, and
// estimatevec's count, :955-975, for sorted rows).
int tile_profile(int nrows, int ncols, const int *rp, const int *ci, bgb_tile_profile *p) {
  p->nrows = nrows;
  p->ncols = ncols;
  p->nnz = rp[nrows];
  for (int t = 0; t < 4; ++t) {
    const int T = kTs[t];
    const long long nbr = (nrows + T - 1) / T;
    long long tiles = 0, nonempty = 0, most = 0, words = 0;
#pragma omp parallel reduction(+ : tiles, nonempty, words) reduction(max : most)
    {
      std::vector<int> buf;
#pragma omp for schedule(dynamic, 256)
      for (long long I = 0; I < nbr; ++I) {
        const int r0 = (int)(I * T), r1 = (int)std::min<long long>((I + 1) * T, nrows);
        buf.clear();
        for (int r = r0; r < r1; ++r) {
          int last = -1;
          for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int J = ci[k] / T;
            buf.push_back(J);
            words += J != last;  // rows are sorted: a new block column starts a word
            last = J;
          }
        }
        if (buf.empty()) continue;
        std::sort(buf.begin(), buf.end());
        const long long d = std::unique(buf.begin(), buf.end()) - buf.begin();
        tiles += d;
        ++nonempty;
        most = std::max(most, d);
      }
    }
    p->tiles[t] = tiles;
    p->nonempty_rows[t] = nonempty;
    p->max_row_tiles[t] = most;
    p->row_words[t] = words;
  }
  return BGB_OK;
}

// ---------------------------------------------------------------------------
// Profile-driven model. Each (op, T, run length) is predicted from the kernel
// path it takes on this matrix -- the short-row kernel (T <= 8, no block row
// over kShortRowMax tiles), the chunked walk with the packed x staged in LDS
// or gathered from memory, the FP sums -- as
//   max(bytes / bandwidth, c0 + a * tiles),
// constants fitted to MI355X runs of these kernels on R-MAT, Erdos-Renyi and
// grid graphs (us, tiles in millions).
constexpr long long kShortRowMax = 32;
struct PathFit {
  double c0;
  double per_mtile[4];
};
const PathFit kBitShort = {5.0, {1.46, 2.2, 0, 0}};
const PathFit kBitLds = {15.0, {3.2, 3.6, 9.0, 0}};
const PathFit kBitMem = {15.0, {4.3, 4.6, 9.3, 24.0}};
const PathFit kFpShort = {4.0, {3.5, 4.4, 0, 0}};
const PathFit kFpChunk = {12.0, {6.0, 7.7, 27.0, 52.0}};

// Bits from memory past the staged first words of a packed x larger than
// LDS (R-MAT 22: about 55 % of the tiles' block columns staged).
const PathFit kBitPart = {15.0, {3.25, 3.7, 8.3, 0}};
// bin_bin_full at T <= 8 with the whole packed x in LDS: the row-piece path
// (R-MAT 20: T = 4 59.4 us, T = 8 74.2 us).
const PathFit kBitPiece = {8.0, {1.89, 2.79, 0, 0}};
// The row-piece path with x larger than LDS: its first words and a bitmap of
// the others' non-zero words in LDS (R-MAT 22: T = 4 334 us, T = 8 381 us).
const PathFit kBitPieceBm = {8.0, {2.82, 3.56, 0, 0}};
// Packed-x words whose bitmap (one bit each, with a staged prefix) fits in
// the piece kernel's 160 KB of LDS.
constexpr long long kBitmapWords = 8ll * 160 * 1024;

// Packed-x placement of the bit op: 0 gathered from memory, 1 all of it in
// LDS, 2 its first words in LDS (the launch rule of the BMV kernels: T <= 16
// and a tile stream at least twice the LDS copies of one staging per CU).
int lds_path(int T, const bgb_tile_profile &p, int cus = 256) {
  const int t = T == 4 ? 0 : T == 8 ? 1 : T == 16 ? 2 : 3;
  if (T > 16) return 0;
  const bool fits = lds_x_bytes(T, p.nrows) <= kLdsForX;
  const double staged = fits ? (double)lds_x_bytes(T, p.nrows) : (double)kLdsForX;
  const double wb = T == 16 ? 2 : 1;
  if ((double)p.tiles[t] * (4.0 + T * wb) < 2.0 * staged * cus) return 0;
  if (!fits && 8.0 * (double)kLdsForX < (double)lds_x_bytes(T, p.nrows)) return 0;  // prefix under 1/8 of x
  return fits ? 1 : 2;
}

// Predicted microseconds of one call of op at tile dim index t, run length S.
double predict_us(int op, int t, int S, const bgb_tile_profile &p, int *x_staging) {
  const int T = kTs[t];
  const double mt = (double)p.tiles[t] / 1e6;
  const long long nbr = (p.nrows + T - 1) / T;
  const double mean = (double)p.tiles[t] / (double)std::max<long long>(p.nonempty_rows[t], 1);
  const bool uniform = (double)p.max_row_tiles[t] <= 4.0 * mean;
  const bool short_rows = T <= 8 && p.max_row_tiles[t] <= kShortRowMax;
  if (x_staging) *x_staging = 1;
  double issue;
  if (op == 0 || op == 1) {
    const PathFit *f;
    if (short_rows) f = op == 0 ? &kBitShort : &kFpShort;
    else if (op == 0 && T <= 8 && !uniform && lds_x_bytes(T, p.nrows) > kLdsForX &&
             (nbr + T - 1) / T * T < kBitmapWords) {
      // (skewed rows: long ones fill their pieces; uniform rows of a few
      // dozen tiles leave pieces mostly empty and keep the chunk walk)
      f = &kBitPieceBm;
      if (x_staging) *x_staging = 2;
    } else if (op == 0 && lds_path(T, p)) {
      f = lds_path(T, p) == 1 ? (T <= 8 ? &kBitPiece : &kBitLds) : &kBitPart;
      if (x_staging) *x_staging = 2;
    } else f = op == 0 ? &kBitMem : &kFpChunk;
    double a = f->per_mtile[t];
    // FP x segments gathered from beyond the L2 with no reuse (uniform
    // columns, x larger than the L2s hold): measured 2.0-2.4x the per-tile cost
    if (op == 1 && !short_rows && T <= 8 && uniform && p.nrows >= (2 << 20)) a *= 2.2;
    issue = f->c0 + a * mt;
    if (!short_rows && S == 16) {
      // runs of 16 tiles: slower at T <= 8 (register budget), faster at T = 16
      // for long uniform rows and at T = 32 for large matrices with rows of
      // more than a few tiles (fewer chunks and carries)
      const bool big = p.tiles[t] >= 5000000;
      const double f16 = T <= 8 ? 1.25 : T == 16 ? ((big && uniform && mean >= 64) ? 0.88 : 1.1)
                                                 : ((big && mean >= 16) ? 0.95 : 1.05);
      issue *= f16;
    }
  } else {
    const Model &m = kModels[op];
    issue = ((double)p.tiles[t] * m.tile_ns[t] + (double)nbr * m.row_ns[t]) / 1e3 + 3.0;
  }
  const double bw = op == 0 ? 5500.0 : 5000.0;
  const double mem_us = op_bytes(op, T, p.nrows, p.tiles[t]) / (bw * 1e3);
  return std::max(issue, mem_us);
}

}  // namespace

extern "C" {

int bgb_autotune_profile(int op, const bgb_tile_profile *prof, int allowed_mask, int *T_out, int *run_length_out,
                         int *x_staging_out, double *cost4_us) {
  return bgb::host_guard([&]() -> int {
    bgb::clear_error();
    BGB_REQUIRE(op >= 0 && op <= 3 && prof && T_out, BGB_ERR_ARG, "bad arguments");
    if (op == 2) allowed_mask &= 0xC;  // the MFMA path takes T = 16, 32
    BGB_REQUIRE(allowed_mask & 0xF, BGB_ERR_ARG, "no tile dim allowed");
    double best = 1e300;
    for (int t = 0; t < 4; ++t) {
      int xs8 = 1, xs16 = 1;
      const double c8 = predict_us(op, t, 8, *prof, &xs8);
      // run lengths are a knob of the bit and FP BMV kernels only
      const double c16 = op <= 1 ? predict_us(op, t, 16, *prof, &xs16) : 1e300;
      const double c = std::min(c8, c16);
      if (cost4_us) cost4_us[t] = c;
      if (((allowed_mask >> t) & 1) && c < best) {
        best = c;
        *T_out = kTs[t];
        if (run_length_out) *run_length_out = c16 < c8 ? 16 : 8;
        if (x_staging_out) *x_staging_out = c16 < c8 ? xs16 : xs8;
      }
    }
    return BGB_OK;
  });
}


int bgb_tile_profile_csr(int nrows, int ncols, const int *h_rowptr, const int *h_colind, bgb_tile_profile *out) {
  return bgb::host_guard([&]() -> int {
    bgb::clear_error();
    BGB_REQUIRE(nrows >= 0 && ncols >= 0 && h_rowptr && out, BGB_ERR_ARG, "bad arguments");
    BGB_REQUIRE(h_rowptr[nrows] == 0 || h_colind, BGB_ERR_ARG, "null colind");
    return tile_profile(nrows, ncols, h_rowptr, h_colind, out);
  });
}

int bgb_tile_stats(int nrows, int ncols, const int *h_rowptr, const int *h_colind, long long *h_nblocks4,
                   long long *h_nonempty4) {
  return bgb::host_guard([&]() -> int {
    bgb::clear_error();
    BGB_REQUIRE(nrows >= 0 && ncols >= 0 && h_rowptr && h_nblocks4, BGB_ERR_ARG, "bad arguments");
    BGB_REQUIRE(h_rowptr[nrows] == 0 || h_colind, BGB_ERR_ARG, "null colind");
    for (int t = 0; t < 4; ++t) {
      const int T = kTs[t];
      const long long nbr = (nrows + T - 1) / T;
      long long tiles = 0, nonempty = 0;
  #pragma omp parallel reduction(+ : tiles, nonempty)
      {
        std::vector<int> buf;
  #pragma omp for schedule(dynamic, 256)
        for (long long I = 0; I < nbr; ++I) {
          const int r0 = (int)(I * T), r1 = (int)std::min<long long>((I + 1) * T, nrows);
          buf.clear();
          for (int k = h_rowptr[r0]; k < h_rowptr[r1]; ++k) buf.push_back(h_colind[k] / T);
          if (buf.empty()) continue;
          std::sort(buf.begin(), buf.end());
          tiles += std::unique(buf.begin(), buf.end()) - buf.begin();
          ++nonempty;
        }
      }
      h_nblocks4[t] = tiles;
      if (h_nonempty4) h_nonempty4[t] = nonempty;
    }
    return BGB_OK;
  });
}

int bgb_autotune(int op, int nrows, const long long *h_nblocks4, int allowed_mask, int *T_out,
                 int *run_length_out, int *x_staging_out, double *cost4_us) {
  return bgb::host_guard([&]() -> int {
    int T = 0;
    const int st = bgb_autotune_tiledim(op, nrows, h_nblocks4, allowed_mask, &T, cost4_us);
    if (st) return st;
    if (T_out) *T_out = T;
    // run length 16 never beat 8 on MI355X (T = 4, 8 slower; 16, 32 equal)
    if (run_length_out) *run_length_out = 8;
    if (x_staging_out) *x_staging_out = (op == 0 && T <= 16 && lds_x_bytes(T, nrows) <= kLdsForX) ? 2 : 1;
    return BGB_OK;
  });
}

int bgb_autotune_tiledim(int op, int nrows, const long long *h_nblocks4, int allowed_mask, int *T_out,
                         double *cost4_us) {
  return bgb::host_guard([&]() -> int {
    bgb::clear_error();
    BGB_REQUIRE(op >= 0 && op <= 3 && nrows >= 0 && h_nblocks4 && T_out, BGB_ERR_ARG, "bad arguments");
    if (op == 2) allowed_mask &= 0xC;  // the MFMA path takes T = 16, 32
    BGB_REQUIRE(allowed_mask & 0xF, BGB_ERR_ARG, "no tile dim allowed");
    const Model &m = kModels[op];
    double best = 1e300;
    for (int t = 0; t < 4; ++t) {
      const int T = kTs[t];
      const double nbr = (double)((nrows + T - 1) / T);
      const double mem_us = op_bytes(op, T, nrows, h_nblocks4[t]) / (m.bw_gbs * 1e3);
      const bool lds = op == 0 && T <= 16 && lds_x_bytes(T, nrows) <= kLdsForX;
      const double tile_ns = lds ? kTileNsLds[t] : m.tile_ns[t];
      const double issue_us = ((double)h_nblocks4[t] * tile_ns + nbr * m.row_ns[t]) / 1e3;
      const double c = std::max(mem_us, issue_us) + 3.0;  // + launch overhead
      if (cost4_us) cost4_us[t] = c;
      if ((allowed_mask >> t) & 1) {
        if (c < best) {
          best = c;
          *T_out = T;
        }
      }
    }
    return BGB_OK;
  });
}

}  // extern "C"
