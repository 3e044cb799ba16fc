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

}  // namespace

extern "C" {

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
