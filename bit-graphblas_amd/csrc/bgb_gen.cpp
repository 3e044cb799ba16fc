// Host-side graph generators and CSR normalisation for libbgb.
//
// The normalisation follows the reference's CSR construction
// (readMtx.hpp:152-218 removeSelfloop, :56-82 customSort, :248-306 coo2csr):
// an undirected input gets the reverse of every non-loop edge, self-loops
// are dropped when asked, duplicates are removed and rows are sorted by
// column. It is written as a parallel bucket scatter + per-row sort/unique,
// which yields the same CSR as the reference's global tuple sort.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <new>
#include <parallel/algorithm>
#include <string>
#include <vector>

#include "bgb_internal.h"

struct bgb_csr {
  int nrows = 0, ncols = 0;
  long long nnz = 0;
  std::vector<int> rowptr;
  std::vector<int> colind;
};

namespace bgb {

static thread_local std::string g_last_error;

int set_error(int status, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}
void clear_error() { g_last_error.clear(); }
const char *last_error() { return g_last_error.c_str(); }

// Builds a normalised CSR from an edge list (see file header).
static int build_csr(int nrows, int ncols, long long nedges, const int *src, const int *dst,
                     bool undirected, bool drop_self, bgb_csr **out) {
  BGB_REQUIRE(nrows >= 0 && ncols >= 0 && nedges >= 0, BGB_ERR_ARG, "negative size");
  BGB_REQUIRE(nedges == 0 || (src && dst), BGB_ERR_ARG, "null edge arrays");
  BGB_REQUIRE(!undirected || nrows == ncols, BGB_ERR_ARG,
              "an undirected graph needs nrows == ncols");
  long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long e = 0; e < nedges; ++e)
    bad += (src[e] < 0 || src[e] >= nrows || dst[e] < 0 || dst[e] >= ncols);
  BGB_REQUIRE(bad == 0, BGB_ERR_ARG, "%lld edge endpoints out of range", bad);

  std::vector<long long> cnt(nrows + 1, 0);
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nedges; ++e) {
    int u = src[e], v = dst[e];
    if (drop_self && u == v) continue;
    __atomic_fetch_add(&cnt[u], 1, __ATOMIC_RELAXED);
    if (undirected && u != v) __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
  }
  std::vector<long long> off(nrows + 1, 0);
  for (int i = 0; i < nrows; ++i) off[i + 1] = off[i] + cnt[i];
  std::vector<int> buf;
  try {
    buf.resize(off[nrows]);
  } catch (const std::bad_alloc &) {
    return set_error(BGB_ERR_ALLOC, "host allocation of %lld entries failed", off[nrows]);
  }
  std::vector<long long> cur(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < nedges; ++e) {
    int u = src[e], v = dst[e];
    if (drop_self && u == v) continue;
    buf[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = v;
    if (undirected && u != v) buf[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = u;
  }
  std::vector<long long> ucnt(nrows, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < nrows; ++i) {
    int *b = buf.data() + off[i], *e = buf.data() + off[i + 1];
    std::sort(b, e);
    ucnt[i] = std::unique(b, e) - b;
  }
  bgb_csr *g = new (std::nothrow) bgb_csr();
  BGB_REQUIRE(g, BGB_ERR_ALLOC, "host allocation failed");
  g->nrows = nrows;
  g->ncols = ncols;
  g->rowptr.assign(nrows + 1, 0);
  long long total = 0;
  for (int i = 0; i < nrows; ++i) {
    if (total > INT32_MAX) break;
    g->rowptr[i] = (int)total;
    total += ucnt[i];
  }
  if (total > INT32_MAX) {
    delete g;
    return set_error(BGB_ERR_SIZE, "nnz %lld exceeds the int32 index range", total);
  }
  g->rowptr[nrows] = (int)total;
  g->nnz = total;
  g->colind.resize(total);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < nrows; ++i)
    std::memcpy(g->colind.data() + g->rowptr[i], buf.data() + off[i], sizeof(int) * ucnt[i]);
  *out = g;
  return BGB_OK;
}

}  // namespace bgb

using namespace bgb;

extern "C" {

const char *bgb_last_error(void) { return bgb::last_error(); }

int bgb_csr_from_coo(int nrows, int ncols, long long nedges, const int *h_src, const int *h_dst,
                     int undirected, int drop_self_loops, bgb_csr **out) {
  return host_guard([&]() -> int {
    BGB_REQUIRE(out, BGB_ERR_ARG, "out is null");
    return build_csr(nrows, ncols, nedges, h_src, h_dst, undirected != 0, drop_self_loops != 0, out);
  });
}

int bgb_gen_er(int n, long long m, uint64_t seed, bgb_csr **out) {
  return host_guard([&]() -> int {
    BGB_REQUIRE(out, BGB_ERR_ARG, "out is null");
    BGB_REQUIRE(n >= 2 && m >= 0, BGB_ERR_ARG, "need n >= 2 and m >= 0");
    BGB_REQUIRE((double)m <= 0.25 * (double)n * (double)(n - 1), BGB_ERR_ARG,
                "m too large for rejection sampling (m <= n(n-1)/4)");
    const uint64_t key = stream_key(seed, 0xE5);
    // Candidate i is a uniform unordered pair without loop; the graph is the
    // first m distinct candidates in index order.
    long long ncand = m + m / 8 + 1024;
    std::vector<long long> keep;
    for (;;) {
      std::vector<std::pair<uint64_t, long long>> cand(ncand);
  #pragma omp parallel for schedule(static)
      for (long long i = 0; i < ncand; ++i) {
        uint64_t r = rand_at(key, (uint64_t)i);
        uint64_t u = ((r >> 32) * (uint64_t)n) >> 32;
        uint64_t v = ((r & 0xffffffffull) * (uint64_t)n) >> 32;
        if (u == v) {
          cand[i] = {~0ull, i};
        } else {
          if (u > v) std::swap(u, v);
          cand[i] = {(u << 32) | v, i};
        }
      }
      __gnu_parallel::sort(cand.begin(), cand.end());
      std::vector<long long> firsts;
      firsts.reserve(ncand);
      for (long long i = 0; i < ncand; ++i)
        if (cand[i].first != ~0ull && (i == 0 || cand[i].first != cand[i - 1].first))
          firsts.push_back(cand[i].second);
      if ((long long)firsts.size() >= m) {
        __gnu_parallel::sort(firsts.begin(), firsts.end());
        firsts.resize(m);
        keep.swap(firsts);
        break;
      }
      ncand *= 2;
    }
    std::vector<int> src(m), dst(m);
  #pragma omp parallel for schedule(static)
    for (long long e = 0; e < m; ++e) {
      uint64_t r = rand_at(key, (uint64_t)keep[e]);
      src[e] = (int)(((r >> 32) * (uint64_t)n) >> 32);
      dst[e] = (int)(((r & 0xffffffffull) * (uint64_t)n) >> 32);
    }
    return build_csr(n, n, m, src.data(), dst.data(), true, true, out);
  });
}

int bgb_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 bgb_csr **out) {
  return host_guard([&]() -> int {
    BGB_REQUIRE(out, BGB_ERR_ARG, "out is null");
    BGB_REQUIRE(scale >= 1 && scale <= 30 && edge_factor >= 1, BGB_ERR_ARG,
                "need 1 <= scale <= 30, edge_factor >= 1");
    BGB_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, BGB_ERR_ARG,
                "bad R-MAT probabilities");
    const long long E = (long long)edge_factor << scale;
    const int n = 1 << scale;
    const uint64_t key = stream_key(seed, 0x3A7);
    const double ab = a + b, abc = a + b + c;
    std::vector<int> src, dst;
    try {
      src.resize(E);
      dst.resize(E);
    } catch (const std::bad_alloc &) {
      return set_error(BGB_ERR_ALLOC, "host allocation of %lld edges failed", E);
    }
  #pragma omp parallel for schedule(static)
    for (long long e = 0; e < E; ++e) {
      uint32_t u = 0, v = 0;
      for (int l = 0; l < scale; ++l) {
        double p = u01(rand_at(key, (uint64_t)e * (uint64_t)scale + (uint64_t)l));
        uint32_t bit = 1u << (scale - 1 - l);
        if (p < a) {
        } else if (p < ab) {
          v |= bit;
        } else if (p < abc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      src[e] = (int)u;
      dst[e] = (int)v;
    }
    return build_csr(n, n, E, src.data(), dst.data(), true, true, out);
  });
}

int bgb_gen_grid2d(int width, int height, bgb_csr **out) {
  return host_guard([&]() -> int {
    BGB_REQUIRE(out, BGB_ERR_ARG, "out is null");
    BGB_REQUIRE(width >= 1 && height >= 1, BGB_ERR_ARG, "grid dims must be >= 1");
    const long long n = (long long)width * height;
    const long long nnz = 4 * n - 2LL * width - 2LL * height;
    BGB_REQUIRE(n <= INT32_MAX && nnz <= INT32_MAX, BGB_ERR_SIZE, "grid too large");
    bgb_csr *g = new (std::nothrow) bgb_csr();
    BGB_REQUIRE(g, BGB_ERR_ALLOC, "host allocation failed");
    g->nrows = g->ncols = (int)n;
    g->nnz = nnz;
    g->rowptr.resize(n + 1);
    g->colind.resize(nnz);
    auto deg = [&](long long v) {
      long long y = v / width, x = v % width;
      return (y > 0) + (x > 0) + (x < width - 1) + (y < height - 1);
    };
    long long pos = 0;
    for (long long v = 0; v < n; ++v) {
      g->rowptr[v] = (int)pos;
      pos += deg(v);
    }
    g->rowptr[n] = (int)pos;
  #pragma omp parallel for schedule(static)
    for (long long v = 0; v < n; ++v) {
      long long y = v / width, x = v % width;
      int *o = g->colind.data() + g->rowptr[v];
      if (y > 0) *o++ = (int)(v - width);
      if (x > 0) *o++ = (int)(v - 1);
      if (x < width - 1) *o++ = (int)(v + 1);
      if (y < height - 1) *o++ = (int)(v + width);
    }
    *out = g;
    return BGB_OK;
  });
}

int bgb_csr_nrows(const bgb_csr *g) { return g ? g->nrows : -1; }
int bgb_csr_ncols(const bgb_csr *g) { return g ? g->ncols : -1; }
long long bgb_csr_nnz(const bgb_csr *g) { return g ? g->nnz : -1; }
const int *bgb_csr_rowptr(const bgb_csr *g) { return g ? g->rowptr.data() : nullptr; }
const int *bgb_csr_colind(const bgb_csr *g) { return g ? g->colind.data() : nullptr; }
void bgb_csr_free(bgb_csr *g) { delete g; }

int bgb_gen_bernoulli_f32(long long n, double p, uint64_t seed, float *h_out) {
  return host_guard([&]() -> int {
    BGB_REQUIRE(n >= 0 && (n == 0 || h_out), BGB_ERR_ARG, "bad output");
    const uint64_t key = stream_key(seed, 0xB1);
  #pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) h_out[i] = u01(rand_at(key, (uint64_t)i)) < p ? 1.0f : 0.0f;
    return BGB_OK;
  });
}

int bgb_gen_uniform_f32(long long n, double lo, double hi, uint64_t seed, float *h_out) {
  return host_guard([&]() -> int {
    BGB_REQUIRE(n >= 0 && (n == 0 || h_out), BGB_ERR_ARG, "bad output");
    const uint64_t key = stream_key(seed, 0x0F);
  #pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i)
      h_out[i] = (float)(lo + (hi - lo) * u01(rand_at(key, (uint64_t)i)));
    return BGB_OK;
  });
}

}  // extern "C"
