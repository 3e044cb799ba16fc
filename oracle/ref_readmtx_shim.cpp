// Exposes the reference's own CSR construction (removeSelfloop + customSort +
// coo2csr from bmv_eval/backend/readMtx.hpp:56-306, compiled where it lies in
// the reference tree) and its MatrixMarket reader (readMtx, :353-422, over
// mmio.hpp) through C functions, so the oracle's restatement and the library's
// reader can be checked against them. Test infrastructure only; built into
// oracle/_ref/.
#include <vector>

#include "readMtx.hpp"

// removeSelfloop's compaction (readMtx.hpp:202-204) reads the column vector
// up to two slots past its end. Reserve those slots and fill the whole
// capacity with -1 (an erased entry) beforehand, so what the reads see does
// not depend on what the heap held there.
template <class V>
static void reserve_filled(std::vector<V> &v, size_t cap, V fill) {
  const size_t n = v.size();
  v.reserve(cap);
  v.resize(cap, fill);
  v.resize(n);  // capacity and the filled slots stay
}

extern "C" int ref_normalize(int nrows, int ncols, long long nedges, const int *src,
                             const int *dst, int undirected, int *rowptr, int *colind,
                             long long *nnz) {
  if (nedges <= 0) return -1;  // removeSelfloop reads element 0 unconditionally
  std::vector<int> r(src, src + nedges), c(dst, dst + nedges);
  std::vector<float> v(nedges, 1.0f);
  reserve_filled(r, 2 * nedges + 4, -1);
  reserve_filled(c, 2 * nedges + 4, -1);
  reserve_filled(v, 2 * nedges + 4, 1.0f);
  int nvals = (int)nedges;
  removeSelfloop<float>(&r, &c, &v, &nvals, undirected != 0);
  customSort<float>(&r, &c, &v);
  std::vector<float> vals(nvals > 0 ? nvals : 1);
  coo2csr<float>(rowptr, colind, vals.data(), r, c, v, nrows, ncols);
  *nnz = nvals;
  return 0;
}

// The reference driver's path (test_bin_bin_full.cu:139-160): readMtx with
// directed = 0 and no .dat name, then coo2csr. rowptr / colind must hold
// nrows+1 / 2*entries ints; returns -1 where readMtx would not return.
extern "C" int ref_read_mtx(const char *path, int directed, int *nrows, int *ncols, long long *nnz, int *rowptr,
                            int *colind, long long cap) {
  std::vector<int> r, c;
  std::vector<float> v;
  int M = 0, N = 0, nv = 0;
  reserve_filled(r, 2 * cap + 4, -1);
  reserve_filled(c, 2 * cap + 4, -1);
  reserve_filled(v, 2 * cap + 4, 1.0f);
  readMtx<float>(path, &r, &c, &v, &M, &N, &nv, directed, false, nullptr);
  if ((long long)r.size() > cap) return -1;
  std::vector<float> vals(r.size() > 0 ? r.size() : 1);
  coo2csr<float>(rowptr, colind, vals.data(), r, c, v, M, N);
  *nrows = M;
  *ncols = N;
  *nnz = (long long)r.size();
  return 0;
}
