// Exposes the reference's own CSR construction (removeSelfloop + customSort +
// coo2csr from bmv_eval/backend/readMtx.hpp:56-306, compiled where it lies in
// the reference tree) through a C function, so the oracle's restatement can be
// checked against it. Test infrastructure only; built into oracle/_ref/.
#include <vector>

#include "readMtx.hpp"

extern "C" int ref_normalize(int nrows, int ncols, long long nedges, const int *src,
                             const int *dst, int undirected, int *rowptr, int *colind,
                             long long *nnz) {
  if (nedges <= 0) return -1;  // removeSelfloop reads element 0 unconditionally
  std::vector<int> r(src, src + nedges), c(dst, dst + nedges);
  std::vector<float> v(nedges, 1.0f);
  // removeSelfloop's compaction may read one element past the end
  // (readMtx.hpp:202-204); reserve so that slot is allocated memory.
  r.reserve(2 * nedges + 1);
  c.reserve(2 * nedges + 1);
  v.reserve(2 * nedges + 1);
  int nvals = (int)nedges;
  removeSelfloop<float>(&r, &c, &v, &nvals, undirected != 0);
  customSort<float>(&r, &c, &v);
  std::vector<float> vals(nvals > 0 ? nvals : 1);
  coo2csr<float>(rowptr, colind, vals.data(), r, c, v, nrows, ncols);
  *nnz = nvals;
  return 0;
}
