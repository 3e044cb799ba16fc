// MatrixMarket coordinate files -> normalised host CSR.
//
// The reference reads its inputs with readMtx (bmv_eval/backend/readMtx.hpp:
// 353-422) over the NIST mmio banner / size parsers (mmio.hpp), then builds
// the CSR in the driver (test_bin_bin_full.cu:139-160, directed = 0):
//   * banner "%%MatrixMarket matrix coordinate <field> <symmetry>", field
//     real | integer | pattern (values are not used by the binary path: the
//     drivers overwrite them with 1, test_bin_bin_full.cu:164);
//   * '%' comment lines, then "M N NNZ", then NNZ 1-based "i j [value]" lines;
//   * undirected iff the symmetry is "symmetric" (directed = 0), or forced
//     directed (1) / undirected (2);
//   * self-loops and duplicates dropped, rows sorted (removeSelfloop,
//     customSort, coo2csr) -- here by bgb_csr_from_coo, which restates them.
// Where the reference exits or reads on (a short file, an index out of
// range, an array-format file), this reader returns BGB_ERR_FORMAT.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bgb_internal.h"

namespace {

std::string lower(std::string s) {
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Next line of buf starting at *pos (without the newline); false at the end.
bool next_line(const std::string &buf, size_t *pos, const char **b, const char **e) {
  if (*pos >= buf.size()) return false;
  const size_t start = *pos;
  size_t nl = buf.find('\n', start);
  if (nl == std::string::npos) nl = buf.size();
  *b = buf.data() + start;
  *e = buf.data() + nl;
  *pos = nl + 1;
  return true;
}

bool blank(const char *b, const char *e) {
  for (; b < e; ++b)
    if (!std::isspace((unsigned char)*b)) return false;
  return true;
}

// Parses a decimal integer at *p (leading blanks skipped); false if none.
bool parse_ll(const char **p, const char *e, long long *v) {
  const char *s = *p;
  while (s < e && (*s == ' ' || *s == '\t' || *s == '\r')) ++s;
  if (s == e) return false;
  char tmp[32];
  size_t k = 0;
  if (*s == '+' || *s == '-') tmp[k++] = *s++;
  while (s < e && std::isdigit((unsigned char)*s) && k < sizeof(tmp) - 1) tmp[k++] = *s++;
  if (k == 0 || (k == 1 && (tmp[0] == '+' || tmp[0] == '-'))) return false;
  tmp[k] = 0;
  *v = std::strtoll(tmp, nullptr, 10);
  *p = s;
  return true;
}

}  // namespace

using namespace bgb;

extern "C" {

int bgb_read_mtx(const char *path, int directed, bgb_csr **out) {
  return host_guard([&]() -> int {
    clear_error();
    BGB_REQUIRE(path && out, BGB_ERR_ARG, "null path or output");
    BGB_REQUIRE(directed >= 0 && directed <= 2, BGB_ERR_ARG, "directed must be 0, 1 or 2 (got %d)", directed);
    FILE *f = std::fopen(path, "rb");
    BGB_REQUIRE(f, BGB_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
    std::string buf;
    {
      char chunk[1 << 16];
      size_t got;
      while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
      std::fclose(f);
    }
    size_t pos = 0;
    const char *b = nullptr, *e = nullptr;
    // banner
    BGB_REQUIRE(next_line(buf, &pos, &b, &e), BGB_ERR_FORMAT, "%s: empty file", path);
    std::vector<std::string> tok;
    {
      std::string line(b, e), t;
      for (char c : line) {
        if (std::isspace((unsigned char)c)) {
          if (!t.empty()) tok.push_back(lower(t)), t.clear();
        } else {
          t += c;
        }
      }
      if (!t.empty()) tok.push_back(lower(t));
    }
    BGB_REQUIRE(tok.size() == 5 && tok[0] == "%%matrixmarket" && tok[1] == "matrix", BGB_ERR_FORMAT,
                "%s: not a MatrixMarket matrix banner", path);
    BGB_REQUIRE(tok[2] == "coordinate", BGB_ERR_FORMAT, "%s: only the coordinate format is read (got %s)", path,
                tok[2].c_str());
    const std::string &field = tok[3], &sym = tok[4];
    BGB_REQUIRE(field == "real" || field == "integer" || field == "pattern", BGB_ERR_FORMAT,
                "%s: field %s not supported (real, integer or pattern)", path, field.c_str());
    BGB_REQUIRE(sym == "general" || sym == "symmetric" || sym == "skew-symmetric" || sym == "hermitian",
                BGB_ERR_FORMAT, "%s: unknown symmetry %s", path, sym.c_str());
    // size line after comments and blank lines
    long long M = -1, N = -1, NNZ = -1;
    for (;;) {
      BGB_REQUIRE(next_line(buf, &pos, &b, &e), BGB_ERR_FORMAT, "%s: no size line", path);
      if (b < e && *b == '%') continue;
      if (blank(b, e)) continue;
      const char *p = b;
      BGB_REQUIRE(parse_ll(&p, e, &M) && parse_ll(&p, e, &N) && parse_ll(&p, e, &NNZ), BGB_ERR_FORMAT,
                  "%s: bad size line", path);
      break;
    }
    BGB_REQUIRE(M >= 0 && N >= 0 && NNZ >= 0 && M <= INT32_MAX && N <= INT32_MAX, BGB_ERR_FORMAT,
                "%s: sizes %lld x %lld, %lld entries out of range", path, M, N, NNZ);
    // an entry line takes at least 3 bytes ("i j"), so the file bounds the
    // count it can hold whatever its size line announces
    const long long fits = (long long)((buf.size() - pos) / 3 + 1);
    std::vector<int> src, dst;
    src.reserve((size_t)(NNZ < fits ? NNZ : fits));
    dst.reserve((size_t)(NNZ < fits ? NNZ : fits));
    long long line_no = 0;
    while ((long long)src.size() < NNZ && next_line(buf, &pos, &b, &e)) {
      ++line_no;
      if (blank(b, e) || *b == '%') continue;
      const char *p = b;
      long long i = 0, j = 0;
      BGB_REQUIRE(parse_ll(&p, e, &i) && parse_ll(&p, e, &j), BGB_ERR_FORMAT, "%s: bad entry line %lld", path,
                  line_no);
      BGB_REQUIRE(i >= 1 && i <= M && j >= 1 && j <= N, BGB_ERR_FORMAT,
                  "%s: entry (%lld, %lld) outside the %lld x %lld matrix", path, i, j, M, N);
      src.push_back((int)(i - 1));
      dst.push_back((int)(j - 1));
    }
    BGB_REQUIRE((long long)src.size() == NNZ, BGB_ERR_FORMAT, "%s: %lld entries announced, %zu found", path, NNZ,
                src.size());
    const bool undirected = directed == 2 || (directed == 0 && sym == "symmetric");
    return bgb_csr_from_coo((int)M, (int)N, NNZ, src.data(), dst.data(), undirected ? 1 : 0, 1, out);
  });
}

}  // extern "C"
