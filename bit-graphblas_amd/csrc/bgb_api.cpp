// Library information entry points of the C ABI.
#include "bgb_internal.h"

extern "C" {

int bgb_version(void) { return 100; }  // 1.0.0

const char *bgb_status_string(int status) {
  switch (status) {
    case BGB_OK: return "ok";
    case BGB_ERR_ARG: return "invalid argument";
    case BGB_ERR_TILEDIM: return "tile dim not in {4, 8, 16, 32, 64} or not taken by this entry point";
    case BGB_ERR_HIP: return "HIP runtime error";
    case BGB_ERR_ALLOC: return "allocation failed";
    case BGB_ERR_SIZE: return "size out of range";
    case BGB_ERR_FORMAT: return "invalid matrix structure";
    default: return "unknown status";
  }
}

long long bgb_packed_vec_words(int T, int nblockrows) {
  if (!bgb::valid_tiledim(T) || nblockrows < 0) return -1;
  return bgb::ceil_div(nblockrows, T) * T;
}

int bgb_word_bytes(int T) { return bgb::word_bytes(T); }

}  // extern "C"
