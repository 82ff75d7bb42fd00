// pxl_host.h -- host-side helpers shared by the translation units of libpxl (not exported).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pxl.h"

#define PXL_HIDDEN __attribute__((visibility("hidden")))

namespace pxl_host {

constexpr int64_t kMaxCells = (int64_t(1) << 31) - 1;
// corner-block entries are indexed entry * 4 + slot in 32 bits
constexpr int64_t kMaxBlockCells = (int64_t(1) << 30) - 1;

// Records the thread-local error message returned by pxl_last_error(); returns code.
PXL_HIDDEN const char* last_error_text();
PXL_HIDDEN int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
PXL_HIDDEN int hip_fail(hipError_t e, const char* what);
// PXL_OK if the last launch succeeded, else the mapped error.
PXL_HIDDEN int launched(const char* what);
PXL_HIDDEN int check_grid(const pxl_grid* g);
// dims positive and X*Y*Z within the kernels' 31-bit cell index.
PXL_HIDDEN int check_dims(int64_t X, int64_t Y, int64_t Z);
PXL_HIDDEN int64_t occupancy_words(int64_t X, int64_t Y, int64_t Z);
PXL_HIDDEN int check_occupancy_words(int64_t X, int64_t Y, int64_t Z);

// Blocks for a grid-stride loop over n items with `threads` threads per block.
inline unsigned stride_blocks(int64_t n, int threads, int64_t cap = 256 * 64) {
  const int64_t b = (n + threads - 1) / threads;
  return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace pxl_host
