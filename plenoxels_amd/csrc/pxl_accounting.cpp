// pxl_accounting.cpp -- host-side traffic accounting for the render kernels.
//
// K_r, the number of distinct grid cells ray r's processed samples read: the nearest cell
// of every kept sample, or the union of the 8 corners of every kept trilinear sample.
// "Kept" means in-grid and not skipped by occupancy; "processed" means sample index
// k <= n_proc[r], the early-termination count pxl_render_fwd stores in aux[r][0].
// The geometry (float32, every op rounded separately; built with -ffp-contract=off) is the
// kernels' own: t = delta*k + near, p = o + d*t, n = (p - min)*inv_pd (or /pd).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "pxl.h"

namespace {

// Occupancy bit of cell (x, y, z) in the 8x8x8-tile layout of pxl_grid.occupancy.
bool occ_bit(const uint32_t* occ, int64_t x, int64_t y, int64_t z, int64_t Y, int64_t Z) {
  if (occ == nullptr) return true;
  const int64_t TY = (Y + 7) / 8, TZ = (Z + 7) / 8;
  const int64_t t = ((x >> 3) * TY + (y >> 3)) * TZ + (z >> 3);
  const int64_t w = t * 16 + (x & 7) * 2 + ((y >> 2) & 1);
  return (occ[w] >> (((y & 3) << 3) | (z & 7))) & 1u;
}

void count_range(const pxl_grid* g, const pxl_march* m, const float* o, const float* d,
                 int64_t r0, int64_t r1, const int32_t* n_proc, int64_t* counts) {
  const int64_t X = g->dims[0], Y = g->dims[1], Z = g->dims[2];
  const float fX = (float)X, fY = (float)Y, fZ = (float)Z;
  const bool exact = (m->flags & PXL_FLAG_EXACT_DIV) != 0;
  std::vector<uint32_t> ids;
  for (int64_t r = r0; r < r1; ++r) {
    ids.clear();
    const int np = n_proc ? std::min(n_proc[r], m->n_samples) : m->n_samples;
    for (int k = 1; k <= np; ++k) {
      float t = m->delta * (float)k;
      t = t + m->near;
      float n[3];
      for (int a = 0; a < 3; ++a) {
        const float p = o[3 * r + a] + d[3 * r + a] * t;
        const float diff = p - m->grid_min[a];
        n[a] = exact ? diff / m->pd : diff * m->inv_pd;
      }
      if (m->interp == PXL_INTERP_NEAREST) {
        const float rx = std::nearbyint(n[0]), ry = std::nearbyint(n[1]), rz = std::nearbyint(n[2]);
        if (!(rx >= 0.f && rx < fX && ry >= 0.f && ry < fY && rz >= 0.f && rz < fZ)) continue;
        const uint32_t cell = (uint32_t)(((int64_t)rx * Y + (int64_t)ry) * Z + (int64_t)rz);
        if (occ_bit(g->occupancy, (int64_t)rx, (int64_t)ry, (int64_t)rz, Y, Z)) ids.push_back(cell);
      } else {
        if (!(n[0] >= 0.f && n[0] < fX && n[1] >= 0.f && n[1] < fY && n[2] >= 0.f && n[2] < fZ)) continue;
        int64_t lo[3], hi[3];
        const int64_t dims[3] = {X, Y, Z};
        for (int a = 0; a < 3; ++a) {
          lo[a] = (int64_t)std::floor(n[a]);
          hi[a] = (int64_t)std::ceil(n[a]);
          if (hi[a] >= dims[a]) hi[a] -= dims[a];
        }
        if (!occ_bit(g->occupancy, lo[0], lo[1], lo[2], Y, Z)) continue;
        for (int j = 0; j < 8; ++j) {
          const int64_t x = (j & 4) ? lo[0] : hi[0];
          const int64_t y = (j & 2) ? lo[1] : hi[1];
          const int64_t z = (j & 1) ? lo[2] : hi[2];
          ids.push_back((uint32_t)((x * Y + y) * Z + z));
        }
      }
    }
    std::sort(ids.begin(), ids.end());
    counts[r] = (int64_t)(std::unique(ids.begin(), ids.end()) - ids.begin());
  }
}

}  // namespace

extern "C" int pxl_count_cells_host(const pxl_grid* grid, const pxl_march* march,
                                    const float* origins, const float* dirs, int64_t n_rays,
                                    const int32_t* n_proc, int64_t* counts, int32_t n_threads) {
  if (!grid || !march || !origins || !dirs || !counts || n_rays < 0) return PXL_E_INVALID;
  if (grid->dims[0] <= 0 || grid->dims[1] <= 0 || grid->dims[2] <= 0) return PXL_E_INVALID;
  if (grid->dims[0] * grid->dims[1] * grid->dims[2] > (int64_t(1) << 31) - 1) return PXL_E_SIZE;
  if (march->interp != PXL_INTERP_NEAREST && march->interp != PXL_INTERP_TRILINEAR) return PXL_E_INVALID;
  const int nt = std::max(1, std::min<int>(n_threads, 256));
  std::vector<std::thread> pool;
  const int64_t per = (n_rays + nt - 1) / nt;
  for (int i = 0; i < nt; ++i) {
    const int64_t r0 = i * per, r1 = std::min(n_rays, r0 + per);
    if (r0 >= r1) break;
    pool.emplace_back(count_range, grid, march, origins, dirs, r0, r1, n_proc, counts);
  }
  for (auto& t : pool) t.join();
  return PXL_OK;
}
