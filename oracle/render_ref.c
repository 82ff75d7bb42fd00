/*
 * FP64 per-ray oracle for the voxel-grid render path (TEST INFRASTRUCTURE ONLY).
 *
 * A scalar restatement of the reference training chain, written as a loop over samples
 * instead of the reference's vectorised torch ops:
 *   sample placement       src/ray_sampling.py:161-167   (float32, reference op order)
 *   grid-index normalise   src/ray_sampling.py:12-13     (float32; divide or reciprocal)
 *   nearest gather         src/grid_functions.py:103-114, scripts/train.py:146-147
 *   trilinear gather       src/grid_functions.py:7-44, :154-170, :220-246
 *   compositing            src/ray_sampling.py:172-192   (FP64, front to back)
 *   gradient               autograd of the above, restated as the reverse recurrence
 *                          S_i = a_{i+1} c_{i+1} + (1 - a_{i+1}) S_{i+1},
 *                          dC/da_i = T_i (c_i - S_i), dC/dc_i = a_i T_i
 * Geometry (sample coordinates, voxel choice, interpolation fractions) is computed in
 * float32 exactly as the reference computes it, so the voxels chosen are the ones an
 * FP32 implementation chooses; everything downstream of the gathered values is FP64.
 *
 * Used only by tests/ (and bench.py's baseline leg) as the checker; never shipped.
 * Build: oracle/Makefile -> oracle/_build/libpxl_oracle.so (gcc, -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { REF_NEAREST = 0, REF_TRILINEAR = 1 };

typedef struct {
  int64_t X, Y, Z;
  int32_t n_samples;
  int32_t interp;      /* REF_NEAREST / REF_TRILINEAR */
  int32_t div_mode;    /* 0 = divide by pd, 1 = multiply by inv_pd */
  int32_t pad;
  float delta, near, pd, inv_pd;
  float gmin[3];
  float t_stop;        /* 0 = off */
} ref_params;

/* One sample's gathered contribution: value, and what the scatter needs. */
typedef struct {
  int keep;
  double val[4];
  int64_t cell[8];     /* linear cell index per corner (nearest: cell[0]) */
  double w[8];         /* interpolation weight per corner (nearest: w[0] = 1) */
  int ncorner;
} ref_sample;

/* Occupancy bit of cell (x, y, z): 8x8x8 tiles of 16 words, tile ((x>>3)*TY + (y>>3))*TZ +
 * (z>>3) (TY = ceil(Y/8), TZ = ceil(Z/8)), bit (y&3)*8 + (z&7) of its word (x&7)*2 +
 * ((y>>2)&1) -- the layout the kernels read. */
static int occ_bit(const ref_params* p, const uint32_t* occ, int64_t x, int64_t y, int64_t z) {
  if (!occ) return 1;
  const int64_t TY = (p->Y + 7) / 8, TZ = (p->Z + 7) / 8;
  const int64_t t = ((x >> 3) * TY + (y >> 3)) * TZ + (z >> 3);
  const int64_t w = t * 16 + (x & 7) * 2 + ((y >> 2) & 1);
  return (int)((occ[w] >> (((y & 3) << 3) | (z & 7))) & 1u);
}

static int64_t wrap(int64_t i, int64_t n) {
  int64_t r = i % n;
  return r < 0 ? r + n : r;
}

static double clip01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

/* float32 sample coordinate in grid-index space: ((o + d*t) - gmin) / pd */
static void sample_coord(const ref_params* p, const float o[3], const float d[3], int k,
                         float n[3]) {
  float t = p->delta * (float)k;
  t = t + p->near;
  for (int a = 0; a < 3; ++a) {
    float dt = d[a] * t;
    float pos = o[a] + dt;
    float diff = pos - p->gmin[a];
    n[a] = p->div_mode ? diff * p->inv_pd : diff / p->pd;
  }
}

static void eval_sample(const ref_params* p, const double* grid, const uint32_t* occ,
                        const float o[3], const float d[3], int k, ref_sample* s) {
  float n[3];
  sample_coord(p, o, d, k, n);
  const int64_t dims[3] = {p->X, p->Y, p->Z};
  memset(s, 0, sizeof(*s));
  if (p->interp == REF_NEAREST) {
    int64_t idx[3];
    int inb = 1;
    for (int a = 0; a < 3; ++a) {
      float r = nearbyintf(n[a]); /* round half to even (default rounding mode) */
      if (!(r >= 0.0f && r < (float)dims[a])) inb = 0;
      idx[a] = inb ? (int64_t)r : 0;
    }
    if (!inb) return;
    int64_t cell = (idx[0] * p->Y + idx[1]) * p->Z + idx[2];
    if (!occ_bit(p, occ, idx[0], idx[1], idx[2])) return;
    s->keep = 1;
    s->ncorner = 1;
    s->cell[0] = cell;
    s->w[0] = 1.0;
    for (int c = 0; c < 4; ++c) s->val[c] = clip01(grid[cell * 4 + c]);
    return;
  }
  /* trilinear: float in-bounds mask on the unrounded coordinate */
  for (int a = 0; a < 3; ++a)
    if (!(n[a] >= 0.0f && n[a] < (float)dims[a])) return;
  int64_t cs[3][2]; /* [axis][0 = ceil, 1 = floor], wrapped */
  double fr[3];
  for (int a = 0; a < 3; ++a) {
    cs[a][0] = wrap((int64_t)ceilf(n[a]), dims[a]);
    cs[a][1] = wrap((int64_t)floorf(n[a]), dims[a]);
    fr[a] = (double)(n[a] - truncf(n[a]));
  }
  if (!occ_bit(p, occ, cs[0][1], cs[1][1], cs[2][1])) return;
  s->keep = 1;
  s->ncorner = 8;
  int j = 0;
  for (int bx = 0; bx < 2; ++bx)
    for (int by = 0; by < 2; ++by)
      for (int bz = 0; bz < 2; ++bz, ++j) {
        int64_t cell = (cs[0][bx] * p->Y + cs[1][by]) * p->Z + cs[2][bz];
        double wx = bx == 0 ? fr[0] : 1.0 - fr[0];
        double wy = by == 0 ? fr[1] : 1.0 - fr[1];
        double wz = bz == 0 ? fr[2] : 1.0 - fr[2];
        s->cell[j] = cell;
        s->w[j] = wx * wy * wz;
        for (int c = 0; c < 4; ++c) s->val[c] += s->w[j] * clip01(grid[cell * 4 + c]);
      }
}

/*
 * Render n_rays rays.  grid: (X,Y,Z,4) raw values as double.  occ: optional packed
 * occupancy bits (nearest: nearest cell; trilinear: floor cell).  out_rgba: (n,4).
 * If grad_rgba != NULL, accumulates dL/dgrid into grad_grid (X,Y,Z,4) (including the
 * clip mask 0 <= v <= 1).  n_proc (optional): processed-sample count per ray.
 * Returns 0, or -1 on allocation failure.
 */
int ref_render(const ref_params* p, const double* grid, const uint32_t* occ,
               const float* origins, const float* dirs, int64_t n_rays, double* out_rgba,
               const double* grad_rgba, double* grad_grid, int32_t* n_proc) {
  const int S = p->n_samples;
  ref_sample* ss = (ref_sample*)malloc(sizeof(ref_sample) * (size_t)(S > 0 ? S : 1));
  double* Ts = (double*)malloc(sizeof(double) * (size_t)(S > 0 ? S : 1));
  if (!ss || !Ts) { free(ss); free(Ts); return -1; }
  for (int64_t r = 0; r < n_rays; ++r) {
    const float* o = origins + 3 * r;
    const float* d = dirs + 3 * r;
    double T = 1.0, C[4] = {0, 0, 0, 0};
    int m = 0; /* processed samples kept for the backward */
    int stopped_at = S;
    for (int k = 1; k <= S; ++k) {
      ref_sample s;
      eval_sample(p, grid, occ, o, d, k, &s);
      if (p->t_stop > 0.0f && T < (double)p->t_stop) { stopped_at = k - 1; break; }
      if (!s.keep) continue;
      double a = s.val[3];
      double w = a * T;
      for (int c = 0; c < 3; ++c) C[c] += w * s.val[c];
      C[3] += w;
      if (grad_rgba) { ss[m] = s; Ts[m] = T; }
      ++m;
      T *= (1.0 - a);
    }
    if (n_proc) n_proc[r] = stopped_at;
    for (int c = 0; c < 4; ++c) out_rgba[4 * r + c] = C[c];
    if (!grad_rgba) continue;
    const double* g = grad_rgba + 4 * r;
    double Q[4] = {0, 0, 0, 0}; /* colour (and alpha) seen from just in front of sample i+1 */
    for (int i = m - 1; i >= 0; --i) {
      const ref_sample* s = &ss[i];
      double a = s->val[3];
      double gv[4];
      for (int c = 0; c < 3; ++c) gv[c] = a * Ts[i] * g[c];
      gv[3] = Ts[i] * (g[0] * (s->val[0] - Q[0]) + g[1] * (s->val[1] - Q[1]) +
                       g[2] * (s->val[2] - Q[2]) + g[3] * (1.0 - Q[3]));
      for (int c = 0; c < 3; ++c) Q[c] = a * s->val[c] + (1.0 - a) * Q[c];
      Q[3] = a + (1.0 - a) * Q[3];
      for (int j = 0; j < s->ncorner; ++j) {
        int64_t cell = s->cell[j];
        for (int c = 0; c < 4; ++c) {
          double raw = grid[cell * 4 + c];
          if (raw >= 0.0 && raw <= 1.0) grad_grid[cell * 4 + c] += s->w[j] * gv[c];
        }
      }
    }
  }
  free(ss);
  free(Ts);
  return 0;
}

/*
 * For each ray, the linear cells its kept samples read (nearest cell, or the 8 corners),
 * tested against a sorted list of query cells: out_hits[r * n_query + q] = 1 when ray r
 * reads query cell q.  Used to restrict finite differences to the rays touching a voxel.
 */
static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

int ref_rays_touching(const ref_params* p, const double* grid, const uint32_t* occ,
                      const float* origins, const float* dirs, int64_t n_rays,
                      const int64_t* query_sorted, int64_t n_query, uint8_t* out_hits) {
  const int S = p->n_samples;
  for (int64_t r = 0; r < n_rays; ++r) {
    const float* o = origins + 3 * r;
    const float* d = dirs + 3 * r;
    for (int k = 1; k <= S; ++k) {
      ref_sample s;
      eval_sample(p, grid, occ, o, d, k, &s);
      if (!s.keep) continue;
      for (int j = 0; j < s.ncorner; ++j) {
        int64_t* hit = (int64_t*)bsearch(&s.cell[j], query_sorted, (size_t)n_query,
                                         sizeof(int64_t), cmp_i64);
        if (hit) out_hits[r * n_query + (hit - query_sorted)] = 1;
      }
    }
  }
  return 0;
}

/* float32 coordinates of sample k of one ray (for bit-level geometry checks). */
void ref_sample_coord(const ref_params* p, const float o[3], const float d[3], int k,
                      float n_out[3]) {
  sample_coord(p, o, d, k, n_out);
}
