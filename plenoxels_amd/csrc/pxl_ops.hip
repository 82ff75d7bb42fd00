// pxl_ops.hip -- the per-function kernels behind the reference's individual call signatures
// (src/ray_sampling.py, src/grid_functions.py, scripts/train.py) and the optimiser step.
//
// These materialise the same intermediate tensors the reference functions return, so that a
// caller composing them (as scripts/train.py does) gets the reference's results; the fused
// path (pxl_render.hip) is what a training loop should call.  Elementwise kernels use
// grid-stride loops; anything that accumulates into a grid uses float atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cmath>

#include "pxl.h"
#include "pxl_device.h"
#include "pxl_host.h"

using pxl_host::fail;
using pxl_host::launched;
using pxl_host::stride_blocks;

namespace pxl {
namespace {

#define PXL_GRID_STRIDE(i, n)                                                       \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)(n); \
       i += (int64_t)gridDim.x * blockDim.x)

// Python's % on int64 (result has the sign of the divisor): src/grid_functions.py:75-77.
__device__ __forceinline__ int64_t pymod(int64_t a, int64_t m) {
  const int64_t r = a % m;
  return r < 0 ? r + m : r;
}

// float -> int64 as torch's .to(torch.long) for the finite values the kernels see;
// non-finite or huge inputs map to a value that fails every bounds test.
__device__ __forceinline__ int64_t to_i64(float x) {
  return (fabsf(x) < 9.0e18f) ? (int64_t)x : INT64_MIN / 2;
}

struct Dims {
  int64_t X, Y, Z;
};

__device__ __forceinline__ int64_t cell_of(const Dims& d, int64_t i, int64_t j, int64_t k) {
  return (i * d.Y + j) * d.Z + k;
}

__device__ __forceinline__ void atomic_add4(float* p, const float4& g) {
  if (g.x != 0.f) unsafeAtomicAdd(p + 0, g.x);
  if (g.y != 0.f) unsafeAtomicAdd(p + 1, g.y);
  if (g.z != 0.f) unsafeAtomicAdd(p + 2, g.z);
  if (g.w != 0.f) unsafeAtomicAdd(p + 3, g.w);
}

// ---------------------------------------------------------------- samples, normalisation

__global__ void k_sample_points(const float* __restrict__ o, int64_t rays_per_origin,
                                const float* __restrict__ d, int64_t n_rays, int S, float delta,
                                float near, float* __restrict__ out) {
  PXL_GRID_STRIDE(i, n_rays * S) {
    const int64_t r = i / S;
    const int k = (int)(i - r * S) + 1;
    const int64_t oi = r / rays_per_origin;
    float t = delta * (float)k;
    t = t + near;
    out[3 * i + 0] = o[3 * oi + 0] + d[3 * r + 0] * t;
    out[3 * i + 1] = o[3 * oi + 1] + d[3 * r + 1] * t;
    out[3 * i + 2] = o[3 * oi + 2] + d[3 * r + 2] * t;
  }
}

__global__ void k_init_min(float* out3) {
  if (threadIdx.x < 3) out3[threadIdx.x] = INFINITY;
}

__device__ __forceinline__ void atomic_min_f32(float* a, float v) {
  unsigned int* ai = reinterpret_cast<unsigned int*>(a);
  unsigned int old = *ai;
  while (v < __uint_as_float(old)) {
    const unsigned int prev = atomicCAS(ai, old, __float_as_uint(v));
    if (prev == old) break;
    old = prev;
  }
}

__global__ void k_coord_min(const float* __restrict__ c, int64_t n, float* out3) {
  float m0 = INFINITY, m1 = INFINITY, m2 = INFINITY;
  PXL_GRID_STRIDE(i, n) {
    m0 = fminf(m0, c[3 * i]);
    m1 = fminf(m1, c[3 * i + 1]);
    m2 = fminf(m2, c[3 * i + 2]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    m0 = fminf(m0, __shfl_xor(m0, off));
    m1 = fminf(m1, __shfl_xor(m1, off));
    m2 = fminf(m2, __shfl_xor(m2, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomic_min_f32(out3 + 0, m0);
    atomic_min_f32(out3 + 1, m1);
    atomic_min_f32(out3 + 2, m2);
  }
}

__global__ void k_normalize(const float* __restrict__ p, int64_t n, const float* __restrict__ m3,
                            float pd, float inv, int exact, float* __restrict__ out) {
  const float g0 = m3[0], g1 = m3[1], g2 = m3[2];
  PXL_GRID_STRIDE(i, n) {
    const float d0 = p[3 * i] - g0, d1 = p[3 * i + 1] - g1, d2 = p[3 * i + 2] - g2;
    out[3 * i + 0] = exact ? __fdiv_rn(d0, pd) : d0 * inv;
    out[3 * i + 1] = exact ? __fdiv_rn(d1, pd) : d1 * inv;
    out[3 * i + 2] = exact ? __fdiv_rn(d2, pd) : d2 * inv;
  }
}

// ---------------------------------------------------------------- bounds, wrap, nearest

__global__ void k_in_bounds_i64(const int64_t* __restrict__ idx, int64_t n, Dims d,
                                uint8_t* __restrict__ mask) {
  PXL_GRID_STRIDE(i, n) {
    const int64_t a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
    mask[i] = (a < d.X) & (a >= 0) & (b < d.Y) & (b >= 0) & (c < d.Z) & (c >= 0);
  }
}

__global__ void k_in_bounds_f32(const float* __restrict__ idx, int64_t n, Dims d,
                                uint8_t* __restrict__ mask) {
  const float fx = (float)d.X, fy = (float)d.Y, fz = (float)d.Z;
  PXL_GRID_STRIDE(i, n) {
    const float a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
    mask[i] = (a < fx) & (a >= 0.f) & (b < fy) & (b >= 0.f) & (c < fz) & (c >= 0.f);
  }
}

__global__ void k_wrap(int64_t* __restrict__ idx, int64_t n, Dims d) {
  PXL_GRID_STRIDE(i, n) {
    idx[3 * i] = pymod(idx[3 * i], d.X);
    idx[3 * i + 1] = pymod(idx[3 * i + 1], d.Y);
    idx[3 * i + 2] = pymod(idx[3 * i + 2], d.Z);
  }
}

// idx = round-half-even(n).long(); mask = in-bounds; cell = wrapped idx
__device__ __forceinline__ int64_t nearest_wrapped(const float* n, int64_t i, const Dims& d,
                                                   bool& inb) {
  const int64_t a = to_i64(__builtin_rintf(n[3 * i])), b = to_i64(__builtin_rintf(n[3 * i + 1])),
                c = to_i64(__builtin_rintf(n[3 * i + 2]));
  inb = a >= 0 && a < d.X && b >= 0 && b < d.Y && c >= 0 && c < d.Z;
  return cell_of(d, pymod(a, d.X), pymod(b, d.Y), pymod(c, d.Z));
}

__global__ void k_nearest_gather(const float* __restrict__ n, int64_t N, Dims d,
                                 const float4* __restrict__ grid, float4* __restrict__ vals,
                                 uint8_t* __restrict__ mask) {
  PXL_GRID_STRIDE(i, N) {
    bool inb;
    const int64_t cell = nearest_wrapped(n, i, d, inb);
    vals[i] = grid[cell];
    if (mask) mask[i] = inb;
  }
}

__global__ void k_nearest_scatter(const float* __restrict__ n, int64_t N, Dims d,
                                  const float4* __restrict__ g, float* __restrict__ grad) {
  PXL_GRID_STRIDE(i, N) {
    bool inb;
    const int64_t cell = nearest_wrapped(n, i, d, inb);
    atomic_add4(grad + 4 * cell, g[i]);
  }
}

// ---------------------------------------------------------------- corners, trilinear

// ceil/floor of each coordinate as int64 (src/grid_functions.py:230-244), [a][0] = ceil.
__device__ __forceinline__ void ceil_floor(const float* n, int64_t i, int64_t cf[3][2]) {
  for (int a = 0; a < 3; ++a) {
    cf[a][0] = to_i64(ceilf(n[3 * i + a]));
    cf[a][1] = to_i64(floorf(n[3 * i + a]));
  }
}

__global__ void k_corner_indices(const float* __restrict__ n, int64_t N, int64_t* __restrict__ out) {
  PXL_GRID_STRIDE(i, N) {
    int64_t cf[3][2];
    ceil_floor(n, i, cf);
    int64_t* o = out + 24 * i;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      o[3 * j + 0] = cf[0][(j >> 2) & 1];
      o[3 * j + 1] = cf[1][(j >> 1) & 1];
      o[3 * j + 2] = cf[2][j & 1];
    }
  }
}

__device__ __forceinline__ bool float_in_bounds(const float* n, int64_t i, const Dims& d) {
  const float a = n[3 * i], b = n[3 * i + 1], c = n[3 * i + 2];
  return a < (float)d.X && a >= 0.f && b < (float)d.Y && b >= 0.f && c < (float)d.Z && c >= 0.f;
}

__global__ void k_collect(const float* __restrict__ n, int64_t N, Dims d,
                          const float4* __restrict__ grid, float4* __restrict__ out,
                          uint8_t* __restrict__ mask) {
  PXL_GRID_STRIDE(i, N) {
    int64_t cf[3][2];
    ceil_floor(n, i, cf);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      out[8 * i + j] = grid[cell_of(d, pymod(cf[0][(j >> 2) & 1], d.X), pymod(cf[1][(j >> 1) & 1], d.Y),
                                    pymod(cf[2][j & 1], d.Z))];
    if (mask) mask[i] = float_in_bounds(n, i, d);
  }
}

__global__ void k_collect_bwd(const float* __restrict__ n, int64_t N, Dims d,
                              const float4* __restrict__ g, float* __restrict__ grad) {
  PXL_GRID_STRIDE(i, N) {
    int64_t cf[3][2];
    ceil_floor(n, i, cf);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      atomic_add4(grad + 4 * cell_of(d, pymod(cf[0][(j >> 2) & 1], d.X),
                                     pymod(cf[1][(j >> 1) & 1], d.Y), pymod(cf[2][j & 1], d.Z)),
                  g[8 * i + j]);
  }
}

__device__ __forceinline__ int64_t corner_cell(const int64_t* c, int64_t i, int j, const Dims& d) {
  const int64_t* p = c + 24 * i + 3 * j;
  return cell_of(d, pymod(p[0], d.X), pymod(p[1], d.Y), pymod(p[2], d.Z));
}

__device__ __forceinline__ float4 lerp_ref(const float4& a, const float4& b, float f, float of) {
  return make_float4((a.x * f) + (b.x * of), (a.y * f) + (b.y * of), (a.z * f) + (b.z * of),
                     (a.w * f) + (b.w * of));
}

// src/grid_functions.py:29-42 in its own op order: x between corners 0-3 (weight frac) and
// 4-7 (1-frac), then y, then z.
__global__ void k_trilinear(const float* __restrict__ n, const int64_t* __restrict__ c, int64_t N,
                            Dims d, const float4* __restrict__ grid, float4* __restrict__ out) {
  PXL_GRID_STRIDE(i, N) {
    const float fx = n[3 * i] - truncf(n[3 * i]);
    const float fy = n[3 * i + 1] - truncf(n[3 * i + 1]);
    const float fz = n[3 * i + 2] - truncf(n[3 * i + 2]);
    float4 q[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
      q[a] = lerp_ref(grid[corner_cell(c, i, a, d)], grid[corner_cell(c, i, 4 + a, d)], fx, 1.f - fx);
    const float4 m0 = lerp_ref(q[0], q[2], fy, 1.f - fy);
    const float4 m1 = lerp_ref(q[1], q[3], fy, 1.f - fy);
    out[i] = lerp_ref(m0, m1, fz, 1.f - fz);
  }
}

__global__ void k_trilinear_bwd(const float* __restrict__ n, const int64_t* __restrict__ c,
                                int64_t N, Dims d, const float4* __restrict__ g,
                                float* __restrict__ grad) {
  PXL_GRID_STRIDE(i, N) {
    const float fx = n[3 * i] - truncf(n[3 * i]);
    const float fy = n[3 * i + 1] - truncf(n[3 * i + 1]);
    const float fz = n[3 * i + 2] - truncf(n[3 * i + 2]);
    const float4 gi = g[i];
    const float wz[2] = {fz, 1.f - fz}, wy[2] = {fy, 1.f - fy}, wx[2] = {fx, 1.f - fx};
    // corner j = 4*bx + 2*by + bz: d out / d corner = wz[bz] * wy[by] * wx[bx] (autograd order)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float sz = wz[j & 1], sy = wy[(j >> 1) & 1], sx = wx[(j >> 2) & 1];
      const float4 gz = make_float4(gi.x * sz, gi.y * sz, gi.z * sz, gi.w * sz);
      const float4 gy = make_float4(gz.x * sy, gz.y * sy, gz.z * sy, gz.w * sy);
      atomic_add4(grad + 4 * corner_cell(c, i, j, d),
                  make_float4(gy.x * sx, gy.y * sx, gy.z * sx, gy.w * sx));
    }
  }
}

// ---------------------------------------------------------------- compositing

// one wave per ray; samples (R,S,4)
__global__ __launch_bounds__(kBlock) void k_composite_fwd(const float4* __restrict__ smp, int64_t R,
                                                           int S, float4* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (r >= R) return;
  float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
  for (int c0 = 0; c0 < S; c0 += kWave) {
    const int k = c0 + lane;
    const float4 v = k < S ? smp[r * S + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float incl = wave_prefix_mul(1.f - v.w);
    const float w = v.w * (T * wave_shift_up(incl, 1.f));
    cr += v.x * w;
    cg += v.y * w;
    cb += v.z * w;
    ca += w;
    T = T * readlane(incl, 63);
  }
  cr = wave_sum(cr);
  cg = wave_sum(cg);
  cb = wave_sum(cb);
  ca = wave_sum(ca);
  if (lane == 0) out[r] = make_float4(cr, cg, cb, ca);
}

constexpr int kMaxCompositeChunks = 1024;  // S <= 65536

__global__ __launch_bounds__(kBlock) void k_composite_bwd(const float4* __restrict__ smp, int64_t R,
                                                           int S, const float4* __restrict__ gout,
                                                           float4* __restrict__ gsmp) {
  __shared__ float s_T[kWavesPerBlock][kMaxCompositeChunks];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * kWavesPerBlock + wv;
  if (r >= R) return;
  const int nch = (S + 63) >> 6;
  float T = 1.f;
  for (int c = 0; c < nch; ++c) {  // transmittance at every chunk start
    const int k = (c << 6) + lane;
    const float a = k < S ? smp[r * S + k].w : 0.f;
    if (lane == 0) s_T[wv][c] = T;
    T = T * readlane(wave_prefix_mul(1.f - a), 63);
  }
  wave_lds_sync();
  const float4 g = gout[r];
  float4 Qn = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int c = nch - 1; c >= 0; --c) {
    const int k = (c << 6) + lane;
    const bool valid = k < S;
    const float4 v = valid ? smp[r * S + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float incl = wave_prefix_mul(1.f - v.w);
    const float Ti = s_T[wv][c] * wave_shift_up(incl, 1.f);
    Affine4 f{v.w * v.x, v.w * v.y, v.w * v.z, v.w, 1.f - v.w};
    if (lane == 63) {
      f.a0 = f.a0 + f.b * Qn.x;
      f.a1 = f.a1 + f.b * Qn.y;
      f.a2 = f.a2 + f.b * Qn.z;
      f.a3 = f.a3 + f.b * Qn.w;
      f.b = 0.f;
    }
    wave_suffix_affine(f, lane);
    const float s0 = wave_shift_down(f.a0, Qn.x), s1 = wave_shift_down(f.a1, Qn.y);
    const float s2 = wave_shift_down(f.a2, Qn.z), s3 = wave_shift_down(f.a3, Qn.w);
    if (valid) {
      const float w = v.w * Ti;
      gsmp[r * S + k] = make_float4(
          w * g.x, w * g.y, w * g.z,
          Ti * (g.x * (v.x - s0) + g.y * (v.y - s1) + g.z * (v.z - s2) + g.w * (1.f - s3)));
    }
    Qn = make_float4(readlane(f.a0, 0), readlane(f.a1, 0), readlane(f.a2, 0), readlane(f.a3, 0));
  }
}

// ---------------------------------------------------------------- grid layout, pooling

__global__ void k_grid_coords(Dims d, float cx, float cy, float cz, float pd,
                              float* __restrict__ coords, int64_t* __restrict__ mesh) {
  PXL_GRID_STRIDE(L, d.X * d.Y * d.Z) {
    const int64_t k = L % d.Z, j = (L / d.Z) % d.Y, i = L / (d.Y * d.Z);
    coords[3 * L + 0] = (((float)i - cx) + 1.f) * pd;
    coords[3 * L + 1] = (((float)j - cy) + 1.f) * pd;
    coords[3 * L + 2] = (((float)k - cz) + 1.f) * pd;
    if (mesh) {
      mesh[3 * L + 0] = i;
      mesh[3 * L + 1] = j;
      mesh[3 * L + 2] = k;
    }
  }
}

// A cubic window is separable: the pooled grid is the mean over k along z, then along y, then
// along x, each pass a thread per output element of the array viewed as [outer][n][inner]
// (consecutive threads take consecutive inner elements, so the y and x passes read and write
// contiguously; the z pass has inner = 1 and its threads read overlapping runs of one row).
// Each pass divides by k, so partial sums stay the size of the values.
__global__ __launch_bounds__(256) void k_box_mean(const float4* __restrict__ in, int64_t outer,
                                                  int64_t n, int64_t inner, int k, int s,
                                                  int64_t n_out, float4* __restrict__ out) {
  const float fk = (float)k;
  PXL_GRID_STRIDE(L, outer * n_out * inner) {
    const int64_t i = L % inner, o = (L / inner) % n_out, a = L / (inner * n_out);
    const float4* src = in + (a * n + o * s) * inner + i;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int d = 0; d < k; ++d) {
      const float4 v = src[(int64_t)d * inner];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    out[L] = make_float4(acc.x / fk, acc.y / fk, acc.z / fk, acc.w / fk);
  }
}

// The z pass (inner = 1): a workgroup loads as many whole rows as fit in LDS (four of 512
// cells) with coalesced reads, then each thread sums a (row, window) pair's cells from LDS in
// order (threads of the generic pass would each read a separate memory segment, their
// windows being s cells apart); with several rows per batch every thread has a window even
// when a row has few (19 at k = 93 over 512 cells).
constexpr int kRowLdsCells = 2048;  // rows up to 2048 cells (32 KiB of LDS)

__global__ __launch_bounds__(256) void k_box_mean_rows(const float4* __restrict__ in, int64_t rows,
                                                       int n, int k, int s, int n_out,
                                                       float4* __restrict__ out) {
  __shared__ float4 s_row[kRowLdsCells];
  const float fk = (float)k;
  const int per = kRowLdsCells / n;  // rows per batch (n <= kRowLdsCells)
  for (int64_t r0 = (int64_t)blockIdx.x * per; r0 < rows; r0 += (int64_t)gridDim.x * per) {
    const int nr = (int)min((int64_t)per, rows - r0);
    for (int i = threadIdx.x; i < nr * n; i += blockDim.x) s_row[i] = in[r0 * n + i];
    __syncthreads();
    for (int t = threadIdx.x; t < nr * n_out; t += blockDim.x) {
      const int rr = t / n_out, o = t - rr * n_out;
      const float4* src = s_row + rr * n + o * s;
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int d = 0; d < k; ++d) {
        const float4 v = src[d];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
      }
      out[(r0 + rr) * n_out + o] = make_float4(acc.x / fk, acc.y / fk, acc.z / fk, acc.w / fk);
    }
    __syncthreads();
  }
}

// windows o with o*s <= x < o*s + k, clipped to [0, n_out)
__device__ __forceinline__ void window_range(int64_t x, int k, int s, int64_t n_out, int64_t& lo,
                                             int64_t& hi) {
  lo = x - k + 1 <= 0 ? 0 : (x - k + 1 + s - 1) / s;
  hi = x / s;
  if (hi > n_out - 1) hi = n_out - 1;
}

// Adjoint of one k_box_mean pass: out[a][x][i] (+)= sum over the windows o containing x of
// in[a][o][i] / k.
__global__ __launch_bounds__(256) void k_box_mean_adjoint(const float4* __restrict__ in, int64_t outer,
                                                          int64_t n, int64_t inner, int k, int s,
                                                          int64_t n_out, int accumulate,
                                                          float4* __restrict__ out) {
  const float fk = (float)k;
  PXL_GRID_STRIDE(L, outer * n * inner) {
    const int64_t i = L % inner, x = (L / inner) % n, a = L / (inner * n);
    int64_t lo, hi;
    window_range(x, k, s, n_out, lo, hi);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t o = lo; o <= hi; ++o) {
      const float4 g = in[(a * n_out + o) * inner + i];
      acc.x += g.x;
      acc.y += g.y;
      acc.z += g.z;
      acc.w += g.w;
    }
    acc = make_float4(acc.x / fk, acc.y / fk, acc.z / fk, acc.w / fk);
    if (accumulate) {
      const float4 c = out[L];
      acc = make_float4(c.x + acc.x, c.y + acc.y, c.z + acc.z, c.w + acc.w);
    }
    out[L] = acc;
  }
}

// ---------------------------------------------------------------- total variation

__device__ __forceinline__ double sq4(const float4& a, const float4& b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z, dw = a.w - b.w;
  return (double)(dx * dx) + (double)(dy * dy) + (double)(dz * dz) + (double)(dw * dw);
}

// The TV kernels walk x columns: a workgroup covers 4 y rows (one wave each) x 64 z (the lanes)
// and a slab of kTvSlab x planes, keeping each lane's x - 1, x and x + 1 cells in registers, so
// the x neighbours (Y*Z cells apart, rarely still in cache when read in linear order) are read
// once; the y and z neighbours are the lines the neighbouring waves and lanes load.
constexpr int kTvSlab = 64;

struct TvUnit {
  uint32_t x0, x1, y, z;
  int64_t col;  // cell (x, y, z) = x * Y*Z + col
  bool active;
};

__device__ __forceinline__ TvUnit tv_unit(Dims d) {
  const uint32_t Y = (uint32_t)d.Y, Z = (uint32_t)d.Z, X = (uint32_t)d.X;
  const uint32_t nzb = (Z + 63u) / 64u, nyb = (Y + 3u) / 4u;
  const uint32_t u = blockIdx.x;
  const uint32_t zb = u % nzb, yb = (u / nzb) % nyb, xs = u / (nzb * nyb);
  TvUnit t;
  t.y = yb * 4u + (threadIdx.x >> 6);
  t.z = zb * 64u + (threadIdx.x & 63u);
  t.x0 = xs * kTvSlab;
  t.x1 = min(X, t.x0 + kTvSlab);
  t.active = t.y < Y && t.z < Z && t.x0 < X;
  t.col = (int64_t)t.y * Z + t.z;
  return t;
}

__host__ __device__ inline int64_t tv_units(int64_t X, int64_t Y, int64_t Z) {
  return ((X + kTvSlab - 1) / kTvSlab) * ((Y + 3) / 4) * ((Z + 63) / 64);
}

// sum over cells of the squared differences to the +y, +z and +x neighbours, in that order
__global__ __launch_bounds__(256) void k_tv_sumsq(const float4* __restrict__ g, Dims d,
                                                  double* __restrict__ sum) {
  const uint32_t Y = (uint32_t)d.Y, Z = (uint32_t)d.Z, X = (uint32_t)d.X, YZ = Y * Z;
  const TvUnit t = tv_unit(d);
  double acc = 0.0;
  if (t.active) {
    float4 v = g[(int64_t)t.x0 * YZ + t.col];
    for (uint32_t x = t.x0; x < t.x1; ++x) {
      const int64_t L = (int64_t)x * YZ + t.col;
      const float4 vx = x + 1 < X ? g[L + YZ] : make_float4(0.f, 0.f, 0.f, 0.f);
      if (t.y + 1 < Y) acc += sq4(v, g[L + Z]);
      if (t.z + 1 < Z) acc += sq4(v, g[L + 1]);
      if (x + 1 < X) acc += sq4(v, vx);
      v = vx;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(sum, acc);
}

// v - neighbour summed over +y, -y, +z, -z, +x, -x (in that order) for the cells of a column
// walk; vm/vp are the x - 1 / x + 1 cells (read by the caller's window)
__device__ __forceinline__ float4 tv_dir(const float4* __restrict__ g, int64_t L, const float4& v, const float4& vm,
                                         const float4& vp, uint32_t x, uint32_t y, uint32_t z, uint32_t X,
                                         uint32_t Y, uint32_t Z, uint32_t YZ) {
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  auto add = [&](const float4& q) {
    acc.x += v.x - q.x;
    acc.y += v.y - q.y;
    acc.z += v.z - q.z;
    acc.w += v.w - q.w;
  };
  if (y + 1 < Y) add(g[L + Z]);
  if (y > 0) add(g[L - Z]);
  if (z + 1 < Z) add(g[L + 1]);
  if (z > 0) add(g[L - 1]);
  if (x + 1 < X) add(vp);
  if (x > 0) add(vm);
  (void)YZ;
  return acc;
}

__global__ __launch_bounds__(256) void k_tv_bwd(const float4* __restrict__ g, Dims d,
                                                const double* __restrict__ sum, float weight,
                                                float4* __restrict__ grad) {
  // d/dx of weight*sqrt(S): weight / sqrt(S) * sum over neighbours of (x - neighbour)
  const float scale = weight / (float)sqrt(*sum);
  const uint32_t Y = (uint32_t)d.Y, Z = (uint32_t)d.Z, X = (uint32_t)d.X, YZ = Y * Z;
  const TvUnit t = tv_unit(d);
  if (!t.active) return;
  float4 vm = t.x0 > 0 ? g[(int64_t)(t.x0 - 1) * YZ + t.col] : make_float4(0.f, 0.f, 0.f, 0.f);
  float4 v = g[(int64_t)t.x0 * YZ + t.col];
  for (uint32_t x = t.x0; x < t.x1; ++x) {
    const int64_t L = (int64_t)x * YZ + t.col;
    const float4 vp = x + 1 < X ? g[L + YZ] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 acc = tv_dir(g, L, v, vm, vp, x, t.y, t.z, X, Y, Z, YZ);
    float4 c = grad[L];
    c.x += scale * acc.x;
    c.y += scale * acc.y;
    c.z += scale * acc.z;
    c.w += scale * acc.w;
    grad[L] = c;
    vm = v;
    v = vp;
  }
}

// ---------------------------------------------------------------- Adam

__device__ __forceinline__ uint16_t to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// One element of torch.optim.Adam's update (see the host entry point for the scalar terms).
__device__ __forceinline__ float adam_elem(float& p, float gi, float& m, float& v, float w1, float beta2,
                                           float w2, float bc2_sqrt, float eps, float neg_step) {
  // exp_avg.lerp_(grad, 1 - beta1): weight < 0.5 form  m + w * (g - m)
  m = m + w1 * (gi - m);
  v = (v * beta2) + (w2 * gi) * gi;  // addcmul: value * t1 * t2
  const float denom = (__builtin_sqrtf(v) / bc2_sqrt) + eps;
  p = p + neg_step * (m / denom);
  return p;
}

__device__ __forceinline__ void store_half(void* half_out, int half_dtype, int64_t i, float x) {
  if (half_dtype == PXL_DTYPE_BF16)
    reinterpret_cast<uint16_t*>(half_out)[i] = to_bf16_rne(x);
  else
    reinterpret_cast<_Float16*>(half_out)[i] = (_Float16)x;
}

// Elements [0, 4*n4) four at a time with 16-byte loads and stores (every pointer 16-byte
// aligned), then the tail [4*n4, n) one at a time.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float* __restrict__ gabs, int64_t n, int64_t n4, float w1,
                       float beta2, float w2, float bc2_sqrt, float eps, float neg_step,
                       void* __restrict__ half_out, int half_dtype) {
  PXL_GRID_STRIDE(i, n4) {
    float4 P = reinterpret_cast<float4*>(p)[i];
    const float4 G = reinterpret_cast<const float4*>(g)[i];
    float4 M = reinterpret_cast<float4*>(m)[i];
    float4 V = reinterpret_cast<float4*>(v)[i];
    adam_elem(P.x, G.x, M.x, V.x, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    adam_elem(P.y, G.y, M.y, V.y, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    adam_elem(P.z, G.z, M.z, V.z, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    adam_elem(P.w, G.w, M.w, V.w, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    reinterpret_cast<float4*>(p)[i] = P;
    reinterpret_cast<float4*>(m)[i] = M;
    reinterpret_cast<float4*>(v)[i] = V;
    if (gabs) {
      float4 A = reinterpret_cast<float4*>(gabs)[i];
      A = make_float4(A.x + fabsf(G.x), A.y + fabsf(G.y), A.z + fabsf(G.z), A.w + fabsf(G.w));
      reinterpret_cast<float4*>(gabs)[i] = A;
    }
    if (half_out) {
      if (half_dtype == PXL_DTYPE_BF16) {
        const uint32_t lo = to_bf16_rne(P.x) | ((uint32_t)to_bf16_rne(P.y) << 16);
        const uint32_t hi = to_bf16_rne(P.z) | ((uint32_t)to_bf16_rne(P.w) << 16);
        reinterpret_cast<uint2*>(half_out)[i] = make_uint2(lo, hi);
      } else {
        store_half(half_out, half_dtype, 4 * i + 0, P.x);
        store_half(half_out, half_dtype, 4 * i + 1, P.y);
        store_half(half_out, half_dtype, 4 * i + 2, P.z);
        store_half(half_out, half_dtype, 4 * i + 3, P.w);
      }
    }
  }
  PXL_GRID_STRIDE(j, n - 4 * n4) {
    const int64_t i = 4 * n4 + j;
    const float gi = g[i];
    float pi = p[i], mi = m[i], vi = v[i];
    adam_elem(pi, gi, mi, vi, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    m[i] = mi;
    v[i] = vi;
    p[i] = pi;
    if (gabs) gabs[i] += fabsf(gi);
    if (half_out) store_half(half_out, half_dtype, i, pi);
  }
}

// Adam on grad + the TV gradient of the parameter, in one pass: the TV term of k_tv_bwd is formed
// in registers from p_in and its six neighbours and added to the gradient exactly as k_tv_bwd
// adds it (same operations, same order), then the Adam update of k_adam follows; |grad| and the
// half copy see the total gradient and the new parameter.  The update goes to p_out, not p_in:
// a neighbour updated in place would change the stencil of the cells read after it.
// Traversal: the TV kernels' column walk (tv_unit), the x - 1 / x + 1 parameters in registers.
__global__ __launch_bounds__(256) void k_adam_tv(const float4* __restrict__ p_in, float4* __restrict__ p_out,
                                                 float4* __restrict__ g, float4* __restrict__ m,
                                                 float4* __restrict__ v, float4* __restrict__ gabs, Dims d,
                                                 const double* __restrict__ sum, float weight, float w1,
                                                 float beta2, float w2, float bc2_sqrt, float eps,
                                                 float neg_step, void* __restrict__ half_out, int half_dtype,
                                                 int zero_grad) {
  const float scale = weight / (float)sqrt(*sum);
  const uint32_t Y = (uint32_t)d.Y, Z = (uint32_t)d.Z, X = (uint32_t)d.X, YZ = Y * Z;
  const TvUnit t = tv_unit(d);
  if (!t.active) return;  // no cross-lane operations below
  float4 Pm = t.x0 > 0 ? p_in[(int64_t)(t.x0 - 1) * YZ + t.col] : make_float4(0.f, 0.f, 0.f, 0.f);
  float4 P = p_in[(int64_t)t.x0 * YZ + t.col];
  for (uint32_t x = t.x0; x < t.x1; ++x) {
    const int64_t L = (int64_t)x * YZ + t.col;
    const float4 Pp = x + 1 < X ? p_in[L + YZ] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 acc = tv_dir(p_in, L, P, Pm, Pp, x, t.y, t.z, X, Y, Z, YZ);
    const float4 G0 = g[L];
    // zero_grad: the scatter's cells go back to zero once read (most cells already are)
    if (zero_grad && (G0.x != 0.f || G0.y != 0.f || G0.z != 0.f || G0.w != 0.f))
      store_stream(g + L, make_float4(0.f, 0.f, 0.f, 0.f));
    float4 G = G0;
    G.x += scale * acc.x;
    G.y += scale * acc.y;
    G.z += scale * acc.z;
    G.w += scale * acc.w;
    float4 Q = P, M = m[L], V = v[L];
    adam_elem(Q.x, G.x, M.x, V.x, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    adam_elem(Q.y, G.y, M.y, V.y, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    adam_elem(Q.z, G.z, M.z, V.z, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    adam_elem(Q.w, G.w, M.w, V.w, w1, beta2, w2, bc2_sqrt, eps, neg_step);
    p_out[L] = Q;
    m[L] = M;
    v[L] = V;
    if (gabs) {
      float4 A = gabs[L];
      A = make_float4(A.x + fabsf(G.x), A.y + fabsf(G.y), A.z + fabsf(G.z), A.w + fabsf(G.w));
      gabs[L] = A;
    }
    if (half_out) {
      if (half_dtype == PXL_DTYPE_BF16) {
        const uint32_t lo = to_bf16_rne(Q.x) | ((uint32_t)to_bf16_rne(Q.y) << 16);
        const uint32_t hi = to_bf16_rne(Q.z) | ((uint32_t)to_bf16_rne(Q.w) << 16);
        reinterpret_cast<uint2*>(half_out)[L] = make_uint2(lo, hi);
      } else {
        store_half(half_out, half_dtype, 4 * L + 0, Q.x);
        store_half(half_out, half_dtype, 4 * L + 1, Q.y);
        store_half(half_out, half_dtype, 4 * L + 2, Q.z);
        store_half(half_out, half_dtype, 4 * L + 3, Q.w);
      }
    }
    Pm = P;
    P = Pp;
  }
}

// out[0] = (float)(scale * sum of v[0, n)) in one launch and the same bits on every call.
// Workgroup b sums, in FP64 and a fixed order, the float4s b*256 + t + k*(gridDim*256) and
// writes partial[b]; the last workgroup to take a ticket adds the partials in index order
// with a fixed tree, writes out and puts the ticket back to zero for the next call.
constexpr int kSumBlocks = 64;

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m);
  return x;
}

__global__ __launch_bounds__(256) void k_scaled_sum(const float4* __restrict__ v4, const float* __restrict__ v,
                                                    int64_t n, double scale, float* __restrict__ out,
                                                    double* __restrict__ partial, uint32_t* __restrict__ ticket) {
  __shared__ double wsum[4];
  __shared__ uint32_t last;
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n4 = n >> 2, stride = (int64_t)gridDim.x * 256;
  int64_t j = (int64_t)blockIdx.x * 256 + tid;
  double s = 0.0;
  for (; j + 3 * stride < n4; j += 4 * stride) {  // four 16-byte loads in flight
    const float4 a = v4[j], b = v4[j + stride], c = v4[j + 2 * stride], d = v4[j + 3 * stride];
    s += (((double)a.x + a.y) + ((double)a.z + a.w)) + (((double)b.x + b.y) + ((double)b.z + b.w));
    s += (((double)c.x + c.y) + ((double)c.z + c.w)) + (((double)d.x + d.y) + ((double)d.z + d.w));
  }
  for (; j < n4; j += stride) {
    const float4 a = v4[j];
    s += ((double)a.x + a.y) + ((double)a.z + a.w);
  }
  if (blockIdx.x == 0 && tid < (int)(n & 3)) s += (double)v[4 * n4 + tid];
  s = wave_sum_f64(s);
  if (lane == 0) wsum[wave] = s;
  __syncthreads();
  if (tid == 0) {
    partial[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    __threadfence();  // the partial is visible device-wide before the ticket is taken
    last = atomicAdd(ticket, 1u) == gridDim.x - 1u ? 1u : 0u;
  }
  __syncthreads();
  if (last == 0u || wave != 0) return;
  __threadfence();
  double p = lane < (int)gridDim.x ? __hip_atomic_load(partial + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                   : 0.0;
  p = wave_sum_f64(p);
  if (lane == 0) {
    out[0] = (float)(p * scale);
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

constexpr int kNoAllocScratchBytes = 1024;

__global__ void k_zero_words(uint32_t* __restrict__ p, int n) {
  const int i = (int)threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace
}  // namespace pxl

// Scratch of pxl_scaled_sum_noalloc: the partials and ticket of k_scaled_sum in the library's own
// device memory, 1 KiB (each GPU holds one copy, at its own address, until the process ends).
// Host code never uses this name as an address -- there it names a host-side stand-in -- only
// the per-device pointer hipGetSymbolAddress returns for it.
static __device__ __align__(256) unsigned char scratch[pxl::kNoAllocScratchBytes];

// ============================================================================= host ABI

namespace {

using pxl::Dims;

int need(bool ok, const char* what) { return ok ? PXL_OK : fail(PXL_E_INVALID, "%s", what); }

int f32_grid(const pxl_grid* g, Dims& d) {
  int rc = pxl_host::check_grid(g);
  if (rc) return rc;
  if (g->dtype != PXL_DTYPE_F32) return fail(PXL_E_DTYPE, "this entry point takes a float32 grid");
  d = Dims{g->dims[0], g->dims[1], g->dims[2]};
  return PXL_OK;
}

int dims3(const int64_t dims[3], Dims& d) {
  if (!dims) return fail(PXL_E_INVALID, "dims is NULL");
  int rc = pxl_host::check_dims(dims[0], dims[1], dims[2]);
  if (rc) return rc;
  d = Dims{dims[0], dims[1], dims[2]};
  return PXL_OK;
}


constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the device that runs `stream` (NULL: the current device),
// looked up once per device with that device current; its first use on a device also clears it
// (the ticket must start at zero) with a fill kernel on `stream`, which may be captured in a
// graph (no memset node).
int scratch_for(hipStream_t stream, void** out) {
  int cur = 0, dev = 0;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return pxl_host::hip_fail(e, "pxl_scaled_sum_noalloc: hipGetDevice");
  dev = cur;
  if (stream != nullptr) {
    e = hipStreamGetDevice(stream, &dev);
    if (e != hipSuccess) return pxl_host::hip_fail(e, "pxl_scaled_sum_noalloc: hipStreamGetDevice");
  }
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(PXL_E_INVALID, "pxl_scaled_sum_noalloc: device %d beyond the %d it keeps scratch for", dev,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    if (dev != cur && (e = hipSetDevice(dev)) != hipSuccess) return pxl_host::hip_fail(e, "pxl_scaled_sum_noalloc: hipSetDevice");
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p != nullptr) {
      hipLaunchKernelGGL(pxl::k_zero_words, dim3(1), dim3(pxl::kNoAllocScratchBytes / 4), 0, stream,
                         reinterpret_cast<uint32_t*>(p), pxl::kNoAllocScratchBytes / 4);
      e = hipGetLastError();
    }
    if (dev != cur) (void)hipSetDevice(cur);
    if (e != hipSuccess) return pxl_host::hip_fail(e, "pxl_scaled_sum_noalloc: scratch set-up");
    if (p == nullptr) return fail(PXL_E_HIP, "pxl_scaled_sum_noalloc: hipGetSymbolAddress gave NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PXL_OK;
}

}  // namespace

extern "C" {

int pxl_sample_points(const float* origins, int64_t rays_per_origin, const float* dirs,
                      int64_t n_rays, int32_t n_samples, float delta, float near, float* out,
                      void* stream) {
  int rc;
  if ((rc = need(n_rays == 0 || n_samples == 0 || (origins && dirs && out), "origins, dirs and out are required")) ||
      (rc = need(n_rays >= 0 && n_samples >= 0 && rays_per_origin > 0, "negative size")))
    return rc;
  if (n_rays == 0 || n_samples == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_sample_points, dim3(stride_blocks(n_rays * n_samples, 256)), dim3(256), 0,
                     (hipStream_t)stream, origins, rays_per_origin, dirs, n_rays, n_samples, delta,
                     near, out);
  return launched("pxl_sample_points");
}

int pxl_coord_min(const float* coords, int64_t n, float* out3, void* stream) {
  int rc;
  if ((rc = need(coords && out3 && n > 0, "coords, out3 and n > 0 are required"))) return rc;
  hipLaunchKernelGGL(pxl::k_init_min, dim3(1), dim3(64), 0, (hipStream_t)stream, out3);
  hipLaunchKernelGGL(pxl::k_coord_min, dim3(stride_blocks(n, 256, 2048)), dim3(256), 0,
                     (hipStream_t)stream, coords, n, out3);
  return launched("pxl_coord_min");
}

int pxl_normalize_points(const float* p, int64_t n, const float* min3, float pd, float inv_pd,
                         int32_t flags, float* out, void* stream) {
  int rc;
  if ((rc = need(n >= 0 && (n == 0 || (p && min3 && out)), "p, min3, out required and n >= 0")) ||
      (rc = need((flags & ~PXL_FLAG_EXACT_DIV) == 0, "unknown flags")))
    return rc;
  if (n == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_normalize, dim3(stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     p, n, min3, pd, inv_pd, (flags & PXL_FLAG_EXACT_DIV) ? 1 : 0, out);
  return launched("pxl_normalize_points");
}

int pxl_in_bounds(const void* idx, int32_t is_float, int64_t n, const int64_t dims[3],
                  uint8_t* mask, void* stream) {
  int rc;
  Dims d;
  if ((rc = need(n >= 0 && (n == 0 || (idx && mask)), "idx, mask required and n >= 0")) || (rc = dims3(dims, d)))
    return rc;
  if (n == 0) return PXL_OK;
  if (is_float)
    hipLaunchKernelGGL(pxl::k_in_bounds_f32, dim3(stride_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float*)idx, n, d, mask);
  else
    hipLaunchKernelGGL(pxl::k_in_bounds_i64, dim3(stride_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, (const int64_t*)idx, n, d, mask);
  return launched("pxl_in_bounds");
}

int pxl_wrap_indices(int64_t* idx, int64_t n, const int64_t dims[3], void* stream) {
  int rc;
  Dims d;
  if ((rc = need(n >= 0 && (n == 0 || idx), "idx required and n >= 0")) || (rc = dims3(dims, d))) return rc;
  if (n == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_wrap, dim3(stride_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, idx,
                     n, d);
  return launched("pxl_wrap_indices");
}

int pxl_nearest_gather(const float* n, int64_t N, const pxl_grid* grid, float* vals,
                       uint8_t* mask, void* stream) {
  int rc;
  Dims d;
  if ((rc = f32_grid(grid, d)) || (rc = need(N >= 0 && (N == 0 || (n && vals)), "n, vals required and N >= 0")))
    return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_nearest_gather, dim3(stride_blocks(N, 256)), dim3(256), 0,
                     (hipStream_t)stream, n, N, d, (const float4*)grid->data, (float4*)vals, mask);
  return launched("pxl_nearest_gather");
}

int pxl_nearest_scatter(const float* n, int64_t N, const pxl_grid* grid, const float* grad_vals,
                        float* grad_grid, void* stream) {
  int rc;
  Dims d;
  if ((rc = f32_grid(grid, d)) ||
      (rc = need(N >= 0 && (N == 0 || (n && grad_vals && grad_grid)), "n, grad_vals, grad_grid required")))
    return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_nearest_scatter, dim3(stride_blocks(N, 256)), dim3(256), 0,
                     (hipStream_t)stream, n, N, d, (const float4*)grad_vals, grad_grid);
  return launched("pxl_nearest_scatter");
}

int pxl_corner_indices(const float* n, int64_t N, int64_t* corners, void* stream) {
  int rc;
  if ((rc = need(N >= 0 && (N == 0 || (n && corners)), "n, corners required and N >= 0"))) return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_corner_indices, dim3(stride_blocks(N, 256)), dim3(256), 0,
                     (hipStream_t)stream, n, N, corners);
  return launched("pxl_corner_indices");
}

int pxl_collect_cells(const float* n, int64_t N, const pxl_grid* grid, float* out, uint8_t* mask,
                      void* stream) {
  int rc;
  Dims d;
  if ((rc = f32_grid(grid, d)) || (rc = need(N >= 0 && (N == 0 || (n && out)), "n, out required and N >= 0")))
    return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_collect, dim3(stride_blocks(N, 256)), dim3(256), 0, (hipStream_t)stream,
                     n, N, d, (const float4*)grid->data, (float4*)out, mask);
  return launched("pxl_collect_cells");
}

int pxl_collect_cells_bwd(const float* n, int64_t N, const pxl_grid* grid, const float* grad_out,
                          float* grad_grid, void* stream) {
  int rc;
  Dims d;
  if ((rc = f32_grid(grid, d)) ||
      (rc = need(N >= 0 && (N == 0 || (n && grad_out && grad_grid)), "n, grad_out, grad_grid required")))
    return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_collect_bwd, dim3(stride_blocks(N, 256)), dim3(256), 0,
                     (hipStream_t)stream, n, N, d, (const float4*)grad_out, grad_grid);
  return launched("pxl_collect_cells_bwd");
}

int pxl_trilinear(const float* n, const int64_t* corners, int64_t N, const pxl_grid* grid,
                  float* out, void* stream) {
  int rc;
  Dims d;
  if ((rc = f32_grid(grid, d)) ||
      (rc = need(N >= 0 && (N == 0 || (n && corners && out)), "n, corners, out required and N >= 0")))
    return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_trilinear, dim3(stride_blocks(N, 256)), dim3(256), 0, (hipStream_t)stream,
                     n, corners, N, d, (const float4*)grid->data, (float4*)out);
  return launched("pxl_trilinear");
}

int pxl_trilinear_bwd(const float* n, const int64_t* corners, int64_t N, const pxl_grid* grid,
                      const float* grad_out, float* grad_grid, void* stream) {
  int rc;
  Dims d;
  if ((rc = f32_grid(grid, d)) ||
      (rc = need(N >= 0 && (N == 0 || (n && corners && grad_out && grad_grid)), "pointers required")))
    return rc;
  if (N == 0) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_trilinear_bwd, dim3(stride_blocks(N, 256)), dim3(256), 0,
                     (hipStream_t)stream, n, corners, N, d, (const float4*)grad_out, grad_grid);
  return launched("pxl_trilinear_bwd");
}

int pxl_composite_fwd(const float* samples, int64_t R, int32_t S, float* out, void* stream) {
  int rc;
  if ((rc = need(R >= 0 && S >= 0 && (R == 0 || ((samples || S == 0) && out)), "samples, out required; sizes >= 0")))
    return rc;
  if (R == 0) return PXL_OK;
  if (S > 64 * pxl::kMaxCompositeChunks) return fail(PXL_E_SIZE, "S > 65536");
  for (int64_t off = 0; off < R; off += int64_t(1) << 26) {
    const int64_t n = std::min<int64_t>(R - off, int64_t(1) << 26);
    hipLaunchKernelGGL(pxl::k_composite_fwd, dim3((unsigned)((n + 3) / 4)), dim3(pxl::kBlock), 0,
                       (hipStream_t)stream, (const float4*)samples + off * S, n, S,
                       (float4*)out + off);
  }
  return launched("pxl_composite_fwd");
}

int pxl_composite_bwd(const float* samples, int64_t R, int32_t S, const float* grad_out,
                      float* grad_samples, void* stream) {
  int rc;
  if ((rc = need(R >= 0 && S >= 0 && (R == 0 || S == 0 || (samples && grad_out && grad_samples)), "pointers required")))
    return rc;
  if (R == 0 || S == 0) return PXL_OK;
  if (S > 64 * pxl::kMaxCompositeChunks) return fail(PXL_E_SIZE, "S > 65536");
  for (int64_t off = 0; off < R; off += int64_t(1) << 26) {
    const int64_t n = std::min<int64_t>(R - off, int64_t(1) << 26);
    hipLaunchKernelGGL(pxl::k_composite_bwd, dim3((unsigned)((n + 3) / 4)), dim3(pxl::kBlock), 0,
                       (hipStream_t)stream, (const float4*)samples + off * S, n, S,
                       (const float4*)grad_out + off, (float4*)grad_samples + off * S);
  }
  return launched("pxl_composite_bwd");
}

int pxl_grid_coords(int64_t X, int64_t Y, int64_t Z, float pd, float* coords, int64_t* mesh,
                    void* stream) {
  int rc;
  if ((rc = pxl_host::check_dims(X, Y, Z)) || (rc = need(coords != nullptr, "coords is NULL")))
    return rc;
  auto c = [](int64_t s) { return (float)((s + 1) / 2); };  // ceil(s/2)
  hipLaunchKernelGGL(pxl::k_grid_coords, dim3(stride_blocks(X * Y * Z, 256)), dim3(256), 0,
                     (hipStream_t)stream, Dims{X, Y, Z}, c(X), c(Y), c(Z), pd, coords, mesh);
  return launched("pxl_grid_coords");
}

int64_t pxl_avg_pool3d_workspace_floats(int64_t X, int64_t Y, int64_t Z, int32_t k, int32_t s) {
  if (X <= 0 || Y <= 0 || Z <= 0 || k < 1 || s < 1 || k > X || k > Y || k > Z) return -1;
  const int64_t oy = (Y - k) / s + 1, oz = (Z - k) / s + 1;
  return (X * Y * oz + X * oy * oz) * 4;
}

int pxl_avg_pool3d(const float* grid, int64_t X, int64_t Y, int64_t Z, int32_t k, int32_t s,
                   float* out, float* workspace, void* stream) {
  int rc;
  if ((rc = pxl_host::check_dims(X, Y, Z)) || (rc = need(grid && out && workspace, "grid, out and workspace required")) ||
      (rc = need(k >= 1 && s >= 1 && k <= X && k <= Y && k <= Z, "need 1 <= k <= dims, s >= 1")))
    return rc;
  const int64_t ox = (X - k) / s + 1, oy = (Y - k) / s + 1, oz = (Z - k) / s + 1;
  float4* a = reinterpret_cast<float4*>(workspace);  // (X, Y, oz)
  float4* b = a + X * Y * oz;                         // (X, oy, oz)
  hipStream_t st = (hipStream_t)stream;
  if (Z <= pxl::kRowLdsCells)
    hipLaunchKernelGGL(pxl::k_box_mean_rows,
                       dim3((unsigned)std::min<int64_t>((X * Y + pxl::kRowLdsCells / Z - 1) / (pxl::kRowLdsCells / Z),
                                                        int64_t(1) << 20)),
                       dim3(256), 0, st, (const float4*)grid, X * Y, (int)Z, k, s, (int)oz, a);
  else
    hipLaunchKernelGGL(pxl::k_box_mean, dim3(stride_blocks(X * Y * oz, 256)), dim3(256), 0, st,
                       (const float4*)grid, X * Y, Z, (int64_t)1, k, s, oz, a);
  hipLaunchKernelGGL(pxl::k_box_mean, dim3(stride_blocks(X * oy * oz, 256)), dim3(256), 0, st,
                     (const float4*)a, X, Y, oz, k, s, oy, b);
  hipLaunchKernelGGL(pxl::k_box_mean, dim3(stride_blocks(ox * oy * oz, 256)), dim3(256), 0, st,
                     (const float4*)b, (int64_t)1, X, oy * oz, k, s, ox, (float4*)out);
  return launched("pxl_avg_pool3d");
}

int pxl_avg_pool3d_bwd(const float* grad_out, int64_t X, int64_t Y, int64_t Z, int32_t k,
                       int32_t s, float* grad_grid, int32_t accumulate, float* workspace, void* stream) {
  int rc;
  if ((rc = pxl_host::check_dims(X, Y, Z)) ||
      (rc = need(grad_out && grad_grid && workspace, "pointers required")) ||
      (rc = need(k >= 1 && s >= 1 && k <= X && k <= Y && k <= Z, "need 1 <= k <= dims, s >= 1")))
    return rc;
  const int64_t ox = (X - k) / s + 1, oy = (Y - k) / s + 1, oz = (Z - k) / s + 1;
  float4* d = reinterpret_cast<float4*>(workspace);  // (X, Y, oz)
  float4* c = d + X * Y * oz;                         // (X, oy, oz)
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(pxl::k_box_mean_adjoint, dim3(stride_blocks(X * oy * oz, 256)), dim3(256), 0, st,
                     (const float4*)grad_out, (int64_t)1, X, oy * oz, k, s, ox, 0, c);
  hipLaunchKernelGGL(pxl::k_box_mean_adjoint, dim3(stride_blocks(X * Y * oz, 256)), dim3(256), 0, st,
                     (const float4*)c, X, Y, oz, k, s, oy, 0, d);
  hipLaunchKernelGGL(pxl::k_box_mean_adjoint, dim3(stride_blocks(X * Y * Z, 256)), dim3(256), 0, st,
                     (const float4*)d, X * Y, Z, (int64_t)1, k, s, oz, accumulate ? 1 : 0, (float4*)grad_grid);
  return launched("pxl_avg_pool3d_bwd");
}

int64_t pxl_scaled_sum_work_bytes(void) { return 1024; }

int pxl_scaled_sum(const float* v, int64_t n, double scale, float* out, void* work, void* stream) {
  if (n < 0) return fail(PXL_E_INVALID, "n must be >= 0");
  if (!out || !work || (n > 0 && !v)) return fail(PXL_E_INVALID, "v, out and work are required");
  if (reinterpret_cast<uintptr_t>(v) % 16 != 0 || reinterpret_cast<uintptr_t>(work) % 16 != 0)
    return fail(PXL_E_INVALID, "v and work must be 16-byte aligned");
  static_assert(pxl::kSumBlocks * sizeof(double) + sizeof(uint32_t) <= 1024, "work layout");
  double* partial = reinterpret_cast<double*>(work);
  uint32_t* ticket = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(work) + pxl::kSumBlocks * sizeof(double));
  hipLaunchKernelGGL(pxl::k_scaled_sum, dim3(pxl::kSumBlocks), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(v), v, n, scale, out, partial, ticket);
  return launched("pxl_scaled_sum");
}

int64_t pxl_scaled_sum_noalloc_scratch_bytes(void) { return pxl::kNoAllocScratchBytes; }

int pxl_scaled_sum_noalloc(const float* v, int64_t n, double scale, float* out, void* stream) {
  if (n < 0) return fail(PXL_E_INVALID, "n must be >= 0");
  if (!out || (n > 0 && !v)) return fail(PXL_E_INVALID, "v and out are required");
  if (reinterpret_cast<uintptr_t>(v) % 16 != 0) return fail(PXL_E_INVALID, "v must be 16-byte aligned");
  static_assert(pxl::kSumBlocks * sizeof(double) + sizeof(uint32_t) <= (size_t)pxl::kNoAllocScratchBytes,
                "scratch layout");
  void* work = nullptr;
  int rc = scratch_for((hipStream_t)stream, &work);
  if (rc) return rc;
  double* partial = reinterpret_cast<double*>(work);
  uint32_t* ticket = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(work) + pxl::kSumBlocks * sizeof(double));
  hipLaunchKernelGGL(pxl::k_scaled_sum, dim3(pxl::kSumBlocks), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(v), v, n, scale, out, partial, ticket);
  return launched("pxl_scaled_sum_noalloc");
}

int pxl_tv_sumsq(const float* grid, int64_t X, int64_t Y, int64_t Z, double* sum_sq, void* stream) {
  int rc;
  if ((rc = pxl_host::check_dims(X, Y, Z)) || (rc = need(grid && sum_sq, "grid and sum_sq required")))
    return rc;
  if (pxl::tv_units(X, Y, Z) > int64_t(0x7fffffff)) return fail(PXL_E_SIZE, "grid too large for pxl_tv_sumsq");
  hipLaunchKernelGGL(pxl::k_tv_sumsq, dim3((unsigned)pxl::tv_units(X, Y, Z)), dim3(256), 0,
                     (hipStream_t)stream, (const float4*)grid, Dims{X, Y, Z}, sum_sq);
  return launched("pxl_tv_sumsq");
}

int pxl_tv_bwd(const float* grid, int64_t X, int64_t Y, int64_t Z, const double* sum_sq,
               float weight, float* grad_grid, void* stream) {
  int rc;
  if ((rc = pxl_host::check_dims(X, Y, Z)) ||
      (rc = need(grid && sum_sq && grad_grid, "grid, sum_sq and grad_grid required")) ||
      (rc = need(((reinterpret_cast<uintptr_t>(grid) | reinterpret_cast<uintptr_t>(grad_grid)) & 15u) == 0,
                 "grid and grad_grid must be 16-byte aligned")))
    return rc;
  if (pxl::tv_units(X, Y, Z) > int64_t(0x7fffffff)) return fail(PXL_E_SIZE, "grid too large for pxl_tv_bwd");
  hipLaunchKernelGGL(pxl::k_tv_bwd, dim3((unsigned)pxl::tv_units(X, Y, Z)), dim3(256), 0,
                     (hipStream_t)stream, (const float4*)grid, Dims{X, Y, Z}, sum_sq, weight,
                     reinterpret_cast<float4*>(grad_grid));
  return launched("pxl_tv_bwd");
}

int pxl_adam_tv_step(const float* param_in, float* param_out, float* grad, float* exp_avg,
                     float* exp_avg_sq, float* grad_abs, int64_t X, int64_t Y, int64_t Z,
                     const double* tv_sum_sq, float tv_weight, double lr, double beta1, double beta2,
                     double eps, int64_t step, void* half_out, int32_t half_dtype, int32_t zero_grad,
                     void* stream) {
  int rc;
  auto al16 = [](const void* q) { return q == nullptr || (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
  if ((rc = pxl_host::check_dims(X, Y, Z)) ||
      (rc = need(param_in && param_out && grad && exp_avg && exp_avg_sq && tv_sum_sq,
                 "param_in, param_out, grad, moments and tv_sum_sq required")) ||
      (rc = need(param_in != param_out, "param_out must not be param_in (the TV stencil reads neighbours)")) ||
      (rc = need(step >= 1, "step counts from 1")) ||
      (rc = need(!half_out || half_dtype == PXL_DTYPE_BF16 || half_dtype == PXL_DTYPE_F16,
                 "half_dtype must be PXL_DTYPE_BF16 or PXL_DTYPE_F16")) ||
      (rc = need(al16(param_in) && al16(param_out) && al16(grad) && al16(exp_avg) && al16(exp_avg_sq) &&
                     al16(grad_abs) && (half_out == nullptr || (reinterpret_cast<uintptr_t>(half_out) & 7u) == 0),
                 "float arrays must be 16-byte aligned (half_out 8-byte)")))
    return rc;
  const double bc1 = 1.0 - std::pow(beta1, (double)step);
  const double bc2 = 1.0 - std::pow(beta2, (double)step);
  const float neg_step = (float)(-(lr / bc1));
  const float bc2_sqrt = (float)std::pow(bc2, 0.5);
  const int64_t units = pxl::tv_units(X, Y, Z);
  if (units > int64_t(0x7fffffff)) return fail(PXL_E_SIZE, "grid too large for pxl_adam_tv_step's launch");
  hipLaunchKernelGGL(pxl::k_adam_tv, dim3((unsigned)units), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(param_in), reinterpret_cast<float4*>(param_out),
                     reinterpret_cast<float4*>(grad), reinterpret_cast<float4*>(exp_avg),
                     reinterpret_cast<float4*>(exp_avg_sq), reinterpret_cast<float4*>(grad_abs), Dims{X, Y, Z},
                     tv_sum_sq, tv_weight, (float)(1.0 - beta1), (float)beta2, (float)(1.0 - beta2), bc2_sqrt,
                     (float)eps, neg_step, half_out, half_dtype, zero_grad ? 1 : 0);
  return launched("pxl_adam_tv_step");
}

int pxl_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                  float* grad_abs, int64_t n, double lr, double beta1, double beta2, double eps,
                  int64_t step, void* half_out, int32_t half_dtype, void* stream) {
  int rc;
  if ((rc = need(n >= 0 && (n == 0 || (param && grad && exp_avg && exp_avg_sq)), "param, grad, moments required")) ||
      (rc = need(step >= 1, "step counts from 1")) ||
      (rc = need(!half_out || half_dtype == PXL_DTYPE_BF16 || half_dtype == PXL_DTYPE_F16,
                 "half_dtype must be PXL_DTYPE_BF16 or PXL_DTYPE_F16")))
    return rc;
  if (n == 0) return PXL_OK;
  // scalar terms in double, as torch.optim.Adam forms them in Python
  const double bc1 = 1.0 - std::pow(beta1, (double)step);
  const double bc2 = 1.0 - std::pow(beta2, (double)step);
  const float neg_step = (float)(-(lr / bc1));
  const float bc2_sqrt = (float)std::pow(bc2, 0.5);
  auto al16 = [](const void* q) { return q == nullptr || (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
  const bool vec = al16(param) && al16(grad) && al16(exp_avg) && al16(exp_avg_sq) && al16(grad_abs) &&
                   (half_out == nullptr || (reinterpret_cast<uintptr_t>(half_out) & 7u) == 0);
  const int64_t n4 = vec ? n / 4 : 0;
  hipLaunchKernelGGL(pxl::k_adam, dim3(stride_blocks(n4 > 0 ? n4 : n, 256)), dim3(256), 0, (hipStream_t)stream,
                     param, grad, exp_avg, exp_avg_sq, grad_abs, n, n4, (float)(1.0 - beta1),
                     (float)beta2, (float)(1.0 - beta2), bc2_sqrt, (float)eps, neg_step, half_out,
                     half_dtype);
  return launched("pxl_adam_step");
}

}  // extern "C"
