// pxl_device.h -- device building blocks of the gfx950 render kernels.
//
// Everything here runs with one 64-lane wavefront per ray: lane l of chunk c handles
// sample k = 64c + l + 1.  Cross-lane work uses DPP (row shifts, row broadcasts, wave
// shifts) and v_readlane; all wave primitives must be called with every lane active.
//
// Floating-point order.  The translation unit is compiled with -ffp-contract=off, so each
// `a*b + c` below rounds twice, exactly as the reference's separate torch ops do; an FMA
// appears only where written as fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pxl.h"

namespace pxl {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// ----------------------------------------------------------------------------- DPP
// GFX9/CDNA DPP control words.
enum : int {
  DPP_ROW_SHL1 = 0x101, DPP_ROW_SHL2 = 0x102, DPP_ROW_SHL4 = 0x104, DPP_ROW_SHL8 = 0x108,
  DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118,
  DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138,
  DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143,
};

// Lane l receives `src` from the lane CTRL selects; lanes with no valid source (or rows
// outside ROW_MASK) receive `old`.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src),
                                                    CTRL, ROW_MASK, 0xf, false));
}

__device__ __forceinline__ float readlane(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// Inclusive product over lanes 0..l (Hillis-Steele within 16-lane rows, then row
// broadcasts of lanes 15 and 31).
__device__ __forceinline__ float wave_prefix_mul(float x) {
  x *= dpp<DPP_ROW_SHR1>(1.f, x);
  x *= dpp<DPP_ROW_SHR2>(1.f, x);
  x *= dpp<DPP_ROW_SHR4>(1.f, x);
  x *= dpp<DPP_ROW_SHR8>(1.f, x);
  x *= dpp<DPP_ROW_BCAST15, 0xa>(1.f, x);
  x *= dpp<DPP_ROW_BCAST31, 0xc>(1.f, x);
  return x;
}

// Exclusive form of an inclusive scan: lane l gets lane l-1's value, lane 0 gets `ident`.
__device__ __forceinline__ float wave_shift_up(float incl, float ident) {
  return dpp<DPP_WAVE_SHR1>(ident, incl);
}

// Lane l gets lane l+1's value, lane 63 gets `fill`.
__device__ __forceinline__ float wave_shift_down(float x, float fill) {
  return dpp<DPP_WAVE_SHL1>(fill, x);
}

// Sum over all 64 lanes, result in every lane.
__device__ __forceinline__ float wave_sum(float x) {
  x += dpp<DPP_ROW_SHR1>(0.f, x);
  x += dpp<DPP_ROW_SHR2>(0.f, x);
  x += dpp<DPP_ROW_SHR4>(0.f, x);
  x += dpp<DPP_ROW_SHR8>(0.f, x);
  x += dpp<DPP_ROW_BCAST15, 0xa>(0.f, x);
  x += dpp<DPP_ROW_BCAST31, 0xc>(0.f, x);
  return readlane(x, 63);
}

// Suffix scan of affine maps f_l(q) = A_l + B_l q under composition f_l o f_{l+1} o ... o
// f_63: on return lane l holds the composed map.  A is 4-wide (RGB and alpha), B scalar.
// Within rows: row_shl Hillis-Steele; across rows: the row totals (lanes 16, 32, 48) are
// read into scalars, composed, and applied as each row's tail.
struct Affine4 {
  float a0, a1, a2, a3, b;
};

template <int CTRL>
__device__ __forceinline__ void affine_step(Affine4& f) {
  const float a0 = dpp<CTRL>(0.f, f.a0), a1 = dpp<CTRL>(0.f, f.a1);
  const float a2 = dpp<CTRL>(0.f, f.a2), a3 = dpp<CTRL>(0.f, f.a3);
  const float b = dpp<CTRL>(1.f, f.b);
  f.a0 = f.a0 + f.b * a0;
  f.a1 = f.a1 + f.b * a1;
  f.a2 = f.a2 + f.b * a2;
  f.a3 = f.a3 + f.b * a3;
  f.b = f.b * b;
}

__device__ __forceinline__ Affine4 affine_readlane(const Affine4& f, int lane) {
  return Affine4{readlane(f.a0, lane), readlane(f.a1, lane), readlane(f.a2, lane),
                 readlane(f.a3, lane), readlane(f.b, lane)};
}

// g o h
__device__ __forceinline__ Affine4 affine_compose(const Affine4& g, const Affine4& h) {
  return Affine4{g.a0 + g.b * h.a0, g.a1 + g.b * h.a1, g.a2 + g.b * h.a2, g.a3 + g.b * h.a3,
                 g.b * h.b};
}

__device__ __forceinline__ void wave_suffix_affine(Affine4& f, int lane) {
  affine_step<DPP_ROW_SHL1>(f);
  affine_step<DPP_ROW_SHL2>(f);
  affine_step<DPP_ROW_SHL4>(f);
  affine_step<DPP_ROW_SHL8>(f);
  const Affine4 r1 = affine_readlane(f, 16), r2 = affine_readlane(f, 32),
                r3 = affine_readlane(f, 48);
  const Affine4 t2 = r3;
  const Affine4 t1 = affine_compose(r2, r3);
  const Affine4 t0 = affine_compose(r1, t1);
  const int row = lane >> 4;
  Affine4 t = row == 0 ? t0 : (row == 1 ? t1 : t2);
  if (row != 3) f = affine_compose(f, t);
}

// Ordering of LDS traffic between lanes of one wavefront (no workgroup barrier: the four
// waves of a block work on different rays and may be at different points).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of m below this lane.
__device__ __forceinline__ int lane_rank(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Streaming 16-byte accesses (the `nt` cache policy): for data touched once per pass -- dense
// gradient rows, cleared block entries, image pixels -- so that it does not evict the lines
// other waves are about to reuse from L2 and the memory-side cache.
typedef float float4_nt __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 load_stream(const float4* p) {
  const float4_nt t = __builtin_nontemporal_load(reinterpret_cast<const float4_nt*>(p));
  return make_float4(t.x, t.y, t.z, t.w);
}

__device__ __forceinline__ void store_stream(float4* p, float4 v) {
  __builtin_nontemporal_store((float4_nt){v.x, v.y, v.z, v.w}, reinterpret_cast<float4_nt*>(p));
}

// ----------------------------------------------------------------------------- grid cells

template <int DT>
__device__ __forceinline__ float4 load_cell(const void* grid, uint32_t cell);

template <>
__device__ __forceinline__ float4 load_cell<PXL_DTYPE_F32>(const void* grid, uint32_t cell) {
  return reinterpret_cast<const float4*>(grid)[cell];
}

template <>
__device__ __forceinline__ float4 load_cell<PXL_DTYPE_BF16>(const void* grid, uint32_t cell) {
  const uint2 w = reinterpret_cast<const uint2*>(grid)[cell];
  return make_float4(__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xffff0000u),
                     __uint_as_float(w.y << 16), __uint_as_float(w.y & 0xffff0000u));
}

__device__ __forceinline__ float half_bits_to_float(uint32_t h) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
}

template <>
__device__ __forceinline__ float4 load_cell<PXL_DTYPE_F16>(const void* grid, uint32_t cell) {
  const uint2 w = reinterpret_cast<const uint2*>(grid)[cell];
  return make_float4(half_bits_to_float(w.x & 0xffffu), half_bits_to_float(w.x >> 16),
                     half_bits_to_float(w.y & 0xffffu), half_bits_to_float(w.y >> 16));
}

// Bit c set when channel c lies in [0,1]: torch.clamp's gradient passes at both bounds.
__device__ __forceinline__ uint32_t clip_bits(const float4& v) {
  return (uint32_t)(v.x >= 0.f && v.x <= 1.f) | ((uint32_t)(v.y >= 0.f && v.y <= 1.f) << 1) |
         ((uint32_t)(v.z >= 0.f && v.z <= 1.f) << 2) | ((uint32_t)(v.w >= 0.f && v.w <= 1.f) << 3);
}

// ----------------------------------------------------------------------------- occupancy
// Occupancy bits are stored in 8x8x8 tiles of 512 bits, one 64-byte line (16 words) each:
// tile (x>>3, y>>3, z>>3) holds words ((tx*TY + ty)*TZ + tz)*16 .. +15 (TY = ceil(Y/8),
// TZ = ceil(Z/8)), and cell (x, y, z) is bit (y&3)*8 + (z&7) of word (x&7)*2 + ((y>>2)&1) of
// its tile.  The samples of a ray, in any direction, stay in one tile for several steps, so
// the occupancy loads of a chunk of 64 samples touch ~12 lines instead of ~34 with a linear
// bit order (sparse256 rays).
struct OccBit {
  uint32_t word, shift;
};

__device__ __forceinline__ OccBit occ_bit(uint32_t x, uint32_t y, uint32_t z, uint32_t TY,
                                          uint32_t TZ) {
  const uint32_t t = ((x >> 3) * TY + (y >> 3)) * TZ + (z >> 3);
  return OccBit{t * 16u + ((x & 7u) << 1) + ((y >> 2) & 1u), ((y & 3u) << 3) | (z & 7u)};
}

__device__ __forceinline__ bool occupied_at(const uint32_t* occ, uint32_t x, uint32_t y, uint32_t z,
                                            uint32_t TY, uint32_t TZ) {
  if (occ == nullptr) return true;
  const OccBit b = occ_bit(x, y, z, TY, TZ);
  return ((occ[b.word] >> b.shift) & 1u) != 0u;
}

// Cell (x, y, z) of occupancy word w, bit b (the inverse of occ_bit); coordinates may lie in
// the padding of a partial tile (>= the grid's dims).
__device__ __forceinline__ void occ_cell(uint32_t w, uint32_t b, uint32_t TY, uint32_t TZ,
                                         uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint32_t t = w >> 4, sub = w & 15u;
  const uint32_t tz = t % TZ, ty = (t / TZ) % TY, tx = t / (TZ * TY);
  x = tx * 8u + (sub >> 1);
  y = ty * 8u + (sub & 1u) * 4u + (b >> 3);
  z = tz * 8u + (b & 7u);
}

// ----------------------------------------------------------------------------- parameters

struct KParams {
  const void* grid;
  const uint32_t* occ;
  int X, Y, Z, n_samples;
  float fX, fY, fZ;
  float delta, near, t_stop, pd, inv_pd;
  float gx, gy, gz;
  int exact_div;
  int cpu_norm;  // ray directions normalised in ATen's CPU summation order (else its GPU order)
  int aux_stride;
  int corner_blocks;  // trilinear gradients go to per-cell corner-block entries (16 floats)
  uint32_t occ_ty, occ_tz;  // occupancy tiles along y and z (ceil(Y/8), ceil(Z/8))
  int64_t n_rays;
  int64_t ray_offset;
  // rays
  const float* origins;
  const float* dirs;
  const int32_t* cam_ids;
  const float* uv;
  const float* c2w;
  const float* imgs;
  int64_t n_cams, img_h, img_w;
  float theta;
};

struct Ray {
  float ox, oy, oz, dx, dy, dz;
  int cam;
  float u, v;
};

// Three-element norms in the summation orders of torch's two reductions (x.norm(dim=-1) on a
// length-3 axis, src/ray_sampling.py:218,262):
//   ATen CPU: sqrt(fma(z,z,fma(y,y,x*x)))
//   ATen GPU: sqrt((x*x + z*z) + y*y) -- the reduce kernel gives each output two threads,
//             thread 0 summing elements 0 and 2 and thread 1 element 1, then combines them.
// __builtin_sqrtf lowers to the correctly rounded sequence (v_sqrt_f32 plus the +-1 ulp
// fma checks); __fsqrt_rn lowers to the bare v_sqrt_f32, which is not correctly rounded.
__device__ __forceinline__ float norm3_cpu(float x, float y, float z) {
  return __builtin_sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
}

__device__ __forceinline__ float norm3_gpu(float x, float y, float z) {
  return __builtin_sqrtf((x * x + z * z) + y * y);
}

__device__ __forceinline__ float norm3(bool cpu_order, float x, float y, float z) {
  return cpu_order ? norm3_cpu(x, y, z) : norm3_gpu(x, y, z);
}

// src/ray_sampling.py:212-264 for one ray of the camera form (every op rounded as torch does).
__device__ __forceinline__ void camera_ray(const KParams& P, int64_t r, Ray& R) {
  // an out-of-range camera index is clamped (the reference raises IndexError instead)
  const int cam = (int)min(max((int64_t)P.cam_ids[r], (int64_t)0), P.n_cams - 1);
  const float u = P.uv[2 * r], v = P.uv[2 * r + 1];
  const float* M = P.c2w + (int64_t)cam * 16;
  const float xx = M[0], xy = M[4], xz = M[8];
  const float yx = M[1], yy = M[5], yz = M[9];
  const float zx = -M[2], zy = -M[6], zz = -M[10];
  const bool cpu = P.cpu_norm != 0;
  const float aspect = __fdiv_rn(norm3(cpu, xx, xy, xz), norm3(cpu, yx, yy, yz));
  const float us = P.theta * (u - 0.5f);
  const float vs = -((P.theta * __fdiv_rn(1.f, aspect)) * (v - 0.5f));
  const float dx = ((us * xx) + (vs * yx)) + zx;
  const float dy = ((us * xy) + (vs * yy)) + zy;
  const float dz = ((us * xz) + (vs * yz)) + zz;
  const float nd = norm3(cpu, dx, dy, dz);
  R.dx = __fdiv_rn(dx, nd);
  R.dy = __fdiv_rn(dy, nd);
  R.dz = __fdiv_rn(dz, nd);
  R.ox = M[3];
  R.oy = M[7];
  R.oz = M[11];
  R.cam = cam;
  R.u = u;
  R.v = v;
}

// Target pixel imgs[cam, min(round(W*v),W-1), min(round(H*u),H-1)] (src/ray_sampling.py:238-248;
// rows from v, columns from u).  Indices are additionally clamped into the image, where the
// reference would raise IndexError instead (only reachable for non-square images).
__device__ __forceinline__ float4 camera_target(const KParams& P, const Ray& R) {
  const float H = (float)P.img_h, W = (float)P.img_w;
  float col = fminf(__builtin_rintf(H * R.u), H - 1.f);
  float row = fminf(__builtin_rintf(W * R.v), W - 1.f);
  col = fminf(fmaxf(col, 0.f), W - 1.f);
  row = fminf(fmaxf(row, 0.f), H - 1.f);
  const int64_t pix = ((int64_t)R.cam * P.img_h + (int64_t)row) * P.img_w + (int64_t)col;
  return load_stream(reinterpret_cast<const float4*>(P.imgs) + pix);
}

// Both forms produce a whole Ray value that is assigned once: branches that stored to
// different fields of R last were merged by the compiler into one store through a selected
// field pointer, which forced the ray into scratch memory.
__device__ __forceinline__ void setup_ray(const KParams& P, int64_t r, Ray& R) {
  Ray q;
  if (P.origins != nullptr) {
    q.ox = P.origins[3 * r];
    q.oy = P.origins[3 * r + 1];
    q.oz = P.origins[3 * r + 2];
    q.dx = P.dirs[3 * r];
    q.dy = P.dirs[3 * r + 1];
    q.dz = P.dirs[3 * r + 2];
    q.cam = 0;
    q.u = q.v = 0.f;
  } else {
    camera_ray(P, r, q);
  }
  R.ox = q.ox; R.oy = q.oy; R.oz = q.oz;
  R.dx = q.dx; R.dy = q.dy; R.dz = q.dz;
  R.cam = q.cam; R.u = q.u; R.v = q.v;
}

// ----------------------------------------------------------------------------- geometry

// (p - g) / points_distance: a true division (EXACT) or a multiply by the rounded reciprocal.
template <bool EXACT>
__device__ __forceinline__ float to_index_t(const KParams& P, float p, float g) {
  const float diff = p - g;
  return EXACT ? __fdiv_rn(diff, P.pd) : diff * P.inv_pd;
}

// Continuous grid-index coordinate of sample k (src/ray_sampling.py:161-167, :13).
template <bool EXACT>
__device__ __forceinline__ void sample_index_t(const KParams& P, const Ray& R, int k, float& nx,
                                               float& ny, float& nz) {
  float t = P.delta * (float)k;
  t = t + P.near;
  nx = to_index_t<EXACT>(P, R.ox + R.dx * t, P.gx);
  ny = to_index_t<EXACT>(P, R.oy + R.dy * t, P.gy);
  nz = to_index_t<EXACT>(P, R.oz + R.dz * t, P.gz);
}

__device__ __forceinline__ void sample_index(const KParams& P, const Ray& R, int k, float& nx,
                                             float& ny, float& nz) {
  if (P.exact_div) sample_index_t<true>(P, R, k, nx, ny, nz);
  else sample_index_t<false>(P, R, k, nx, ny, nz);
}

// Nearest voxel: round half to even, integer in-bounds test (src/grid_functions.py:111-112).
__device__ __forceinline__ bool nearest_cell(const KParams& P, float nx, float ny, float nz,
                                             uint32_t& cell) {
  const float rx = __builtin_rintf(nx), ry = __builtin_rintf(ny), rz = __builtin_rintf(nz);
  // branch-free: the six tests combined with '&', the index computed for every lane and
  // selected (a lane outside the grid, or with a NaN coordinate, gets cell 0)
  const bool inb = (rx >= 0.f) & (rx < P.fX) & (ry >= 0.f) & (ry < P.fY) & (rz >= 0.f) & (rz < P.fZ);
  const uint32_t c = ((uint32_t)rx * (uint32_t)P.Y + (uint32_t)ry) * (uint32_t)P.Z + (uint32_t)rz;
  cell = inb ? c : 0u;
  return inb;
}

// Trilinear cell: float in-bounds test on the unrounded coordinate (src/grid_functions.py:162),
// ceil/floor corners wrapped (:168; ceil == X wraps to 0), fractions n - trunc(n) (:29).
struct TriCell {
  uint32_t xs[2], ys[2], zs[2];  // [0] = ceil, [1] = floor; xs pre-multiplied by Y*Z, ys by Z
  float fx, fy, fz;
};

__device__ __forceinline__ bool trilinear_cell(const KParams& P, float nx, float ny, float nz,
                                               TriCell& t) {
  const bool inb = nx >= 0.f && nx < P.fX && ny >= 0.f && ny < P.fY && nz >= 0.f && nz < P.fZ;
  if (!inb) return false;
  const uint32_t x0 = (uint32_t)floorf(nx), y0 = (uint32_t)floorf(ny), z0 = (uint32_t)floorf(nz);
  uint32_t x1 = (uint32_t)ceilf(nx), y1 = (uint32_t)ceilf(ny), z1 = (uint32_t)ceilf(nz);
  x1 = x1 >= (uint32_t)P.X ? x1 - (uint32_t)P.X : x1;
  y1 = y1 >= (uint32_t)P.Y ? y1 - (uint32_t)P.Y : y1;
  z1 = z1 >= (uint32_t)P.Z ? z1 - (uint32_t)P.Z : z1;
  const uint32_t yz = (uint32_t)P.Y * (uint32_t)P.Z;
  t.xs[0] = x1 * yz;
  t.xs[1] = x0 * yz;
  t.ys[0] = y1 * (uint32_t)P.Z;
  t.ys[1] = y0 * (uint32_t)P.Z;
  t.zs[0] = z1;
  t.zs[1] = z0;
  t.fx = nx - truncf(nx);
  t.fy = ny - truncf(ny);
  t.fz = nz - truncf(nz);
  return true;
}

// ----------------------------------------------------------------------------- samples

// Whether sample k lies in the grid (nearest: its rounded index, src/grid_functions.py:111-112;
// trilinear: the unrounded coordinate, :162), and the occupancy bit that decides it (nearest:
// its cell; trilinear: its floor cell); bit 0 of word 0 when it lies outside.  Branch-free:
// the tests are combined with '&' and the bit is computed for every lane and selected.
template <int INTERP, bool EXACT>
__device__ __forceinline__ bool locate(const KParams& P, const Ray& R, int k, OccBit& b) {
  float nx, ny, nz;
  sample_index_t<EXACT>(P, R, k, nx, ny, nz);
  float fx, fy, fz;
  bool inb;
  if constexpr (INTERP == PXL_INTERP_NEAREST) {
    fx = __builtin_rintf(nx);
    fy = __builtin_rintf(ny);
    fz = __builtin_rintf(nz);
    inb = (fx >= 0.f) & (fx < P.fX) & (fy >= 0.f) & (fy < P.fY) & (fz >= 0.f) & (fz < P.fZ);
  } else {
    inb = (nx >= 0.f) & (nx < P.fX) & (ny >= 0.f) & (ny < P.fY) & (nz >= 0.f) & (nz < P.fZ);
    fx = floorf(nx);
    fy = floorf(ny);
    fz = floorf(nz);
  }
  const OccBit o = occ_bit((uint32_t)fx, (uint32_t)fy, (uint32_t)fz, P.occ_ty, P.occ_tz);
  b.word = inb ? o.word : 0u;
  b.shift = inb ? o.shift : 0u;
  return inb;
}

// True when sample k lies in the grid and is not skipped by occupancy: the samples whose
// values the reference chain can make nonzero.  Geometry only -- no grid value is read.
// Every lane loads an occupancy word (word 0 for a sample outside the grid).
template <int INTERP>
__device__ __forceinline__ bool classify(const KParams& P, const Ray& R, int k) {
  OccBit b;
  const bool inb = P.exact_div ? locate<INTERP, true>(P, R, k, b) : locate<INTERP, false>(P, R, k, b);
  if (P.occ == nullptr) return inb;
  return inb & (((P.occ[b.word] >> b.shift) & 1u) != 0u);
}

// classify() for NC chunks of each of NR rays (sample k = 64c + lane + 1 of ray r): every sample
// is placed first (the division mode chosen once, outside the loops), then all occupancy words
// are loaded back to back and only then tested, so the loads of all chunks share one memory
// round trip.  kp[r][c]: in the grid, occupied and k <= n_samples.
template <int INTERP, int NR, int NC>
__device__ __forceinline__ void classify_chunks(const KParams& P, const Ray (&R)[NR], int lane,
                                                bool (&kp)[NR][NC]) {
  OccBit ob[NR][NC];
  bool inb[NR][NC];
  if (P.exact_div) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) inb[r][c] = locate<INTERP, true>(P, R[r], (c << 6) + lane + 1, ob[r][c]);
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) inb[r][c] = locate<INTERP, false>(P, R[r], (c << 6) + lane + 1, ob[r][c]);
  }
  if (P.occ != nullptr) {
    uint32_t w[NR][NC];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) w[r][c] = P.occ[ob[r][c].word];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) inb[r][c] = inb[r][c] & (((w[r][c] >> ob[r][c].shift) & 1u) != 0u);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) kp[r][c] = inb[r][c] & ((c << 6) + lane + 1 <= P.n_samples);
}

// torch.clamp(v, 0, 1): NaN stays NaN (v_med3 alone would return a bound)
__device__ __forceinline__ float clamp01(float v) {
  return __builtin_isnan(v) ? v : __builtin_amdgcn_fmed3f(v, 0.f, 1.f);
}

// Clipped value and clip-gradient bit of one channel (bit set iff 0 <= raw <= 1).
__device__ __forceinline__ float clip_ch(float raw, uint32_t& bits, int bit) {
  const float c = clamp01(raw);
  bits |= (uint32_t)(c == raw) << bit;
  return c;
}

// Value of a kept sample: nearest cell, or trilinear interpolation of the clipped corners
// (sum of the eight corner weights (wx*wy)*wz times the corner; the reference's nested lerps
// src/grid_functions.py:34-42 compute the same polynomial).  clip: bit c (nearest) or
// 4j + c (trilinear corner j = 4*bx + 2*by + bz, b = 0 ceil / 1 floor, the corner order of
// src/grid_functions.py:238-243).
template <int DT, int INTERP>
__device__ __forceinline__ void eval_kept(const KParams& P, const Ray& R, int k, float4& v,
                                          uint32_t& clip) {
  float nx, ny, nz;
  sample_index(P, R, k, nx, ny, nz);
  clip = 0u;
  if constexpr (INTERP == PXL_INTERP_NEAREST) {
    uint32_t cell;
    nearest_cell(P, nx, ny, nz, cell);
    const float4 raw = load_cell<DT>(P.grid, cell);
    v.x = clip_ch(raw.x, clip, 0);
    v.y = clip_ch(raw.y, clip, 1);
    v.z = clip_ch(raw.z, clip, 2);
    v.w = clip_ch(raw.w, clip, 3);
  } else {
    TriCell t;
    trilinear_cell(P, nx, ny, nz, t);
    float4 c[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      c[j] = load_cell<DT>(P.grid, t.xs[j >> 2] + t.ys[(j >> 1) & 1] + t.zs[j & 1]);
    const float ofx = 1.f - t.fx, ofy = 1.f - t.fy, ofz = 1.f - t.fz;
    v = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float w = ((j & 4 ? ofx : t.fx) * (j & 2 ? ofy : t.fy)) * (j & 1 ? ofz : t.fz);
      v.x = fmaf(w, clip_ch(c[j].x, clip, 4 * j + 0), v.x);
      v.y = fmaf(w, clip_ch(c[j].y, clip, 4 * j + 1), v.y);
      v.z = fmaf(w, clip_ch(c[j].z, clip, 4 * j + 2), v.z);
      v.w = fmaf(w, clip_ch(c[j].w, clip, 4 * j + 3), v.w);
    }
  }
}

// Append the kept samples among chunk k = 64c + lane + 1 (k <= S) to `list` in sample order;
// returns the new length.  Every lane must be active.
template <int INTERP>
__device__ __forceinline__ int append_kept(const KParams& P, const Ray& R, int k, int S,
                                           int* list, int n) {
  const bool keep = k <= S && classify<INTERP>(P, R, k);
  const uint64_t km = __ballot(keep);
  if (keep) list[n + lane_rank(km)] = k;
  return n + __popcll(km);
}

// ----------------------------------------------------------------------------- scatter

__device__ __forceinline__ void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

// Gradient items per sample, and the LDS staging capacity: the items of kStageSamples kept
// samples are staged and flushed at a time.  Trilinear stages 32 samples (256 items, 5 KiB
// per wave; with the second ray's sample list the pair kernel's wave holds 6 KiB), so LDS
// leaves room for six waves per SIMD, one more than the train step's registers allow.
template <int INTERP>
struct ScatterItems {
  static constexpr int kPerSample = INTERP == PXL_INTERP_NEAREST ? 1 : 8;
  static constexpr int kStageSamples = INTERP == PXL_INTERP_NEAREST ? kWave : 32;
  static constexpr int kPerChunk = kPerSample * kStageSamples;
};

// Per-sample gradient items: the cells sample k read and d loss / d cell value for each
// (clip-gradient mask applied), written to LDS at slot `base`.  The item's index is the cell
// (the flush adds into grad + 4*index), or, with corner blocks, entry*4 + slot: the entry is
// the floor cell's (x, y) row at the corner's z (floor z, or the wrapped ceil z), the slot
// 2bx + by, so the sample's eight items fill two 64-byte entries (adjacent unless z wraps).
template <int INTERP>
__device__ __forceinline__ void write_items(const KParams& P, const Ray& R, int k, uint32_t clip,
                                            const float4& gv, uint32_t* s_cell, float4* s_val,
                                            int base) {
  float nx, ny, nz;
  sample_index(P, R, k, nx, ny, nz);
  if constexpr (INTERP == PXL_INTERP_NEAREST) {
    uint32_t cell;
    nearest_cell(P, nx, ny, nz, cell);
    s_cell[base] = cell;
    s_val[base] = make_float4((clip & 1u) ? gv.x : 0.f, (clip & 2u) ? gv.y : 0.f,
                              (clip & 4u) ? gv.z : 0.f, (clip & 8u) ? gv.w : 0.f);
  } else {
    TriCell t;
    trilinear_cell(P, nx, ny, nz, t);
    const float ofx = 1.f - t.fx, ofy = 1.f - t.fy, ofz = 1.f - t.fz;
    const uint32_t row = t.xs[1] + t.ys[1];  // floor (x, y) row of the corner-block entries
#pragma unroll
    for (int bz = 0; bz < 2; ++bz) {
      const float wz = bz == 0 ? t.fz : ofz;
      const float4 gz = make_float4(gv.x * wz, gv.y * wz, gv.z * wz, gv.w * wz);
#pragma unroll
      for (int by = 0; by < 2; ++by) {
        const float wy = by == 0 ? t.fy : ofy;
        const float4 gy = make_float4(gz.x * wy, gz.y * wy, gz.z * wy, gz.w * wy);
#pragma unroll
        for (int bx = 0; bx < 2; ++bx) {
          const float wx = bx == 0 ? t.fx : ofx;
          const int j = 4 * bx + 2 * by + bz;
          const uint32_t cb = clip >> (4 * j);
          s_cell[base + j] = P.corner_blocks ? (row + t.zs[bz]) * 4u + (uint32_t)(2 * bx + by)
                                             : t.xs[bx] + t.ys[by] + t.zs[bz];
          s_val[base + j] = make_float4((cb & 1u) ? gy.x * wx : 0.f, (cb & 2u) ? gy.y * wx : 0.f,
                                        (cb & 4u) ? gy.z * wx : 0.f, (cb & 8u) ? gy.w * wx : 0.f);
        }
      }
    }
  }
}

// Scatter one chunk's gradients: kept lanes stage their items in this wave's LDS slice
// (compacted by rank), then the whole wave flushes them four lanes per cell, so one atomic
// instruction covers 16 cells and each cell's four channels go out as one 16-byte group.
// With RUNS, trilinear stages where at least a quarter of the samples share the previous
// sample's items (coarse grids: the pooled grids of progressive growing) are flushed by runs
// instead: each half-wave walks its 16 staged samples with one lane per (corner, channel),
// sums a run of samples with the same eight item indices in sample order and adds the sum
// once -- two memory requests per run instead of per sample.  A sample's items are fixed by
// its floor cell and by which coordinates are exact integers (there the ceil index equals the
// floor index), so a run is keyed on item 0 (all-ceil corner) and item NI - 1 (all-floor
// corner) together: the floor corner alone misses the integer case, the ceil corner alone
// confuses a sample at integer z = c with its neighbour of floor c - 1.
// Segmented inclusive sum over the wave's lanes in lane order: a lane with `head` starts a new
// segment.  Hillis-Steele within 16-lane rows, then the row-15 and row-31 broadcasts, with the
// operator (v, f) o (v', f') = (f' ? v' : v + v', f | f') -- each lane ends with the sum from its
// segment's head to itself.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ void seg_step(float4& v, int& f) {
  const int lf = __builtin_amdgcn_update_dpp(0, f, CTRL, ROW_MASK, 0xf, false);
  const float4 lv = make_float4(dpp<CTRL, ROW_MASK>(0.f, v.x), dpp<CTRL, ROW_MASK>(0.f, v.y),
                                dpp<CTRL, ROW_MASK>(0.f, v.z), dpp<CTRL, ROW_MASK>(0.f, v.w));
  v.x = f ? v.x : lv.x + v.x;  // selects, not a branch: every lane stays active for the next DPP
  v.y = f ? v.y : lv.y + v.y;
  v.z = f ? v.z : lv.z + v.z;
  v.w = f ? v.w : lv.w + v.w;
  f = f | lf;
}

__device__ __forceinline__ void wave_segmented_sum(float4& v, bool head) {
  int f = head ? 1 : 0;
  seg_step<DPP_ROW_SHR1>(v, f);
  seg_step<DPP_ROW_SHR2>(v, f);
  seg_step<DPP_ROW_SHR4>(v, f);
  seg_step<DPP_ROW_SHR8>(v, f);
  seg_step<DPP_ROW_BCAST15, 0xa>(v, f);
  seg_step<DPP_ROW_BCAST31, 0xc>(v, f);
}

// Nearest staging area (ns items in sample order, one cell each): when at least a quarter of
// the items repeat the previous item's cell (grids whose cells are longer than the sample
// spacing: dense scenes, the pooled grids of progressive growing), each run of equal cells is
// summed by a segmented wave scan and compacted to one item, in place.  Returns the number of
// items left to flush.
__device__ __forceinline__ int merge_nearest_runs(int lane, int ns, uint32_t* s_cell, float4* s_val) {
  const bool live = lane < ns;
  const uint32_t cell = live ? s_cell[lane] : 0xffffffffu;
  const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)0xfffffffeu, (int)cell, DPP_WAVE_SHR1, 0xf, 0xf,
                                                               false);
  const bool dup = live && lane > 0 && cell == prev;
  const uint64_t dups = __ballot(dup);
  if (4 * __popcll(dups) < ns) return ns;
  float4 v = live ? s_val[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
  wave_segmented_sum(v, !dup);
  // a run ends where the next item starts a new one (or at the last item)
  const bool next_dup = lane + 1 < ns && ((dups >> (lane + 1)) & 1ull) != 0ull;
  const bool tail = live && !next_dup;
  const uint64_t tails = __ballot(tail);
  wave_lds_sync();  // every item is read before any is rewritten
  if (tail) {
    const int r = lane_rank(tails);
    s_cell[r] = cell;
    s_val[r] = v;
  }
  wave_lds_sync();
  return __popcll(tails);
}

template <int INTERP, bool RUNS = false>
__device__ __forceinline__ void scatter_chunk(const KParams& P, const Ray& R, int lane, int k,
                                              bool keep, uint32_t clip, const float4& gv,
                                              float* grad, uint32_t* s_cell, float4* s_val) {
  constexpr int NI = ScatterItems<INTERP>::kPerSample;
  constexpr int NS = ScatterItems<INTERP>::kStageSamples;
  const uint64_t km = __ballot(keep);
  if (km == 0ull) return;
  const int nk = __popcll(km);
  const int rank = lane_rank(km);
  const int ch = lane & 3;
  const float* sv = reinterpret_cast<const float*>(s_val);
  for (int h = 0; h < nk; h += NS) {  // kept samples of rank h .. h+NS-1 per stage
    if (keep && rank >= h && rank < h + NS) write_items<INTERP>(P, R, k, clip, gv, s_cell, s_val, (rank - h) * NI);
    wave_lds_sync();
    const int ns = min(nk - h, NS);
    int n = ns * NI;
    if constexpr (INTERP == PXL_INTERP_NEAREST) n = merge_nearest_runs(lane, ns, s_cell, s_val);
    bool by_runs = false;
    if constexpr (RUNS && INTERP == PXL_INTERP_TRILINEAR) {
      // same items as the previous sample: items 0 and NI - 1 both equal
      auto same_items = [&](int a, int b) {
        return s_cell[a * NI] == s_cell[b * NI] && s_cell[a * NI + NI - 1] == s_cell[b * NI + NI - 1];
      };
      const bool dup = lane >= 1 && lane < ns && same_items(lane, lane - 1);
      by_runs = 4 * __popcll(__ballot(dup)) >= ns;
      if (by_runs) {
        const int j = (lane >> 2) & 7, r0 = (lane >> 5) * (NS / 2), r1 = min(ns, r0 + NS / 2);
        float acc = 0.f;
        for (int r = r0; r < r1; ++r) {
          acc += sv[4 * (r * NI + j) + ch];
          if (r + 1 == r1 || !same_items(r + 1, r)) {  // end of a run
            if (acc != 0.f) atomic_add_f32(grad + (size_t)s_cell[r * NI + j] * 4u + ch, acc);
            acc = 0.f;
          }
        }
      }
    }
    if (!by_runs) {
      for (int base = 0; base < n; base += kWave / 4) {
        const int it = base + (lane >> 2);
        if (it < n) {
          const float v = sv[4 * it + ch];
          if (v != 0.f) atomic_add_f32(grad + (size_t)s_cell[it] * 4u + ch, v);
        }
      }
    }
    wave_lds_sync();
  }
}

}  // namespace pxl
