// pxl_render.hip -- fused render / backward / training-step kernels for gfx950 and the
// C ABI declared in include/pxl.h.
//
// Work mapping: one 64-lane wavefront per ray, four rays per 256-thread block (the train step:
// two rays per wavefront, sharing one group of 64 lanes when each keeps at most 32 samples).
//   classify   every sample k = 64c + lane + 1 is placed and tested (in grid, occupied) in
//              chunks of 64; kept samples are compacted, in order, into a per-wave LDS list
//   groups     the expensive work (cell loads, clip, interpolation, compositing, scatter)
//              runs on groups of 64 consecutive kept samples, so lanes are not left idle
//              on skipped samples.  Skipped samples have alpha = 0, so the transmittance
//              T_i = T_group * exclusive wave product of (1 - alpha) over the group's lanes
//              equals the reference's cumprod over all samples.
//   backward   suffix colour S_i by a wave suffix scan of the affine maps
//              q -> alpha_i (c_i, 1) + (1 - alpha_i) q, so dC/dalpha_i = T_i (c_i - S_i)
//              needs no division by (1 - alpha_i) and stays exact at alpha = 1.
// Gradients reach the grid through float atomics (global_atomic_add_f32), staged per group
// in LDS and flushed four lanes per cell.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "pxl.h"
#include "pxl_device.h"
#include "pxl_host.h"

namespace pxl {

// ----------------------------------------------------------------------------- kernels

__device__ __forceinline__ int64_t wave_ray(const KParams& P) {
  return P.ray_offset + (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
}

// aux per ray (pxl_aux_floats(S) floats): [0] n_proc -- samples k <= n_proc may contribute
// (S unless early termination stopped the ray); [1] number of groups; then per group g:
// [2+2g] transmittance before the group, [3+2g] sample index of its first member.
__device__ __forceinline__ void accumulate(bool proc, float Ti, const float4& v, float& cr,
                                           float& cg, float& cb, float& ca) {
  if (proc) {
    const float w = v.w * Ti;
    cr += v.x * w;
    cg += v.y * w;
    cb += v.z * w;
    ca += w;
  }
}

// Early termination within a group: lanes from the first one whose transmittance is below
// t_stop on are dropped.  Returns true (and the sample bound) when the ray stops here.
__device__ __forceinline__ bool stop_check(const KParams& P, bool valid, float Ti, int k,
                                           int lane, bool& proc, int& n_proc) {
  if (!(P.t_stop > 0.f)) return false;
  const uint64_t dead = __ballot(valid && Ti < P.t_stop);
  if (!dead) return false;
  const int first = __builtin_ctzll(dead);
  n_proc = __builtin_amdgcn_readlane(k, first) - 1;
  proc = proc && lane < first;
  return true;
}

// Beta regulariser of scripts/train.py:170-174: f(a) = log(a + eps) - log(1 - a + eps) per
// sample, eps = 1e-4, and its derivative.
constexpr float kBetaEps = 1e-4f;
__device__ __forceinline__ float beta_term(float a) {
  return logf(a + kBetaEps) - logf(1.f - a + kBetaEps);
}
__device__ __forceinline__ float beta_dloss(float a) {
  return 1.f / (a + kBetaEps) + 1.f / (1.f - a + kBetaEps);
}

// Forward sweep of one ray (one wave): its kept samples stream through `buf` (2 * kWave ints
// of LDS) in groups of 64; returns the composite (every lane holds it).  When `grp` is set,
// group g's start state goes to grp[2g] (transmittance before the group) and grp[2g + 1] (its
// first sample index, as int bits), `ng` receives the group count.  n_proc: the bound of the
// samples that may contribute (S unless early termination stopped the ray).  With `beta`,
// bsum (every lane) = sum over processed samples of f(a) - f(0).  Any S is handled.
template <int DT, int INTERP>
__device__ __forceinline__ float4 forward_sweep(const KParams& P, const Ray& R, int lane, int* buf,
                                                float* grp, int& ng, int& n_proc, bool beta, float& bsum) {
  const int S = P.n_samples, nch = (S + 63) >> 6;
  const float f0 = beta_term(0.f);
  float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
  float bs = 0.f;
  int nbuf = 0;
  bool stop = false;
  n_proc = S;
  ng = 0;
  for (int c = 0; c < nch && !stop; ++c) {
    nbuf = append_kept<INTERP>(P, R, (c << 6) + lane + 1, S, buf, nbuf);
    const bool last = c == nch - 1;
    while (!stop && (nbuf >= kWave || (last && nbuf > 0))) {
      wave_lds_sync();
      const int m = min(nbuf, kWave);
      const bool valid = lane < m;
      const int k = valid ? buf[lane] : 0;
      if (grp && lane == 0) {
        grp[2 * ng] = T;
        grp[2 * ng + 1] = __int_as_float(k);
      }
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      uint32_t clip;
      if (valid) eval_kept<DT, INTERP>(P, R, k, v, clip);
      const float incl = wave_prefix_mul(1.f - v.w);
      const float Ti = T * wave_shift_up(incl, 1.f);
      bool proc = valid;
      stop = stop_check(P, valid, Ti, k, lane, proc, n_proc);
      accumulate(proc, Ti, v, cr, cg, cb, ca);
      if (beta && proc) bs += beta_term(v.w) - f0;
      T = T * readlane(incl, 63);
      ++ng;
      const int rest = nbuf - m;  // move the unprocessed tail to the front
      const int carry = lane < rest ? buf[kWave + lane] : 0;
      wave_lds_sync();
      if (lane < rest) buf[lane] = carry;
      nbuf = rest;
      wave_lds_sync();
    }
  }
  if (beta) bsum = wave_sum(bs);
  return make_float4(wave_sum(cr), wave_sum(cg), wave_sum(cb), wave_sum(ca));
}

// Forward: out_rgba (n,4) and, when aux != NULL, the per-group state the backward needs.
template <int DT, int INTERP>
__global__ __launch_bounds__(kBlock) void k_render_fwd(KParams P, float* __restrict__ out,
                                                        float* __restrict__ aux) {
  __shared__ int s_buf[kWavesPerBlock][2 * kWave];
  const int lane = threadIdx.x & 63;
  const int64_t ray = wave_ray(P);
  if (ray >= P.n_rays) return;
  Ray R;
  setup_ray(P, ray, R);
  float* ax = aux ? aux + ray * P.aux_stride : nullptr;
  int ng, n_proc;
  float bsum;
  const float4 c = forward_sweep<DT, INTERP>(P, R, lane, s_buf[threadIdx.x >> 6], ax ? ax + 2 : nullptr, ng,
                                             n_proc, false, bsum);
  if (lane == 0) {
    reinterpret_cast<float4*>(out)[ray] = c;
    if (ax) {
      ax[0] = __int_as_float(n_proc);
      ax[1] = __int_as_float(ng);
    }
  }
}


// One group of the reverse sweep: given the group's clipped values, transmittance and Qn
// (the suffix value just behind the group), emit dL/dvalue per sample, scatter it, and
// return in Qn the suffix value in front of the group.
template <int INTERP, bool RUNS = false>
__device__ __forceinline__ void reverse_chunk(const KParams& P, const Ray& R, int lane, int k,
                                              bool keep, const float4& v, uint32_t clip,
                                              float Ti, const float4& g, float beta_scale,
                                              float4& Qn, float* grad, uint32_t* s_cell,
                                              float4* s_val) {
  const float a = keep ? v.w : 0.f;
  Affine4 f;
  f.a0 = a * v.x;
  f.a1 = a * v.y;
  f.a2 = a * v.z;
  f.a3 = a;
  f.b = 1.f - a;
  if (lane == 63) {  // fold the tail: constant map q -> Q of the first sample after the group
    f.a0 = f.a0 + f.b * Qn.x;
    f.a1 = f.a1 + f.b * Qn.y;
    f.a2 = f.a2 + f.b * Qn.z;
    f.a3 = f.a3 + f.b * Qn.w;
    f.b = 0.f;
  }
  wave_suffix_affine(f, lane);  // f.a = Q_i, the colour seen just in front of sample i
  const float s0 = wave_shift_down(f.a0, Qn.x), s1 = wave_shift_down(f.a1, Qn.y);
  const float s2 = wave_shift_down(f.a2, Qn.z), s3 = wave_shift_down(f.a3, Qn.w);
  float4 gv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (keep) {
    const float w = a * Ti;
    gv.x = w * g.x;
    gv.y = w * g.y;
    gv.z = w * g.z;
    gv.w = Ti * (g.x * (v.x - s0) + g.y * (v.y - s1) + g.z * (v.z - s2) + g.w * (1.f - s3));
    if (beta_scale != 0.f) gv.w += beta_scale * beta_dloss(v.w);
  }
  scatter_chunk<INTERP, RUNS>(P, R, lane, k, keep, clip, gv, grad, s_cell, s_val);
  Qn = make_float4(readlane(f.a0, 0), readlane(f.a1, 0), readlane(f.a2, 0), readlane(f.a3, 0));
}

// Per-wave LDS of the kernels that scatter: the kept-sample list and, once the list has been
// read into registers, the gradient staging area of scatter_chunk.
template <int INTERP, int LIST>
struct WaveLds {
  static constexpr int kItems = ScatterItems<INTERP>::kPerChunk;
  static constexpr int kStage = kItems * 20;
  static constexpr int kBytes = (LIST * 4 > kStage ? LIST * 4 : kStage);
  static __device__ __forceinline__ int* list(unsigned char* base) { return reinterpret_cast<int*>(base); }
  static __device__ __forceinline__ uint32_t* cells(unsigned char* base) {
    return reinterpret_cast<uint32_t*>(base);
  }
  static __device__ __forceinline__ float4* vals(unsigned char* base) {
    return reinterpret_cast<float4*>(base + kItems * 4);
  }
};

// Reverse sweep of one ray (one wave) from its per-group start state grp (as forward_sweep
// writes it; ng groups, samples k <= n_proc): group by group from the back, its kept samples
// are found again (`list`, kWave ints of LDS, may alias the staging area), re-gathered, and
// their gradients -- the loss gradient g through the compositing, plus the beta term's
// per-sample derivative when beta_scale != 0 -- scattered into grad.
template <int DT, int INTERP, bool RUNS>
__device__ __forceinline__ void reverse_sweep(const KParams& P, const Ray& R, int lane, int* list,
                                              const float* grp, int ng, int n_proc, const float4& g,
                                              float beta_scale, float* grad, uint32_t* s_cell, float4* s_val) {
  float4 Qn = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int gi = ng - 1; gi >= 0; --gi) {
    const float T0 = grp[2 * gi];
    const int k0 = __float_as_int(grp[2 * gi + 1]);
    int n = 0;  // the group: the first 64 kept samples from k0 on (k <= n_proc)
    for (int kb = k0; n < kWave && kb <= n_proc; kb += kWave) {
      const int k = kb + lane;
      const bool keep = k <= n_proc && classify<INTERP>(P, R, k);
      const uint64_t km = __ballot(keep);
      const int r = n + lane_rank(km);
      if (keep && r < kWave) list[r] = k;
      n += __popcll(km);
    }
    n = min(n, kWave);
    wave_lds_sync();
    const bool valid = lane < n;
    const int k = valid ? list[lane] : 0;
    wave_lds_sync();
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t clip = 0u;
    if (valid) eval_kept<DT, INTERP>(P, R, k, v, clip);
    const float incl = wave_prefix_mul(1.f - v.w);
    const float Ti = T0 * wave_shift_up(incl, 1.f);
    reverse_chunk<INTERP, RUNS>(P, R, lane, k, valid, v, clip, Ti, g, beta_scale, Qn, grad, s_cell, s_val);
  }
}

template <int DT, int INTERP>
__global__ __launch_bounds__(kBlock) void k_render_bwd(KParams P, const float* __restrict__ grad_out,
                                                        const float* __restrict__ aux,
                                                        float* __restrict__ grad) {
  using L = WaveLds<INTERP, kWave>;
  __shared__ __align__(16) unsigned char s_mem[kWavesPerBlock][L::kBytes];
  const int lane = threadIdx.x & 63;
  const int64_t ray = wave_ray(P);
  if (ray >= P.n_rays) return;
  const float4 g = reinterpret_cast<const float4*>(grad_out)[ray];
  if (g.x == 0.f && g.y == 0.f && g.z == 0.f && g.w == 0.f) return;
  unsigned char* mem = s_mem[threadIdx.x >> 6];
  Ray R;
  setup_ray(P, ray, R);
  const float* ax = aux + ray * P.aux_stride;
  reverse_sweep<DT, INTERP, false>(P, R, lane, L::list(mem), ax + 2, __float_as_int(ax[1]), __float_as_int(ax[0]), g,
                                   0.f, grad, L::cells(mem), L::vals(mem));
}

// Training step for S <= 64*NG: every kept sample's value stays in registers between the
// forward sweep and the reverse sweep, the loss gradient is formed in place -- one gather of
// each kept sample per step.
struct LossParams {
  const float* targets;
  float grad_scale, beta_scale;
  float* sq_err;
  float* beta;
};

// The train step of one ray after classification: its kept samples (list[0..nk)) in groups of
// 64, loss and its gradient, the reverse sweep and the scatter.  `list` may alias the staging
// area (s_cell, s_val): it is read completely before the first scatter.
// The run flush of scatter_chunk goes into the train kernels for more than 256 samples per ray
// (three waves per SIMD, registers to spare): the kernels for up to 256 samples keep five
// waves per SIMD only without it (with it they spill or lose a wave).
template <int NG>
constexpr bool kRunFlush = NG > 4;

template <int DT, int INTERP, int NG>
__device__ __forceinline__ void train_ray(const KParams& P, const LossParams& LP, const Ray& R,
                                          int64_t ray, const float4& tgt, const int* list, int nk,
                                          int lane, uint32_t* s_cell, float4* s_val,
                                          float* __restrict__ out, float* __restrict__ grad) {
  const int S = P.n_samples;
  const int ng = (nk + 63) >> 6;
  float4 V[NG];
  uint32_t CL[NG];
  float TI[NG];
  int KI[NG];
  uint64_t KEEP[NG];
  float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f, bsum = 0.f;
  int n_proc = S;
  bool stopped = false;
  const float f0 = beta_term(0.f);
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    V[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    CL[g] = 0u;
    TI[g] = 0.f;
    KI[g] = 0;
    KEEP[g] = 0ull;
    if (g < ng && !stopped) {
      const int idx = (g << 6) + lane;
      const bool valid = idx < nk;
      const int k = valid ? list[idx] : 0;
      if (valid) eval_kept<DT, INTERP>(P, R, k, V[g], CL[g]);
      const float incl = wave_prefix_mul(1.f - V[g].w);
      const float Ti = T * wave_shift_up(incl, 1.f);
      bool proc = valid;
      stopped = stop_check(P, valid, Ti, k, lane, proc, n_proc);
      accumulate(proc, Ti, V[g], cr, cg, cb, ca);
      if (LP.beta && proc) bsum += beta_term(V[g].w) - f0;
      TI[g] = Ti;
      KI[g] = k;
      KEEP[g] = __ballot(proc);
      T = T * readlane(incl, 63);
    }
  }
  wave_lds_sync();  // the list area may become the scatter staging area
  cr = wave_sum(cr);
  cg = wave_sum(cg);
  cb = wave_sum(cb);
  ca = wave_sum(ca);
  if (LP.beta) bsum = wave_sum(bsum);
  const float d0 = cr - tgt.x, d1 = cg - tgt.y, d2 = cb - tgt.z, d3 = ca - tgt.w;
  if (lane == 0) {
    if (out) reinterpret_cast<float4*>(out)[ray] = make_float4(cr, cg, cb, ca);
    if (LP.sq_err) LP.sq_err[ray] = ((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3;
    if (LP.beta) LP.beta[ray] = bsum;
  }
  const float gs = LP.grad_scale;
  const float4 gl = make_float4(gs * d0, gs * d1, gs * d2, gs * d3);
  float4 Qn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int g = NG - 1; g >= 0; --g) {
    // a group with no processed sample is the identity map: Qn passes through unchanged
    if (g < ng && KEEP[g] != 0ull) {
      const bool keep = (KEEP[g] >> lane) & 1ull;
      reverse_chunk<INTERP, kRunFlush<NG>>(P, R, lane, KI[g], keep, V[g], CL[g], TI[g], gl, LP.beta_scale, Qn,
                            grad, s_cell, s_val);
    }
  }
}

// Two rays per wavefront.  Both rays are classified together (the occupancy loads of all their
// chunks in flight at once); when each keeps at most 32 samples (most rays of sparse scenes)
// ray 0 runs in lanes 0-31 and ray 1 in lanes 32-63 of one group: the row-wise scans and sums
// stop before the step that joins the two halves (half_prefix_mul, half_suffix_affine,
// half_sums), which evaluates, for each half, the same operation tree the one-ray kernel uses
// on lanes 0-31 -- so every value is bit-identical to the one-ray path's (and to
// k_render_fwd's).  Otherwise the two rays run one after the other through the one-ray path
// (train_ray).  The register budget is capped at five waves per SIMD (96 VGPRs for trilinear,
// S <= 256), which the one-ray path alone needed; S > 256 and FP16 trilinear would spill.
__device__ __forceinline__ float half_prefix_mul(float x) {
  x *= dpp<DPP_ROW_SHR1>(1.f, x);
  x *= dpp<DPP_ROW_SHR2>(1.f, x);
  x *= dpp<DPP_ROW_SHR4>(1.f, x);
  x *= dpp<DPP_ROW_SHR8>(1.f, x);
  x *= dpp<DPP_ROW_BCAST15, 0xa>(1.f, x);
  return x;
}

// Row-wise partial sums up to the row-15 broadcast: lane 31 holds the sum of lanes 0-31 and
// lane 63 that of lanes 32-63, in wave_sum's order.
__device__ __forceinline__ float half_sums(float x) {
  x += dpp<DPP_ROW_SHR1>(0.f, x);
  x += dpp<DPP_ROW_SHR2>(0.f, x);
  x += dpp<DPP_ROW_SHR4>(0.f, x);
  x += dpp<DPP_ROW_SHR8>(0.f, x);
  x += dpp<DPP_ROW_BCAST15, 0xa>(0.f, x);
  return x;
}

// wave_suffix_affine within each 32-lane half (rows 0-1 and rows 2-3).
__device__ __forceinline__ void half_suffix_affine(Affine4& f, int lane) {
  affine_step<DPP_ROW_SHL1>(f);
  affine_step<DPP_ROW_SHL2>(f);
  affine_step<DPP_ROW_SHL4>(f);
  affine_step<DPP_ROW_SHL8>(f);
  const Affine4 r1 = affine_readlane(f, 16), r3 = affine_readlane(f, 48);
  const int row = lane >> 4;
  if (row == 0) f = affine_compose(f, r1);
  if (row == 2) f = affine_compose(f, r3);
}

__device__ __forceinline__ Ray select_ray(bool second, const Ray& a, const Ray& b) {
  Ray r;
  r.ox = second ? b.ox : a.ox;
  r.oy = second ? b.oy : a.oy;
  r.oz = second ? b.oz : a.oz;
  r.dx = second ? b.dx : a.dx;
  r.dy = second ? b.dy : a.dy;
  r.dz = second ? b.dz : a.dz;
  r.cam = second ? b.cam : a.cam;
  r.u = second ? b.u : a.u;
  r.v = second ? b.v : a.v;
  return r;
}

// Rays ray0 (lanes 0-31, list0[0..n0)) and ray0 + 1 (lanes 32-63, list1[0..n1)), n0, n1 <= 32,
// one group of each; the steps of train_ray with T = 1 and Qn = 0.
template <int DT, int INTERP, int NG>
__device__ __forceinline__ void train_half_pair(const KParams& P, const LossParams& LP,
                                                const Ray& R0, const Ray& R1, int64_t ray0,
                                                bool has1, const float4& tgt0, const float4& tgt1,
                                                const int* list0, int n0, const int* list1, int n1,
                                                int lane, uint32_t* s_cell, float4* s_val,
                                                float* __restrict__ out, float* __restrict__ grad) {
  const bool hi = lane >= 32;
  const int h = lane & 31;
  const bool valid = h < (hi ? n1 : n0);
  const int k = valid ? (hi ? list1[h] : list0[h]) : 0;
  const Ray R = select_ray(hi, R0, R1);
  float4 V = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t CL = 0u;
  if (valid) eval_kept<DT, INTERP>(P, R, k, V, CL);
  const float incl = half_prefix_mul(1.f - V.w);
  // lane 32 starts ray 1 with transmittance 1: the boundary value goes into the source lane
  // (31) before the shift, so every lane uses the DPP result.  (Selecting on the shifted value
  // instead lets the compiler run the DPP under a lane mask, and a DPP read of a disabled
  // lane returns the fill value.)
  const float Ti = wave_shift_up(lane == 31 ? 1.f : incl, 1.f);
  bool proc = valid;
  if (P.t_stop > 0.f) {  // stop_check per half
    const uint64_t dead = __ballot(valid && Ti < P.t_stop);
    const uint32_t dh = hi ? (uint32_t)(dead >> 32) : (uint32_t)dead;
    if (dh) proc = proc && h < __builtin_ctz(dh);
  }
  float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f, bsum = 0.f;
  accumulate(proc, Ti, V, cr, cg, cb, ca);
  if (LP.beta && proc) bsum += beta_term(V.w) - beta_term(0.f);
  const uint64_t keep_mask = __ballot(proc);
  wave_lds_sync();  // list0 becomes the scatter staging area
  cr = half_sums(cr);
  cg = half_sums(cg);
  cb = half_sums(cb);
  ca = half_sums(ca);
  if (LP.beta) bsum = half_sums(bsum);
  // each ray's totals and loss gradient, formed in the lane that writes its outputs (31, 63)
  const float d0 = cr - (hi ? tgt1.x : tgt0.x), d1 = cg - (hi ? tgt1.y : tgt0.y);
  const float d2 = cb - (hi ? tgt1.z : tgt0.z), d3 = ca - (hi ? tgt1.w : tgt0.w);
  if (h == 31 && (!hi || has1)) {
    const int64_t ray = ray0 + (hi ? 1 : 0);
    if (out) reinterpret_cast<float4*>(out)[ray] = make_float4(cr, cg, cb, ca);
    if (LP.sq_err) LP.sq_err[ray] = ((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3;
    if (LP.beta) LP.beta[ray] = bsum;
  }
  if (keep_mask == 0ull) return;
  const float gs = LP.grad_scale;
  const float4 gl0 = make_float4(gs * readlane(d0, 31), gs * readlane(d1, 31), gs * readlane(d2, 31),
                                 gs * readlane(d3, 31));
  const float4 gl1 = make_float4(gs * readlane(d0, 63), gs * readlane(d1, 63), gs * readlane(d2, 63),
                                 gs * readlane(d3, 63));
  const float4 gl = hi ? gl1 : gl0;
  // reverse sweep of the one group (reverse_chunk with Qn = 0 and the halves kept apart)
  const float a = proc ? V.w : 0.f;
  Affine4 f;
  f.a0 = a * V.x;
  f.a1 = a * V.y;
  f.a2 = a * V.z;
  f.a3 = a;
  f.b = 1.f - a;
  half_suffix_affine(f, lane);
  // what lies behind each sample: the next lane's suffix, nothing behind lane 31 (end of ray 0)
  // or lane 63; lane 32 feeds 0 to lane 31 for the same reason as above
  const bool first1 = lane == 32;
  const float s0 = wave_shift_down(first1 ? 0.f : f.a0, 0.f), s1 = wave_shift_down(first1 ? 0.f : f.a1, 0.f);
  const float s2 = wave_shift_down(first1 ? 0.f : f.a2, 0.f), s3 = wave_shift_down(first1 ? 0.f : f.a3, 0.f);
  float4 gv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (proc) {
    const float w = a * Ti;
    gv.x = w * gl.x;
    gv.y = w * gl.y;
    gv.z = w * gl.z;
    gv.w = Ti * (gl.x * (V.x - s0) + gl.y * (V.y - s1) + gl.z * (V.z - s2) + gl.w * (1.f - s3));
    if (LP.beta_scale != 0.f) gv.w += LP.beta_scale * beta_dloss(V.w);
  }
  scatter_chunk<INTERP, kRunFlush<NG>>(P, R, lane, k, proc, CL, gv, grad, s_cell, s_val);
}

template <int DT, int INTERP, int NG>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NG <= 4 && DT != PXL_DTYPE_F16 ? 5 : 1))) void k_train_pair(KParams P, LossParams LP,
                                                        float* __restrict__ out,
                                                        float* __restrict__ grad) {
  // per wave: [ray 0's list, then the scatter staging area | ray 1's list]
  using L = WaveLds<INTERP, kWave * NG>;
  constexpr int kList1 = kWave * NG * 4;
  __shared__ __align__(16) unsigned char s_mem[kWavesPerBlock][L::kBytes + kList1];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t ray0 = P.ray_offset + 2 * ((int64_t)blockIdx.x * kWavesPerBlock + wave);
  if (ray0 >= P.n_rays) return;
  const bool has1 = ray0 + 1 < P.n_rays;
  unsigned char* mem = s_mem[wave];
  int* list0 = L::list(mem);
  int* list1 = reinterpret_cast<int*>(mem + L::kBytes);
  Ray R0, R1;
  setup_ray(P, ray0, R0);
  if (has1) setup_ray(P, ray0 + 1, R1);
  else R1 = R0;
  const float* targets = LP.targets;
  const float4 tgt0 = targets ? reinterpret_cast<const float4*>(targets)[ray0] : camera_target(P, R0);
  const float4 tgt1 = !has1 ? tgt0
                    : targets ? reinterpret_cast<const float4*>(targets)[ray0 + 1] : camera_target(P, R1);
  const Ray RR[2] = {R0, R1};
  bool kp[2][NG];
  classify_chunks<INTERP, 2, NG>(P, RR, lane, kp);
  int n0 = 0, n1 = 0;
#pragma unroll
  for (int c = 0; c < NG; ++c) {
    const bool k1 = kp[1][c] & has1;
    const uint64_t km0 = __ballot(kp[0][c]), km1 = __ballot(k1);
    if (kp[0][c]) list0[n0 + lane_rank(km0)] = (c << 6) + lane + 1;
    if (k1) list1[n1 + lane_rank(km1)] = (c << 6) + lane + 1;
    n0 += __popcll(km0);
    n1 += __popcll(km1);
  }
  wave_lds_sync();
  if (n0 <= 32 && n1 <= 32) {
    train_half_pair<DT, INTERP, NG>(P, LP, R0, R1, ray0, has1, tgt0, tgt1, list0, n0, list1, n1, lane,
                                L::cells(mem), L::vals(mem), out, grad);
  } else {
    train_ray<DT, INTERP, NG>(P, LP, R0, ray0, tgt0, list0, n0, lane, L::cells(mem), L::vals(mem), out, grad);
    if (has1)
      train_ray<DT, INTERP, NG>(P, LP, R1, ray0 + 1, tgt1, list1, n1, lane, L::cells(mem), L::vals(mem), out,
                                grad);
  }
}

// Training step for more than 512 samples per ray (up to 64 * kLongGroups): one ray per
// wavefront.  The forward sweep keeps each group's start state (transmittance, first sample)
// in LDS, the loss gradient is formed, and the reverse sweep re-gathers group by group from
// the back, as pxl_render_fwd + pxl_render_bwd would with the aux in LDS instead of memory.
// Coarse grids (the pooled steps of progressive growing, where this sample count is the
// reference's default of 600) keep tens of consecutive samples in one cell: the scatter
// flushes such stages by runs.
constexpr int kLongGroups = PXL_TRAIN_MAX_SAMPLES / kWave;

template <int DT, int INTERP>
__global__ __launch_bounds__(kBlock) void k_train_long(KParams P, LossParams LP, float* __restrict__ out,
                                                        float* __restrict__ grad) {
  using L = WaveLds<INTERP, 2 * kWave>;  // the forward's sample buffer, then the scatter staging
  __shared__ __align__(16) unsigned char s_mem[kWavesPerBlock][L::kBytes];
  __shared__ float s_grp[kWavesPerBlock][2 * kLongGroups];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t ray = wave_ray(P);
  if (ray >= P.n_rays) return;
  unsigned char* mem = s_mem[wave];
  float* grp = s_grp[wave];
  Ray R;
  setup_ray(P, ray, R);
  const float4 tgt = LP.targets ? reinterpret_cast<const float4*>(LP.targets)[ray] : camera_target(P, R);
  int ng, n_proc;
  float bsum = 0.f;
  const float4 c = forward_sweep<DT, INTERP>(P, R, lane, L::list(mem), grp, ng, n_proc, LP.beta != nullptr, bsum);
  const float d0 = c.x - tgt.x, d1 = c.y - tgt.y, d2 = c.z - tgt.z, d3 = c.w - tgt.w;
  if (lane == 0) {
    if (out) reinterpret_cast<float4*>(out)[ray] = c;
    if (LP.sq_err) LP.sq_err[ray] = ((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3;
    if (LP.beta) LP.beta[ray] = bsum;
  }
  wave_lds_sync();  // lane 0's group records are read by every lane
  const float gs = LP.grad_scale;
  const float4 gl = make_float4(gs * d0, gs * d1, gs * d2, gs * d3);
  reverse_sweep<DT, INTERP, true>(P, R, lane, L::list(mem), grp, ng, n_proc, gl, LP.beta_scale, grad,
                                  L::cells(mem), L::vals(mem));
}

__global__ void k_camera_rays(KParams P, float* __restrict__ o, float* __restrict__ d,
                              float* __restrict__ tgt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + P.ray_offset;
  if (r >= P.n_rays) return;
  Ray R;
  camera_ray(P, r, R);
  o[3 * r] = R.ox;
  o[3 * r + 1] = R.oy;
  o[3 * r + 2] = R.oz;
  d[3 * r] = R.dx;
  d[3 * r + 1] = R.dy;
  d[3 * r + 2] = R.dz;
  if (tgt) reinterpret_cast<float4*>(tgt)[r] = camera_target(P, R);
}

// Dilated occupancy: out bit of cell (x, y, z) = OR of the in bits of (x+a, y+b, z+c), a, b, c in
// {0, 1}, wrapped -- the cells a trilinear sample with floor cell (x, y, z) reads.  One thread
// per output word (32 cells of one tile); padding cells of partial tiles stay 0.
__global__ void k_dilate(const uint32_t* __restrict__ in, uint32_t X, uint32_t Y, uint32_t Z,
                         uint32_t* __restrict__ out, int64_t n_words) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint32_t TY = (Y + 7u) >> 3, TZ = (Z + 7u) >> 3;
  uint32_t word = 0u;
  for (uint32_t b = 0; b < 32u; ++b) {
    uint32_t i, j, k;
    occ_cell((uint32_t)w, b, TY, TZ, i, j, k);
    if (i >= X || j >= Y || k >= Z) continue;
    const uint32_t i1 = i + 1 == X ? 0u : i + 1, j1 = j + 1 == Y ? 0u : j + 1, k1 = k + 1 == Z ? 0u : k + 1;
    const uint32_t is[2] = {i, i1}, js[2] = {j, j1}, ks[2] = {k, k1};
    bool any = false;
    for (int c = 0; c < 8 && !any; ++c) any = occupied_at(in, is[c >> 2], js[(c >> 1) & 1], ks[c & 1], TY, TZ);
    word |= (uint32_t)any << b;
  }
  out[w] = word;
}

// Occupancy from the grid's own alpha: bit of cell (i, j, k) set when raw alpha >= 0 there
// (DILATE: at any of the corners (i+a, j+b, k+c), a,b,c in {0,1}, wrapped -- the cells a
// trilinear sample with floor cell (i, j, k) reads).  A sample whose cells all have raw alpha
// < 0 is exactly zero in the reference's chain: clamp gives alpha 0, so its colour weight and
// colour gradients vanish, and the clamp's gradient mask is 0 for every alpha it read.  NaN
// alpha counts as empty.
// A workgroup covers a y quad (y = 4q .. 4q+3, one wave each) and 64 consecutive z (the lanes,
// so each wave reads whole runs of cells), and walks a slab of kAlphaOccSlab x planes: with
// DILATE, plane x's "some corner (x, y+b, z+c) has alpha >= 0" is carried from the previous
// step, so each plane's cells are read once.  Each wave's ballot holds one byte-row (y & 3) of
// the 8 tile words along z; eight lanes assemble the words from the four ballots in LDS.
// Units cover the padded dims (multiples of 8), so every word, padding included, is written.
constexpr int kAlphaOccThreads = 256;
constexpr uint32_t kAlphaOccSlab = 32;  // a multiple of 8: slabs cover whole tiles along x

template <int DT, bool DILATE>
__global__ __launch_bounds__(kAlphaOccThreads) void k_alpha_occupancy(const void* __restrict__ grid, uint32_t X,
                                                                      uint32_t Y, uint32_t Z,
                                                                      uint32_t* __restrict__ out) {
  __shared__ uint64_t s_m[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t TX = (X + 7u) >> 3, TY = (Y + 7u) >> 3, TZ = (Z + 7u) >> 3;
  const uint32_t ZB = (TZ + 7u) >> 3, YQ = TY * 2u, XS = (TX * 8u + kAlphaOccSlab - 1u) / kAlphaOccSlab;
  const uint64_t n_units = (uint64_t)XS * YQ * ZB;
  for (uint64_t u = blockIdx.x; u < n_units; u += gridDim.x) {
    const uint32_t zb = (uint32_t)(u % ZB), yq = (uint32_t)((u / ZB) % YQ), xs = (uint32_t)(u / ((uint64_t)ZB * YQ));
    const uint32_t j = yq * 4u + (uint32_t)wave, k = zb * 64u + (uint32_t)lane;
    const bool cell_ok = j < Y && k < Z;
    const uint32_t j1 = j + 1 == Y ? 0u : j + 1, k1 = k + 1 == Z ? 0u : k + 1;
    // any alpha >= 0 among plane i's corners (i, j+b, k+c) (DILATE) or at (i, j, k)
    auto plane_any = [&](uint32_t i) {
      if constexpr (DILATE) {
        const uint32_t js[2] = {j, j1}, ks[2] = {k, k1};
        bool any = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) any |= load_cell<DT>(grid, (i * Y + js[c >> 1]) * Z + ks[c & 1]).w >= 0.f;
        return any;
      } else {
        return load_cell<DT>(grid, (i * Y + j) * Z + k).w >= 0.f;
      }
    };
    const uint32_t x0 = xs * kAlphaOccSlab, x1 = min(TX * 8u, x0 + kAlphaOccSlab);
    bool cur = DILATE && cell_ok && x0 < X ? plane_any(x0) : false;
    for (uint32_t i = x0; i < x1; ++i) {
      bool occ = false;
      if (cell_ok && i < X) {
        if constexpr (DILATE) {
          const uint32_t i1 = i + 1 == X ? 0u : i + 1;
          const bool nxt = plane_any(i1);
          occ = cur | nxt;
          cur = nxt;
        } else {
          occ = plane_any(i);
        }
      }
      const uint64_t m = __ballot(occ);
      if (lane == 0) s_m[wave] = m;
      __syncthreads();
      if (threadIdx.x < 8u) {  // tile word of z tile 8 zb + t
        const uint32_t t = threadIdx.x, tz = zb * 8u + t;
        if (tz < TZ) {
          uint32_t word = 0u;
#pragma unroll
          for (int w = 0; w < 4; ++w) word |= (uint32_t)((s_m[w] >> (8u * t)) & 0xffu) << (8 * w);
          const uint32_t tile = ((i >> 3) * TY + (yq >> 1)) * TZ + tz;
          out[(size_t)tile * 16u + (i & 7u) * 2u + (yq & 1u)] = word;
        }
      }
      __syncthreads();
    }
  }
}

// Bytes (X,Y,Z), nonzero = occupied, to occupancy tiles: one thread per output word.
__global__ void k_pack_occupancy(const uint8_t* __restrict__ occ, uint32_t X, uint32_t Y, uint32_t Z,
                                 uint32_t* __restrict__ out, int64_t n_words) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint32_t TY = (Y + 7u) >> 3, TZ = (Z + 7u) >> 3;
  uint32_t word = 0u;
  for (uint32_t b = 0; b < 32u; ++b) {
    uint32_t i, j, k;
    occ_cell((uint32_t)w, b, TY, TZ, i, j, k);
    if (i < X && j < Y && k < Z) word |= (uint32_t)(occ[((size_t)i * Y + j) * Z + k] != 0) << b;
  }
  out[w] = word;
}

__global__ void k_fill(float4* __restrict__ dst, int64_t n4, float value) {
  const float4 v = make_float4(value, value, value, value);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

// Fill of the cells a scatter can reach (pxl_fill_reachable_f32): one wavefront per 8x8x8 tile
// (workgroups of one wave: a workgroup of four waves per tile took 19.8 us at 256^3 with ~5%
// occupancy, in latency-bound rounds; one wave per tile keeps four times as many tiles in
// flight: 10.8 us).  The
// lanes read the tile's own occupancy words (TRI: also those of the seven tiles at -1 along any
// subset of x, y, z, wrapped -- a floor cell there has corners in this tile); if any bit is set
// the wave fills the tile, cell (x, y, z) = (j, lane >> 3, lane & 7) in pass j, a z row of eight
// cells (128 bytes) per eight lanes: every cell (TRI), or only the occupied ones (nearest: a
// sample adds only to its own occupied cell; lane i < 16 holds word i and passes it on).  Cells
// beyond the dims (partial tiles) are skipped.
template <bool TRI>
__global__ __launch_bounds__(64) void k_fill_reachable(const uint32_t* __restrict__ occ, uint32_t X, uint32_t Y,
                                                       uint32_t Z, float4* __restrict__ grad, float value) {
  const uint32_t TX = (X + 7u) >> 3, TY = (Y + 7u) >> 3, TZ = (Z + 7u) >> 3;
  const uint32_t t = blockIdx.x;
  const uint32_t tz = t % TZ, ty = (t / TZ) % TY, tx = t / (TZ * TY);
  const int lane = threadIdx.x;
  constexpr int kWords = TRI ? 8 * 16 : 16;  // word i: tile offset j = i >> 4 in {0, -1}^3, word i & 15
  uint32_t w = 0u, any = 0u;
#pragma unroll
  for (int i0 = 0; i0 < kWords; i0 += 64) {
    const int i = i0 + lane;
    if (i < kWords) {
      const int j = i >> 4;
      const uint32_t ax = (j & 4) ? (tx == 0 ? TX - 1u : tx - 1u) : tx;
      const uint32_t ay = (j & 2) ? (ty == 0 ? TY - 1u : ty - 1u) : ty;
      const uint32_t az = (j & 1) ? (tz == 0 ? TZ - 1u : tz - 1u) : tz;
      const uint32_t v = occ[((ax * TY + ay) * TZ + az) * 16u + (uint32_t)(i & 15)];
      if (i0 == 0) w = v;  // lanes 0..15: this tile's own words
      any |= v;
    }
  }
  if (__ballot(any != 0u) == 0ull) return;
  const float4 v = make_float4(value, value, value, value);
  const uint32_t y = ty * 8u + ((uint32_t)lane >> 3), z = tz * 8u + ((uint32_t)lane & 7u);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t x = tx * 8u + (uint32_t)j;
    bool reach = x < X && y < Y && z < Z;
    if (!TRI) {  // the cell's own bit: word 2j + ((y>>2)&1) of the tile, bit (y&3)*8 + (z&7)
      const uint32_t wj = (uint32_t)__shfl((int)w, 2 * j + ((lane >> 5) & 1));
      reach = reach && ((wj >> ((((uint32_t)lane >> 3) & 3u) * 8u + ((uint32_t)lane & 7u))) & 1u) != 0u;
    }
    if (reach) store_stream(grad + ((size_t)x * Y + y) * Z + z, v);
  }
}

__global__ void k_fill_tail(float* __restrict__ dst, int64_t n, float value) {
  const int64_t i = threadIdx.x;
  if (i < n) dst[i] = value;
}

// Corner blocks -> gradient.  Entry E = (x, y, z) (4 float4) holds in slot s = 2bx + by the
// four channel gradients of corner (bx, by) of the samples whose floor cell is (x, y, *) and
// whose corner z is z (b = 0: the ceil index, wrapped; b = 1: the floor index).  Corner (bx, by)
// of row (x, y) names cell (x + 1 - bx, y + 1 - by), so output cell (x, y, z) sums slot m of
// entry (x - 1 + bx, y - 1 + by, z), wrapped, for m = 2bx + by = 0..3 in that order.  Entry
// (x, y, z) is written only by samples with floor cell (x, y, z) or (x, y, z - 1 wrapped), so
// with occupancy the entries of cells where neither bit is set are skipped.
//
// A workgroup covers up to 256 cells of NY consecutive (x, y) rows, one thread per cell along z
// (a lane handles its z in all NY rows); its four waves work independently (no LDS, no
// barrier).  Output row y0 + i reads slot m = 2bx + by from source row (x - 1 + bx, y0 - 1 + i
// + by), so the NY rows share NY + 1 source rows per bx: each lane tests its own z in each
// (the z - 1 bits come from the lane below by a DPP shift; the first lane of each wave reads its
// own), then issues all 4 NY entry loads (16 bytes each) before summing, so one wave keeps
// NY times the loads in flight behind a single occupancy round trip.  Units are numbered so
// that the workgroups in flight on one XCD (workgroup id mod 8) hold neighbouring y rows,
// which read the same entries.  `clear` (a second block buffer, not read in this pass) has the
// written entries of each output row zeroed in the same pass by coalesced stores, ready for
// the next step.  Gradient rows and cleared entries are stored with the streaming cache
// policy: written once here, they would otherwise evict the entry lines that the workgroup of
// the next x row reads (measured: 167 -> 88 us for this pass alone at 256^3 with ~5%
// occupancy, and the following train step's gathers run faster too).  Entry loads stay
// ordinary for that reuse.
constexpr int kReduceCells = 256;
constexpr int kReduceRows = 4;  // NY

__device__ __forceinline__ uint32_t shift_up_u32(uint32_t v, uint32_t fill) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, DPP_WAVE_SHR1, 0xf, 0xf, false);
}

template <int NY>
__global__ __launch_bounds__(kReduceCells) void k_blocks_reduce(const float4* __restrict__ blocks,
                                                                const uint32_t* __restrict__ occ,
                                                                int X, int Y, int Z, int accumulate,
                                                                float4* __restrict__ clear,
                                                                float4* __restrict__ grad) {
  constexpr int NR = NY + 1;  // source rows per bx
  const uint32_t nzc = ((uint32_t)Z + kReduceCells - 1) / kReduceCells;
  const uint32_t nyg = ((uint32_t)Y + NY - 1) / NY;
  const uint32_t n_units = (uint32_t)X * nyg * nzc;
  const uint32_t n_wg = ((n_units + 7u) / 8u) * 8u, per_xcd = n_wg / 8u;
  const int tid = (int)threadIdx.x, lane = tid & 63;
  for (uint32_t wg = blockIdx.x; wg < n_wg; wg += gridDim.x) {
    const uint32_t u = (wg % 8u) * per_xcd + wg / 8u;  // XCD-contiguous unit order
    if (u >= n_units) continue;
    const uint32_t zc = u % nzc, rxy = u / nzc;
    const uint32_t x = rxy / nyg, y0 = (rxy % nyg) * NY;
    const uint32_t z0 = zc * kReduceCells;
    const uint32_t nz = min((uint32_t)kReduceCells, (uint32_t)Z - z0);
    const uint32_t xm = x == 0 ? (uint32_t)X - 1u : x - 1u;
    // source row r (0..NY) is y0 - 1 + r (r = 0 wraps at y0 = 0); rows past Y - 1 are unused
    uint32_t base[2][NR], yrow[NR];
    bool rv[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t yr = r == 0 ? (y0 == 0 ? (uint32_t)Y - 1u : y0 - 1u) : y0 - 1u + (uint32_t)r;
      rv[r] = r == 0 || yr < (uint32_t)Y;
      yrow[r] = yr;
      base[0][r] = (xm * (uint32_t)Y + yr) * (uint32_t)Z;
      base[1][r] = (x * (uint32_t)Y + yr) * (uint32_t)Z;
    }
    const uint32_t xrow[2] = {xm, x};
    const uint32_t TY = ((uint32_t)Y + 7u) >> 3, TZ = ((uint32_t)Z + 7u) >> 3;
    const uint32_t z = z0 + (uint32_t)tid;
    const bool active = (uint32_t)tid < nz;
    const uint32_t zp = z == 0 ? (uint32_t)Z - 1u : z - 1u;
    // occupancy bit (bx, r) at bit bx * NR + r: own for z; the lane below supplies z - 1 (the
    // first lane of each wave reads it).  Every word index is formed first (rows past Y - 1 and
    // lanes past the row's end read word 0; lanes other than 0 repeat their own word for z - 1,
    // which coalesces), then all loads are issued before any is tested: one memory round trip.
    uint32_t own = 0u, own_p = 0u;
    if (occ == nullptr) {
#pragma unroll
      for (int bx = 0; bx < 2; ++bx)
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (active && rv[r]) own |= 1u << (bx * NR + r);
      own_p = lane == 0 ? own : 0u;
    } else {
      OccBit ob[2][NR], obp[2][NR];
      const uint32_t zc = active ? z : 0u, zpc = active && lane == 0 ? zp : zc;
#pragma unroll
      for (int bx = 0; bx < 2; ++bx)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const uint32_t yc = rv[r] ? yrow[r] : 0u;
          ob[bx][r] = occ_bit(xrow[bx], yc, zc, TY, TZ);
          obp[bx][r] = occ_bit(xrow[bx], yc, zpc, TY, TZ);
        }
      uint32_t w[2][NR], wp[2][NR];
#pragma unroll
      for (int bx = 0; bx < 2; ++bx)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          w[bx][r] = occ[ob[bx][r].word];
          wp[bx][r] = occ[obp[bx][r].word];
        }
#pragma unroll
      for (int bx = 0; bx < 2; ++bx)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const bool ok = active && rv[r];
          own |= (uint32_t)(ok && ((w[bx][r] >> ob[bx][r].shift) & 1u)) << (bx * NR + r);
          own_p |= (uint32_t)(ok && lane == 0 && ((wp[bx][r] >> obp[bx][r].shift) & 1u)) << (bx * NR + r);
        }
    }
    // entry (bx, r, z) may hold sums; lanes past the row's end (z >= Z) touch nothing
    const uint32_t written = active ? own | shift_up_u32(own, own_p) : 0u;
    float4 v[NY][4];  // all loads first (independent, in flight together), then the sums
#pragma unroll
    for (int i = 0; i < NY; ++i)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int bx = m >> 1, r = i + (m & 1);
        v[i][m] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (y0 + (uint32_t)i < (uint32_t)Y && ((written >> (bx * NR + r)) & 1u))
          v[i][m] = blocks[(size_t)(base[bx][r] + z) * 4u + m];
      }
#pragma unroll
    for (int i = 0; i < NY; ++i) {
      if (y0 + (uint32_t)i >= (uint32_t)Y) break;
      float4 acc = v[i][0];
#pragma unroll
      for (int m = 1; m < 4; ++m) {
        acc.x += v[i][m].x;
        acc.y += v[i][m].y;
        acc.z += v[i][m].z;
        acc.w += v[i][m].w;
      }
      // the cell's four source entries (bx, row i + by), any of which a step may have written
      const uint32_t reach = (3u << i) | (3u << (NR + i));
      if (active && (accumulate != PXL_REDUCE_OVERWRITE_REACHABLE || (written & reach) != 0u)) {
        const uint32_t c = base[1][i + 1] + z;
        if (accumulate == PXL_REDUCE_ACCUMULATE) {
          const float4 g = grad[c];
          acc = make_float4(g.x + acc.x, g.y + acc.y, g.z + acc.z, g.w + acc.w);
        }
        store_stream(grad + c, acc);
      }
      if (clear) {  // row i's 64 entries of this wave = 256 float4: lane l of pass k clears 64k + l
        const uint32_t wz0 = z0 + (uint32_t)(tid & ~63);
        float4* cb = clear + (size_t)(base[1][i + 1] + wz0) * 4u;
        const uint32_t wr = (written >> (NR + i + 1)) & 1u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int cell = 16 * k + (lane >> 2);  // lane within the wave whose entry this is
          if (__shfl(wr, cell)) store_stream(cb + 64 * k + lane, make_float4(0.f, 0.f, 0.f, 0.f));
        }
      }
    }
  }
}

// Zero every entry that a step with this occupancy may have written (entry (x, y, z) when the
// bit of (x, y, z) or (x, y, z - 1 wrapped) is set; every entry without occupancy).  Four lanes
// per 64-byte entry, so each store instruction covers sixteen whole consecutive entries.
__global__ __launch_bounds__(256) void k_blocks_zero(float4* __restrict__ blocks,
                                                     const uint32_t* __restrict__ occ,
                                                     int Y, int Z, int64_t n_cells) {
  const int64_t n = n_cells * 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const uint32_t TY = ((uint32_t)Y + 7u) >> 3, TZ = ((uint32_t)Z + 7u) >> 3;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const uint32_t e = (uint32_t)(t >> 2);
    const uint32_t ez = e % (uint32_t)Z, ey = (e / (uint32_t)Z) % (uint32_t)Y, ex = e / ((uint32_t)Z * (uint32_t)Y);
    const uint32_t zp = ez == 0 ? (uint32_t)Z - 1u : ez - 1u;
    if (occupied_at(occ, ex, ey, ez, TY, TZ) || occupied_at(occ, ex, ey, zp, TY, TZ))
      store_stream(blocks + t, make_float4(0.f, 0.f, 0.f, 0.f));
  }
}

}  // namespace pxl

// ============================================================================= host ABI

namespace pxl_host {

namespace {
thread_local std::string g_err;
}

const char* last_error_text() { return g_err.c_str(); }

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    return fail(PXL_E_NO_DEVICE, "%s: no HIP device (%s)", what, hipGetErrorString(e));
  return fail(PXL_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int launched(const char* what) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PXL_OK : hip_fail(e, what);
}

int check_dims(int64_t X, int64_t Y, int64_t Z) {
  if (X <= 0 || Y <= 0 || Z <= 0)
    return fail(PXL_E_INVALID, "grid dims must be positive (%lld, %lld, %lld)", (long long)X,
                (long long)Y, (long long)Z);
  if (X > kMaxCells || Y > kMaxCells || Z > kMaxCells || X * Y > kMaxCells || X * Y * Z > kMaxCells)
    return fail(PXL_E_SIZE, "grid has more than 2^31-1 cells");
  return PXL_OK;
}

// Words of the occupancy tiles of an (X,Y,Z) grid: 16 per 8x8x8 tile (dims rounded up to 8).
int64_t occupancy_words(int64_t X, int64_t Y, int64_t Z) {
  return ((X + 7) / 8) * ((Y + 7) / 8) * ((Z + 7) / 8) * 16;
}

// The kernels index occupancy words in 32 bits.
int check_occupancy_words(int64_t X, int64_t Y, int64_t Z) {
  if (occupancy_words(X, Y, Z) > int64_t(0xffffffff))
    return fail(PXL_E_SIZE, "occupancy tiles of a (%lld, %lld, %lld) grid exceed 2^32 words", (long long)X,
                (long long)Y, (long long)Z);
  return PXL_OK;
}

int check_grid(const pxl_grid* g) {
  if (!g) return fail(PXL_E_INVALID, "grid descriptor is NULL");
  if (!g->data) return fail(PXL_E_INVALID, "grid data is NULL");
  if (g->channels != 4) return fail(PXL_E_DTYPE, "grid must have 4 channels, got %d", g->channels);
  if (g->dtype != PXL_DTYPE_F32 && g->dtype != PXL_DTYPE_BF16 && g->dtype != PXL_DTYPE_F16)
    return fail(PXL_E_DTYPE, "unsupported grid dtype %d", g->dtype);
  int rc;
  if ((rc = check_dims(g->dims[0], g->dims[1], g->dims[2]))) return rc;
  return g->occupancy ? check_occupancy_words(g->dims[0], g->dims[1], g->dims[2]) : PXL_OK;
}

}  // namespace pxl_host

namespace {

using pxl_host::check_grid;
using pxl_host::fail;
using pxl_host::hip_fail;
using pxl_host::kMaxCells;
using pxl_host::kMaxBlockCells;
using pxl_host::check_dims;
using pxl_host::check_occupancy_words;
using pxl_host::occupancy_words;
using pxl_host::launched;
using pxl_host::stride_blocks;

constexpr int64_t kMaxRaysPerLaunch = int64_t(1) << 26;  // at most 2^24 blocks per launch

int check_march(const pxl_march* m) {
  if (!m) return fail(PXL_E_INVALID, "march descriptor is NULL");
  if (m->n_samples < 0 || m->n_samples > (1 << 24))
    return fail(PXL_E_INVALID, "n_samples %d outside [0, 2^24]", m->n_samples);
  if (m->interp != PXL_INTERP_NEAREST && m->interp != PXL_INTERP_TRILINEAR)
    return fail(PXL_E_INVALID, "unknown interp %d", m->interp);
  if ((m->flags & ~PXL_FLAG_EXACT_DIV) != 0) return fail(PXL_E_INVALID, "unknown flags 0x%x", m->flags);
  if (!(m->pd > 0.f) || !std::isfinite(m->pd) || !std::isfinite(m->inv_pd))
    return fail(PXL_E_INVALID, "points_distance must be positive and finite");
  if (!std::isfinite(m->delta) || !std::isfinite(m->near))
    return fail(PXL_E_INVALID, "delta/near must be finite");
  if (!(m->t_stop >= 0.f) || m->t_stop > 1.f) return fail(PXL_E_INVALID, "t_stop must be in [0,1]");
  return PXL_OK;
}

int check_rays(const pxl_rays* r, bool need_imgs) {
  if (!r) return fail(PXL_E_INVALID, "rays descriptor is NULL");
  if (r->n_rays < 0) return fail(PXL_E_INVALID, "n_rays is negative");
  if ((r->flags & ~PXL_RAYS_CPU_NORM) != 0) return fail(PXL_E_INVALID, "unknown ray flags 0x%x", r->flags);
  if (r->n_rays == 0) return PXL_OK;  // an empty batch reads nothing (torch gives it NULL buffers)
  if (r->origins) {
    if (!r->dirs) return fail(PXL_E_INVALID, "origins given without dirs");
  } else {
    if (!r->cam_ids || !r->uv || !r->c2w)
      return fail(PXL_E_INVALID, "camera-form rays need cam_ids, uv and c2w");
    if (r->n_cams <= 0) return fail(PXL_E_INVALID, "n_cams must be positive");
  }
  if (need_imgs) {
    if (r->origins) return fail(PXL_E_INVALID, "targets from images need camera-form rays");
    if (!r->imgs || r->img_h <= 0 || r->img_w <= 0)
      return fail(PXL_E_INVALID, "camera targets need imgs and positive img_h/img_w");
  }
  return PXL_OK;
}

pxl::KParams make_params(const pxl_grid* g, const pxl_march* m, const pxl_rays* r) {
  pxl::KParams P{};
  if (g) {
    P.grid = g->data;
    P.occ = g->occupancy;
    P.X = (int)g->dims[0];
    P.Y = (int)g->dims[1];
    P.Z = (int)g->dims[2];
    P.fX = (float)g->dims[0];
    P.fY = (float)g->dims[1];
    P.fZ = (float)g->dims[2];
    P.occ_ty = (uint32_t)((g->dims[1] + 7) / 8);
    P.occ_tz = (uint32_t)((g->dims[2] + 7) / 8);
  }
  if (m) {
    P.n_samples = m->n_samples;
    P.delta = m->delta;
    P.near = m->near;
    P.t_stop = m->t_stop;
    P.pd = m->pd;
    P.inv_pd = m->inv_pd;
    P.gx = m->grid_min[0];
    P.gy = m->grid_min[1];
    P.gz = m->grid_min[2];
    P.exact_div = (m->flags & PXL_FLAG_EXACT_DIV) ? 1 : 0;
    P.aux_stride = (int)pxl_aux_floats(m->n_samples);
  }
  P.n_rays = r->n_rays;
  P.origins = r->origins;
  P.dirs = r->dirs;
  P.cam_ids = r->cam_ids;
  P.uv = r->uv;
  P.c2w = r->c2w;
  P.imgs = r->imgs;
  P.n_cams = r->n_cams;
  P.img_h = r->img_h;
  P.img_w = r->img_w;
  P.theta = r->camera_angle_x;
  P.cpu_norm = (r->flags & PXL_RAYS_CPU_NORM) ? 1 : 0;
  return P;
}

// Launch `fn(P, blocks)` over all rays in pieces of at most kMaxRaysPerLaunch.
template <typename F>
int for_ray_pieces(pxl::KParams P, int64_t rays_per_block, F&& fn, const char* what) {
  for (int64_t off = 0; off < P.n_rays; off += kMaxRaysPerLaunch) {
    P.ray_offset = off;
    const int64_t n = std::min<int64_t>(kMaxRaysPerLaunch, P.n_rays - off);
    fn(P, (unsigned)((n + rays_per_block - 1) / rays_per_block));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
  }
  return PXL_OK;
}

#define PXL_DISPATCH_DT_INTERP(DT, INTERP, BODY)                                     \
  do {                                                                               \
    if ((DT) == PXL_DTYPE_F32) {                                                     \
      constexpr int kDT = PXL_DTYPE_F32;                                             \
      if ((INTERP) == PXL_INTERP_NEAREST) {                                          \
        constexpr int kIN = PXL_INTERP_NEAREST;                                      \
        BODY;                                                                        \
      } else {                                                                       \
        constexpr int kIN = PXL_INTERP_TRILINEAR;                                    \
        BODY;                                                                        \
      }                                                                              \
    } else if ((DT) == PXL_DTYPE_BF16) {                                             \
      constexpr int kDT = PXL_DTYPE_BF16;                                            \
      if ((INTERP) == PXL_INTERP_NEAREST) {                                          \
        constexpr int kIN = PXL_INTERP_NEAREST;                                      \
        BODY;                                                                        \
      } else {                                                                       \
        constexpr int kIN = PXL_INTERP_TRILINEAR;                                    \
        BODY;                                                                        \
      }                                                                              \
    } else {                                                                         \
      constexpr int kDT = PXL_DTYPE_F16;                                             \
      if ((INTERP) == PXL_INTERP_NEAREST) {                                          \
        constexpr int kIN = PXL_INTERP_NEAREST;                                      \
        BODY;                                                                        \
      } else {                                                                       \
        constexpr int kIN = PXL_INTERP_TRILINEAR;                                    \
        BODY;                                                                        \
      }                                                                              \
    }                                                                                \
  } while (0)

}  // namespace

extern "C" {

int pxl_abi_version(void) { return PXL_ABI_VERSION; }

const char* pxl_last_error(void) { return pxl_host::last_error_text(); }

int64_t pxl_aux_floats(int32_t n_samples) {
  return 2 + 2 * (n_samples > 0 ? ((int64_t)n_samples + 63) / 64 : 0);
}

int pxl_render_fwd(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                   float* out_rgba, float* aux, void* stream) {
  int rc;
  if ((rc = check_grid(grid)) || (rc = check_march(march)) || (rc = check_rays(rays, false))) return rc;
  if (rays->n_rays == 0) return PXL_OK;
  if (!out_rgba) return fail(PXL_E_INVALID, "out_rgba is NULL");
  pxl::KParams P = make_params(grid, march, rays);
  hipStream_t s = (hipStream_t)stream;
  PXL_DISPATCH_DT_INTERP(grid->dtype, march->interp, {
    rc = for_ray_pieces(P, pxl::kWavesPerBlock, [&](const pxl::KParams& Q, unsigned nb) {
      hipLaunchKernelGGL((pxl::k_render_fwd<kDT, kIN>), dim3(nb), dim3(pxl::kBlock), 0, s, Q,
                         out_rgba, aux);
    }, "pxl_render_fwd");
  });
  return rc;
}

int pxl_render_bwd(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                   const float* grad_rgba, const float* aux, float* grad_grid, void* stream) {
  int rc;
  if ((rc = check_grid(grid)) || (rc = check_march(march)) || (rc = check_rays(rays, false))) return rc;
  if (rays->n_rays == 0) return PXL_OK;
  if (!grad_rgba || !aux || !grad_grid) return fail(PXL_E_INVALID, "grad_rgba, aux and grad_grid are required");
  if (march->n_samples == 0) return PXL_OK;
  pxl::KParams P = make_params(grid, march, rays);
  hipStream_t s = (hipStream_t)stream;
  PXL_DISPATCH_DT_INTERP(grid->dtype, march->interp, {
    rc = for_ray_pieces(P, pxl::kWavesPerBlock, [&](const pxl::KParams& Q, unsigned nb) {
      hipLaunchKernelGGL((pxl::k_render_bwd<kDT, kIN>), dim3(nb), dim3(pxl::kBlock), 0, s, Q,
                         grad_rgba, aux, grad_grid);
    }, "pxl_render_bwd");
  });
  return rc;
}

static int train_step(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                      const pxl_loss* loss, float* out_rgba, float* grad_grid, bool blocks,
                      void* stream) {
  int rc;
  if (!loss) return fail(PXL_E_INVALID, "loss descriptor is NULL");
  if ((rc = check_grid(grid)) || (rc = check_march(march)) ||
      (rc = check_rays(rays, loss->targets == nullptr)))
    return rc;
  if (!grad_grid) return fail(PXL_E_INVALID, blocks ? "blocks is NULL" : "grad_grid is NULL");
  if (blocks && march->interp != PXL_INTERP_TRILINEAR)
    return fail(PXL_E_INVALID, "corner blocks hold trilinear gradients (interp must be trilinear)");
  if (blocks && grid->dims[0] * grid->dims[1] * grid->dims[2] > kMaxBlockCells)
    return fail(PXL_E_SIZE, "corner blocks index entry*4 + slot in 32 bits: at most %lld cells",
                (long long)kMaxBlockCells);
  if (!std::isfinite(loss->grad_scale) || !std::isfinite(loss->beta_scale))
    return fail(PXL_E_INVALID, "grad_scale and beta_scale must be finite");
  if (march->n_samples > PXL_TRAIN_MAX_SAMPLES)
    return fail(PXL_E_INVALID, "pxl_train_step handles n_samples <= %d (got %d); use "
                "pxl_render_fwd + pxl_render_bwd", PXL_TRAIN_MAX_SAMPLES, march->n_samples);
  if (rays->n_rays == 0) return PXL_OK;
  pxl::KParams P = make_params(grid, march, rays);
  P.corner_blocks = blocks ? 1 : 0;
  pxl::LossParams LP{loss->targets, loss->grad_scale, loss->beta_scale, loss->ray_sq_err,
                     loss->ray_beta};
  hipStream_t s = (hipStream_t)stream;
  const int S = march->n_samples;
  if (S > 512) {  // one ray per wavefront, group state in LDS (k_train_long)
    PXL_DISPATCH_DT_INTERP(grid->dtype, march->interp, {
      rc = for_ray_pieces(P, pxl::kWavesPerBlock, [&](const pxl::KParams& Q, unsigned nb) {
        hipLaunchKernelGGL((pxl::k_train_long<kDT, kIN>), dim3(nb), dim3(pxl::kBlock), 0, s, Q, LP, out_rgba,
                           grad_grid);
      }, blocks ? "pxl_train_step_blocks" : "pxl_train_step");
    });
    return rc;
  }
  PXL_DISPATCH_DT_INTERP(grid->dtype, march->interp, {
    // two rays per wavefront (k_train_pair), eight per block
    rc = for_ray_pieces(P, 2 * pxl::kWavesPerBlock, [&](const pxl::KParams& Q, unsigned nb) {
      const pxl::LossParams LQ = LP;  // per-ray arrays are indexed by the global ray number
      if (S <= 64)
        hipLaunchKernelGGL((pxl::k_train_pair<kDT, kIN, 1>), dim3(nb), dim3(pxl::kBlock), 0, s, Q,
                           LQ, out_rgba, grad_grid);
      else if (S <= 128)
        hipLaunchKernelGGL((pxl::k_train_pair<kDT, kIN, 2>), dim3(nb), dim3(pxl::kBlock), 0, s, Q,
                           LQ, out_rgba, grad_grid);
      else if (S <= 256)
        hipLaunchKernelGGL((pxl::k_train_pair<kDT, kIN, 4>), dim3(nb), dim3(pxl::kBlock), 0, s, Q,
                           LQ, out_rgba, grad_grid);
      else
        hipLaunchKernelGGL((pxl::k_train_pair<kDT, kIN, 8>), dim3(nb), dim3(pxl::kBlock), 0, s, Q,
                           LQ, out_rgba, grad_grid);
    }, blocks ? "pxl_train_step_blocks" : "pxl_train_step");
  });
  return rc;
}

int pxl_train_step(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                   const pxl_loss* loss, float* out_rgba, float* grad_grid, void* stream) {
  return train_step(grid, march, rays, loss, out_rgba, grad_grid, false, stream);
}

int pxl_train_step_blocks(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                          const pxl_loss* loss, float* out_rgba, float* blocks, void* stream) {
  return train_step(grid, march, rays, loss, out_rgba, blocks, true, stream);
}

int64_t pxl_blocks_floats(const pxl_grid* grid) {
  if (!grid || grid->dims[0] <= 0 || grid->dims[1] <= 0 || grid->dims[2] <= 0) return -1;
  return grid->dims[0] * grid->dims[1] * grid->dims[2] * 16;
}

int pxl_blocks_reduce(const pxl_grid* grid, float* blocks, float* grad_grid, int32_t accumulate,
                      float* clear, void* stream) {
  int rc;
  if (!grid) return fail(PXL_E_INVALID, "grid descriptor is NULL");
  if ((rc = check_dims(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  if (!blocks || !grad_grid) return fail(PXL_E_INVALID, "blocks and grad_grid are required");
  if (accumulate != PXL_REDUCE_OVERWRITE && accumulate != PXL_REDUCE_ACCUMULATE &&
      accumulate != PXL_REDUCE_OVERWRITE_REACHABLE)
    return fail(PXL_E_INVALID, "unknown reduce mode %d", accumulate);
  if (grid->occupancy && (rc = check_occupancy_words(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  const int64_t n = grid->dims[0] * grid->dims[1] * grid->dims[2];
  if (n > kMaxBlockCells)
    return fail(PXL_E_SIZE, "corner blocks index entry*4 + slot in 32 bits: at most %lld cells",
                (long long)kMaxBlockCells);
  hipStream_t s = (hipStream_t)stream;
  const int64_t units = grid->dims[0] * ((grid->dims[1] + pxl::kReduceRows - 1) / pxl::kReduceRows) *
                        ((grid->dims[2] + pxl::kReduceCells - 1) / pxl::kReduceCells);
  const unsigned n_wg = (unsigned)std::min<int64_t>(((units + 7) / 8) * 8, int64_t(1) << 24);
  const bool separate = clear == blocks;  // the pass reads `blocks`: clear them afterwards
  hipLaunchKernelGGL(pxl::k_blocks_reduce<pxl::kReduceRows>, dim3(n_wg), dim3(pxl::kReduceCells), 0, s,
                     reinterpret_cast<const float4*>(blocks), grid->occupancy, (int)grid->dims[0],
                     (int)grid->dims[1], (int)grid->dims[2], accumulate,
                     separate ? nullptr : reinterpret_cast<float4*>(clear),
                     reinterpret_cast<float4*>(grad_grid));
  if ((rc = launched("pxl_blocks_reduce"))) return rc;
  if (!separate) return PXL_OK;
  hipLaunchKernelGGL(pxl::k_blocks_zero, dim3(stride_blocks(n * 4, 256)), dim3(256), 0, s,
                     reinterpret_cast<float4*>(blocks), grid->occupancy, (int)grid->dims[1],
                     (int)grid->dims[2], n);
  return launched("pxl_blocks_reduce");
}

int pxl_blocks_clear(const pxl_grid* grid, float* blocks, void* stream) {
  int rc;
  if (!grid) return fail(PXL_E_INVALID, "grid descriptor is NULL");
  if ((rc = check_dims(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  if (grid->occupancy && (rc = check_occupancy_words(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  if (!blocks) return fail(PXL_E_INVALID, "blocks is NULL");
  const int64_t n = grid->dims[0] * grid->dims[1] * grid->dims[2];
  if (n > kMaxBlockCells)
    return fail(PXL_E_SIZE, "corner blocks index entry*4 + slot in 32 bits: at most %lld cells",
                (long long)kMaxBlockCells);
  hipLaunchKernelGGL(pxl::k_blocks_zero, dim3(stride_blocks(n * 4, 256)), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<float4*>(blocks), grid->occupancy, (int)grid->dims[1],
                     (int)grid->dims[2], n);
  return launched("pxl_blocks_clear");
}

int pxl_camera_rays(const pxl_rays* rays, float* origins, float* dirs, float* targets,
                    void* stream) {
  int rc;
  if ((rc = check_rays(rays, targets != nullptr))) return rc;
  if (rays->n_rays == 0) return PXL_OK;
  if (rays->origins) return fail(PXL_E_INVALID, "pxl_camera_rays needs camera-form rays");
  if (!origins || !dirs) return fail(PXL_E_INVALID, "origins and dirs outputs are required");
  pxl::KParams P = make_params(nullptr, nullptr, rays);
  hipStream_t s = (hipStream_t)stream;
  return for_ray_pieces(P, 256, [&](const pxl::KParams& Q, unsigned nb) {
    hipLaunchKernelGGL(pxl::k_camera_rays, dim3(nb), dim3(256), 0, s, Q, origins, dirs, targets);
  }, "pxl_camera_rays");
}

int pxl_occupancy_dilate(const uint32_t* occ_in, int64_t X, int64_t Y, int64_t Z,
                         uint32_t* occ_out, void* stream) {
  if (!occ_in || !occ_out) return fail(PXL_E_INVALID, "occupancy buffers are required");
  if (occ_in == occ_out) return fail(PXL_E_INVALID, "occ_out must not alias occ_in");
  if (X <= 0 || Y <= 0 || Z <= 0) return fail(PXL_E_INVALID, "dims must be positive");
  if (X * Y > kMaxCells || X * Y * Z > kMaxCells) return fail(PXL_E_SIZE, "grid has more than 2^31-1 cells");
  int rc;
  if ((rc = check_occupancy_words(X, Y, Z))) return rc;
  const int64_t words = occupancy_words(X, Y, Z);
  hipLaunchKernelGGL(pxl::k_dilate, dim3((unsigned)((words + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, occ_in, (uint32_t)X, (uint32_t)Y, (uint32_t)Z, occ_out, words);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PXL_OK : hip_fail(e, "pxl_occupancy_dilate");
}

int pxl_alpha_occupancy(const pxl_grid* grid, int32_t dilate, uint32_t* bits, void* stream) {
  int rc;
  if ((rc = check_grid(grid))) return rc;
  if (!bits) return fail(PXL_E_INVALID, "bits is NULL");
  if ((rc = check_occupancy_words(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  const int64_t units = (((grid->dims[0] + 7) / 8 * 8 + pxl::kAlphaOccSlab - 1) / pxl::kAlphaOccSlab) *
                        ((grid->dims[1] + 7) / 8 * 2) * (((grid->dims[2] + 7) / 8 + 7) / 8);
  const unsigned nb = (unsigned)std::min<int64_t>(units, int64_t(1) << 20);  // workgroup per unit
  hipStream_t s = (hipStream_t)stream;
  const uint32_t X = (uint32_t)grid->dims[0], Y = (uint32_t)grid->dims[1], Z = (uint32_t)grid->dims[2];
#define PXL_ALPHA_OCC(DTV)                                                                           \
  if (dilate)                                                                                        \
    hipLaunchKernelGGL((pxl::k_alpha_occupancy<DTV, true>), dim3(nb), dim3(pxl::kAlphaOccThreads), 0, s,     \
                       grid->data, X, Y, Z, bits);                                                   \
  else                                                                                               \
    hipLaunchKernelGGL((pxl::k_alpha_occupancy<DTV, false>), dim3(nb), dim3(pxl::kAlphaOccThreads), 0, s,    \
                       grid->data, X, Y, Z, bits);
  if (grid->dtype == PXL_DTYPE_F32) {
    PXL_ALPHA_OCC(PXL_DTYPE_F32)
  } else if (grid->dtype == PXL_DTYPE_BF16) {
    PXL_ALPHA_OCC(PXL_DTYPE_BF16)
  } else {
    PXL_ALPHA_OCC(PXL_DTYPE_F16)
  }
#undef PXL_ALPHA_OCC
  return launched("pxl_alpha_occupancy");
}

int pxl_pack_occupancy(const uint8_t* occ, int64_t X, int64_t Y, int64_t Z, uint32_t* out, void* stream) {
  int rc;
  if (!occ || !out) return fail(PXL_E_INVALID, "occ and out are required");
  if ((rc = check_dims(X, Y, Z)) || (rc = check_occupancy_words(X, Y, Z))) return rc;
  const int64_t words = occupancy_words(X, Y, Z);
  hipLaunchKernelGGL(pxl::k_pack_occupancy, dim3((unsigned)((words + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, occ, (uint32_t)X, (uint32_t)Y, (uint32_t)Z, out, words);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PXL_OK : hip_fail(e, "pxl_pack_occupancy");
}

int64_t pxl_occupancy_words(int64_t X, int64_t Y, int64_t Z) {
  if (X <= 0 || Y <= 0 || Z <= 0 || X > kMaxCells || Y > kMaxCells || Z > kMaxCells) return -1;
  return occupancy_words(X, Y, Z);
}

int pxl_fill_f32(float* dst, int64_t n, float value, void* stream) {
  if (n < 0) return fail(PXL_E_INVALID, "n is negative");
  if (n == 0) return PXL_OK;
  if (!dst) return fail(PXL_E_INVALID, "dst is NULL");
  hipStream_t s = (hipStream_t)stream;
  const uintptr_t addr = (uintptr_t)dst;
  if (addr % 16 != 0) return fail(PXL_E_INVALID, "dst must be 16-byte aligned");
  const int64_t n4 = n / 4;
  if (n4 > 0) {
    const int64_t want = (n4 + 255) / 256;
    const unsigned blocks = (unsigned)std::min<int64_t>(want, 256 * 8 * 16);
    hipLaunchKernelGGL(pxl::k_fill, dim3(blocks), dim3(256), 0, s, reinterpret_cast<float4*>(dst), n4, value);
  }
  if (n % 4) hipLaunchKernelGGL(pxl::k_fill_tail, dim3(1), dim3(64), 0, s, dst + n4 * 4, n % 4, value);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PXL_OK : hip_fail(e, "pxl_fill_f32");
}

int pxl_fill_reachable_f32(const pxl_grid* grid, int32_t interp, float* grad, float value, void* stream) {
  int rc;
  if (!grid) return fail(PXL_E_INVALID, "grid descriptor is NULL");
  if ((rc = check_dims(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  if (!grid->occupancy || !grad) return fail(PXL_E_INVALID, "grid->occupancy and grad are required");
  if (interp != PXL_INTERP_NEAREST && interp != PXL_INTERP_TRILINEAR) return fail(PXL_E_INVALID, "unknown interp %d", interp);
  if ((rc = check_occupancy_words(grid->dims[0], grid->dims[1], grid->dims[2]))) return rc;
  if (reinterpret_cast<uintptr_t>(grad) % 16 != 0) return fail(PXL_E_INVALID, "grad must be 16-byte aligned");
  const int64_t tiles = occupancy_words(grid->dims[0], grid->dims[1], grid->dims[2]) / 16;
  if (tiles > int64_t(0x7fffffff)) return fail(PXL_E_SIZE, "too many occupancy tiles for one launch");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t X = (uint32_t)grid->dims[0], Y = (uint32_t)grid->dims[1], Z = (uint32_t)grid->dims[2];
  if (interp == PXL_INTERP_TRILINEAR)
    hipLaunchKernelGGL(pxl::k_fill_reachable<true>, dim3((unsigned)tiles), dim3(64), 0, s, grid->occupancy, X, Y, Z,
                       reinterpret_cast<float4*>(grad), value);
  else
    hipLaunchKernelGGL(pxl::k_fill_reachable<false>, dim3((unsigned)tiles), dim3(64), 0, s, grid->occupancy, X, Y,
                       Z, reinterpret_cast<float4*>(grad), value);
  return launched("pxl_fill_reachable_f32");
}

}  // extern "C"
