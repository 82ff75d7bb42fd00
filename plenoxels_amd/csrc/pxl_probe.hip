// pxl_probe.hip -- measurement and test kernels outside the render path (include/pxl_probe.h):
// the wave primitives the render kernels are built from, run on one wavefront for the tests, and
// a memory-request rate microbenchmark that prices the gradient scatter's atomics against plain
// stores and loads of the same access pattern.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pxl.h"
#include "pxl_device.h"
#include "pxl_probe.h"

namespace pxl {
namespace {

__global__ void k_wave_ops(const float* __restrict__ in, float* __restrict__ out) {
  const int lane = threadIdx.x;
  const float x = in[lane];
  const float incl = wave_prefix_mul(x);
  out[lane] = incl;
  out[64 + lane] = wave_shift_up(incl, 1.f);
  out[128 + lane] = wave_sum(x);
  Affine4 f{in[64 + lane], in[128 + lane], in[192 + lane], in[256 + lane], in[320 + lane]};
  wave_suffix_affine(f, lane);
  out[192 + lane] = f.a0;
  out[256 + lane] = f.a3;
  out[320 + lane] = f.b;
  out[384 + lane] = wave_shift_down(x, -1.f);
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// One wave = four per 256-thread block.  Lane l works on segment slot l / L of each instruction,
// dword (l % L) * (16 / L) of it; the segment is a hash of (wave, instruction, slot) within the
// span (or within this XCD's eighth of it).  Every address lies in [0, span * 16) floats.
template <int OP>
__global__ __launch_bounds__(256) void k_memory_rate(float* __restrict__ buf, uint64_t span, int L, int iters,
                                                     int64_t n_waves, int xcd_local, float* __restrict__ sink) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= n_waves) return;
  const int slot = lane / L;
  const int dw = (lane % L) * (16 / L);
  uint64_t base = 0, range = span;
  if (xcd_local) {
    range = span / 8u;
    base = (uint64_t)(blockIdx.x & 7u) * range;
  }
  const uint32_t wkey = mix32((uint32_t)wave * 0x9e3779b9u + 0x632be5abu);
  float acc = 0.f;
  for (int it = 0; it < iters; ++it) {
    const uint32_t h = mix32(wkey ^ (uint32_t)(it * 64 + slot));
    float* p = buf + (base + (uint64_t)h % range) * 16u + (uint64_t)dw;
    if constexpr (OP == 0) {
      unsafeAtomicAdd(p, 1.0f);
    } else if constexpr (OP == 1) {
      *p = (float)it;
    } else {
      acc += *p;
    }
  }
  if constexpr (OP == 2) {
    if (lane == 0) sink[wave] = acc;
  }
}

}  // namespace
}  // namespace pxl

extern "C" {

int pxl_probe_wave_ops(const float* in, float* out, void* stream) {
  if (!in || !out) return PXL_E_INVALID;
  hipLaunchKernelGGL(pxl::k_wave_ops, dim3(1), dim3(64), 0, (hipStream_t)stream, in, out);
  return hipGetLastError() == hipSuccess ? PXL_OK : PXL_E_HIP;
}

int pxl_probe_memory_rate(float* buf, int64_t span_segments, int32_t op, int32_t lanes_per_segment,
                          int32_t iters, int64_t n_waves, int32_t xcd_local, float* sink, void* stream) {
  const int L = lanes_per_segment;
  if (!buf || span_segments <= 0 || iters < 0 || n_waves < 0 || n_waves > (int64_t(1) << 26)) return PXL_E_INVALID;
  if (L != 1 && L != 2 && L != 4 && L != 8 && L != 16) return PXL_E_INVALID;
  if (op < 0 || op > 2 || (op == 2 && !sink)) return PXL_E_INVALID;
  if (xcd_local && span_segments < 8) return PXL_E_INVALID;
  if (n_waves == 0 || iters == 0) return PXL_OK;
  const unsigned blocks = (unsigned)((n_waves + 3) / 4);
  hipStream_t s = (hipStream_t)stream;
  const uint64_t span = (uint64_t)span_segments;
  if (op == 0)
    hipLaunchKernelGGL(pxl::k_memory_rate<0>, dim3(blocks), dim3(256), 0, s, buf, span, L, iters, n_waves,
                       xcd_local, sink);
  else if (op == 1)
    hipLaunchKernelGGL(pxl::k_memory_rate<1>, dim3(blocks), dim3(256), 0, s, buf, span, L, iters, n_waves,
                       xcd_local, sink);
  else
    hipLaunchKernelGGL(pxl::k_memory_rate<2>, dim3(blocks), dim3(256), 0, s, buf, span, L, iters, n_waves,
                       xcd_local, sink);
  return hipGetLastError() == hipSuccess ? PXL_OK : PXL_E_HIP;
}

}  // extern "C"
