/*
 * pxl_probe.h -- C ABI of libpxl_probe, measurement and test kernels that are not part of the
 * render path (no counterparts in the reference).  libpxl_probe is built beside libpxl from
 * plenoxels_amd/csrc/pxl_probe.hip; bench.py and the tests load it, the package never does.
 * Conventions as in pxl.h: device pointers, caller-owned buffers, `stream` a hipStream_t,
 * 0 or a negative PXL_E_* code.
 */
#ifndef PXL_PROBE_H
#define PXL_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the wave primitives the render kernels are built from on one wavefront.  in: 6*64 floats
 * (x, then affine-map fields a0, a1, a2, a3, b); out: 7*64 floats (inclusive product, exclusive
 * product, wave sum, suffix-composed a0, a3, b, and x shifted down one lane with -1 filling
 * lane 63). */
int pxl_probe_wave_ops(const float* in, float* out, void* stream);

/* Memory-request rate microbenchmark.  n_waves wavefronts each issue `iters` instructions of
 * kind `op` (0: global_atomic_add_f32 without return, 1: global_store_dword, 2:
 * global_load_dword) on buf (span_segments 64-byte segments, 16 floats each).  In each
 * instruction the 64 lanes touch 64 / lanes_per_segment segments: lanes_per_segment (1, 2, 4,
 * 8 or 16) consecutive lanes write distinct dwords of one segment, and the segments are chosen
 * by a hash of (wave, instruction, slot) -- spread over the whole span, or, with
 * xcd_local != 0, over the 1/8 of it assigned to the XCD the wave's workgroup runs on
 * (workgroup id mod 8).  Loads accumulate into sink[wave].  Each memory request is one segment
 * of one instruction; time the launch to get requests per second. */
int pxl_probe_memory_rate(float* buf, int64_t span_segments, int32_t op, int32_t lanes_per_segment,
                          int32_t iters, int64_t n_waves, int32_t xcd_local, float* sink, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PXL_PROBE_H */
