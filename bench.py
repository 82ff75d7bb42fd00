#!/usr/bin/env python
"""Benchmark of the fused render-and-optimise step on one MI355X.

Metric: fused fwd+bwd rays/s at a 256^3 grid, beside the plain PyTorch-ROCm path on the
same GPU, with the dominant kernel's fraction of HBM peak.

Workload "sparse256": 256^3 FP32 grid (blob field inside ~5% value-noise occupancy), 2^18
rays drawn uniformly from 100 cameras (800x800 images resident in HBM, focal 1111 px) on
the upper hemisphere of radius 4, 256 samples per ray from t = 2 + 4k/256, trilinear
lookup, occupancy skipping, early termination at transmittance 1e-4.  Synthetic, seeded
(grid 0, cameras 1, ray uniforms 2, targets 3).

One step = pxl_camera_rays (ray generation from (camera, u, v) and target gather) +
pxl_train_step_blocks (sample placement, gather, compositing, MSE gradient, backward scatter
into per-cell corner-block entries) + pxl_blocks_reduce (writes the FP32 (X,Y,Z,4) grid gradient -- no separate
zero-fill -- and zeroes the other of two alternating block buffers) + the loss value (pxl_scaled_sum).  --scatter direct times the
zero-fill + pxl_train_step with atomics straight into the gradient instead (nearest always
scatters directly: one cell per sample).  Adam is not part of the step (reference
scripts/train.py:182 runs it after the timed chain).

Baselines (outside the timed region of the fused path):
  plain_gpu     the plain PyTorch chain (oracle.plain, the reference's op sequence with
                autograd backward) on the same GPU, same rays and semantics, in ray chunks
  cpu_baseline  the same plain chain on the host CPU on the first --cpu-rays rays
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from plenoxels_amd import (CameraRays, CornerBlockPair, LossMean, MarchConfig, _lib, blocks_reduce,  # noqa: E402
                           camera_rays, fill_, occupancy_dilate, pack_occupancy, train_step)
from plenoxels_amd import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "fused fwd+bwd rays/s at 256^3 grid vs plain PyTorch-ROCm path; % HBM peak"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--interp", default="trilinear", choices=["trilinear", "nearest"])
    ap.add_argument("--dtype", default="float32", choices=["float32", "bfloat16", "float16"])
    ap.add_argument("--rays", type=int, default=1 << 18)
    ap.add_argument("--scatter", default="blocks", choices=["blocks", "direct"],
                    help="trilinear gradient path: corner blocks + reduce, or atomics into the grid")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--plain-chunk", type=int, default=32768)
    ap.add_argument("--cpu-rays", type=int, default=16384)
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-accounting", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scales rocprofv3's "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


class Setup:
    """Device-resident inputs of the sparse256 workload."""

    def __init__(self, args, dev):
        t0 = time.time()
        w = scenes.make_workload("sparse256", n_rays=args.rays, interp=args.interp)
        self.w = w
        self.dev = dev
        dtype = getattr(torch, args.dtype)
        self.grid32 = w.grid.to(dev)
        self.grid = self.grid32 if dtype == torch.float32 else self.grid32.to(dtype)
        pts = pack_occupancy(w.occupancy.to(dev))
        self.occ_bits = occupancy_dilate(pts, w.dims) if w.interp == "trilinear" else pts
        B = w.c2w.shape[0]
        H, W = w.img_hw
        gen = torch.Generator(device=dev).manual_seed(3)
        self.imgs = torch.rand(B, H, W, 4, device=dev, generator=gen)
        self.camera = CameraRays(w.cam_ids.to(dev), w.uv.to(dev), w.c2w.to(dev),
                                 w.camera_angle_x, self.imgs)
        self.cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance,
                               w.interp, w.near, w.t_stop, "reciprocal")
        R = w.cam_ids.shape[0]
        self.R = R
        self.grad = torch.zeros(self.grid.shape, dtype=torch.float32, device=dev)
        self.use_blocks = args.scatter == "blocks" and w.interp == "trilinear"
        self.pair = CornerBlockPair(self.grid) if self.use_blocks else None
        self.out = torch.empty(R, 4, device=dev)
        self.sq = torch.empty(R, device=dev)
        self.loss_mean = LossMean(dev)
        torch.cuda.synchronize()
        log(f"setup {time.time() - t0:.1f}s: grid {tuple(self.grid.shape)} {self.grid.dtype}, "
            f"occupancy {w.occupancy.float().mean().item():.4f}, rays {R}, S {w.num_samples}, {w.interp}")

    def step(self, ev=None, ev_reduce=None):
        """One step; ev = (train start, train end) events around the train-step kernel, ev_reduce
        = (start, end) around the blocks reduce, or None (each recorded event costs the
        stream a few microseconds, so the timed loop records only the train-step pair)."""
        if not self.use_blocks:
            fill_(self.grad, 0.0)
        o, d, tgt = camera_rays(self.camera)
        if ev is not None:
            ev[0].record()
        train_step(self.grid, None if self.use_blocks else self.grad, cfg=self.cfg, origins=o, dirs=d,
                   targets=tgt, occupancy=self.occ_bits, out=self.out, ray_sq_err=self.sq,
                   blocks=self.pair.blocks if self.use_blocks else None)
        if ev is not None:
            ev[1].record()
        if self.use_blocks:
            if ev_reduce is not None:
                ev_reduce[0].record()
            blocks_reduce(self.grid, self.pair.blocks, self.grad, occupancy=self.occ_bits, accumulate=False,
                          clear=self.pair.other)
            if ev_reduce is not None:
                ev_reduce[1].record()
            self.pair.flip()
        return self.loss_mean(self.sq)


def timed_fused(S, steps, warmup):
    """Wall time of `steps` steps (train-step kernel events on each), then the train-step and
    reduce kernel durations from events."""
    for _ in range(warmup):
        S.step()
    torch.cuda.synchronize()
    ev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(2)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        S.step(ev[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [e[0].elapsed_time(e[1]) for e in ev]
    red_ms = [0.0]
    if S.use_blocks:  # the reduce's duration, in steps of their own outside the timed loop
        evr = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(2)) for _ in range(steps)]
        for i in range(steps):
            S.step(None, evr[i])
        torch.cuda.synchronize()
        red_ms = [e[0].elapsed_time(e[1]) for e in evr]
    return wall, float(np.mean(kern_ms)), float(np.min(kern_ms)), float(np.mean(red_ms))


def algorithmic_bytes(S, args):
    """Per-launch algorithmic bytes of pxl_train_step: per ray 60 B (origin + direction 24 and
    target 16 in, RGBA 16 and squared error 4 out) plus K_r cells x 4 channels x (s_g read + 8
    for the FP32 gradient read-modify-write).  K_r counted exactly on the host from the seeded
    batch (pxl_count_cells_host) using the termination counts of the forward kernel."""
    w = S.w
    L = _lib.lib()
    o, d, _ = camera_rays(S.camera, want_targets=False)
    aux = torch.empty(S.R, int(L.pxl_aux_floats(w.num_samples)), device=S.dev)
    out = torch.empty(S.R, 4, device=S.dev)
    from plenoxels_amd.render import _grid_struct, _rays_struct
    g = _grid_struct(S.grid, S.occ_bits)
    m = S.cfg.to_struct()
    r = _rays_struct(o, d)
    _lib.check(L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), _lib.ptr(out),
                                _lib.ptr(aux), _lib.stream_handle(S.dev)))
    nproc = aux[:, 0].contiguous().view(torch.int32).cpu().numpy().copy()
    occ_host = S.occ_bits.cpu().numpy().view(np.uint32).copy()
    gh = _lib.Grid()
    for a in range(3):
        gh.dims[a] = w.dims[a]
    gh.channels = 4
    gh.occupancy = occ_host.ctypes.data
    oh = o.cpu().numpy().copy()
    dh = d.cpu().numpy().copy()
    counts = np.zeros(S.R, np.int64)
    nt = min(16, os.cpu_count() or 1)
    t0 = time.time()
    _lib.check(L.pxl_count_cells_host(ctypes.byref(gh), ctypes.byref(m), oh.ctypes.data, dh.ctypes.data,
                                      S.R, nproc.ctypes.data, counts.ctypes.data, nt))
    s_g = S.grid.element_size()
    per_cell = 4 * (s_g + 8)
    total = S.R * 60 + int(counts.sum()) * per_cell
    log(f"accounting {time.time() - t0:.1f}s: mean processed samples/ray {nproc.mean():.1f}, "
        f"mean distinct cells/ray {counts.mean():.1f}")
    return total, float(nproc.mean()), float(counts.mean())


def plain_gpu(S, args):
    """Plain PyTorch chain on the same GPU: rays -> samples -> gather -> composite -> MSE ->
    autograd backward into grid.grad, chunked over rays, same semantics as the fused step."""
    from oracle import plain
    w = S.w
    dev = S.dev
    occ = w.occupancy.to(dev)
    occ_used = plain.dilate_cells(occ) if w.interp == "trilinear" else occ
    cam = w.cam_ids.to(dev).long()
    u, v = w.uv[:, 0].to(dev).contiguous(), w.uv[:, 1].to(dev).contiguous()
    gmin = torch.tensor(w.grid_min, device=dev)
    grid = S.grid32.clone().requires_grad_(True)

    def run():
        grid.grad = None
        o, d, tgt = plain.camera_rays(S.imgs, S.camera.c2w, w.camera_angle_x, cam, u, v)
        plain.render_chunked_backward(grid, o, d, tgt, chunk_rays=args.plain_chunk,
                                      num_samples=w.num_samples, delta_step=w.delta_step,
                                      gmin=gmin, points_distance=w.points_distance,
                                      interp=w.interp, near=w.near, div_mode="reciprocal",
                                      occupancy=occ_used, t_stop=w.t_stop)
    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.plain_steps
    # agreement of the plain path with the fused step on this batch (same inputs)
    S.step()
    torch.cuda.synchronize()
    gerr = ((grid.grad - S.grad).abs().max() / grid.grad.abs().max().clamp_min(1e-30)).item()
    # the same comparison with the plain chain fed the fused path's rays (pxl_camera_rays, whose
    # direction norm follows torch's CPU order; torch's GPU norm differs in the last bit on some
    # rays, which moves samples across the occupancy and termination tests)
    o, d, tgt = camera_rays(S.camera)
    grid.grad = None
    plain.render_chunked_backward(grid, o, d, tgt, chunk_rays=args.plain_chunk, num_samples=w.num_samples,
                                  delta_step=w.delta_step, gmin=gmin, points_distance=w.points_distance,
                                  interp=w.interp, near=w.near, div_mode="reciprocal", occupancy=occ_used,
                                  t_stop=w.t_stop)
    torch.cuda.synchronize()
    gerr_same = ((grid.grad - S.grad).abs().max() / grid.grad.abs().max().clamp_min(1e-30)).item()
    return S.R / dt, dt, gerr, gerr_same


def cpu_baseline(S, args):
    """The plain chain on the host CPU for the first --cpu-rays rays (kind 'port': the
    oracle's restatement of the reference ops)."""
    from oracle import plain
    w = S.w
    n = min(args.cpu_rays, S.R)
    cores = min(16, os.cpu_count() or 1)
    torch.set_num_threads(cores)
    occ = w.occupancy
    occ_used = plain.dilate_cells(occ) if w.interp == "trilinear" else occ
    grid = w.grid.clone().requires_grad_(True)
    cam = w.cam_ids[:n].long()
    u, v = w.uv[:n, 0].contiguous(), w.uv[:n, 1].contiguous()
    imgs = S.imgs[: w.c2w.shape[0]].cpu()
    gmin = torch.tensor(w.grid_min)

    def run():
        grid.grad = None
        o, d, tgt = plain.camera_rays(imgs, w.c2w, w.camera_angle_x, cam, u, v, norm_mode="gpu")
        plain.render_chunked_backward(grid, o, d, tgt, chunk_rays=n, num_samples=w.num_samples,
                                      delta_step=w.delta_step, gmin=gmin,
                                      points_distance=w.points_distance, interp=w.interp,
                                      near=w.near, div_mode="reciprocal", occupancy=occ_used,
                                      t_stop=w.t_stop)
    run()
    reps = 3
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return {"value": n / dt, "unit": "rays/s", "cores": cores, "kind": "port",
            "sample": f"first {n} rays of the {S.R}-ray sparse256 batch ({w.interp}, S={w.num_samples}, "
                      f"fwd+bwd, plain PyTorch chain on CPU, median of {reps} after 1 warm-up)"}


def load_traffic(args):
    """HBM bytes per launch of the train-step kernel from the committed PMC summary, if any."""
    scatter = args.scatter if args.interp == "trilinear" else "direct"
    path = os.path.join(ROOT, "profiles", f"traffic_sparse256_{args.interp}_{args.dtype}_{scatter}.json")
    if not os.path.exists(path) or args.rays != (1 << 18):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {args.gpus}); the path shards "
              "as data parallelism over rays, which this build does not launch.", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty(n, device=dev)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}))
        return
    S = Setup(args, dev)
    wall, kern_ms, kern_min_ms, red_ms = timed_fused(S, args.steps, args.warmup)
    ms_per_step = wall * 1e3 / args.steps
    value = S.R / (wall / args.steps)
    log(f"fused: {ms_per_step:.3f} ms/step, train-step kernel {kern_ms:.3f} ms (min {kern_min_ms:.3f}), "
        f"blocks reduce {red_ms:.3f} ms")

    roofline = None
    extra = {}
    if not args.no_accounting:
        b_alg, mean_proc, mean_cells = algorithmic_bytes(S, args)
        achieved = b_alg / (kern_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(args),
                    "kernel": "k_train_pair", "kernel_ms": round(kern_ms, 4),
                    "algorithmic_bytes": b_alg}
        extra.update({"mean_processed_samples_per_ray": round(mean_proc, 2),
                      "mean_distinct_cells_per_ray": round(mean_cells, 2)})
    if S.use_blocks:
        extra["blocks_reduce_ms"] = round(red_ms, 4)
    plain = None
    cpu = None
    if not args.no_baselines:
        rps, dt, gerr, gerr_same = plain_gpu(S, args)
        plain = {"value": round(rps, 1), "unit": "rays/s", "s_per_step": round(dt, 4),
                 "grad_max_rel_diff_vs_fused": gerr, "grad_max_rel_diff_vs_fused_same_rays": gerr_same}
        log(f"plain GPU: {rps:.0f} rays/s ({dt * 1e3:.1f} ms/step), fused speed-up {value / rps:.1f}x")
        cpu = cpu_baseline(S, args)
        log(f"cpu baseline: {cpu['value']:.0f} rays/s on {cpu['cores']} threads")
    rec = {
        "metric": METRIC, "value": round(value, 1), "unit": "rays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"float32": "f32", "bfloat16": "bf16 storage, f32 compute",
                  "float16": "f16 storage, f32 compute"}[args.dtype],
        "data": "synthetic (seeded blob field, value-noise occupancy, random targets)",
        "config": {"workload": "sparse256", "grid": [256, 256, 256], "rays": S.R, "samples_per_ray": 256,
                   "interp": args.interp, "occupancy": "~5%", "t_stop": 1e-4, "cameras": 100,
                   "image": [800, 800],
                   "scatter": "corner blocks + reduce" if S.use_blocks else "direct atomics + zero-fill"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "speedup_vs_plain_gpu": round(value / plain["value"], 2) if plain else None,
    }
    rec.update(extra)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
