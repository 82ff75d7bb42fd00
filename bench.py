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
into per-cell corner-block entries) + pxl_blocks_reduce (writes the FP32 (X,Y,Z,4) grid
gradient and zeroes the other of two alternating block buffers) + the loss value
(pxl_scaled_sum).  The gradient buffer is zeroed once at setup; each step's reduce writes every
cell the occupancy lets a step reach (PXL_REDUCE_OVERWRITE_REACHABLE) -- the other cells are
zero in every step's gradient and stay so.  --scatter direct (and nearest, one cell per sample)
times the zero-fill of those same reachable cells (pxl_fill_reachable_f32) + pxl_train_step
with atomics straight into the gradient.  Adam is not part of the step (reference
scripts/train.py:182 runs it after the timed chain).  Step times: HIP events at the step
boundaries on the stream the kernels run on; the headline is the median step.

The default run also measures the same workload with nearest lookup (the reference's own
training chain, scripts/train.py:146), reported under "nearest".  `--workload blobs64` (64^3,
8 cameras x 64x64 even-spread rays, 128 trilinear samples, no occupancy) and `--workload view800`
(128^3, one 800x800 view, 256 nearest samples) print the same fields for those configurations
(dense zero-fill and reduce; no CPU baseline leg).  `--memory-rate` (the atomic/store/load request
rates of the probe library) and `--stream-rate` (torch's copy/add/fill rates) measure the device
itself and exit.

Baselines (outside the timed region of the fused path):
  plain_gpu     the plain PyTorch chain (oracle.plain, the reference's op sequence with
                autograd backward) on the same GPU, same rays and semantics, in ray chunks
  cpu_baseline  the same plain chain on the host CPU: the first --cpu-rays rays of the batch
                (trilinear and nearest) and the 64^3 blob workload in full, on the CPU threads
                this process is granted
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from plenoxels_amd import (CameraRays, CornerBlockPair, LossMean, MarchConfig, _lib, blocks_reduce,  # noqa: E402
                           camera_rays, fill_, fill_reachable_, occupancy_dilate, pack_occupancy, train_step)
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
    ap.add_argument("--workload", default="sparse256", choices=["sparse256", "blobs64", "view800"],
                    help="sparse256 (the headline); blobs64 (64^3, 32,768 rays x 128 samples, trilinear) and "
                         "view800 (128^3, one 800x800 view, 256 samples, nearest) print the same fields for theirs")
    ap.add_argument("--interp", default=None, choices=["trilinear", "nearest"],
                    help="lookup (default: the workload's own, trilinear for sparse256)")
    ap.add_argument("--dtype", default="float32", choices=["float32", "bfloat16", "float16"])
    ap.add_argument("--rays", type=int, default=1 << 18)
    ap.add_argument("--scatter", default="blocks", choices=["blocks", "direct"],
                    help="trilinear gradient path: corner blocks + reduce, or atomics into the grid")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--plain-chunk", type=int, default=32768)
    ap.add_argument("--cpu-rays", type=int, default=16384)
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-accounting", action="store_true")
    ap.add_argument("--no-nearest", action="store_true", help="skip the nearest-lookup measurement")
    ap.add_argument("--memory-rate", action="store_true",
                    help="only run the memory-request rate microbenchmark (libpxl_probe: atomics, stores and "
                         "loads of 1, 4 or 16 dwords per 64-byte segment over small and large spans) and exit")
    ap.add_argument("--stream-rate", action="store_true",
                    help="only measure the device's streaming rates with torch's own kernels (copy, add, fill "
                         "of 2 GiB float32 tensors) -- the practical HBM ceiling beside the 8 TB/s nominal -- and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (a known-size reference for "
                         "rocprofv3's memory counters) and exit")
    return ap.parse_args()


class Setup:
    """Device-resident inputs of a workload (sparse256 unless --workload says otherwise)."""

    def __init__(self, args, dev, interp, imgs=None):
        t0 = time.time()
        name = getattr(args, "workload", "sparse256")
        w = scenes.make_workload(name, n_rays=args.rays if name == "sparse256" else None, interp=interp)
        self.w = w
        self.dev = dev
        dtype = getattr(torch, args.dtype)
        self.grid32 = w.grid.to(dev)
        self.grid = self.grid32 if dtype == torch.float32 else self.grid32.to(dtype)
        self.occ_bits = None  # blobs64 and view800 have no occupancy: dense fill and reduce
        if w.occupancy is not None:
            pts = pack_occupancy(w.occupancy.to(dev))
            self.occ_bits = occupancy_dilate(pts, w.dims) if w.interp == "trilinear" else pts
        B = w.c2w.shape[0]
        H, W = w.img_hw
        if imgs is None:
            gen = torch.Generator(device=dev).manual_seed(3)
            imgs = torch.rand(B, H, W, 4, device=dev, generator=gen)
        self.imgs = imgs
        self.camera = CameraRays(w.cam_ids.to(dev), w.uv.to(dev), w.c2w.to(dev),
                                 w.camera_angle_x, self.imgs)
        self.cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance,
                               w.interp, w.near, w.t_stop, "reciprocal")
        R = w.cam_ids.shape[0]
        self.R = R
        self.grad = torch.zeros(self.grid.shape, dtype=torch.float32, device=dev)
        self.use_blocks = args.scatter == "blocks" and w.interp == "trilinear"
        self.scatter = "corner blocks + reduce" if self.use_blocks else "direct atomics + zero-fill"
        self.scatter_key = "blocks" if self.use_blocks else "direct"
        self.pair = CornerBlockPair(self.grid) if self.use_blocks else None
        self.out = torch.empty(R, 4, device=dev)
        self.sq = torch.empty(R, device=dev)
        self.loss_mean = LossMean(dev)
        torch.cuda.synchronize()
        log(f"setup {time.time() - t0:.1f}s: grid {tuple(self.grid.shape)} {self.grid.dtype}, "
            f"occupancy {'none' if w.occupancy is None else round(w.occupancy.float().mean().item(), 4)}, "
            f"rays {R}, S {w.num_samples}, {w.interp}")

    def step(self, ev=None, ev_reduce=None):
        """One step; ev = (train start, train end) events around the train-step kernel, ev_reduce
        = (start, end) around the blocks reduce, or None (each recorded event costs the
        stream a few microseconds, so the headline loop records only step boundaries)."""
        if not self.use_blocks:  # the zero-fill of the cells this occupancy lets the scatter reach
            if self.occ_bits is None:
                fill_(self.grad, 0.0)
            else:
                fill_reachable_(self.grad, self.occ_bits, self.w.interp)
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
                          clear=self.pair.other, reachable_only=self.occ_bits is not None)
            if ev_reduce is not None:
                ev_reduce[1].record()
            self.pair.flip()
        return self.loss_mean(self.sq)


def _events(n):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def timed_fused(S, steps, warmup):
    """Step durations from HIP events at the step boundaries (the headline: their median), the
    wall time of those steps, then in further steps the train-step kernel's and the reduce's
    durations from events around each."""
    for _ in range(warmup):
        S.step()
    torch.cuda.synchronize()
    ev = _events(steps + 1)
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(steps):
        S.step()
        ev[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    evk = [tuple(_events(2)) for _ in range(steps)]
    for i in range(steps):
        S.step(evk[i])
    torch.cuda.synchronize()
    kern_ms = [e[0].elapsed_time(e[1]) for e in evk]
    red_ms = [0.0]
    if S.use_blocks:
        evr = [tuple(_events(2)) for _ in range(steps)]
        for i in range(steps):
            S.step(None, evr[i])
        torch.cuda.synchronize()
        red_ms = [e[0].elapsed_time(e[1]) for e in evr]
    return {"step_ms": step_ms, "wall": wall, "kern_ms": kern_ms, "red_ms": red_ms}


def algorithmic_bytes(S):
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
    occ_host = None if S.occ_bits is None else S.occ_bits.cpu().numpy().view(np.uint32).copy()
    gh = _lib.Grid()
    for a in range(3):
        gh.dims[a] = w.dims[a]
    gh.channels = 4
    gh.occupancy = None if occ_host is None else occ_host.ctypes.data
    oh = o.cpu().numpy().copy()
    dh = d.cpu().numpy().copy()
    counts = np.zeros(S.R, np.int64)
    t0 = time.time()
    _lib.check(L.pxl_count_cells_host(ctypes.byref(gh), ctypes.byref(m), oh.ctypes.data, dh.ctypes.data,
                                      S.R, nproc.ctypes.data, counts.ctypes.data, cpu_threads()))
    s_g = S.grid.element_size()
    per_cell = 4 * (s_g + 8)
    total = S.R * 60 + int(counts.sum()) * per_cell
    log(f"accounting {time.time() - t0:.1f}s: mean processed samples/ray {nproc.mean():.1f}, "
        f"mean distinct cells/ray {counts.mean():.1f}")
    return total, float(nproc.mean()), float(counts.mean())


def profile_record(S, args):
    """Per-launch memory counters of the train-step kernel from the committed rocprofv3 --pmc
    summary of this configuration (profiles/pmc_*.json), if any: raw TCC->EA read requests by
    size (32, 64, 128 B), non-atomic write requests by size, atomic requests (each a 64-byte
    segment executed at the memory side) and DRAM-bound reads."""
    name = f"pmc_{S.w.name}_{S.w.interp}_{args.dtype}_{S.scatter_key}.json"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path) or (S.w.name == "sparse256" and args.rays != (1 << 18)):
        return None
    with open(path) as f:
        rec = json.load(f)
    rec["file"] = "profiles/" + name
    return rec


def atomic_peak(dev):
    """The memory-side atomic request rate of this GPU, measured now with the probe kernel
    (pxl_probe_memory_rate: global_atomic_add_f32, 16 lanes per 64-byte segment) over a small
    (4 MiB) and a 1 GiB span: requests per second, the median of 5 launches each."""
    P = _lib.probe_lib()
    stream = _lib.stream_handle(dev)
    buf = torch.zeros((1 << 24) * 16, device=dev)
    n_waves, iters = 1 << 16, 64
    out = {}
    for name, span in (("4MiB", 1 << 16), ("1GiB", 1 << 24)):
        def run():
            _lib.check(P.pxl_probe_memory_rate(_lib.ptr(buf), span, 0, 16, iters, n_waves, 0, None, stream))
        run()
        times = []
        for _ in range(5):
            e0, e1 = _events(2)
            e0.record()
            run()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) * 1e-3)
        out[name] = n_waves * iters * 4 / float(np.median(times))
    del buf
    torch.cuda.empty_cache()
    return out


def measure(S, args, peaks):
    """The fused step of one Setup: headline figures, the train-step kernel's roofline."""
    t = timed_fused(S, args.steps, args.warmup)
    ms_step = float(np.median(t["step_ms"]))
    kern_ms = float(np.median(t["kern_ms"]))
    red_ms = float(np.median(t["red_ms"]))
    log(f"{S.w.interp}: {ms_step:.3f} ms/step (median; wall {t['wall'] * 1e3 / args.steps:.3f} ms/step), "
        f"train-step kernel {kern_ms:.3f} ms, reduce {red_ms:.3f} ms")
    res = {"value": round(S.R / (ms_step * 1e-3), 1), "ms_per_step": round(ms_step, 4),
           "ms_per_step_wall_mean": round(t["wall"] * 1e3 / args.steps, 4),
           "ms_per_step_min": round(float(np.min(t["step_ms"])), 4)}
    if S.use_blocks:
        res["blocks_reduce_ms"] = round(red_ms, 4)
    if args.no_accounting:
        return res
    b_alg, mean_proc, mean_cells = algorithmic_bytes(S)
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    prof = profile_record(S, args)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_train_pair", "kernel_ms": round(kern_ms, 4), "algorithmic_bytes": b_alg}
    if prof is not None:
        roof["traffic"] = prof["hbm_bytes_per_launch"]
        roof["traffic_detail"] = {k: prof[k] for k in ("read_bytes", "write_bytes", "atomic_bytes",
                                                       "memory_side_read_bytes", "file") if k in prof}
        if prof.get("trace_avg_ns"):  # the profiled kernel's mean duration, beside this run's
            roof["traffic_detail"]["profile_kernel_ms"] = round(prof["trace_avg_ns"] * 1e-6, 4)
        req = prof.get("atomic_requests")
        if req and peaks:
            rate = req / (kern_ms * 1e-3)
            roof["atomic"] = {"achieved": round(rate / 1e9, 2), "peak": round(peaks["4MiB"] / 1e9, 2),
                              "peak_1GiB_span": round(peaks["1GiB"] / 1e9, 2), "unit": "G requests/s",
                              "frac": round(rate / peaks["4MiB"], 4), "requests_per_launch": req,
                              "note": "float atomics execute at the memory side (TCC_EA0_ATOMIC = TCC_ATOMIC); "
                                      "peak measured now by pxl_probe_memory_rate"}
    res["roofline"] = roof
    res["mean_processed_samples_per_ray"] = round(mean_proc, 2)
    res["mean_distinct_cells_per_ray"] = round(mean_cells, 2)
    return res


def plain_gpu(S, args):
    """Plain PyTorch chain on the same GPU: rays -> samples -> gather -> composite -> MSE ->
    autograd backward into grid.grad, chunked over rays, same semantics as the fused step, and
    the fused gradient's largest difference from it, relative to its largest element."""
    from oracle import plain
    w = S.w
    dev = S.dev
    occ_used = None
    if w.occupancy is not None:
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
    times = []
    for _ in range(args.plain_steps):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    S.step()  # the fused gradient of the same batch
    torch.cuda.synchronize()
    gerr = ((grid.grad - S.grad).abs().max() / grid.grad.abs().max().clamp_min(1e-30)).item()
    log(f"plain GPU {w.interp}: {S.R / dt:.0f} rays/s ({dt * 1e3:.1f} ms/step), grad max rel diff {gerr:.3g}")
    return {"value": round(S.R / dt, 1), "unit": "rays/s", "s_per_step": round(dt, 4),
            "grad_max_rel_diff_vs_fused": gerr}


def cpu_threads():
    """CPU threads this process is granted: OMP_NUM_THREADS when the host sets it (the GPU
    hosts here grant each GPU's jobs a share of the cores that way), else the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = len(os.sched_getaffinity(0))
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(args, S):
    """The plain chain (oracle.plain, the reference's ops) on the host CPU, kind "port": fwd +
    MSE + autograd backward, 1 warm-up then the median of 3.  Headline: the first --cpu-rays rays
    of the sparse256 batch (trilinear, as the fused headline); beside it the same rays with
    nearest lookup and the 64^3 blob workload (32,768 rays x 128 trilinear samples) in full."""
    from oracle import plain
    threads = cpu_threads()
    torch.set_num_threads(threads)
    imgs = S.imgs.cpu()

    def timed(fn):
        fn()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        return float(np.median(times))

    def sparse(interp):
        w = scenes.make_workload("sparse256", n_rays=args.rays, interp=interp)
        n = min(args.cpu_rays, w.cam_ids.shape[0])
        occ = plain.dilate_cells(w.occupancy) if interp == "trilinear" else w.occupancy
        grid = w.grid.clone().requires_grad_(True)
        cam = w.cam_ids[:n].long()
        u, v = w.uv[:n, 0].contiguous(), w.uv[:n, 1].contiguous()

        def run():
            grid.grad = None
            o, d, tgt = plain.camera_rays(imgs, w.c2w, w.camera_angle_x, cam, u, v, norm_mode="gpu")
            plain.render_chunked_backward(grid, o, d, tgt, chunk_rays=n, num_samples=w.num_samples,
                                          delta_step=w.delta_step, gmin=torch.tensor(w.grid_min),
                                          points_distance=w.points_distance, interp=interp, near=w.near,
                                          div_mode="reciprocal", occupancy=occ, t_stop=w.t_stop)
        return n / timed(run), n

    tri, n = sparse("trilinear")
    near, _ = sparse("nearest")
    w = scenes.make_workload("blobs64")
    cam = w.cam_ids.long()
    R = cam.shape[0]
    grid = w.grid.clone().requires_grad_(True)
    tgt_b = scenes.random_targets(R)

    def blobs():
        grid.grad = None
        o, d, _ = plain.camera_rays(torch.zeros(w.c2w.shape[0], 1, 1, 4), w.c2w, w.camera_angle_x, cam,
                                    w.uv[:, 0].contiguous(), w.uv[:, 1].contiguous(), norm_mode="gpu")
        plain.render_chunked_backward(grid, o, d, tgt_b, chunk_rays=8192, num_samples=w.num_samples,
                                      delta_step=w.delta_step, gmin=torch.tensor(w.grid_min),
                                      points_distance=w.points_distance, interp="trilinear", near=w.near,
                                      div_mode="reciprocal")
    blobs_rps = R / timed(blobs)
    log(f"cpu baseline ({threads} threads of {os.cpu_count()}, {cpu_model()}): sparse256 trilinear {tri:.0f}, "
        f"nearest {near:.0f}, blobs64 {blobs_rps:.0f} rays/s")
    return {"value": round(tri, 1), "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"first {n} rays of the sparse256 batch (trilinear, S=256, occupancy, t_stop 1e-4), fwd + "
                      "MSE + autograd backward of the plain PyTorch chain on the CPU, median of 3 after 1 warm-up",
            "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "nearest": {"value": round(near, 1), "unit": "rays/s", "sample": f"the same {n} rays, nearest lookup"},
            "blobs64_full": {"value": round(blobs_rps, 1), "unit": "rays/s",
                             "sample": f"64^3 blob workload in full: {R} rays x 128 trilinear samples"}}


def memory_rate(dev):
    """Requests per second of the probe kernel (pxl_probe_memory_rate) for each op, lanes per
    64-byte segment and span; one JSON line per case.  A request is one segment touched by one
    wave instruction -- the unit the corner-block scatter issues two of per kept sample."""
    P = _lib.probe_lib()
    stream = _lib.stream_handle(dev)
    spans = {"4MiB": 1 << 16, "256MiB": 1 << 22, "1GiB": 1 << 24}
    buf = torch.zeros((1 << 24) * 16, device=dev)
    n_waves, iters = 1 << 16, 64
    sink = torch.zeros(n_waves, device=dev)
    for op, name in ((0, "atomic_add_f32"), (1, "store_dword"), (2, "load_dword")):
        for L in (16, 4, 1):
            for span_name, span in spans.items():
                for xcd in ((0, 1) if op == 0 and span_name != "4MiB" else (0,)):
                    def run():
                        _lib.check(P.pxl_probe_memory_rate(_lib.ptr(buf), span, op, L, iters, n_waves, xcd,
                                                           _lib.ptr(sink), stream))
                    run()
                    torch.cuda.synchronize()
                    times = []
                    for _ in range(5):
                        e0, e1 = _events(2)
                        e0.record()
                        run()
                        e1.record()
                        torch.cuda.synchronize()
                        times.append(e0.elapsed_time(e1) * 1e-3)
                    t = float(np.median(times))
                    req = n_waves * iters * (64 // L)
                    print(json.dumps({"probe": "memory_rate", "op": name, "lanes_per_segment": L,
                                      "segments_per_instruction": 64 // L, "span": span_name,
                                      "xcd_local": bool(xcd), "requests": req, "seconds": round(t, 6),
                                      "requests_per_s": round(req / t / 1e9, 3), "unit": "G requests/s",
                                      "bytes_touched_per_s": round(req * 4 * L / t / 1e9, 1)}), flush=True)


def stream_rate(dev):
    """Bytes per second of torch's own streaming kernels on 2 GiB float32 tensors: copy (read +
    write), add (two reads + write), fill (write); the mean of 10 launches after one warm-up."""
    n = (1 << 31) // 4
    a = torch.rand(n, device=dev)
    b = torch.rand(n, device=dev)
    c = torch.empty(n, device=dev)

    def rate(fn, nbytes):
        fn()
        torch.cuda.synchronize()
        e0, e1 = _events(2)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return nbytes / (e0.elapsed_time(e1) * 1e-4) / 1e12

    out = {"probe": "stream_rate", "unit": "TB/s", "bytes_per_tensor": n * 4,
           "copy": round(rate(lambda: c.copy_(a), 2 * n * 4), 3),
           "add": round(rate(lambda: torch.add(a, b, out=c), 3 * n * 4), 3),
           "fill": round(rate(lambda: c.fill_(1.0), n * 4), 3)}
    print(json.dumps(out), flush=True)


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
    if args.memory_rate:
        memory_rate(dev)
        return
    if args.stream_rate:
        stream_rate(dev)
        return
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty(n, device=dev)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}))
        return
    peaks = atomic_peak(dev) if not args.no_accounting else None
    S = Setup(args, dev, args.interp)
    main_res = measure(S, args, peaks)
    near = None
    S2 = None
    if args.workload == "sparse256" and S.w.interp == "trilinear" and not args.no_nearest:
        S2 = Setup(args, dev, "nearest", imgs=S.imgs)
        near = measure(S2, args, peaks)
        near["scatter"] = S2.scatter
    plain = cpu = None
    if not args.no_baselines:
        plain = plain_gpu(S, args)
        if S2 is not None:
            near["plain_gpu"] = plain_gpu(S2, args)
            near["speedup_vs_plain_gpu"] = round(near["value"] / near["plain_gpu"]["value"], 2)
        cpu = cpu_baseline(args, S) if args.workload == "sparse256" else None
    value = main_res.pop("value")
    metric = METRIC if args.workload == "sparse256" else METRIC.replace("256^3", f"{S.w.dims[0]}^3")
    rec = {
        "metric": metric, "value": value, "unit": "rays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main_res.pop("ms_per_step"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"float32": "f32", "bfloat16": "bf16 storage, f32 compute",
                  "float16": "f16 storage, f32 compute"}[args.dtype],
        "data": "synthetic (seeded blob field{}, random targets)".format(
            ", value-noise occupancy" if S.w.occupancy is not None else ""),
        "config": {"workload": S.w.name, "grid": list(S.w.dims), "rays": S.R, "samples_per_ray": S.w.num_samples,
                   "interp": S.w.interp, "occupancy": "~5%" if S.w.occupancy is not None else "none",
                   "t_stop": S.w.t_stop, "cameras": int(S.w.c2w.shape[0]), "image": list(S.w.img_hw),
                   "scatter": S.scatter},
        "roofline": main_res.pop("roofline", None),
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "speedup_vs_plain_gpu": round(value / plain["value"], 2) if plain else None,
    }
    rec.update(main_res)
    if near is not None:
        if cpu is not None and "nearest" in cpu:  # the nearest line's own CPU baseline
            near["cpu_baseline"] = {"value": cpu["nearest"]["value"], "unit": "rays/s", "cores": cpu["cores"],
                                    "kind": cpu["kind"], "sample": cpu["nearest"]["sample"]}
        rec["nearest"] = near
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
