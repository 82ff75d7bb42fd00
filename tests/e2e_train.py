"""End-to-end optimisation run (the 512^3 BF16 training configuration and scaled-down versions).

Ground truth: a seeded scene of 12 coloured spheres and boxes (radii / half-extents 0.1-0.5)
in [-1.2,1.2]^3, voxelised at the training resolution (alpha 1 inside) and rendered by the FP64
C oracle (trilinear, samples in [near, near + S*delta]) to `views` training views and `heldout`
held-out views of res x res pixels from cameras on the radius-4 upper hemisphere.

Training (plenoxels_amd.train.Trainer, the reference loop scripts/train.py:79-184 on the fused
kernels): grid initialised to colour 0.5 and density 0.1, trilinear lookup, Adam lr 0.1 with FP32
state, TV weight 1e-5, no beta term, progressive growing as in the reference, `rays` random rays
per iteration (each from a uniformly drawn camera; --per-camera 1: rays // views from every camera, the
reference's batching), early termination at t_stop 1e-4.  Runs once per
storage dtype (BF16 and FP32 by default).

Reports (one JSON line, also written to --out): ground-truth render time; the fused FP32 forward
of the ground-truth grid against the oracle's images (parity at this size); per dtype the wall
time, ms/iteration, rays/s, final loss, held-out and training-view PSNR (RGB and RGBA); the
BF16-vs-FP32 held-out PSNR delta; and the plain FP32 chain (the reference ops with autograd and
torch.optim.Adam, chunked over rays) timed on full-grid steps of the same batch.

Usage: python tests/e2e_train.py [--grid 512 --iters 2000 --rays 524288 --samples 512 ...]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import cref, plain  # noqa: E402
from plenoxels_amd import CameraRays, MarchConfig, scenes  # noqa: E402
from plenoxels_amd.render import render_camera, render_rays  # noqa: E402
from plenoxels_amd.train import Trainer  # noqa: E402


def log(*a):
    print(*a, flush=True)


def psnr(a, b):
    mse = float(((a.double() - b.double()) ** 2).mean())
    return 10 * math.log10(1.0 / max(mse, 1e-30))


def view_rays(c2w, theta, res):
    """Origins and directions (CPU float32) of every pixel of every view, view-major."""
    uv = scenes.pixel_uv(res, res)
    B = c2w.shape[0]
    cam = torch.repeat_interleave(torch.arange(B), res * res)
    uvs = uv.repeat(B, 1)
    o, d, _ = plain.camera_rays(torch.zeros(B, 1, 1, 4), c2w, theta, cam, uvs[:, 0].contiguous(), uvs[:, 1].contiguous(),
                                norm_mode="gpu")
    return o.contiguous(), d.contiguous()


def render_views_fused(grid, c2w, theta, res, cfg, dev):
    """(B,res,res,4) float32 images of `grid` from the fused forward kernel."""
    uv = scenes.pixel_uv(res, res).to(dev)
    out = []
    c2w_d = c2w.to(dev).contiguous()
    for b in range(c2w.shape[0]):
        cam = torch.full((res * res,), b, dtype=torch.int32, device=dev)
        cr = CameraRays(cam, uv, c2w_d, theta)
        out.append(render_camera(grid, cr, cfg).reshape(res, res, 4))
    return torch.stack(out)


def plain_step_time(grid0, imgs, c2w, theta, R_per_cam, S, delta, near, pd, gmin, lr, tv, steps, chunk, dev):
    """Seconds per full-grid step of the plain FP32 chain (reference ops, autograd, Adam, TV)."""
    grid = grid0.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([grid], lr=lr)
    B = c2w.shape[0]
    kw = dict(num_samples=S, delta_step=delta, gmin=torch.tensor(gmin, device=dev), points_distance=pd,
              interp="trilinear", near=near, div_mode="reciprocal")
    times = []
    for s in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.time()
        uv = torch.rand(B, R_per_cam, 2, device=dev)
        cam = torch.repeat_interleave(torch.arange(B, device=dev), R_per_cam)
        uvf = uv.reshape(-1, 2)
        o, d, tgt = plain.camera_rays(imgs, c2w.to(dev), theta, cam, uvf[:, 0].contiguous(), uvf[:, 1].contiguous())
        opt.zero_grad()
        plain.render_chunked_backward(grid, o, d, tgt, chunk_rays=chunk, **kw)
        t = tv * plain.tv_loss(grid)
        t.backward()
        opt.step()
        torch.cuda.synchronize()
        times.append(time.time() - t0)
        log(f"  plain FP32 step {s}: {times[-1]:.2f} s")
    return float(np.mean(times[1:])) if steps > 0 else times[0]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--views", type=int, default=100)
    ap.add_argument("--heldout", type=int, default=20)
    ap.add_argument("--res", type=int, default=800)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--rays", type=int, default=1 << 19)
    ap.add_argument("--per-camera", type=int, default=0,
                    help="1: rays // views rays from every camera per iteration (the reference's batching); "
                         "0: exactly --rays rays per iteration, each from a uniformly drawn camera")
    ap.add_argument("--samples", type=int, default=512)
    ap.add_argument("--near", type=float, default=2.0)
    ap.add_argument("--lr", type=float, default=0.1)
    ap.add_argument("--tv", type=float, default=1e-5)
    ap.add_argument("--t-stop", type=float, default=1e-4)
    ap.add_argument("--dtypes", default="bfloat16,float32")
    ap.add_argument("--corner-blocks", type=int, default=0, help="Trainer(corner_blocks=...)")
    ap.add_argument("--skip-empty", type=int, default=1, help="Trainer(skip_empty=...)")
    ap.add_argument("--progressive", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--plain-chunk", type=int, default=4096)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--log-every", type=int, default=100)
    ap.add_argument("--out", default=os.path.join(os.environ.get("PXL_REPORT_DIR") or os.path.join(ROOT, "reports"),
                                                  "e2e.json"))
    a = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    X, S = a.grid, a.samples
    delta = 4.0 / S
    pd = scenes.points_distance_for(X)
    gmin = scenes.grid_min(X, X, X, pd)
    theta = 800 / 1111
    res = {"config": {"grid": X, "views": a.views, "heldout": a.heldout, "res": a.res, "iters": a.iters,
                      "rays_per_iter": (a.rays // a.views) * a.views if a.per_camera else a.rays,
                      "ray_batching": "rays // views per camera" if a.per_camera else "uniformly drawn camera per ray",
                      "samples": S, "delta": delta, "near": a.near,
                      "lr": a.lr, "tv": a.tv, "beta": 0.0, "t_stop": a.t_stop, "interp": "trilinear",
                      "progressive_growing": bool(a.progressive), "init": "colour 0.5, density 0.1",
                      "corner_blocks": bool(a.corner_blocks), "skip_empty": bool(a.skip_empty)}}

    # ground truth
    t0 = time.time()
    prims = scenes.primitive_scene(12, seed=0)
    gt = scenes.voxelize_primitives(prims, X)
    occ_bits = plain.occupancy_to_bits(plain.dilate_cells(gt[..., 3] > 0)).numpy()
    c2w_train = scenes.sphere_cameras(a.views, seed=1, hemisphere=True)
    c2w_held = scenes.sphere_cameras(a.heldout, seed=7, hemisphere=True)
    log(f"scene voxelised at {X}^3 ({float((gt[..., 3] > 0).float().mean()) * 100:.2f}% occupied) in {time.time() - t0:.1f} s")
    kw = dict(num_samples=S, delta_step=delta, gmin=gmin, points_distance=pd, interp="trilinear", near=a.near,
              div_mode="reciprocal", t_stop=a.t_stop)
    gt64 = gt.double().numpy()

    def gt_images(c2w, name):
        o, d = view_rays(c2w, theta, a.res)
        t1 = time.time()
        last = [0.0]

        def prog(done, total):
            if time.time() - last[0] > 20:
                last[0] = time.time()
                log(f"  {name}: {done}/{total} rays, {time.time() - t1:.0f} s")
        rgba = cref.render_parallel(gt64, o.numpy(), d.numpy(), threads=a.threads, chunk=8192,
                                    occupancy_bits=occ_bits, progress=prog, **kw)
        return torch.from_numpy(rgba).float().reshape(c2w.shape[0], a.res, a.res, 4)

    t0 = time.time()
    imgs_train = gt_images(c2w_train, "train views")
    imgs_held = gt_images(c2w_held, "held-out views")
    res["gt_render_s"] = round(time.time() - t0, 1)
    log(f"ground truth: {a.views}+{a.heldout} views rendered by the FP64 oracle in {res['gt_render_s']} s")

    # parity at this size: fused FP32 forward of the ground-truth grid vs the oracle's images
    cfg_full = MarchConfig(S, delta, gmin, pd, "trilinear", a.near, a.t_stop)
    gt_dev = gt.to(dev)
    fused_gt = render_views_fused(gt_dev, c2w_held, theta, a.res, cfg_full, dev)
    err = (fused_gt.cpu() - imgs_held).abs()
    res["gt_forward_parity"] = {"max_abs_err": float(err.max()), "mean_abs_err": float(err.mean())}
    log("fused forward of the ground-truth grid vs oracle:", res["gt_forward_parity"])
    del gt_dev, fused_gt

    imgs_dev = imgs_train.to(dev)
    held_dev = imgs_held.to(dev)
    init = torch.tensor([0.5, 0.5, 0.5, 0.1])
    # a fixed batch of training rays on which the final loss is evaluated, by the fused kernel and
    # by the FP64 oracle on the trained FP32 grid
    ge = torch.Generator().manual_seed(11)
    n_ev = min(8192, a.rays)
    ev_cam = torch.randint(0, a.views, (n_ev,), generator=ge)
    ev_uv = torch.rand(n_ev, 2, generator=ge)
    ev_o, ev_d, ev_t = plain.camera_rays(imgs_train, c2w_train, theta, ev_cam, ev_uv[:, 0].contiguous(),
                                         ev_uv[:, 1].contiguous(), norm_mode="gpu")
    R_per_cam = a.rays // a.views
    rays_per_iter = R_per_cam * a.views if a.per_camera else a.rays
    res["fused"] = {}
    for name in a.dtypes.split(","):
        dt = getattr(torch, name)
        tr = Trainer(imgs_dev, c2w_train, theta, gridsize=[X] * 3, points_distance=pd, number_of_rays=R_per_cam,
                     total_rays=None if a.per_camera else a.rays,
                     num_samples=S, delta_step=delta, lr=a.lr, tv=a.tv, beta=0.0, device=dev,
                     progressive_growing=bool(a.progressive), interp="trilinear", storage_dtype=dt,
                     init_grid=init.to(dev), near=a.near, t_stop=a.t_stop, corner_blocks=bool(a.corner_blocks),
                     skip_empty=bool(a.skip_empty))
        rgrid = lambda: tr.grid if tr.render_copy is None else tr.render_copy  # noqa: E731
        p0 = psnr(render_views_fused(rgrid(), c2w_held, theta, a.res, cfg_full, dev)[..., :3], held_dev[..., :3])
        # every dtype trains on the same ray draws (torch.rand on the device, re-seeded), so the
        # runs differ only in the storage precision
        torch.manual_seed(1234)
        torch.cuda.synchronize()
        t0 = time.time()
        hist = []
        for i in range(a.iters):
            out = tr.step(i)
            if a.log_every and (i % a.log_every == 0 or i == a.iters - 1):
                loss = float(out["loss"])
                hist.append([i, loss])
                log(f"  {name} iter {i}: loss {loss:.6f} ({time.time() - t0:.1f} s)")
        torch.cuda.synchronize()
        wall = time.time() - t0
        held = render_views_fused(rgrid(), c2w_held, theta, a.res, cfg_full, dev)
        train_v = render_views_fused(rgrid(), c2w_train[:5], theta, a.res, cfg_full, dev)
        with torch.no_grad():
            out_ev = render_rays(rgrid(), ev_o.to(dev), ev_d.to(dev), num_samples=S, delta_step=delta, grid_min=gmin,
                                 points_distance=pd, interp="trilinear", near=a.near, t_stop=a.t_stop).cpu()
        loss_fused = float(((out_ev.double() - ev_t.double()) ** 2).mean())
        rgba64 = cref.render_parallel(tr.grid.double().cpu().numpy(), ev_o.numpy(), ev_d.numpy(), threads=a.threads,
                                      chunk=4096, **kw)
        loss64 = float(np.mean((rgba64 - ev_t.double().numpy()) ** 2))
        r = {"wall_s": round(wall, 2), "ms_per_iter": round(wall / a.iters * 1e3, 3),
             "rays_per_s": round(a.iters * rays_per_iter / wall, 1), "rays_per_iter": rays_per_iter, "final_loss": hist[-1][1] if hist else None,
             "heldout_psnr_rgb_init": round(p0, 3),
             "heldout_psnr_rgb": round(psnr(held[..., :3], held_dev[..., :3]), 3),
             "heldout_psnr_rgba": round(psnr(held, held_dev), 3),
             "train5_psnr_rgb": round(psnr(train_v[..., :3], imgs_dev[:5, ..., :3]), 3),
             "final_loss_eval": {"rays": n_ev, "fused": loss_fused, "fp64_of_fp32_master": loss64,
                                 "rel_err": abs(loss_fused - loss64) / loss64},
             "loss_history": hist}
        res["fused"][name] = r
        log(f"{name}: " + json.dumps({k: v for k, v in r.items() if k != "loss_history"}))
        del tr, held, train_v
        torch.cuda.empty_cache()
    f = res["fused"]
    if "bfloat16" in f and "float32" in f:
        res["heldout_psnr_delta_bf16_vs_fp32_db"] = round(f["bfloat16"]["heldout_psnr_rgb"] - f["float32"]["heldout_psnr_rgb"], 3)

    if a.plain_steps > 0:
        g0 = init.to(dev).expand(X, X, X, 4).contiguous()
        s = plain_step_time(g0, imgs_dev, c2w_train, theta, R_per_cam, S, delta, a.near, pd, gmin, a.lr, a.tv,
                            a.plain_steps, a.plain_chunk, dev)
        res["plain_fp32"] = {"s_per_step": round(s, 3), "rays_per_s": round(R_per_cam * a.views / s, 1),
                             "steps_timed": a.plain_steps, "note": "full-grid steps (no pooling), same batch size"}
        log("plain FP32:", res["plain_fp32"])
    line = json.dumps(res)
    print(line, flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as fh:
        fh.write(line + "\n")
    return res


if __name__ == "__main__":
    main()
