"""Parity at the benchmark configurations' own sizes (seeded synthetic scenes).

blobs64  : 64^3 blob field, 8 cameras x 64x64 even-spread rays (32,768), S=128 in [2,6],
           trilinear -- full batch vs the plain chain on the CPU.
view800  : 128^3 blob field, one 800x800 view (640,000 rays), S=256, nearest -- fused
           gradient vs PyTorch autograd of the plain chain on the same GPU, and, on 512 voxels
           away from the clamp kinks, vs the FP64 C oracle's analytic gradient, which is itself
           checked against FP64 central finite differences (eps 1e-3).
sparse256: 256^3 with ~5% occupancy, BF16/FP16 storage -- error report of the kernel against
           the FP64 C oracle on 4096 rays (max/mean abs RGBA error, gradient relative error,
           PSNR against the targets and its change against the FP32 kernel).
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import cref
from oracle import plain
from plenoxels_amd import (CameraRays, LossMean, MarchConfig, camera_rays, occupancy_dilate, pack_occupancy, scenes,
                           train_step)
from util import grad_close, say

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
# the precision reports go to $PXL_REPORT_DIR (default: reports/ at the repository root)
OUT = os.environ.get("PXL_REPORT_DIR") or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                        "reports")


def _rays(w, n=None):
    cam = w.cam_ids if n is None else w.cam_ids[:n]
    uv = w.uv if n is None else w.uv[:n]
    imgs = torch.zeros(w.c2w.shape[0], 1, 1, 4)
    o, d, _ = plain.camera_rays(imgs, w.c2w, w.camera_angle_x, cam.long(), uv[:, 0].contiguous(), uv[:, 1].contiguous(),
                                norm_mode="gpu")
    return o.contiguous(), d.contiguous()


def test_blobs64_full_batch_vs_plain_cpu():
    say("blobs64 full batch")
    w = scenes.make_workload("blobs64")
    o, d = _rays(w)
    R = o.shape[0]
    tgt = scenes.random_targets(R)
    kw = dict(num_samples=w.num_samples, delta_step=w.delta_step, gmin=torch.tensor(w.grid_min),
              points_distance=w.points_distance, interp="trilinear", near=w.near, div_mode="reciprocal")
    g = w.grid.clone().requires_grad_(True)
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    want, _ = plain.render_chunked_backward(g, o, d, tgt, chunk_rays=4096, **kw)
    cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance, "trilinear", w.near)
    gg = torch.zeros(w.grid.shape, device=DEV)
    out = torch.empty(R, 4, device=DEV)
    train_step(w.grid.to(DEV), gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out)
    err = (out.cpu() - want).abs().max().item()
    ok, info = grad_close(gg.cpu(), g.grad)
    say("  rgba max err", err, "grad", info)
    assert err <= 1e-5
    assert ok, info
    # the camera-form kernel (in-kernel ray generation) gives the same rays
    cr = CameraRays(w.cam_ids.to(DEV), w.uv.to(DEV), w.c2w.to(DEV), w.camera_angle_x)
    oc, dc, _ = camera_rays(cr, want_targets=False)
    assert torch.equal(oc.cpu(), o) and torch.equal(dc.cpu(), d)


def test_view800_gradient_vs_autograd_and_fp64_finite_differences():
    say("view800")
    w = scenes.make_workload("view800")
    o, d = _rays(w)
    R = o.shape[0]
    tgt = scenes.random_targets(R)
    gmin = torch.tensor(w.grid_min)
    S, delta, pd = w.num_samples, w.delta_step, w.points_distance
    # fused
    cfg = MarchConfig(S, delta, w.grid_min, pd, "nearest", w.near)
    gg = torch.zeros(w.grid.shape, device=DEV)
    out = torch.empty(R, 4, device=DEV)
    train_step(w.grid.to(DEV), gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out)
    # plain chain + autograd on the same GPU (same rays; GPU division semantics)
    ga = w.grid.to(DEV).requires_grad_(True)
    want, _ = plain.render_chunked_backward(ga, o.to(DEV), d.to(DEV), tgt.to(DEV), chunk_rays=65536,
                                            num_samples=S, delta_step=delta, gmin=gmin.to(DEV), points_distance=pd,
                                            interp="nearest", near=w.near, div_mode="reciprocal")
    err = (out - want).abs().max().item()
    ok, info = grad_close(gg, ga.grad)
    say("  rgba max err vs plain GPU", err, "grad vs autograd", info)
    assert err <= 1e-5
    assert ok, info
    # FP64 central differences on 512 seeded voxels whose 4 values lie in [1e-3, 1-1e-3]
    grid64 = w.grid.double().numpy()
    cand = np.flatnonzero(((grid64 >= 1e-3) & (grid64 <= 1 - 1e-3)).all(-1).reshape(-1)
                          & (gg.abs().sum(-1).reshape(-1).cpu().numpy() > 0))
    rng = np.random.default_rng(0)
    vox = np.sort(rng.choice(cand, size=min(512, cand.size), replace=False))
    kw = dict(num_samples=S, delta_step=delta, gmin=w.grid_min, points_distance=pd, interp="nearest",
              near=w.near, div_mode="reciprocal")
    hits = cref.rays_touching(grid64, o.numpy(), d.numpy(), vox, **kw)
    tgt64 = tgt.double().numpy()
    # FP64 analytic gradient of the full loss at the chosen voxels: only rays touching them
    # contribute, so one oracle pass over the union of those rays gives it exactly
    union = np.flatnonzero(hits.any(1))
    rgba_u, _, _ = cref.render(grid64, o.numpy()[union], d.numpy()[union], **kw)
    up = 2 * (rgba_u - tgt64[union]) / (4 * R)
    _, g64, _ = cref.render(grid64, o.numpy()[union], d.numpy()[union], grad_rgba=up, **kw)
    g64 = g64.reshape(-1, 4)[vox]
    gk = gg.cpu().double().numpy().reshape(-1, 4)[vox]
    eps = 1e-3
    fd = np.zeros_like(g64)
    for q, L in enumerate(vox):
        rays = np.flatnonzero(hits[:, q])
        for c in range(4):
            gp, gm = grid64.copy().reshape(-1, 4), grid64.copy().reshape(-1, 4)
            gp[L, c] += eps
            gm[L, c] -= eps
            lp = ((cref.render(gp.reshape(grid64.shape), o.numpy()[rays], d.numpy()[rays], **kw)[0] - tgt64[rays]) ** 2).sum()
            lm = ((cref.render(gm.reshape(grid64.shape), o.numpy()[rays], d.numpy()[rays], **kw)[0] - tgt64[rays]) ** 2).sum()
            fd[q, c] = (lp - lm) / (2 * eps) / (4 * R)
    scale = np.abs(g64).max()
    floor = 1e-2 * scale  # relative error measured against max(|g|, 1% of the largest), as grad_close
    fd_vs_64 = (np.abs(fd - g64) / np.maximum(np.abs(g64), floor)).max()
    k_vs_64 = (np.abs(gk - g64) / np.maximum(np.abs(g64), floor)).max()
    k_vs_fd = (np.abs(gk - fd) / np.maximum(np.abs(fd), floor)).max()
    say("  over", len(vox), "voxels x 4: FD vs FP64 analytic", fd_vs_64, "| fused vs FP64 analytic", k_vs_64,
        "| fused vs FD", k_vs_fd)
    assert fd_vs_64 <= 1e-4   # the oracle's analytic gradient is the derivative of its loss
    assert k_vs_64 <= 1e-4    # the FP32 kernel's gradient within 1e-4 of FP64
    assert k_vs_fd <= 1e-4


@pytest.mark.parametrize("interp", ["trilinear", "nearest"])
def test_sparse256_full_batch_vs_plain_gpu(interp):
    """The benchmark's whole batch (256^3, ~5% occupancy, 2^18 rays, 256 samples, termination
    at 1e-4) through the fused step (trilinear: corner blocks; nearest: direct scatter), against
    the plain chain's autograd on the same GPU and the same rays."""
    say("sparse256 full batch vs plain chain on the GPU", interp)
    from plenoxels_amd import corner_blocks
    w = scenes.make_workload("sparse256", interp=interp)
    o, d = _rays(w)
    R = o.shape[0]
    assert R == 1 << 18
    tgt = scenes.random_targets(R).to(DEV)
    od, dd = o.to(DEV), d.to(DEV)
    occ_dev = w.occupancy.to(DEV)
    g = w.grid.to(DEV).requires_grad_(True)
    want, _ = plain.render_chunked_backward(g, od, dd, tgt, chunk_rays=32768, num_samples=w.num_samples,
                                            delta_step=w.delta_step, gmin=torch.tensor(w.grid_min, device=DEV),
                                            points_distance=w.points_distance, interp=interp, near=w.near,
                                            div_mode="reciprocal", t_stop=w.t_stop,
                                            occupancy=plain.dilate_cells(occ_dev) if interp == "trilinear" else occ_dev)
    cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance, interp, w.near, w.t_stop)
    gd = w.grid.to(DEV)
    out = torch.empty(R, 4, device=DEV)
    if interp == "trilinear":
        bits = occupancy_dilate(pack_occupancy(occ_dev), w.dims)
        gg = torch.full(w.grid.shape, 7.0, device=DEV)  # overwritten by the reduce
        train_step(gd, gg, cfg=cfg, origins=od, dirs=dd, targets=tgt, out=out, occupancy=bits,
                   blocks=corner_blocks(gd), overwrite=True)
    else:
        gg = torch.zeros(w.grid.shape, device=DEV)
        train_step(gd, gg, cfg=cfg, origins=od, dirs=dd, targets=tgt, out=out, occupancy=pack_occupancy(occ_dev))
    err = (out - want).abs().max().item()
    ok, info = grad_close(gg, g.grad)
    say("  rgba max err", err, "grad", info)
    assert err <= 1e-5
    assert ok, info


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_sparse256_half_storage_error_report(dtype):
    """Half-precision storage on the benchmark scene (4096 rays, trilinear), against the FP64 C
    oracle on the FP32 grid: per-result errors of the render, the gradient and the step's loss.
    The PSNR bar is taken where PSNR means something: the targets are the FP64 oracle's render
    plus seeded noise at 35 dB (a well-trained reconstruction), so the half-precision render's
    PSNR against them may drop by at most 0.1 dB from the FP32 render's.  The render error itself
    is bounded by the storage rounding: a pixel is a convex combination of clamped cells, so its
    error stays within a small multiple of the format's half-ulp at 1 (2^-9 BF16, 2^-12 FP16)."""
    say("sparse256 precision report", dtype)
    w = scenes.make_workload("sparse256")
    n = 4096
    o, d = _rays(w, n)
    occ_cells = plain.dilate_cells(w.occupancy)
    bits_np = plain.occupancy_to_bits(occ_cells).numpy()
    kw = dict(num_samples=w.num_samples, delta_step=w.delta_step, gmin=w.grid_min, points_distance=w.points_distance,
              interp="trilinear", near=w.near, div_mode="reciprocal")
    rgba64, _, _ = cref.render(w.grid.numpy(), o.numpy(), d.numpy(), occupancy_bits=bits_np, **kw)
    noise = np.random.default_rng(9).normal(size=rgba64.shape) * 10 ** (-35 / 20)
    tgt64 = np.clip(rgba64 + noise, 0.0, 1.0)
    tgt = torch.from_numpy(tgt64).float()
    tgt64 = tgt.double().numpy()
    up = 2 * (rgba64 - tgt64) / (4 * n)
    _, g64, _ = cref.render(w.grid.numpy(), o.numpy(), d.numpy(), occupancy_bits=bits_np, grad_rgba=up, **kw)
    bits = occupancy_dilate(pack_occupancy(w.occupancy.to(DEV)), w.dims)
    cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance, "trilinear", w.near)
    loss64 = float(np.mean((rgba64 - tgt64) ** 2))
    loss_mean = LossMean(DEV)
    res = {}
    for name, dt in (("f32", torch.float32), ("half", dtype)):
        gg = torch.zeros(w.grid.shape, device=DEV)
        out = torch.empty(n, 4, device=DEV)
        sq = torch.empty(n, device=DEV)
        train_step(w.grid.to(dt).to(DEV), gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV),
                   out=out, occupancy=bits, ray_sq_err=sq)
        loss = float(loss_mean(sq))
        e = np.abs(out.cpu().double().numpy() - rgba64)
        gk = gg.cpu().double().numpy()
        gerr = np.abs(gk - g64).max() / np.abs(g64).max()
        psnr = 10 * np.log10(1.0 / np.mean((out.cpu().double().numpy() - tgt64) ** 2))
        res[name] = {"rgba_max_abs_err": float(e.max()), "rgba_mean_abs_err": float(e.mean()),
                     "grad_rel_err": float(gerr), "psnr_vs_targets_db": float(psnr),
                     "loss": loss, "loss_rel_err_vs_fp64": abs(loss - loss64) / loss64}
    psnr64 = 10 * np.log10(1.0 / np.mean((rgba64 - tgt64) ** 2))
    res["targets"] = "FP64 render + N(0, 10^(-35/20)) noise, clipped to [0, 1]"
    res["fp64_psnr_vs_targets_db"] = float(psnr64)
    res["fp64_loss"] = loss64
    res["psnr_delta_half_vs_f32_db"] = res["f32"]["psnr_vs_targets_db"] - res["half"]["psnr_vs_targets_db"]
    say("  ", json.dumps(res))
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, f"precision_sparse256_{str(dtype).split('.')[-1]}.json"), "w") as f:
        json.dump(res, f, indent=1)
    assert 30.0 <= psnr64 <= 40.0  # the bar is taken at a realistic reconstruction quality
    assert res["f32"]["rgba_max_abs_err"] <= 1e-5
    assert res["f32"]["grad_rel_err"] <= 1e-4
    assert res["f32"]["loss_rel_err_vs_fp64"] <= 1e-5
    half_ulp = 2.0 ** -9 if dtype == torch.bfloat16 else 2.0 ** -12
    assert res["half"]["rgba_max_abs_err"] <= 4 * half_ulp
    assert res["half"]["rgba_mean_abs_err"] <= half_ulp / 4
    assert abs(res["psnr_delta_half_vs_f32_db"]) <= 0.1
