"""The callers either side of the hot path: the forward-only inference render
(src/visualization.py:111-154, visulize_3d_in_2d), the NeRF-synthetic loader
(src/data_processing.py:18-29, :51-60) and the whole fit() (scripts/train.py:68-210) ending in
the grid_cells_trained.pth file, each against the reference's ops run with torch on the GPU
(the reference's default device)."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import plain
from plenoxels_amd import checkpoint
from plenoxels_amd.data import load_data, load_image_data_from_path
from plenoxels_amd.visualization import visulize_3d_in_2d
from util import random_cameras, say

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _reference_visualization(data, c2w, theta, imgs, gi, do_threshold, thr, R, S):
    """src/visualization.py:111-154 restated with the oracle's ops on whatever device the tensors
    are on; returns the float RGBA before the uint8 conversion and the uint8 image."""
    cells = data["grid"].detach().to(c2w.device).clip(0.0, 1.0)
    if do_threshold:
        alphas = cells[..., -1]
        cells[..., -1][alphas < thr] = 0.0
    pd, delta = data["param"]["points_distance"], data["param"]["delta_step"]
    res = int(np.round(np.sqrt(R)))
    R = res * res
    uni = plain.even_spread_uniforms(R, 1, device=c2w.device)
    d, _ = plain.generate_rays_batched(imgs, R, c2w, theta, uni)
    o = torch.repeat_interleave(c2w[:, :3, 3], R, 0)
    n = plain.normalize_samples(plain.sample_positions(o, d, S, delta), plain.grid_min(gi), pd, "divide")
    vals, _ = plain.nearest_values(n, cells)
    col = plain.compute_alpha_weighted_pixels(vals.reshape(1, R, S, 4)).cpu().numpy()
    img = (col * 255).round().clip(0, 255).astype(np.uint8).reshape(res, res, 4)
    return col.reshape(res, res, 4), np.transpose(img, (1, 0, 2))


@pytest.mark.parametrize("do_threshold", [False, True])
def test_visulize_3d_in_2d_matches_reference_render(do_threshold):
    say("inference render, threshold", do_threshold)
    X, pd, delta, R, S = 28, 0.07, 0.03, 48 * 48, 160
    g = torch.Generator().manual_seed(70)
    grid = torch.rand(X, X, X, 4, generator=g) * 1.3 - 0.15
    grid[..., 3] *= 0.3
    data = {"grid": grid, "grid_grad": torch.zeros_like(grid),
            "param": {"device": "cuda", "number_of_rays": R, "num_samples": S, "delta_step": delta,
                      "even_spread": True, "camera_ray": False, "points_distance": pd, "gridsize": [X, X, X]}}
    c2w = random_cameras(1, 71, radius=2.2).to(DEV)
    imgs = torch.rand(1, 32, 32, 4, generator=g).to(DEV)
    gi, _, _, _ = plain.generate_grid(X, X, X, points_distance=pd, device=DEV)
    got = visulize_3d_in_2d(data, c2w, 0.69, imgs, gi, do_threshold, 0.1, R, S, device="cuda")
    col, want = _reference_visualization(data, c2w, 0.69, imgs, gi, do_threshold, 0.1, R, S)
    assert got.dtype == np.uint8 and got.shape == want.shape == (48, 48, 4)
    # uint8-exact, except a channel whose value * 255 lies within rounding noise of a .5 boundary
    # (the kernel's compositing rounds its products in another order than torch's cumprod)
    near_tie = np.abs(np.transpose(col * 255, (1, 0, 2)) % 1.0 - 0.5) < 1e-3
    diff = np.abs(got.astype(int) - want.astype(int))
    say("  channels differing:", int((diff > 0).sum()), "near a rounding tie:", int(near_tie.sum()))
    assert diff.max() <= 1 and not (diff[~near_tie] > 0).any()
    assert (want > 0).any()  # the view sees the grid


def _write_nerf_folder(root, B, H, seed):
    """A NeRF-synthetic layout: transforms_train.json + train/r_<i>.png (RGBA)."""
    from PIL import Image
    os.makedirs(os.path.join(root, "train"))
    c2w = random_cameras(B, seed, radius=2.0)
    g = np.random.default_rng(seed)
    frames = []
    pix = []
    for i in range(B):
        img = g.integers(0, 256, size=(H, H, 4), dtype=np.uint8)
        Image.fromarray(img, mode="RGBA").save(os.path.join(root, "train", f"r_{i}.png"))
        pix.append(img)
        frames.append({"file_path": f"./train/r_{i}", "rotation": 0.0125,
                       "transform_matrix": c2w[i].tolist()})
    with open(os.path.join(root, "transforms_train.json"), "w") as f:
        json.dump({"camera_angle_x": 0.6911112070083618, "frames": frames}, f)
    return np.stack(pix), c2w


def test_loader_reads_the_nerf_layout(tmp_path):
    say("NeRF-synthetic loader")
    pix, c2w = _write_nerf_folder(str(tmp_path), 3, 12, 80)
    data, imgs = load_image_data_from_path(str(tmp_path / "train"), str(tmp_path / "transforms_train.json"))
    assert imgs.dtype == torch.float32 and tuple(imgs.shape) == (3, 12, 12, 4)
    assert torch.equal(imgs, torch.tensor(pix, dtype=torch.float) / 255)
    mats, paths, theta = load_data(data)
    assert torch.equal(mats, c2w) and paths == ["./train/r_0", "./train/r_1", "./train/r_2"]
    assert theta == 0.6911112070083618


def test_fit_writes_the_reference_checkpoint_and_matches_reference_loop(tmp_path):
    """fit() end to end: loader, the reference schedule's first steps (pooled grids), the .pth
    file; the losses and the saved grid equal the reference loop's on the same draws."""
    from plenoxels_amd.train import fit
    say("fit() on a NeRF folder")
    root = str(tmp_path)
    _write_nerf_folder(root, 3, 16, 81)
    save = os.path.join(root, "grid_cells_trained.pth")
    # 128^3: the schedule's first receptive fields (93, 91; stride 23, 22) pool it to 2^3 cells,
    # as the reference requires (a grid under 93 cells cannot be pooled by its first steps)
    kw = dict(gridsize=[128, 128, 128], points_distance_original=0.02, number_of_rays=40, num_samples=64,
              delta_step=0.05, lr=0.01, tv=1e-5, beta=5e-3, steps=6, even_spread=False)
    torch.manual_seed(82)
    tr, hist = fit(path=os.path.join(root, "train"), transform_path=os.path.join(root, "transforms_train.json"),
                   save_path=save, device="cuda", **kw)
    d = torch.load(save, weights_only=True)
    assert set(d) == {"grid", "grid_grad", "param"} and tuple(d["param"]) == checkpoint.PARAM_KEYS
    assert d["param"]["number_of_rays"] == 40 and d["param"]["num_samples"] == 64
    assert d["param"]["gridsize"] == [128, 128, 128] and d["param"]["points_distance"] == 0.02
    assert d["param"]["device"] == "cuda" and d["param"]["camera_ray"] is False
    assert torch.equal(d["grid"], tr.grid.cpu()) and torch.equal(d["grid_grad"], tr.grad_abs.cpu())
    # the same run through the reference loop restated, same uniforms (fit draws one
    # torch.rand(B, R, 2) per step on the device)
    data, imgs = load_image_data_from_path(os.path.join(root, "train"), os.path.join(root, "transforms_train.json"))
    mats, _, theta = load_data(data)
    torch.manual_seed(82)
    uniforms = [torch.rand(3, 40, 2, device=DEV) for _ in range(kw["steps"])]
    grid_w, gabs_w, hist_w = plain.reference_fit(imgs.to(DEV), mats.to(DEV), theta, gridsize=kw["gridsize"],
                                                 points_distance=0.02, num_samples=64, delta_step=0.05, lr=0.01,
                                                 tv=1e-5, beta=5e-3, uniforms=uniforms, norm_mode=None)
    assert len(hist["closs"]) == kw["steps"]
    for i, rec in enumerate(hist_w):
        assert abs(hist["closs"][i] - rec["closs"]) <= 1e-4 * rec["closs"], (i, hist["closs"][i], rec["closs"])
        assert abs(hist["betaloss"][i] - rec["betaloss"]) <= 1e-4 * abs(rec["betaloss"])
    diff = (d["grid"] - grid_w.cpu()).abs()
    say("  grid max diff", diff.max().item())
    assert (diff > 1e-5).float().mean().item() <= 1e-3
