"""The reference training loop (scripts/train.py:16-236) on the fused kernels.

Per step i (scripts/train.py:104-191):
  * progressive growing: for the first 5*46 steps the grid rendered is
    average_pool3d_grid(grid, k, stride=max(1, k//4)) with k = 93, 91, ..., 3 (5 steps each),
    its geometry the full grid's coordinates from index k//2 with points distance pd*stride;
  * B*R rays (R per camera, uniforms from torch.rand(B,R,2) on the device, or an even
    spread), S samples each, nearest voxel of grid.clip(0,1), compositing, mean-MSE loss
    against the image pixels, plus beta * mean per-sample log-odds term (beta > 0) and
    tv * tv_loss(grid) (tv > 0 and k < 19);
  * Adam step on the full grid and |grad| accumulation.
The render, loss gradients and backward are one pxl_train_step pass; pooling, TV and Adam
are their own kernels.  Loss values stay on the device (one host sync per logged step).
"""
import argparse
import math

import numpy as np
import torch

from . import _lib
from .checkpoint import save_grid
from .grid_functions import average_pool3d_grid, pool_backward_
from .optim import adam_step_, adam_tv_step_, tv_backward_, tv_sumsq
from .render import (CameraRays, CornerBlockPair, LossMean, MarchConfig, alpha_occupancy, beta_term_zero, blocks_reduce,
                     fill_, train_step)

RECEPTIVE_FIELDS = [i for i in range(93, 2, -2) if i % 2 != 0]  # scripts/train.py:91
STEPS_PER_FIELD = 5                                              # scripts/train.py:92


def grid_coordinate(i, size, points_distance):
    """float32 world coordinate of index i along an axis of `size` cells
    ((i - ceil(size/2) + 1) * pd, each op rounded to float32; src/grid_functions.py:205-209)."""
    c = np.float32(i) - np.float32(math.ceil(size / 2))
    c = c + np.float32(1)
    return float(c * np.float32(points_distance))


# Pooled trilinear grids up to this many cells scatter through corner blocks (two 64-byte
# entries per cell and buffer: 128 MiB each at the limit).
POOLED_BLOCKS_MAX_CELLS = 128 ** 3


def receptive_field(i):
    """Receptive-field size of step i (scripts/train.py:108), 0 after the schedule.  It gates
    pooling (with progressive growing, scripts/train.py:110) and the TV term (k < 19,
    scripts/train.py:163) -- the latter also when progressive growing is off."""
    return RECEPTIVE_FIELDS[i // STEPS_PER_FIELD] if i // STEPS_PER_FIELD < len(RECEPTIVE_FIELDS) else 0


class Trainer:
    """State of one fit: FP32 master grid, Adam moments, |grad| accumulator, rays setup.

    storage_dtype bfloat16/float16 renders the full-resolution steps from a reduced-precision
    copy of the grid that Adam writes in the same pass (FP32 master and moments).
    uniforms_fn(step) -> (B,R,2) uniforms overrides torch.rand (tests feed fixed ones);
    init_grid replaces the all-zero initial grid; near (first sample at near + delta) and
    t_stop (early ray termination) are the march extensions, both 0 in the reference.
    corner_blocks (trilinear only) routes the gradient through corner blocks (a 64-byte entry
    per cell) and their reduce (pxl_train_step_blocks), which replaces the zero-fill; it costs
    two such buffers, 128 B per cell of device memory, and a dense reduce pass per step.
    Pooled trilinear grids of up to POOLED_BLOCKS_MAX_CELLS cells use corner blocks in any case
    (reduced and cleared densely, so the occupancy may change from step to step).
    skip_empty (default on) renders each step with the occupancy of the cells whose raw alpha
    is >= 0 (pxl_alpha_occupancy): the samples it skips are exactly zero in the reference chain
    (clamped alpha 0, zero clamp gradient), so finite results are unchanged while the empty
    space the training drives below zero costs nothing.
    total_rays (extension): draw that many rays per step, each from a camera chosen uniformly
    at random, instead of number_of_rays per camera (the reference's scheme) -- for batch sizes
    the camera count does not divide, e.g. 2^19 rays over 100 views; uniforms_fn(step) then
    returns (cam_ids (N,) int32, uv (N,2)).
    """

    def __init__(self, imgs, transform_matricies, camera_angle_x, *, gridsize, points_distance,
                 number_of_rays, num_samples, delta_step, lr, tv, beta, even_spread=False, device="cuda",
                 progressive_growing=True, interp="nearest", storage_dtype=torch.float32, uniforms_fn=None,
                 init_grid=None, near=0.0, t_stop=0.0, corner_blocks=False, skip_empty=True, total_rays=None):
        self.dev = torch.device(device)
        self.imgs = imgs.to(self.dev).float().contiguous()
        self.c2w = transform_matricies.to(self.dev).float().contiguous()
        self.theta = float(camera_angle_x)
        self.dims = tuple(int(x) for x in gridsize)
        self.pd = float(points_distance)
        self.number_of_rays = number_of_rays  # as given: the .pth records it (scripts/train.py:201)
        self.R = int(np.round(np.sqrt(number_of_rays)) ** 2) if even_spread else int(number_of_rays)
        self.S = int(num_samples)
        self.delta = float(delta_step)
        self.lr, self.tv, self.beta = float(lr), float(tv), float(beta)
        self.even_spread = bool(even_spread)
        self.progressive = bool(progressive_growing)
        self.interp = interp
        self.uniforms_fn = uniforms_fn
        self.near, self.t_stop = float(near), float(t_stop)
        if corner_blocks and interp != "trilinear":
            raise ValueError("corner_blocks needs interp='trilinear'")
        self.use_blocks = bool(corner_blocks)
        self.skip_empty = bool(skip_empty)
        if self.use_blocks and self.skip_empty:
            raise ValueError("corner_blocks needs a fixed occupancy; use skip_empty=False with it")
        self._blocks = {}  # grid shape -> CornerBlockPair (each reduce zeroes the other buffer)
        self.B = self.imgs.shape[0]
        X, Y, Z = self.dims
        self.grid = torch.zeros((X, Y, Z, 4), dtype=torch.float32, device=self.dev)
        if init_grid is not None:  # the reference starts from zeros (src/grid_functions.py:214)
            self.grid.copy_(init_grid)
        self.grad = torch.zeros_like(self.grid)
        self._grad_zero = True  # self.grad is all zero (the fused TV + Adam pass can leave it so)
        self.exp_avg = torch.zeros_like(self.grid)
        self.exp_avg_sq = torch.zeros_like(self.grid)
        self.grad_abs = torch.zeros_like(self.grid)
        self.render_copy = None if storage_dtype == torch.float32 else self.grid.to(storage_dtype)
        self.step_count = 0
        self.total_rays = None if total_rays is None else int(total_rays)
        if self.total_rays is not None and self.even_spread:
            raise ValueError("total_rays draws random rays; it does not combine with even_spread")
        n = self.B * self.R if self.total_rays is None else self.total_rays
        self.n_rays = n
        self.cam_ids = torch.repeat_interleave(torch.arange(self.B, dtype=torch.int32, device=self.dev), self.R)
        self.sq = torch.empty(n, device=self.dev)
        self.loss_mean = LossMean(self.dev)
        self.bsum = torch.empty(n, device=self.dev)
        self.tvs = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self._grid_next = None  # second FP32 grid of the fused TV + Adam update (allocated on first use)
        self._pooled_pair = None  # (pooled dims, CornerBlockPair) of the current receptive field
        if self.even_spread:
            m = int(round(math.sqrt(self.R)))
            ls = torch.linspace(0, end=1, steps=m, device=self.dev)
            self.even_uv = torch.cartesian_prod(ls, ls).repeat([self.B, 1, 1])

    def _uniforms(self, i):
        if self.uniforms_fn is not None:
            return self.uniforms_fn(i).to(self.dev).float()
        if self.even_spread:
            return self.even_uv
        return torch.rand(self.B, self.R, 2, device=self.dev)

    def _rays(self, i):
        """(cam_ids (N,) int32, uv (N,2) float32) of step i."""
        if self.total_rays is None:
            return self.cam_ids, self._uniforms(i).reshape(self.B * self.R, 2).contiguous()
        if self.uniforms_fn is not None:
            cam, uv = self.uniforms_fn(i)
            return cam.to(self.dev).to(torch.int32).contiguous(), uv.to(self.dev).float().contiguous()
        cam = torch.randint(0, self.B, (self.total_rays,), device=self.dev, dtype=torch.int32)
        return cam, torch.rand(self.total_rays, 2, device=self.dev)

    def step(self, i):
        """One training step; returns device scalars {"closs", "tvloss", "betaloss", "loss"}."""
        X, Y, Z = self.dims
        k = receptive_field(i)
        pooled = k > 1 and self.progressive
        if pooled:
            stride = max(1, k // 4)
            start = int(k / 2)
            gmin = tuple(grid_coordinate(start, s, self.pd) for s in self.dims)
            pd = self.pd * stride
            grid = average_pool3d_grid(self.grid, receptive_field_size=k, stride=stride).detach()
            # small pooled grids go through corner blocks (half the atomic requests of the direct
            # scatter, which bounds these steps while every sample is still kept), reduced and
            # cleared densely so the per-step occupancy may change
            pooled_blocks = (not self.use_blocks and self.interp == "trilinear"
                             and grid.shape[0] * grid.shape[1] * grid.shape[2] <= POOLED_BLOCKS_MAX_CELLS)
            grad = torch.empty_like(grid) if self.use_blocks or pooled_blocks else torch.zeros_like(grid)
        else:
            pooled_blocks = False
            gmin = tuple(grid_coordinate(0, s, self.pd) for s in self.dims)
            pd = self.pd
            grid = self.grid if self.render_copy is None else self.render_copy
            grad = self.grad
            if not self.use_blocks and not self._grad_zero:
                fill_(grad, 0.0)
            self._grad_zero = False
        cam_ids, uv = self._rays(i)
        cam = CameraRays(cam_ids, uv, self.c2w, self.theta, self.imgs)
        cfg = MarchConfig(self.S, self.delta, gmin, pd, self.interp, self.near, self.t_stop)
        n_rays = self.n_rays
        pair = None
        if self.use_blocks:
            key = tuple(grid.shape[:3])
            if key not in self._blocks:
                self._blocks[key] = CornerBlockPair(grid)
            pair = self._blocks[key]
        occ = alpha_occupancy(grid, self.interp) if self.skip_empty else None
        if pooled_blocks:
            key = tuple(grid.shape[:3])
            if self._pooled_pair is None or self._pooled_pair[0] != key:
                self._pooled_pair = (key, CornerBlockPair(grid))
            pair = self._pooled_pair[1]
            train_step(grid, None, cfg=cfg, camera=cam, occupancy=occ, ray_sq_err=self.sq,
                       beta=self.beta if self.beta > 0 else 0.0, ray_beta=self.bsum if self.beta > 0 else None,
                       blocks=pair.blocks)
            blocks_reduce(grid, pair.blocks, grad, occupancy=None, accumulate=False, clear=pair.other)
        else:
            train_step(grid, grad, cfg=cfg, camera=cam, occupancy=occ, ray_sq_err=self.sq,
                       beta=self.beta if self.beta > 0 else 0.0,
                       ray_beta=self.bsum if self.beta > 0 else None, blocks=pair.blocks if pair else None,
                       overwrite=True, clear=pair.other if pair else None)
        if pair:
            pair.flip()
        closs = self.loss_mean(self.sq, n_rays)
        out = {"closs": closs}
        loss = closs
        # full-grid steps with the TV term fold its backward into the Adam pass (pxl_adam_tv_step,
        # the same bits), which writes the updated grid into a second buffer
        fused_tv = self.tv > 0 and k < 19 and not pooled
        if self.tv > 0 and k < 19:
            tv_grid = grid if grid.dtype == torch.float32 else self.grid
            self.tvs.zero_()
            tv_sumsq(tv_grid, out=self.tvs)
            if not fused_tv:
                tv_backward_(tv_grid, self.tvs, self.tv, grad)
            tvl = (self.tv * torch.sqrt(self.tvs)).float().reshape(())
            out["tvloss"] = tvl
            loss = loss + tvl
        if self.beta > 0:
            n_s = n_rays * self.S
            bl = self.beta * (self.bsum.sum() + n_s * beta_term_zero()) / n_s
            out["betaloss"] = bl
            loss = loss + bl
        if pooled:  # the full grid's gradient, every cell written
            pool_backward_(grad, (X, Y, Z), k, max(1, k // 4), self.grad, accumulate=False)
            self._grad_zero = False
        self.step_count += 1
        if fused_tv:
            if self._grid_next is None:
                self._grid_next = torch.empty_like(self.grid)
            # the pass leaves the gradient zero for the next step (no separate zero-fill)
            adam_tv_step_(self.grid, self._grid_next, self.grad, self.exp_avg, self.exp_avg_sq, self.step_count,
                          tv_sumsq=self.tvs, tv_weight=self.tv, lr=self.lr, grad_abs=self.grad_abs,
                          half_out=self.render_copy, zero_grad=True)
            self._grad_zero = True
            self.grid, self._grid_next = self._grid_next, self.grid
        else:
            adam_step_(self.grid, self.grad, self.exp_avg, self.exp_avg_sq, self.step_count, lr=self.lr,
                       grad_abs=self.grad_abs, half_out=self.render_copy)
        out["loss"] = loss
        return out

    def save(self, save_path, device_name):
        save_grid(save_path, self.grid, self.grad_abs, device=device_name, number_of_rays=self.number_of_rays,
                  num_samples=self.S, delta_step=self.delta, even_spread=self.even_spread, camera_ray=False,
                  points_distance=self.pd, gridsize=list(self.dims))


def fit(gridsize, points_distance_original, number_of_rays, num_samples, delta_step, lr, tv, beta, steps,
        even_spread, path, transform_path, save_path, device, progressive_growing=True, log_every=1):
    """scripts/train.py:68-210 on the fused kernels: trains on the NeRF-synthetic images at
    `path` / `transform_path` and writes the grid_cells_trained.pth dict to save_path."""
    from tqdm import tqdm

    from .data import load_data, load_image_data_from_path
    data, imgs = load_image_data_from_path(path, transform_path)
    transform_matricies, _, camera_angle_x = load_data(data)
    tr = Trainer(imgs, transform_matricies, camera_angle_x, gridsize=gridsize,
                 points_distance=points_distance_original, number_of_rays=number_of_rays,
                 num_samples=num_samples, delta_step=delta_step, lr=lr, tv=tv, beta=beta,
                 even_spread=even_spread, device=device, progressive_growing=progressive_growing)
    if steps < STEPS_PER_FIELD * len(RECEPTIVE_FIELDS):
        print(f"not enough steps to for frequency regularization {steps=} < "
              f"{(STEPS_PER_FIELD * len(RECEPTIVE_FIELDS))=}")
    hist = {"closs": [], "tvloss": [], "betaloss": []}
    pbar = tqdm(total=steps, desc="Processing items")
    for i in range(steps):
        out = tr.step(i)
        if log_every and (i % log_every == 0 or i == steps - 1):
            desc = [f"grid size: {tuple(tr.dims)}, kernel size: {receptive_field(i)}"]
            for key in ("closs", "tvloss", "betaloss"):
                if key in out:
                    v = float(out[key])
                    hist[key].append(v)
                    desc.append(f"{key}:{v:10.6f}")
            pbar.set_description(" ".join(desc))
        pbar.update(1)
    pbar.close()
    tr.save(save_path, device)
    return tr, hist


def main(args):
    """scripts/train.py:16-41 (argument handling)."""
    device = args.device
    if device == "auto":
        device = "cuda:0" if torch.cuda.is_available() else "cpu"
    gridsize = [args.gridsize] * 3
    if -1 in gridsize:
        gridsize = [int(n) for n in args.grid_dim]
    return fit(gridsize, args.points_distance, args.ray_num, args.sample_num, args.delta_step, args.lr, args.tv,
               args.beta, args.steps, args.even_spread, args.path, args.transform_path, args.save_path, device)


def build_parser():
    """The reference CLI flags (scripts/train.py:215-233), same names and defaults."""
    parser = argparse.ArgumentParser(description='train a voxel grid on NeRF-synthetic images')
    parser.add_argument('--device', default="auto", help='device to run on (cuda, auto)')
    parser.add_argument('--save_path', default="src/grid_cells_trained.pth", help='path to trained grid')
    parser.add_argument('--ray_num', default=128, type=int, help='amount of rays from each camera during training')
    parser.add_argument('--sample_num', default=600, type=int, help='number of samples along a ray')
    parser.add_argument('--path', default="data/mic/train", help='path to dataset')
    parser.add_argument('--transform_path', default="data/mic/transforms_train.json", help='path to json file')
    parser.add_argument('--gridsize', default=-1, type=int, help='specify cubic size of grid as a single number')
    parser.add_argument('--grid_dim', default=[256, 256, 256], nargs='+', help='size of grid, write as "x y z"')
    parser.add_argument('--points_distance', default=0.0125, type=float, help='delta distance between points in grid')
    parser.add_argument('--delta_step', default=0.0125, type=float, help='distance between samples along a ray')
    parser.add_argument('--lr', default=0.0075, type=float, help='learning rate')
    parser.add_argument('--tv', default=1e-5, type=float, help='tv loss strength (set to 0 to disable)')
    parser.add_argument('--beta', default=5e-3, type=float, help='beta loss strength (set to 0 to disable)')
    parser.add_argument('--steps', default=500, type=int, help='training steps')
    parser.add_argument('--progressive_growing', type=bool, default=True, help="whether to grow the grid while training")
    parser.add_argument('--even_spread', default=False, type=bool, help='whether to set spread rays evenly')
    return parser


if __name__ == "__main__":
    main(build_parser().parse_args())
