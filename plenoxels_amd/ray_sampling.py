"""Drop-in counterparts of the reference's src/ray_sampling.py hot-path functions.

Same names, arguments, shapes and dtypes as the reference; the arithmetic runs in libpxl's
HIP kernels (include/pxl.h).  Differences, all at the reference's failure points:
  * camera_ray=True raises NotImplementedError: the reference branch is shape-broken
    (src/ray_sampling.py:150-151) and no caller uses it;
  * image indices from uniforms outside [0,1] (or non-square images) are clamped into the
    image where the reference would raise IndexError;
  * images must have 4 channels (RGBA, what the reference's loss compares against).
Device semantics: the reference's default device is 'cuda', and these drop-ins reproduce what
its ops compute there -- directions normalised with torch's GPU 3-norm order, division by
points_distance as ATen's GPU kernel does it (a multiply by the float32 reciprocal).
Gradients: compute_alpha_weighted_pixels is differentiable w.r.t. its samples, as in the
reference's training chain; the ray/sample geometry outputs carry no gradient.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .render import CameraRays, camera_rays


def _dev(t):
    if not t.is_cuda:
        raise ValueError("plenoxels_amd runs on a HIP device; move the tensors there first")
    return t.device


def normalize_samples_for_indecies(grid_indices, samples_interval, points_distance):
    """src/ray_sampling.py:12-13: (samples - grid_indices.min(0)[0]) / points_distance.

    As on a GPU in the reference, the division by the Python float is a multiplication by
    float32(1/float32(points_distance)).
    """
    dev = _dev(samples_interval)
    gi = grid_indices.float().contiguous()
    p = samples_interval.float().contiguous()
    if gi.dim() != 2 or gi.shape[1] != 3 or p.dim() != 2 or p.shape[1] != 3:
        raise ValueError("grid_indices and samples_interval must be (M,3) and (N,3)")
    L = _lib.lib()
    s = _lib.stream_handle(dev)
    m3 = torch.empty(3, dtype=torch.float32, device=dev)
    _lib.check(L.pxl_coord_min(_lib.ptr(gi), gi.shape[0], _lib.ptr(m3), s))
    out = torch.empty_like(p)
    pd = np.float32(points_distance)
    _lib.check(L.pxl_normalize_points(_lib.ptr(p), p.shape[0], _lib.ptr(m3), float(pd),
                                      float(np.float32(1.0) / pd), 0, _lib.ptr(out), s))
    return out


def generate_rays_batched(imgs, number_of_rays, transform_matricies, camera_angle_x, even_spread=False,
                          device='cuda'):
    """src/ray_sampling.py:195-264: (ray directions (B*R,3), target pixels (B*R,4)).

    The uniforms are drawn with the reference's own torch calls (torch.rand(B,R,2) on
    `device`, or cartesian_prod(linspace(0,1,sqrt R)) for even_spread), so a seeded run
    draws the same rays; directions and pixel gathers run in the pxl_camera_rays kernel.
    """
    B = transform_matricies.shape[0]
    if even_spread:
        n = int(np.round(np.sqrt(number_of_rays)))
        if n * n != number_of_rays:
            raise ValueError("even_spread needs a square number_of_rays (the reference reshapes "
                             "n*n rays into number_of_rays)")
        ls = torch.linspace(0, end=1, steps=n, device=device)
        uv = torch.cartesian_prod(ls, ls).repeat([imgs.shape[0], 1, 1])
    else:
        uv = torch.rand(B, number_of_rays, 2, device=device)
    return rays_from_uniforms(imgs, transform_matricies, camera_angle_x, uv)


def rays_from_uniforms(imgs, transform_matricies, camera_angle_x, uv, norm_mode="gpu"):
    """generate_rays_batched on given uniforms uv (B,R,2): (dirs (B*R,3), pixels (B*R,4)).
    norm_mode: "gpu" (torch's GPU 3-norm, the reference on 'cuda') or "cpu" (its CPU norm)."""
    B, R = uv.shape[0], uv.shape[1]
    dev = _dev(transform_matricies)
    cam = CameraRays(torch.repeat_interleave(torch.arange(B, dtype=torch.int32, device=dev), R),
                     uv.reshape(B * R, 2).float().contiguous(),
                     transform_matricies.float().contiguous(), float(camera_angle_x),
                     imgs.float().contiguous(), norm_mode=norm_mode)
    _, d, tgt = camera_rays(cam, want_targets=True)
    return d, tgt


def sample_camera_rays_batched(transform_matrices, camera_angle_x, imgs, number_of_rays, num_samples,
                               delta_step, even_spread, camera_ray, device='cuda'):
    """src/ray_sampling.py:128-169: (samples_interval (B*R*S,3), pixels_to_rays (B*R,4),
    camera_positions (B,3), ray_directions (B*R,3)); samples o + d*delta*k, k = 1..S,
    sample-fastest, then ray, then camera."""
    if even_spread:
        number_of_rays = int(np.round(np.sqrt(number_of_rays)) ** 2)
    if camera_ray:
        raise NotImplementedError("camera_ray=True is shape-broken in the reference "
                                  "(src/ray_sampling.py:150-151) and unused by its callers")
    dirs, pixels = generate_rays_batched(imgs, number_of_rays, transform_matrices, camera_angle_x,
                                         even_spread=even_spread, device=device)
    camera_positions = transform_matrices[:, :3, 3]
    B = imgs.shape[0]
    n_rays = B * number_of_rays
    out = torch.empty((n_rays * num_samples, 3), dtype=torch.float32, device=dirs.device)
    cp = camera_positions.float().contiguous()
    _lib.check(_lib.lib().pxl_sample_points(_lib.ptr(cp), number_of_rays, _lib.ptr(dirs), n_rays,
                                            int(num_samples), float(np.float32(delta_step)), 0.0,
                                            _lib.ptr(out), _lib.stream_handle(dirs.device)))
    return out, pixels, camera_positions, dirs


class _CompositeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, samples):
        shape = samples.shape
        S = shape[-2]
        x = samples.float().contiguous()
        out = torch.empty(shape[:-2] + (4,), dtype=torch.float32, device=x.device)
        R = out.numel() // 4  # S = 0 composites to zeros, as the reference's empty sums
        _lib.check(_lib.lib().pxl_composite_fwd(_lib.ptr(x), R, S, _lib.ptr(out),
                                                _lib.stream_handle(x.device)))
        ctx.save_for_backward(x)
        ctx.R, ctx.S = R, S
        return out

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        gs = torch.empty_like(x)
        gc = g.float().contiguous()
        _lib.check(_lib.lib().pxl_composite_bwd(_lib.ptr(x), ctx.R, ctx.S, _lib.ptr(gc), _lib.ptr(gs),
                                                _lib.stream_handle(x.device)))
        return gs


def compute_alpha_weighted_pixels(samples):
    """src/ray_sampling.py:172-192: (..., S, 4) samples -> (..., 4) RGBA by front-to-back
    alpha compositing (alpha = last channel, no background); differentiable."""
    _dev(samples)
    if samples.dim() < 2 or samples.shape[-1] != 4:
        raise ValueError("samples must be (..., S, 4)")
    return _CompositeFn.apply(samples)
