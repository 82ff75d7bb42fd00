"""Forward-only inference render of a trained grid: the counterpart of
src/visualization.py:111-154 (visulize_3d_in_2d) on the fused forward kernel."""
import numpy as np
import torch

from .render import CameraRays, MarchConfig, grid_min_of, render_camera


def visulize_3d_in_2d(grid_cells_data, transform_matrices, camera_angle_x, imgs, grid_indices, do_threshold,
                      transparency_threshold, number_of_rays, num_samples, device="cuda"):
    """Render the even-spread rays of the first camera through grid_cells_data["grid"]:
    returns a (res,res,4) uint8 image transposed as the reference returns it.

    grid: clip(0,1) and optional alpha threshold as src/visualization.py:114-118; samples o +
    d*delta*k with delta and points_distance from the checkpoint's "param"; nearest voxel;
    front-to-back compositing.  The reference reshapes to one camera (:145), so imgs and
    transform_matrices hold one camera.
    """
    dev = torch.device(device)
    grid = grid_cells_data["grid"].detach().to(dev).float().clip(0.0, 1.0).contiguous()
    if do_threshold:
        alphas = grid[..., -1]
        grid[..., -1][alphas < transparency_threshold] = 0.0
    pd = grid_cells_data["param"]["points_distance"]
    delta = grid_cells_data["param"]["delta_step"]
    n = int(np.round(np.sqrt(number_of_rays)))
    number_of_rays = n * n
    ls = torch.linspace(0, end=1, steps=n, device=dev)
    uv = torch.cartesian_prod(ls, ls).contiguous()
    tm = transform_matrices.to(dev).float().contiguous()
    if tm.dim() == 2:
        tm = tm.unsqueeze(0)
    cam = CameraRays(torch.zeros(number_of_rays, dtype=torch.int32, device=dev), uv, tm[:1].contiguous(),
                     float(camera_angle_x))
    cfg = MarchConfig(int(num_samples), float(delta), grid_min_of(grid_indices.to(dev)), float(pd), "nearest")
    out = render_camera(grid, cam, cfg)
    pixels = out.cpu().numpy()
    pixels = (pixels * 255).round().clip(0, 255).astype(np.uint8)
    image = pixels.reshape(n, n, 4)
    return np.transpose(image, (1, 0, 2))
