"""The grid_cells_trained.pth format of scripts/train.py:197-210.

A torch.save'd dict: "grid" (X,Y,Z,4) float32 CPU tensor, "grid_grad" (X,Y,Z,4) float32
CPU tensor (sum over steps of |d loss / d grid|), "param" dict with device,
number_of_rays, num_samples, delta_step, even_spread, camera_ray, points_distance, gridsize.
Read by the reference's src/visualization.py:114-122, scripts/visulize_grid.py:23-30 and
scripts/compare_inference_to_image.py:41-50.  BF16/FP16 grids are saved upcast to float32.
"""
import torch

PARAM_KEYS = ("device", "number_of_rays", "num_samples", "delta_step", "even_spread", "camera_ray",
              "points_distance", "gridsize")


def save_grid(path, grid, grid_grad, *, device, number_of_rays, num_samples, delta_step, even_spread,
              camera_ray, points_distance, gridsize):
    torch.save({
        "grid": grid.detach().float().cpu(),
        "grid_grad": grid_grad.detach().float().cpu(),
        "param": {
            "device": device,
            "number_of_rays": number_of_rays,
            "num_samples": num_samples,
            "delta_step": delta_step,
            "even_spread": even_spread,
            "camera_ray": camera_ray,
            "points_distance": points_distance,
            "gridsize": list(gridsize),
        },
    }, path)


def load_grid(path, device="cpu"):
    """Load a .pth written by the reference or by save_grid (weights_only: no code runs)."""
    d = torch.load(path, map_location=device, weights_only=True)
    missing = [k for k in ("grid", "param") if k not in d]
    if missing:
        raise KeyError(f"{path} lacks {missing}")
    return d
