"""plenoxels_amd -- MI355X-native voxel-grid render-and-optimise path.

Drop-in for the hot path of DanJbk/Plenoxels (src/ray_sampling.py, src/grid_functions.py as
called from scripts/train.py and src/visualization.py), computed by hand-written gfx950 HIP
kernels in libpxl.so (C ABI: include/pxl.h).  There is no CPU fallback: every entry point
raises if the library is not built or no HIP device is present.
"""
from .render import (CameraRays, CornerBlockPair, LossMean, MarchConfig, alpha_occupancy, blocks_clear, blocks_reduce, camera_rays, corner_blocks, fill_,  # noqa: F401
                     fill_reachable_, grid_min_of, occupancy_dilate, pack_occupancy, render_rays, train_step)

__all__ = ["CameraRays", "CornerBlockPair", "LossMean", "MarchConfig", "alpha_occupancy", "blocks_clear", "blocks_reduce", "camera_rays", "corner_blocks", "fill_", "fill_reachable_", "grid_min_of",
           "occupancy_dilate", "pack_occupancy", "render_rays", "train_step"]
