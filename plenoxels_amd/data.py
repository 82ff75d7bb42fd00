"""NeRF-synthetic loading as src/data_processing.py:7-60 does it: transforms_*.json camera
matrices (the c2w convention the ray kernels use) and RGBA PNGs as float32 / 255."""
import json

import numpy as np
import torch


def load_data(data):
    """src/data_processing.py:18-29: (transform_matricies (B,4,4), file_paths, camera_angle_x)."""
    mats = [torch.tensor(f["transform_matrix"]).unsqueeze(0) for f in data["frames"]]
    paths = [f["file_path"] for f in data["frames"]]
    return torch.cat(mats, 0), paths, data["camera_angle_x"]


def load_image_data_from_path(path, transformpath):
    """src/data_processing.py:51-60: (json dict, imgs (B,H,W,C) float32 in [0,1])."""
    from PIL import Image
    with open(transformpath, "r") as f:
        data = json.load(f)
    imgs = [np.array(Image.open(f'{path}/{frame["file_path"].split("/")[-1]}.png')) for frame in data["frames"]]
    imgs = torch.tensor(np.array(imgs), dtype=torch.float)
    return data, imgs / 255
