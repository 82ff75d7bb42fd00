"""Seeded synthetic scenes for the benchmark workloads and tests.

A scene is a dense (X,Y,Z,4) grid in the reference layout (src/grid_functions.py:214-215,
channels [R,G,B,alpha]) spanning the cube [-1.5,1.5]^3 (points_distance 3/(X-1), cell
world position (i - ceil(X/2) + 1) * pd as src/grid_functions.py:205-209 places it), a set
of NeRF/Blender-convention camera-to-world matrices (camera looks along -Z, src/ray_sampling.py:215-217)
and per-ray (camera, u, v) uniforms.

Workloads (sizes as used by bench.py and the tests):
  blobs64   : 64^3, 8 cameras on a radius-4 sphere, 64x64 even-spread rays (32,768), S=128,
              near 2, delta 4/128, focal 1.1 W (camera_angle_x = 1/1.1), trilinear
  view800   : 128^3, one 800x800 view (640,000 rays), S=256, focal 1111 px
  sparse256 : 256^3 with ~5% occupancy from thresholded value noise, 2^18 rays drawn
              uniformly from 100 cameras on the upper hemisphere (800x800, focal 1111 px),
              S=256, near 2, delta 4/256
Seeds: grid 0, cameras 1, ray uniforms 2, targets 3.
"""
import math
from dataclasses import dataclass

import numpy as np
import torch

RADIUS = 4.0


def points_distance_for(X):
    return 3.0 / (X - 1)


def world_axis(X, pd):
    """World coordinate of grid index i along one axis (float64)."""
    return (np.arange(X) - math.ceil(X / 2) + 1) * np.float32(pd)


def grid_min(X, Y, Z, pd):
    """grid_indices.min(0)[0] of the reference's generate_grid, as float32 values: the float32
    coordinate of index 0 computed as fl(fl(0 - ceil(S/2) + 1) * fl32(pd))."""
    out = []
    for S in (X, Y, Z):
        c = np.float32(0 - math.ceil(S / 2) + 1)
        out.append(float(np.float32(c * np.float32(pd))))
    return tuple(out)


def blob_field(X, seed=0, n_blobs=6, peak=20.0):
    """channel 3 = ReLU(sum_b peak * exp(-|x - mu_b|^2 / (2 sigma_b^2))), mu ~ U[-1,1]^3,
    sigma ~ U[0.1,0.3]; channels 0-2 ~ U[0,1].  float32 (X,X,X,4) on the CPU."""
    g = torch.Generator().manual_seed(seed)
    mu = torch.rand(n_blobs, 3, generator=g, dtype=torch.float64) * 2 - 1
    sig = torch.rand(n_blobs, generator=g, dtype=torch.float64) * 0.2 + 0.1
    pd = points_distance_for(X)
    ax = torch.from_numpy(world_axis(X, pd))
    dens = torch.zeros(X, X, X, dtype=torch.float64)
    for b in range(n_blobs):
        dx = (ax - mu[b, 0]) ** 2
        dy = (ax - mu[b, 1]) ** 2
        dz = (ax - mu[b, 2]) ** 2
        r2 = dx[:, None, None] + dy[None, :, None] + dz[None, None, :]
        dens += peak * torch.exp(-r2 / (2 * sig[b] ** 2))
    rgb = torch.rand(X, X, X, 3, generator=g)
    return torch.cat([rgb, dens.clamp_min(0).float().unsqueeze(-1)], -1).contiguous()


def value_noise_occupancy(X, seed=0, lattice=32, frac=0.05):
    """Seeded value noise on a lattice^3 grid, trilinearly upsampled to X^3 and thresholded at
    its (1 - frac) quantile: (X,X,X) bool with ~frac occupied cells."""
    g = torch.Generator().manual_seed(seed + 1000)
    lat = torch.rand(1, 1, lattice, lattice, lattice, generator=g)
    up = torch.nn.functional.interpolate(lat, size=(X, X, X), mode="trilinear", align_corners=True)[0, 0]
    thr = torch.quantile(up.reshape(-1)[:: max(1, (X ** 3) // 2_000_000)], 1 - frac)
    return up > thr


def look_at(pos):
    """c2w (4,4) float32 for a camera at pos looking at the origin, world up +Z."""
    pos = np.asarray(pos, np.float64)
    z = pos / np.linalg.norm(pos)            # camera +Z points away from the target
    up = np.array([0.0, 0.0, 1.0])
    if abs(np.dot(up, z)) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = x, y, z, pos
    return m.astype(np.float32)


def sphere_cameras(n, seed=1, hemisphere=False, radius=RADIUS):
    """n cameras at seeded positions on a sphere (or upper hemisphere) of `radius`."""
    g = np.random.default_rng(seed)
    mats = []
    for _ in range(n):
        v = g.normal(size=3)
        if hemisphere:
            v[2] = abs(v[2]) + 0.05
        v = v / np.linalg.norm(v) * radius
        mats.append(look_at(v))
    return torch.from_numpy(np.stack(mats))


def primitive_scene(n=12, seed=0, lo=-1.2, hi=1.2, rmin=0.1, rmax=0.5):
    """n seeded coloured primitives inside [lo,hi]^3, alternating spheres (radius ~ U[rmin,rmax])
    and axis-aligned boxes (half-extents ~ U[rmin,rmax]^3): list of dicts with kind, center (3,),
    half-size (3,) and RGB colour (3,), all float64."""
    g = np.random.default_rng(seed + 2000)
    prims = []
    for i in range(n):
        kind = "sphere" if i % 2 == 0 else "box"
        half = np.full(3, g.uniform(rmin, rmax)) if kind == "sphere" else g.uniform(rmin, rmax, size=3)
        center = g.uniform(lo + half, hi - half)
        prims.append({"kind": kind, "center": center, "half": half, "colour": g.uniform(0.0, 1.0, size=3)})
    return prims


def voxelize_primitives(prims, X, alpha=1.0):
    """(X,X,X,4) float32 grid of the scene at the grid's cell positions (world_axis): a cell
    inside a primitive takes its colour (later primitives over earlier ones) and `alpha`;
    cells outside every primitive are 0."""
    pd = points_distance_for(X)
    ax = world_axis(X, pd)
    grid = np.zeros((X, X, X, 4), np.float32)
    for p in prims:
        c, h = p["center"], p["half"]
        sl = []
        for a in range(3):  # index range of the primitive's bounding box
            idx = np.flatnonzero(np.abs(ax - c[a]) <= h[a])
            sl.append(slice(int(idx[0]), int(idx[-1]) + 1) if idx.size else slice(0, 0))
        if any(t.stop <= t.start for t in sl):
            continue
        dx, dy, dz = (ax[sl[a]] - c[a] for a in range(3))
        if p["kind"] == "sphere":
            inside = (dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2) <= h[0] ** 2
        else:
            inside = np.ones((dx.size, dy.size, dz.size), bool)
        sub = grid[sl[0], sl[1], sl[2]]
        sub[inside, :3] = p["colour"].astype(np.float32)
        sub[inside, 3] = alpha
    return torch.from_numpy(grid)


def pixel_uv(H, W):
    """(H*W, 2) float32 uniforms whose target pixel (src/ray_sampling.py:238-248: row =
    round(W*v), col = round(H*u)) is (row, col), row-major: u = col/H, v = row/W."""
    rows = torch.arange(H, dtype=torch.float32) / W
    cols = torch.arange(W, dtype=torch.float32) / H
    v = rows[:, None].expand(H, W)
    u = cols[None, :].expand(H, W)
    return torch.stack([u, v], -1).reshape(-1, 2).contiguous()


@dataclass
class Workload:
    name: str
    grid: torch.Tensor            # (X,X,X,4) float32 (CPU)
    occupancy: torch.Tensor       # (X,X,X) bool or None
    c2w: torch.Tensor             # (B,4,4)
    camera_angle_x: float
    img_hw: tuple
    cam_ids: torch.Tensor         # (R,) int32
    uv: torch.Tensor              # (R,2) float32
    num_samples: int
    delta_step: float
    near: float
    points_distance: float
    interp: str
    t_stop: float = 0.0

    @property
    def dims(self):
        return tuple(self.grid.shape[:3])

    @property
    def grid_min(self):
        return grid_min(*self.dims, self.points_distance)


def even_uv(n_side, n_cams):
    """cartesian_prod(linspace(0,1,n), linspace(0,1,n)) per camera (src/ray_sampling.py:220-223)."""
    ls = torch.linspace(0, 1, n_side)
    uv = torch.cartesian_prod(ls, ls)
    return uv.repeat(n_cams, 1).contiguous(), torch.repeat_interleave(
        torch.arange(n_cams, dtype=torch.int32), n_side * n_side)


def make_workload(name, *, n_rays=None, interp=None, t_stop=None):
    if name == "blobs64":
        X = 64
        grid = blob_field(X, seed=0)
        c2w = sphere_cameras(8, seed=1)
        uv, cam = even_uv(64, 8)
        w = Workload(name, grid, None, c2w, 1 / 1.1, (64, 64), cam, uv, 128, 4 / 128, 2.0,
                     points_distance_for(X), "trilinear")
    elif name == "view800":
        X = 128
        grid = blob_field(X, seed=0)
        c2w = sphere_cameras(1, seed=1)
        uv, cam = even_uv(800, 1)
        w = Workload(name, grid, None, c2w, 800 / 1111, (800, 800), cam, uv, 256, 4 / 256, 2.0,
                     points_distance_for(X), "nearest")
    elif name == "sparse256":
        X = 256
        occ = value_noise_occupancy(X, seed=0)
        grid = blob_field(X, seed=0)
        grid[..., 3] *= occ
        c2w = sphere_cameras(100, seed=1, hemisphere=True)
        R = 1 << 18 if n_rays is None else n_rays
        g = torch.Generator().manual_seed(2)
        cam = torch.randint(0, 100, (R,), generator=g, dtype=torch.int64).to(torch.int32)
        uv = torch.rand(R, 2, generator=g)
        w = Workload(name, grid, occ, c2w, 800 / 1111, (800, 800), cam, uv, 256, 4 / 256, 2.0,
                     points_distance_for(X), "trilinear", 1e-4)
    else:
        raise ValueError(name)
    if n_rays is not None and name != "sparse256":
        w.cam_ids, w.uv = w.cam_ids[:n_rays].contiguous(), w.uv[:n_rays].contiguous()
    if interp is not None:
        w.interp = interp
    if t_stop is not None:
        w.t_stop = t_stop
    return w


def random_targets(R, seed=3):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(R, 4, generator=g)
