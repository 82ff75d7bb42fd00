"""Writes tests/golden/reference_vectors.npz from the reference implementation itself.

Run in a container where the reference checkout exists (read-only is fine):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py /root/reference

Only tiny seeded CPU inputs are used; the outputs (inputs and expected outputs, nothing
else) are committed as data.  These vectors pin the composed conventions that no closed
form fixes on its own: the pixel/target mapping and signs of generate_rays_batched
(src/ray_sampling.py:195-264), the flattened sample layout of sample_camera_rays_batched
(:128-169), the corner order of get_grid_points_indices (src/grid_functions.py:220-246) as
trilinear_interpolation (:7-44) consumes it, the channel packing of
compute_alpha_weighted_pixels (src/ray_sampling.py:172-192), and the composed nearest-voxel
training chain of scripts/train.py:142-156 including its autograd gradient.
"""
import os
import sys

import numpy as np

sys.dont_write_bytecode = True


def main(ref_root):
    sys.path.insert(0, ref_root)
    import torch
    from src import grid_functions as gf
    from src import ray_sampling as rs
    from scripts import train as tr

    torch.set_num_threads(1)
    out = {}

    def rot(seed):
        g = torch.Generator().manual_seed(seed)
        q, _ = torch.linalg.qr(torch.randn(3, 3, generator=g, dtype=torch.float64))
        return q.float()

    def cams(B, seed, radius=2.0, scale_y=1.0):
        m = torch.zeros(B, 4, 4)
        for b in range(B):
            R = rot(seed + b)
            R[:, 1] *= scale_y
            m[b, :3, :3] = R
            m[b, :3, 3] = R[:, 2] * radius
            m[b, 3, 3] = 1
        return m

    # ---- generate_rays_batched (random and even spread) + sample_camera_rays_batched
    g = torch.Generator().manual_seed(11)
    B, H, W = 3, 6, 6
    imgs = torch.rand(B, H, W, 4, generator=g)
    c2w = cams(B, 100, scale_y=1.25)  # non-unit Y axis: aspect != 1
    theta = 0.6911112070083618
    out["rays_imgs"], out["rays_c2w"], out["rays_theta"] = imgs.numpy(), c2w.numpy(), np.float64(theta)
    R = 7
    torch.manual_seed(5)
    uni = torch.rand(B, R, 2)
    torch.manual_seed(5)
    dirs, pix = rs.generate_rays_batched(imgs, R, c2w, theta, even_spread=False, device="cpu")
    out["rays_rand_uniforms"], out["rays_rand_dirs"], out["rays_rand_pixels"] = (
        uni.numpy(), dirs.numpy(), pix.numpy())
    dirs, pix = rs.generate_rays_batched(imgs, 9, c2w, theta, even_spread=True, device="cpu")
    out["rays_even_dirs"], out["rays_even_pixels"] = dirs.numpy(), pix.numpy()

    S, delta = 5, 0.3
    torch.manual_seed(6)
    uni = torch.rand(B, R, 2)
    torch.manual_seed(6)
    smp, pix, cpos, dirs = rs.sample_camera_rays_batched(c2w, theta, imgs, R, S, delta,
                                                         even_spread=False, camera_ray=False,
                                                         device="cpu")
    out["scr_uniforms"], out["scr_samples"], out["scr_pixels"] = uni.numpy(), smp.numpy(), pix.numpy()
    out["scr_campos"], out["scr_dirs"], out["scr_S"], out["scr_delta"] = (
        cpos.numpy(), dirs.numpy(), np.int64(S), np.float64(delta))

    # ---- generate_grid + normalize_samples_for_indecies
    gi, gc, mg, gg = gf.generate_grid(5, 6, 7, points_distance=0.25, info_size=4, device="cpu")
    out["grid_coords"], out["grid_meshgrid"], out["grid_grid"] = gi.numpy(), mg.numpy(), gg.detach().numpy()
    p = (torch.rand(40, 3, generator=g) - 0.5) * 2.5
    out["norm_in"] = p.numpy()
    out["norm_out"] = rs.normalize_samples_for_indecies(gi, p, 0.25).numpy()

    # ---- nearest voxels on crafted ties / bounds
    X, Y, Z = 4, 5, 6
    grid = torch.rand(X, Y, Z, 4, generator=g) * 1.4 - 0.2
    ties = torch.tensor([0.5, 1.5, 2.5, -0.5, -0.6, 3.5, 3.49, 4.5, -1.0, 0.0, 2.0, 5.5])
    n = torch.stack([ties, ties.roll(3) * 1.2, ties.roll(7) * 1.4], 1)
    n = torch.cat([n, (torch.rand(30, 3, generator=g) * 8 - 1.5)], 0)
    vals, mask = gf.get_nearest_voxels(n, grid.clip(0, 1))
    out["near_grid"], out["near_n"], out["near_vals"], out["near_mask"] = (
        grid.numpy(), n.numpy(), vals.numpy(), mask.numpy())

    # ---- corners + collect + trilinear (wrapped corners)
    corners = gf.get_grid_points_indices(n)
    out["tri_corners"] = corners.numpy()
    cells, tmask = gf.collect_cell_information_via_indices(n, grid.clip(0, 1))
    out["tri_cells"], out["tri_mask"] = cells.numpy(), tmask.numpy()
    c = corners.reshape(-1, 3).clone()
    gf.fix_out_of_bounds(c, grid)
    tri = gf.trilinear_interpolation(n, c.reshape(-1, 8, 3), grid.clip(0, 1))
    out["tri_vals"] = tri.numpy()

    # ---- compositing forward + autograd gradient
    smp = torch.rand(2, 3, 9, 4, generator=g)
    smp[0, 1, 4, 3] = 1.0
    smp[1, 2, :, 3] = 0.0
    smp[1, 0, 2, 3] = 1.0
    smp = smp.requires_grad_(True)
    res = rs.compute_alpha_weighted_pixels(smp)
    up = torch.randn(res.shape, generator=g)
    (res * up).sum().backward()
    out["comp_in"], out["comp_out"], out["comp_up"], out["comp_grad"] = (
        smp.detach().numpy(), res.detach().numpy(), up.numpy(), smp.grad.numpy())

    # ---- composed nearest training chain (scripts/train.py:142-156) with grid gradient
    Xc = 8
    gi, cells, _, _ = gf.generate_grid(Xc, Xc, Xc, points_distance=0.3, info_size=4, device="cpu")
    with torch.no_grad():
        cells.copy_(torch.rand(Xc, Xc, Xc, 4, generator=g) * 1.3 - 0.15)
    c2 = cams(2, 300, radius=2.2)
    im2 = torch.rand(2, 6, 6, 4, generator=g)
    Rr, Ss, dd = 6, 16, 0.27
    torch.manual_seed(8)
    uni = torch.rand(2, Rr, 2)
    torch.manual_seed(8)
    smp, pix, _, dirs = rs.sample_camera_rays_batched(c2, theta, im2, Rr, Ss, dd, even_spread=False,
                                                      camera_ray=False, device="cpu")
    nrm = rs.normalize_samples_for_indecies(gi, smp, 0.3)
    near, mask = gf.get_nearest_voxels(nrm, cells.clip(0, 1))
    near = near * mask.unsqueeze(-1)
    near = near.reshape(2, Rr, Ss, 4)
    col = rs.compute_alpha_weighted_pixels(near)
    col = col.reshape(2 * Rr, 4)
    loss = torch.nn.functional.mse_loss(col, pix)
    loss.backward()
    out.update({
        "chain_grid": cells.detach().numpy(), "chain_c2w": c2.numpy(), "chain_imgs": im2.numpy(),
        "chain_uniforms": uni.numpy(), "chain_R": np.int64(Rr), "chain_S": np.int64(Ss),
        "chain_delta": np.float64(dd), "chain_pd": np.float64(0.3), "chain_theta": np.float64(theta),
        "chain_rgba": col.detach().numpy(), "chain_loss": loss.detach().numpy(),
        "chain_grad": cells.grad.numpy(), "chain_dirs": dirs.numpy(),
    })

    # ---- next rows: pooled grid and TV loss
    pg = torch.rand(9, 9, 9, 4, generator=g)
    out["pool_in"] = pg.numpy()
    out["pool_k5_s1"] = gf.average_pool3d_grid(pg, receptive_field_size=5, stride=1).numpy()
    out["pool_k7_s2"] = gf.average_pool3d_grid(pg, receptive_field_size=7, stride=2).numpy()
    tvin = torch.rand(5, 6, 7, 4, generator=g).requires_grad_(True)
    tv = tr.tv_loss(tvin)
    tv.backward()
    out["tv_in"], out["tv_out"], out["tv_grad"] = (
        tvin.detach().numpy(), tv.detach().numpy(), tvin.grad.numpy())

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
