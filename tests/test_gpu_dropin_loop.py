"""The drop-in promise, tested end to end on the GPU.

1. scripts/train.py:104-184 written out with the plenoxels_amd drop-ins in place of the
   reference's src/ functions (imports switched, nothing else), with torch.optim.Adam, against
   the reference loop restated op for op (oracle.plain.reference_fit) run with torch's own ops
   on the same GPU -- the reference on its default device.  The drop-ins draw their uniforms
   with torch.rand on the device as the reference does; the oracle gets the same draws by
   re-seeding the generator.  Steps cover pooled grids (progressive growing) and full-grid
   steps, with the TV and beta terms.
2. The reference's own golden vectors (tests/golden/reference_vectors.npz, made by running the
   reference on the CPU) fed to the HIP drop-ins in CPU semantics (exact division, CPU 3-norm
   order): bitwise where the reference's arithmetic is elementwise float32.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import plain
from plenoxels_amd import grid_functions as gf
from plenoxels_amd import ray_sampling as rs
from plenoxels_amd import MarchConfig, render_rays, _lib
from plenoxels_amd.optim import tv_loss
from util import grad_close, random_cameras, say

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
T = torch.from_numpy


def _dropin_fit(imgs, c2w, theta, *, gridsize, pd0, R, S, delta, lr, tv, beta, steps, start, seeds, init):
    """scripts/train.py:79-184 with the drop-ins (its code path, its variable names)."""
    grid_indices, grid_cells_full, meshgrid, grid_grid_original = gf.generate_grid(*gridsize, points_distance=pd0,
                                                                                   info_size=4, device="cuda")
    with torch.no_grad():
        grid_cells_full.copy_(init)
    grid_cells_full_grad = torch.zeros_like(grid_cells_full)
    optimizer = torch.optim.Adam([grid_cells_full], lr=lr)
    receptive_fields_sizes = [i for i in range(93, 2, -2) if i % 2 != 0]
    hist = []
    for i in range(start, start + steps):
        k = receptive_fields_sizes[i // 5] if i // 5 < len(receptive_fields_sizes) else 0
        if k > 1:
            stride = max(1, k // 4)
            start_index = int(k / 2)
            grid_grid = grid_grid_original[start_index::stride, start_index::stride, start_index::stride]
            grid_indices = grid_grid.reshape(-1, 3)
            points_distance = pd0 * stride
            grid_cells = gf.average_pool3d_grid(grid_cells_full, receptive_field_size=k, stride=stride)
        else:
            grid_cells = grid_cells_full
            grid_indices = grid_grid_original.reshape(-1, 3)
            points_distance = pd0
        torch.manual_seed(seeds[i - start])
        samples_interval, pixels_to_rays, camera_positions, ray_directions = rs.sample_camera_rays_batched(
            transform_matrices=c2w, camera_angle_x=theta, imgs=imgs, number_of_rays=R, num_samples=S,
            delta_step=delta, even_spread=False, camera_ray=False, device="cuda")
        n = rs.normalize_samples_for_indecies(grid_indices, samples_interval, points_distance)
        nearest, mask = gf.get_nearest_voxels(n, grid_cells.clip(0, 1))
        nearest = nearest * mask.unsqueeze(-1)
        nearest = nearest.reshape(imgs.shape[0], R, S, 4)
        pixels_color = rs.compute_alpha_weighted_pixels(nearest)
        cs = pixels_color.shape
        pixels_color = pixels_color.reshape([cs[0] * cs[1], cs[2]])
        mseloss = F.mse_loss(pixels_color, pixels_to_rays)
        loss = mseloss
        rec = {"closs": float(mseloss.detach())}
        if tv > 0 and k < 19:
            tvloss = tv * tv_loss(grid_cells)
            loss = loss + tvloss
            rec["tvloss"] = float(tvloss.detach())
        if beta > 0:
            eps = 0.0001
            betaloss = beta * (torch.log(nearest[:, :, :, -1] + eps) - torch.log(1 - nearest[:, :, :, -1] + eps)).mean()
            loss = loss + betaloss
            rec["betaloss"] = float(betaloss.detach())
        optimizer.zero_grad()
        loss.backward()
        optimizer.step()
        grid_cells_full_grad += torch.abs(grid_cells_full.grad)
        rec["grad"] = grid_cells_full.grad.clone()
        hist.append(rec)
    return grid_cells_full.detach(), grid_cells_full_grad, hist


@pytest.mark.parametrize("start", [220, 228])
def test_dropin_training_loop_matches_reference_fit_on_gpu(start):
    """Steps start..start+5: k = 5 and 3 pooled grids (TV on), then full-grid steps."""
    say("drop-in training loop from step", start)
    B, H, R, S = 3, 20, 48, 64
    dims, pd0, delta, lr, tv, beta = (20, 20, 20), 0.07, 0.045, 0.01, 1e-3, 5e-3
    c2w = random_cameras(B, 40, radius=2.0).to(DEV)
    imgs = torch.rand(B, H, H, 4, generator=torch.Generator().manual_seed(41)).to(DEV)
    init = (torch.rand(*dims, 4, generator=torch.Generator().manual_seed(42)) * 0.7 - 0.1).to(DEV)
    steps = 5
    seeds = [1000 + s for s in range(steps)]
    grid_d, gabs_d, hist_d = _dropin_fit(imgs, c2w, 0.69, gridsize=dims, pd0=pd0, R=R, S=S, delta=delta, lr=lr,
                                         tv=tv, beta=beta, steps=steps, start=start, seeds=seeds, init=init)
    uniforms = []
    for sd in seeds:  # the draws torch.rand(B, R, 2) made inside the drop-ins
        torch.manual_seed(sd)
        uniforms.append(torch.rand(B, R, 2, device=DEV))
    grid_w, gabs_w, hist_w = plain.reference_fit(imgs, c2w, 0.69, gridsize=dims, points_distance=pd0, num_samples=S,
                                                 delta_step=delta, lr=lr, tv=tv, beta=beta, uniforms=uniforms,
                                                 start_step=start, init_grid=init, norm_mode=None)
    for i, (a, b) in enumerate(zip(hist_d, hist_w)):
        assert set(a) - {"grad"} == set(b), (i, a, b)
        for key in b:
            assert abs(a[key] - b[key]) <= 1e-4 * abs(b[key]) + 1e-9, (start + i, key, a[key], b[key])
    # the first step's gradient (both loops from the same grid) at the 1e-4 bar; later steps start
    # from grids ~1e-5 apart, where a value crossing a clamp kink at 0 or 1 switches its
    # gradient on or off, so the |grad| sums are held to 1e-3 of the largest
    _, g1_w, _ = plain.reference_fit(imgs, c2w, 0.69, gridsize=dims, points_distance=pd0, num_samples=S,
                                     delta_step=delta, lr=lr, tv=tv, beta=beta, uniforms=uniforms[:1],
                                     start_step=start, init_grid=init, norm_mode=None)
    ok, info = grad_close(hist_d[0]["grad"].abs().cpu(), g1_w.cpu())  # reference_fit returns sum |grad|
    say("  first-step gradient", info)
    assert ok, info
    err = (gabs_d - gabs_w).abs().max().item()
    diff = (grid_d - grid_w).abs()
    say("  |grad| sums max err", err, "of", gabs_w.abs().max().item(), "grid max diff", diff.max().item())
    assert err <= 1e-3 * gabs_w.abs().max().item()
    assert (diff > 1e-5).float().mean().item() <= 1e-3


def test_golden_vectors_through_hip_dropins(golden):
    """The reference's outputs (made on the CPU) from the HIP drop-ins in CPU semantics."""
    say("golden vectors through the HIP drop-ins")
    g = {k: T(v) if isinstance(v, np.ndarray) and v.ndim > 0 else v for k, v in golden.items()}
    imgs, c2w, theta = g["rays_imgs"].to(DEV), g["rays_c2w"].to(DEV), float(golden["rays_theta"])
    d, pix = rs.rays_from_uniforms(imgs, c2w, theta, g["rays_rand_uniforms"].to(DEV), norm_mode="cpu")
    assert torch.equal(d.cpu(), g["rays_rand_dirs"]) and torch.equal(pix.cpu(), g["rays_rand_pixels"])
    d, pix = rs.rays_from_uniforms(imgs, c2w, theta, plain.even_spread_uniforms(9, 3).to(DEV), norm_mode="cpu")
    assert torch.equal(d.cpu(), g["rays_even_dirs"]) and torch.equal(pix.cpu(), g["rays_even_pixels"])
    # sample_camera_rays_batched: rays from the golden uniforms, then the sample kernel
    uni = g["scr_uniforms"].to(DEV)
    B, R = uni.shape[:2]
    S, delta = int(golden["scr_S"]), float(golden["scr_delta"])
    d, pix = rs.rays_from_uniforms(imgs, c2w, theta, uni, norm_mode="cpu")
    cp = c2w[:, :3, 3].contiguous()
    smp = torch.empty(B * R * S, 3, device=DEV)
    _lib.check(_lib.lib().pxl_sample_points(_lib.ptr(cp), R, _lib.ptr(d), B * R, S, float(np.float32(delta)), 0.0,
                                            _lib.ptr(smp), _lib.stream_handle(DEV)))
    assert torch.equal(smp.cpu(), g["scr_samples"]) and torch.equal(d.cpu(), g["scr_dirs"])
    assert torch.equal(pix.cpu(), g["scr_pixels"])
    # generate_grid + normalize (pd = 0.25: the reciprocal 4 is exact, so both divisions agree)
    gi, gc, mg, gg = gf.generate_grid(5, 6, 7, points_distance=0.25)
    assert torch.equal(gi.cpu(), g["grid_coords"]) and torch.equal(mg.cpu(), g["grid_meshgrid"])
    assert torch.equal(gg.cpu(), g["grid_grid"])
    n = rs.normalize_samples_for_indecies(gi, g["norm_in"].to(DEV), 0.25)
    assert torch.equal(n.cpu(), g["norm_out"])
    # nearest voxels on ties / bounds, corners, collected cells, trilinear
    grid, n = g["near_grid"].to(DEV), g["near_n"].to(DEV)
    vals, mask = gf.get_nearest_voxels(n, grid.clip(0, 1))
    assert torch.equal(vals.cpu(), g["near_vals"]) and torch.equal(mask.cpu(), g["near_mask"])
    corners = gf.get_grid_points_indices(n)
    assert torch.equal(corners.cpu(), g["tri_corners"])
    cells, tmask = gf.collect_cell_information_via_indices(n, grid.clip(0, 1))
    assert torch.equal(cells.cpu(), g["tri_cells"]) and torch.equal(tmask.cpu(), g["tri_mask"])
    c = corners.reshape(-1, 3).clone()
    gf.fix_out_of_bounds(c, grid)
    tri = gf.trilinear_interpolation(n, c.reshape(-1, 8, 3), grid.clip(0, 1))
    assert torch.equal(tri.cpu(), g["tri_vals"])
    # compositing: the wave scan's products round in another order than torch's cumprod
    x = g["comp_in"].to(DEV).requires_grad_(True)
    res = rs.compute_alpha_weighted_pixels(x)
    assert (res.detach().cpu() - g["comp_out"]).abs().max().item() <= 1e-6
    (res * g["comp_up"].to(DEV)).sum().backward()
    assert (x.grad.cpu() - g["comp_grad"]).abs().max().item() <= 1e-6
    # the composed nearest training chain (scripts/train.py:142-156) with its grid gradient
    cg, c2 = g["chain_grid"].to(DEV), g["chain_c2w"].to(DEV)
    uni = g["chain_uniforms"].to(DEV)
    Rr, Ss, dd, pd = int(golden["chain_R"]), int(golden["chain_S"]), float(golden["chain_delta"]), float(golden["chain_pd"])
    d, pix = rs.rays_from_uniforms(g["chain_imgs"].to(DEV), c2, float(golden["chain_theta"]), uni, norm_mode="cpu")
    assert torch.equal(d.cpu(), g["chain_dirs"])
    o = torch.repeat_interleave(c2[:, :3, 3], Rr, 0).contiguous()
    gi8, _, _, _ = plain.generate_grid(8, 8, 8, points_distance=pd)
    leaf = cg.clone().requires_grad_(True)
    out = render_rays(leaf, o, d, num_samples=Ss, delta_step=dd, grid_min=tuple(plain.grid_min(gi8).tolist()),
                      points_distance=pd, div_mode="divide")
    loss = F.mse_loss(out, pix)
    loss.backward()
    assert (out.detach().cpu() - g["chain_rgba"]).abs().max().item() <= 1e-6
    assert abs(loss.item() - float(golden["chain_loss"])) <= 1e-6
    ok, info = grad_close(leaf.grad.cpu(), g["chain_grad"])
    assert ok, info
    # pooled grid and TV loss (the "next" rows)
    pin = g["pool_in"].to(DEV)
    for (k, s), key in (((5, 1), "pool_k5_s1"), ((7, 2), "pool_k7_s2")):
        got = gf.average_pool3d_grid(pin, receptive_field_size=k, stride=s).cpu()
        assert (got - g[key]).abs().max().item() <= 1e-6, key
    tin = g["tv_in"].to(DEV).requires_grad_(True)
    tvv = tv_loss(tin)
    tvv.backward()
    # (the drop-in sums the squares in FP64, the reference in float32)
    assert abs(tvv.item() - float(golden["tv_out"])) <= 1e-5 * float(golden["tv_out"])
    assert (tin.grad.cpu() - g["tv_grad"]).abs().max().item() <= 1e-5 * g["tv_grad"].abs().max().item()
