"""Occupancy from the grid's own alpha (pxl_alpha_occupancy) and its exactness.

The bits equal a torch restatement (raw alpha >= 0, dilated over the 8 corners for trilinear)
for every storage dtype and odd shapes; rendering and training with them give the results of
the unskipped kernels (the skipped samples are exactly zero in the reference chain: clamped
alpha 0, zero clamp gradient) -- compared at FP32 summation-order tolerance, since the kept
samples regroup -- and the Trainer with skip_empty follows the reference loop."""
import pytest
import torch

from oracle import plain
from plenoxels_amd import MarchConfig, alpha_occupancy, train_step
from util import grad_close, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _want_bits(alpha_ge0, interp):
    occ = plain.dilate_cells(alpha_ge0) if interp == "trilinear" else alpha_ge0
    return plain.occupancy_to_bits(occ)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("dims", [(20, 20, 20), (7, 5, 70), (3, 4, 33)])
def test_alpha_occupancy_bits(dtype, dims):
    say("alpha occupancy", dtype, dims)
    g = torch.Generator().manual_seed(sum(dims))
    grid = torch.rand(*dims, 4, generator=g) - 0.6  # ~60% of alphas negative
    grid[0, 0, 0, 3] = -0.0                          # signed zero counts as occupied
    grid[1, 1, 1, 3] = float("nan")                  # NaN counts as empty
    gd = grid.to(dtype)
    a = gd.float()[..., 3]
    ge0 = (a >= 0) & ~torch.isnan(a)
    for interp in ("nearest", "trilinear"):
        got = alpha_occupancy(gd.to(DEV), interp).cpu()
        assert torch.equal(got, _want_bits(ge0, interp)), interp


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_skipping_negative_alpha_changes_nothing(interp):
    say("alpha-occupancy skip exact", interp)
    grid, o, d, gmin, pd = small_scene(110, X=24, R=768)
    grid[..., 3] = grid[..., 3] - 0.06  # many cells below zero
    R = o.shape[0]
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(111))
    cfg = MarchConfig(150, 0.026, tuple(gmin.tolist()), pd, interp)
    gd = grid.to(DEV)
    outs, grads = [], []
    for occ in (None, alpha_occupancy(gd, interp)):
        gg = torch.zeros_like(gd)
        out = torch.empty(R, 4, device=DEV)
        train_step(gd, gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out,
                   occupancy=occ)
        outs.append(out.cpu())
        grads.append(gg.cpu())
    assert (outs[0] - outs[1]).abs().max().item() <= 1e-6
    ok, info = grad_close(grads[1], grads[0], rel=1e-5)
    say("  skipped vs all samples", info)
    assert ok, info
    # and against the plain chain
    g = grid.clone().requires_grad_(True)
    want = plain.render(g, o, d, num_samples=150, delta_step=0.026, gmin=gmin, points_distance=pd,
                        interp=interp, div_mode="reciprocal")
    ((want - tgt) ** 2).mean().backward()
    assert (outs[1] - want.detach()).abs().max().item() <= 1e-5
    ok, info = grad_close(grads[1], g.grad)
    assert ok, info


def test_trainer_skip_empty_matches_reference_loop():
    """Trainer(skip_empty=True) vs the reference loop on a grid with negative alphas."""
    from plenoxels_amd.train import Trainer
    from util import random_cameras
    say("trainer skip_empty vs reference loop")
    B, H, R, steps, dims = 3, 24, 64, 4, (24, 24, 24)
    c2w = random_cameras(B, 20, radius=2.0)
    imgs = torch.rand(B, H, H, 4, generator=torch.Generator().manual_seed(21))
    g = torch.Generator().manual_seed(22)
    uniforms = [torch.rand(B, R, 2, generator=g) for _ in range(steps)]
    init = torch.rand(*dims, 4, generator=torch.Generator().manual_seed(30)) * 0.6 - 0.3
    kw = dict(gridsize=dims, points_distance=0.06, num_samples=48, delta_step=0.05, lr=0.01, tv=1e-3,
              beta=5e-3, init_grid=init)
    grid_w, gabs_w, hist_w = plain.reference_fit(imgs, c2w, 0.69, uniforms=uniforms, div_mode="reciprocal",
                                                 progressive_growing=False, **kw)
    kw["init_grid"] = init.to(DEV)
    tr = Trainer(imgs, c2w, 0.69, number_of_rays=R, device="cuda", progressive_growing=False,
                 uniforms_fn=lambda i: uniforms[i], skip_empty=True, **kw)
    for i in range(steps):
        out = tr.step(i)
        for key, v in hist_w[i].items():
            assert abs(float(out[key]) - v) <= 1e-4 * abs(v) + 1e-9, (i, key, float(out[key]), v)
    diff = (tr.grid.cpu() - grid_w).abs()
    say("  grid: max diff", diff.max().item(), "frac > 1e-5:", (diff > 1e-5).float().mean().item())
    assert (diff > 1e-5).float().mean().item() <= 1e-3
