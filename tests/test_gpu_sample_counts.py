"""Every dispatch boundary of the sample count, against the plain chain's autograd on the GPU.

The training step picks its kernel by the number of samples per ray (k_train_pair with one,
two, four or eight 64-sample chunks per ray up to 512, k_train_long above), and the render
kernels stream 64-sample groups; the counts on either side of each boundary (1, 63-65,
127-129, 255-257, 511-513 and 1024) go through pxl_train_step (direct scatter, and corner
blocks for trilinear) and through render_rays' forward and hand-written backward, with
occupancy and early termination in some cases.  Bars: RGBA within 1e-5 absolute, the grid
gradient within 1e-4 relative (of the largest element, and per element above 1% of it).
"""
import pytest
import torch

from oracle import plain
from plenoxels_amd import MarchConfig, corner_blocks, occupancy_dilate, pack_occupancy, render_rays, train_step
from util import grad_close, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
COUNTS = [1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1024]


def _case(S, interp):
    X, R = 14, 130
    grid, o, d, gmin, pd = small_scene(900 + S, X=X, R=R, pd=0.11)
    near = 1.0
    delta = (1.4 + (X - 1) * pd + 0.6 - near) / S  # the march crosses the whole grid whatever S
    occupied = S % 3 == 0
    t_stop = 1e-3 if S % 2 else 0.0
    cells = bits = None
    if occupied:
        occ = torch.rand(X, X, X, generator=torch.Generator().manual_seed(S)) < 0.4
        cells = plain.dilate_cells(occ) if interp == "trilinear" else occ
        bits = pack_occupancy(occ.to(DEV))
        if interp == "trilinear":
            bits = occupancy_dilate(bits, (X, X, X))
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(S + 1))
    kw = dict(num_samples=S, delta_step=delta, gmin=gmin.to(DEV), points_distance=pd, interp=interp, near=near,
              t_stop=t_stop, div_mode="reciprocal", occupancy=None if cells is None else cells.to(DEV))
    g = grid.to(DEV).requires_grad_(True)
    want = plain.render(g, o.to(DEV), d.to(DEV), **kw)
    plain.mse_loss(want, tgt.to(DEV)).backward()
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, interp, near, t_stop)
    return grid, o, d, tgt, cfg, bits, want.detach(), g.grad.detach(), occupied, t_stop


@pytest.mark.parametrize("S", COUNTS)
@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_train_step_at_every_sample_count(S, interp):
    grid, o, d, tgt, cfg, bits, want, gwant, occupied, t_stop = _case(S, interp)
    say("train step S", S, interp, "occupancy" if occupied else "dense", "t_stop", t_stop)
    R = o.shape[0]
    modes = [False, True] if interp == "trilinear" else [False]
    for blocks in modes:
        out = torch.empty(R, 4, device=DEV)
        gg = torch.zeros(grid.shape, device=DEV)
        gd = grid.to(DEV)
        train_step(gd, gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=bits, out=out,
                   blocks=corner_blocks(gd) if blocks else None, overwrite=blocks)
        err = (out - want).abs().max().item()
        ok, info = grad_close(gg.cpu(), gwant.cpu())
        say("  ", "blocks" if blocks else "direct", "rgba", err, "grad", info)
        assert err <= 1e-5
        assert ok, info


@pytest.mark.parametrize("S", COUNTS)
@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_render_fwd_bwd_at_every_sample_count(S, interp):
    grid, o, d, tgt, cfg, bits, want, gwant, occupied, t_stop = _case(S, interp)
    say("render fwd/bwd S", S, interp, "occupancy" if occupied else "dense", "t_stop", t_stop)
    leaf = grid.to(DEV).requires_grad_(True)
    out = render_rays(leaf, o.to(DEV), d.to(DEV), num_samples=S, delta_step=cfg.delta_step, grid_min=cfg.grid_min,
                      points_distance=cfg.points_distance, interp=interp, near=cfg.near, t_stop=t_stop,
                      occupancy=bits)
    plain.mse_loss(out, tgt.to(DEV)).backward()
    err = (out.detach() - want).abs().max().item()
    ok, info = grad_close(leaf.grad.cpu(), gwant.cpu())
    say("  rgba", err, "grad", info)
    assert err <= 1e-5
    assert ok, info
