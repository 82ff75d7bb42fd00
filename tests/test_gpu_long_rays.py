"""Training steps with more than 512 samples per ray (the reference's default --sample_num is 600,
scripts/train.py:218): pxl_train_step runs them one ray per wavefront with each ray's group
state in LDS (k_train_long).  Parity against the plain chain's autograd on the same inputs, at
the bars of the other render tests (RGBA 1e-5 absolute, gradient 1e-4 relative), with and
without corner blocks, early termination, occupancy, the beta term and half storage.

Also the regression test of the run-flushed scatter (taken by the kernels for more than 256
samples per ray when consecutive samples share their cells): samples placed exactly on integer
grid coordinates, where a cell's ceil index equals its floor index, must not be merged into
the previous cell's run.
"""
import pytest
import torch

from oracle import plain
from plenoxels_amd import MarchConfig, corner_blocks, occupancy_dilate, pack_occupancy, render_rays, train_step
from util import grad_close, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _plain_step(grid, o, d, tgt, beta=0.0, **kw):
    g = grid.clone().requires_grad_(True)
    out = plain.render(g, o, d, **kw)
    loss = ((out - tgt) ** 2).mean()
    if beta:
        S = kw["num_samples"]
        n = plain.normalize_samples(plain.sample_positions(o, d, S, kw["delta_step"], kw.get("near", 0.0)),
                                    kw["gmin"], kw["points_distance"], kw["div_mode"])
        vals, _ = (plain.nearest_values if kw["interp"] == "nearest" else plain.trilinear_values)(
            n, g, kw.get("occupancy"))
        a = vals.reshape(o.shape[0], S, 4)[..., -1]
        if kw.get("t_stop", 0.0) > 0:  # terminated samples count as a = 0, as in the compositing
            with torch.no_grad():
                T = torch.cat([torch.ones_like(a[:, :1]), 1 - a[:, :-1]], dim=1).cumprod(dim=1)
            a = a * (T >= kw["t_stop"])
        loss = loss + beta * (torch.log(a + 1e-4) - torch.log(1 - a + 1e-4)).mean()
    loss.backward()
    return out.detach(), g.grad


def _fused_step(grid, o, d, tgt, cfg, blocks=False, occupancy=None, beta=0.0, dtype=torch.float32):
    R = o.shape[0]
    gd = grid.to(dtype).to(DEV)
    gg = torch.zeros(grid.shape, device=DEV)
    out = torch.empty(R, 4, device=DEV)
    sq = torch.empty(R, device=DEV)
    rb = torch.empty(R, device=DEV) if beta else None
    train_step(gd, gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out, ray_sq_err=sq,
               occupancy=occupancy, beta=beta, ray_beta=rb, blocks=corner_blocks(gd) if blocks else None,
               overwrite=True)
    return out.cpu(), gg.cpu(), sq.cpu()


@pytest.mark.parametrize("interp,blocks", [("nearest", False), ("trilinear", False), ("trilinear", True)])
@pytest.mark.parametrize("S,t_stop,beta", [(600, 0.0, 0.0), (600, 0.02, 5e-3), (1100, 0.0, 5e-3)])
def test_long_ray_train_step_vs_plain(interp, blocks, S, t_stop, beta):
    say("long rays", interp, "blocks" if blocks else "direct", S, t_stop, beta)
    grid, o, d, gmin, pd = small_scene(900 + S, X=24, R=333)
    R = o.shape[0]
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(901))
    delta = 4.5 / S  # every ray crosses the grid within its samples
    kw = dict(num_samples=S, delta_step=delta, points_distance=pd, interp=interp, gmin=gmin,
              div_mode="reciprocal", t_stop=t_stop)
    want, gwant = _plain_step(grid, o, d, tgt, beta=beta, **kw)
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, interp, t_stop=t_stop)
    out, gg, sq = _fused_step(grid, o, d, tgt, cfg, blocks=blocks, beta=beta)
    err = (out - want).abs().max().item()
    ok, info = grad_close(gg, gwant)
    say("  rgba max err", err, "grad", info)
    assert err <= 1e-5
    assert abs(sq.sum().item() / (4 * R) - ((want - tgt) ** 2).mean().item()) <= 1e-6
    assert ok, info
    # the forward kernel (pxl_render_fwd) gives the same bits as the train step's forward sweep
    ref = render_rays(grid.to(DEV), o.to(DEV), d.to(DEV), num_samples=S, delta_step=delta,
                      grid_min=tuple(gmin.tolist()), points_distance=pd, interp=interp, t_stop=t_stop).cpu()
    assert torch.equal(ref, out)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_long_ray_half_storage_with_occupancy(dtype):
    say("long rays half storage", dtype)
    grid, o, d, gmin, pd = small_scene(950, X=24, R=256)
    R = o.shape[0]
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(951))
    occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(952)) < 0.2
    cells = plain.dilate_cells(occ)
    bits = occupancy_dilate(pack_occupancy(occ.to(DEV)), grid.shape[:3])
    rounded = grid.to(dtype).float()
    S, delta = 700, 4.5 / 700
    kw = dict(num_samples=S, delta_step=delta, points_distance=pd, interp="trilinear", gmin=gmin,
              div_mode="reciprocal", occupancy=cells)
    want, gwant = _plain_step(rounded, o, d, tgt, **kw)
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, "trilinear")
    out, gg, _ = _fused_step(grid, o, d, tgt, cfg, occupancy=bits, dtype=dtype)
    ok, info = grad_close(gg, gwant)
    say("  rgba max err", (out - want).abs().max().item(), "grad", info)
    assert (out - want).abs().max().item() <= 1e-5
    assert ok, info


def _integer_rays(X, pd, gmin, R, seed):
    """Rays along +z through the grid whose samples land on exact grid-index coordinates every
    fourth sample (delta = pd / 4, origin z on the lattice), x and y on the lattice for some rays
    and between lattice points for others; every quantity is a short binary fraction, so the
    kernel's and the oracle's float32 arithmetic place samples exactly there."""
    g = torch.Generator().manual_seed(seed)
    ix = torch.randint(1, X - 2, (R, 2), generator=g).float()
    frac = torch.tensor([0.0, 0.25, 0.5, 0.75])[torch.randint(0, 4, (R, 2), generator=g)]
    o = torch.empty(R, 3)
    o[:, :2] = gmin[:2] + (ix + frac) * pd
    o[:, 2] = gmin[2] - 2 * pd  # two cells below the grid: sample 8 is z index 0
    d = torch.zeros(R, 3)
    d[:, 2] = 1.0
    return o, d


# Z = 7: 28 in-grid samples per ray, so two rays share a wavefront's group (the pair kernel's
# half-wave path); Z = 10: 40 per ray, one ray after the other
@pytest.mark.parametrize("Z", [7, 10])
@pytest.mark.parametrize("S", [400, 600])
@pytest.mark.parametrize("blocks", [False, True])
def test_run_flush_keeps_integer_samples_apart(S, blocks, Z):
    say("run flush at integer coordinates", S, "blocks" if blocks else "direct", "Z", Z)
    X, pd = 10, 0.25
    g = torch.Generator().manual_seed(960)
    grid = torch.rand(X, X, Z, 4, generator=g) * 0.9 + 0.05
    grid[..., 3] *= 0.05  # light: every ray keeps its samples to the far side
    gi, _, _, _ = plain.generate_grid(X, X, Z, points_distance=pd)
    gmin = plain.grid_min(gi)
    o, d = _integer_rays(X, pd, gmin, 200, 961)
    R = o.shape[0]
    delta = pd / 4
    n = plain.normalize_samples(plain.sample_positions(o, d, S, delta), gmin, pd, "reciprocal").reshape(R, S, 3)
    on_lattice = (n[..., 2] == torch.round(n[..., 2])) & (n[..., 2] >= 0) & (n[..., 2] < Z)
    assert int(on_lattice.sum()) >= R * (Z - 1)  # the case under test occurs on every ray
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(962))
    kw = dict(num_samples=S, delta_step=delta, points_distance=pd, interp="trilinear", gmin=gmin,
              div_mode="reciprocal")
    want, gwant = _plain_step(grid, o, d, tgt, **kw)
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, "trilinear")
    out, gg, _ = _fused_step(grid, o, d, tgt, cfg, blocks=blocks)
    ok, info = grad_close(gg, gwant)
    say("  rgba max err", (out - want).abs().max().item(), "grad", info)
    assert (out - want).abs().max().item() <= 1e-5
    assert ok, info
