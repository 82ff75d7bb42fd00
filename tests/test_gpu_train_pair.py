"""The train step runs two rays per wavefront (k_train_pair): a pair whose rays keep at most 32
samples each shares one group, ray 0 in lanes 0-31 and ray 1 in lanes 32-63; other pairs run
one ray after the other.  Both forms must give, per ray, the bits of the forward kernel
(pxl_render_fwd, one ray per wave), and the gradient of the plain chain's autograd.  The
scenes here mix pairs of both forms (occupancy thinned to ~4, ~20 and ~70 kept samples per
ray), odd ray counts (a last wave with one ray), early termination inside a half, and the
beta term.
"""
import pytest
import torch

from oracle import plain
from plenoxels_amd import MarchConfig, pack_occupancy, occupancy_dilate, render_rays, train_step
from util import grad_close, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _scene(interp, density, R, seed):
    grid, o, d, gmin, pd = small_scene(seed, X=24, R=R)
    occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(seed + 1)) < density
    cells = plain.dilate_cells(occ) if interp == "trilinear" else occ
    bits = pack_occupancy(occ.to(DEV))
    if interp == "trilinear":
        bits = occupancy_dilate(bits, grid.shape[:3])
    return grid, o, d, gmin, pd, cells, bits


def _kept_counts(o, d, gmin, pd, cells, interp, S, delta):
    n = plain.normalize_samples(plain.sample_positions(o, d, S, delta), gmin, pd, "reciprocal")
    n = n.reshape(o.shape[0], S, 3)
    X = cells.shape[0]
    if interp == "nearest":
        idx = torch.round(n)
        inb = ((idx >= 0) & (idx < X)).all(2)
    else:
        idx = torch.floor(n)
        inb = ((n >= 0) & (n < X)).all(2)
    idx = idx.long().clamp(0, X - 1)
    return (inb & cells[idx[..., 0], idx[..., 1], idx[..., 2]]).sum(1)


# occupancy densities giving mostly shared-group pairs, a mix, and no shared group (every
# in-grid sample kept); trilinear reads the mask dilated over the 8 corners
DENSITY = {"nearest": {"sparse": 0.05, "mixed": 0.3, "dense": 1.0},
           "trilinear": {"sparse": 0.006, "mixed": 0.04, "dense": 1.0}}


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
@pytest.mark.parametrize("level,t_stop", [("sparse", 0.0), ("mixed", 0.0), ("mixed", 0.05), ("dense", 0.05)])
def test_pairs_match_one_ray_kernel_and_autograd(interp, level, t_stop):
    density = DENSITY[interp][level]
    say("train pairs", interp, level, t_stop)
    R, S, delta = 777, 150, 0.026  # odd: the last wave holds one ray
    grid, o, d, gmin, pd, cells, bits = _scene(interp, density, R, 300 + len(level))
    kept = _kept_counts(o, d, gmin, pd, cells, interp, S, delta)
    pairs = kept[: R - 1].reshape(-1, 2)
    half = ((pairs <= 32).all(1)).float().mean().item()
    say("  kept/ray mean", kept.float().mean().item(), "pairs in one group", half)
    if level != "dense":
        assert half > 0.0, "no pair takes the shared-group form"
    if level != "sparse":
        assert half < 1.0, "every pair takes the shared-group form"
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, interp, 0.0, t_stop)
    gd = grid.to(DEV)
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(7))
    out = torch.empty(R, 4, device=DEV)
    sq = torch.empty(R, device=DEV)
    gg = torch.zeros_like(gd)
    train_step(gd, gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=bits, out=out,
               ray_sq_err=sq)
    # forward: the bits of the one-ray forward kernel, every ray
    g_ref = gd.clone().requires_grad_(True)
    want = render_rays(g_ref, o.to(DEV), d.to(DEV), num_samples=S, delta_step=delta, grid_min=tuple(gmin.tolist()),
                       points_distance=pd, interp=interp, t_stop=t_stop, occupancy=bits)
    assert torch.equal(out, want.detach())
    e = (want.detach() - tgt.to(DEV))
    sq_want = ((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]) + e[:, 3] * e[:, 3]
    assert torch.equal(sq, sq_want)
    # gradient: the forward kernel's autograd backward of the same MSE, at the usual 1e-4 bar
    # (FP32 atomics in another order; dense cells sum hundreds of terms)
    ((want - tgt.to(DEV)) ** 2).mean().backward()
    ok, info = grad_close(gg.cpu(), g_ref.grad.cpu())
    say("  grad vs render autograd", info)
    assert ok, info


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_pairs_beta_term_vs_plain(interp):
    """The beta regulariser (scripts/train.py:170-177, mean over all samples of
    log(a + 1e-4) - log(1 - a + 1e-4); skipped samples have a = 0) through shared-group pairs."""
    from plenoxels_amd.render import beta_term_zero
    say("train pairs beta", interp)
    R, S, delta, beta = 301, 150, 0.026, 5e-3
    grid, o, d, gmin, pd, cells, bits = _scene(interp, DENSITY[interp]["mixed"], R, 330)
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(8))
    go = grid.clone().requires_grad_(True)
    n = plain.normalize_samples(plain.sample_positions(o, d, S, delta), gmin, pd, "reciprocal")
    vals, _ = (plain.nearest_values if interp == "nearest" else plain.trilinear_values)(n, go, cells)
    vals = vals.reshape(R, S, 4)
    out_w = plain.compute_alpha_weighted_pixels(vals)
    a = vals[..., -1]
    bl = beta * (torch.log(a + 1e-4) - torch.log(1 - a + 1e-4)).mean()
    (plain.mse_loss(out_w, tgt) + bl).backward()
    gg = torch.zeros(grid.shape, device=DEV)
    rb = torch.empty(R, device=DEV)
    train_step(grid.to(DEV), gg, cfg=MarchConfig(S, delta, tuple(gmin.tolist()), pd, interp), origins=o.to(DEV),
               dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=bits, beta=beta, ray_beta=rb)
    bl_k = beta * (rb.sum().item() + R * S * beta_term_zero()) / (R * S)
    assert abs(bl_k - float(bl.detach())) <= 1e-5 * abs(float(bl.detach()))
    ok, info = grad_close(gg.cpu(), go.grad)
    say("  grad", info)
    assert ok, info


@pytest.mark.parametrize("S,storage", [(50, torch.float32), (100, torch.bfloat16), (400, torch.float32),
                                       (400, torch.float16)])
def test_pairs_every_chunk_count(S, storage):
    """Train kernel instantiations for S <= 64, 128, 512 (one, two and eight 64-sample chunks
    per ray) and half storage: per-ray outputs equal the one-ray forward kernel's bits, and the
    gradient the forward kernel's autograd backward."""
    say("train pairs S", S, storage)
    R, delta = 301, 3.9 / S  # the march crosses the whole grid whatever S
    grid, o, d, gmin, pd, cells, bits = _scene("trilinear", 0.02, R, 400 + S)
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, "trilinear", 0.0, 1e-3)
    gd = grid.to(storage).to(DEV)
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(9))
    out = torch.empty(R, 4, device=DEV)
    gg = torch.zeros(grid.shape, device=DEV)
    train_step(gd, gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=bits, out=out)
    # the one-ray kernels through the C ABI, FP32 gradient whatever the storage (an autograd
    # leaf of half dtype would round its gradient to that dtype)
    import ctypes
    from plenoxels_amd import _lib
    from plenoxels_amd.render import _grid_struct, _rays_struct
    L = _lib.lib()
    od, dd = o.to(DEV).contiguous(), d.to(DEV).contiguous()
    g, m, r = _grid_struct(gd, bits), cfg.to_struct(), _rays_struct(od, dd)
    aux = torch.empty(R, int(L.pxl_aux_floats(S)), device=DEV)
    want = torch.empty(R, 4, device=DEV)
    st = _lib.stream_handle(DEV)
    _lib.check(L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), _lib.ptr(want), _lib.ptr(aux), st))
    assert torch.equal(out, want)
    grad_rgba = ((2.0 / (4 * R)) * (want - tgt.to(DEV))).contiguous()
    gb = torch.zeros(grid.shape, device=DEV)
    _lib.check(L.pxl_render_bwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), _lib.ptr(grad_rgba),
                                _lib.ptr(aux), _lib.ptr(gb), st))
    ok, info = grad_close(gg.cpu(), gb.cpu())
    say("  grad vs render_bwd", info)
    assert ok, info
