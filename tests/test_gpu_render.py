"""GPU parity of the fused kernels against the oracle, on the same seeded inputs.

Bars: rendered RGBA within 1e-5 absolute of the plain PyTorch path; grid gradients within
1e-4 relative (of the largest element, and per element above 1% of it) of the plain path's
autograd gradient and of the FP64 C oracle.
"""
import numpy as np
import pytest
import torch

from oracle import cref
from oracle import plain
from plenoxels_amd import CameraRays, MarchConfig, _lib, camera_rays, render_rays, train_step
from util import grad_close, random_cameras, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _oracle(grid, o, d, tgt, **kw):
    g = grid.clone().requires_grad_(True)
    out = plain.render(g, o, d, **kw)
    ((out - tgt) ** 2).mean().backward()
    return out.detach(), g.grad


def test_wave_primitives():
    say("wave primitives")
    g = torch.Generator().manual_seed(0)
    x = torch.rand(64, generator=g) * 0.5 + 0.75
    aff = torch.rand(5, 64, generator=g)
    inp = torch.cat([x, aff.reshape(-1)]).to(DEV)
    out = torch.empty(7 * 64, device=DEV)
    _lib.check(_lib.probe_lib().pxl_probe_wave_ops(_lib.ptr(inp), _lib.ptr(out), _lib.stream_handle(DEV)))
    out = out.cpu().double().reshape(7, 64)
    xd = x.double()
    incl = torch.cumprod(xd, 0)
    assert torch.allclose(out[0], incl, rtol=1e-6)
    assert torch.allclose(out[1], torch.cat([torch.ones(1, dtype=torch.float64), incl[:-1]]), rtol=1e-6)
    assert torch.allclose(out[2], xd.sum().expand(64), rtol=1e-6)
    a = aff.double()
    A0, A3, B = torch.zeros(64, dtype=torch.float64), torch.zeros(64, dtype=torch.float64), torch.zeros(64, dtype=torch.float64)
    acc0, acc3, accb = 0.0, 0.0, 1.0  # composition f_l o ... o f_63 as (a, b)
    for l in range(63, -1, -1):
        acc0 = a[0, l] + a[4, l] * acc0
        acc3 = a[3, l] + a[4, l] * acc3
        accb = a[4, l] * accb
        A0[l], A3[l], B[l] = acc0, acc3, accb
    assert torch.allclose(out[3], A0, rtol=1e-5)
    assert torch.allclose(out[4], A3, rtol=1e-5)
    assert torch.allclose(out[5], B, rtol=1e-5, atol=1e-30)
    assert torch.equal(out[6, :63].float(), x[1:])
    assert out[6, 63] == -1.0


def test_torch_gpu_scalar_division_is_reciprocal_multiply():
    """Pins oracle.plain's div_mode='reciprocal': ATen's GPU tensor / python-float computes
    x * float32(1/float32(pd))."""
    say("torch GPU division semantics")
    g = torch.Generator().manual_seed(1)
    x = (torch.rand(1 << 20, generator=g) * 8 - 4)
    for pd in (0.0125, 0.1, 3 / 255, 3 / 127, 0.3):
        got = (x.to(DEV) / pd).cpu()
        assert torch.equal(got, x * plain.reciprocal_f32(pd)), pd


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
@pytest.mark.parametrize("div_mode", ["reciprocal", "divide"])
def test_render_fwd_bwd_vs_plain(interp, div_mode):
    say("render fwd/bwd", interp, div_mode)
    grid, o, d, gmin, pd = small_scene(10)
    R = o.shape[0]
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(11))
    kw = dict(num_samples=96, delta_step=0.04, points_distance=pd, interp=interp)
    want, gwant = _oracle(grid, o, d, tgt, gmin=gmin, div_mode=div_mode, **kw)
    gd = grid.to(DEV).requires_grad_(True)
    out = render_rays(gd, o.to(DEV), d.to(DEV), grid_min=gmin.tolist(), div_mode=div_mode, **kw)
    ((out - tgt.to(DEV)) ** 2).mean().backward()
    err = (out.detach().cpu() - want).abs().max().item()
    say("  rgba max abs err", err)
    assert err <= 1e-5
    ok, info = grad_close(gd.grad.cpu(), gwant)
    say("  grad (max err, scale, per-elem rel)", info)
    assert ok
    # and against the FP64 C oracle
    up = (2 * (out.detach().cpu().double() - tgt.double()) / (4 * R)).numpy()
    rgba64, g64, _ = cref.render(grid.numpy(), o.numpy(), d.numpy(), grad_rgba=up, gmin=gmin.tolist(),
                                 div_mode=div_mode, **kw)
    assert np.abs(out.detach().cpu().double().numpy() - rgba64).max() <= 1e-5
    ok, info = grad_close(gd.grad.cpu(), torch.from_numpy(g64))
    say("  grad vs fp64 (max err, scale, per-elem rel)", info)
    assert ok


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_train_step_equals_autograd_pair(interp):
    say("train_step vs pair", interp)
    grid, o, d, gmin, pd = small_scene(20)
    R = o.shape[0]
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(21))
    kw = dict(num_samples=150, delta_step=0.026, points_distance=pd, interp=interp)
    want, gwant = _oracle(grid, o, d, tgt, gmin=gmin, div_mode="reciprocal", **kw)
    cfg = MarchConfig(150, 0.026, tuple(gmin.tolist()), pd, interp)
    gd = grid.to(DEV)
    gg = torch.zeros_like(gd)
    out = torch.empty(R, 4, device=DEV)
    sq = torch.empty(R, device=DEV)
    train_step(gd, gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out, ray_sq_err=sq)
    assert (out.cpu() - want).abs().max().item() <= 1e-5
    loss = sq.sum().item() / (4 * R)
    assert abs(loss - ((want - tgt) ** 2).mean().item()) <= 1e-6
    ok, info = grad_close(gg.cpu(), gwant)
    say("  grad", info)
    assert ok


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_half_storage_vs_plain_on_rounded_grid(dtype, interp):
    say("half storage", dtype, interp)
    grid, o, d, gmin, pd = small_scene(30)
    R = o.shape[0]
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(31))
    rounded = grid.to(dtype).float()
    kw = dict(num_samples=96, delta_step=0.04, points_distance=pd, interp=interp)
    want, gwant = _oracle(rounded, o, d, tgt, gmin=gmin, div_mode="reciprocal", **kw)
    cfg = MarchConfig(96, 0.04, tuple(gmin.tolist()), pd, interp)
    gg = torch.zeros(grid.shape, device=DEV)
    out = torch.empty(R, 4, device=DEV)
    train_step(grid.to(dtype).to(DEV), gg, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV),
               targets=tgt.to(DEV), out=out)
    assert (out.cpu() - want).abs().max().item() <= 1e-5
    ok, info = grad_close(gg.cpu(), gwant)
    say("  grad", info)
    assert ok


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_occupancy_skip_vs_oracle(interp):
    say("occupancy", interp)
    grid, o, d, gmin, pd = small_scene(40)
    R = o.shape[0]
    occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(41)) < 0.4
    bits = plain.occupancy_to_bits(occ)
    bits_dev = bits.to(DEV)
    if interp == "trilinear":
        from plenoxels_amd import occupancy_dilate
        cells = occupancy_dilate(bits_dev, grid.shape[:3])
        want_cells = plain.occupancy_to_bits(plain.dilate_cells(occ))
        assert torch.equal(cells.cpu(), want_cells)
        occ_used, bits_dev = plain.dilate_cells(occ), cells
    else:
        occ_used = occ
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(42))
    kw = dict(num_samples=96, delta_step=0.04, points_distance=pd, interp=interp)
    want, gwant = _oracle(grid, o, d, tgt, gmin=gmin, div_mode="reciprocal", occupancy=occ_used, **kw)
    gd = grid.to(DEV).requires_grad_(True)
    out = render_rays(gd, o.to(DEV), d.to(DEV), grid_min=gmin.tolist(), occupancy=bits_dev, **kw)
    ((out - tgt.to(DEV)) ** 2).mean().backward()
    assert (out.detach().cpu() - want).abs().max().item() <= 1e-5
    ok, info = grad_close(gd.grad.cpu(), gwant)
    say("  grad", info)
    assert ok


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_early_termination_vs_fp64_oracle(interp):
    say("t_stop", interp)
    grid, o, d, gmin, pd = small_scene(50, alpha_scale=0.5, p_opaque=0.0)
    R = o.shape[0]
    t_stop = 0.02
    kw = dict(num_samples=96, delta_step=0.04, points_distance=pd, interp=interp)
    rgba64, _, nproc = cref.render(grid.numpy(), o.numpy(), d.numpy(), gmin=gmin.tolist(),
                                   div_mode="reciprocal", t_stop=t_stop, **kw)
    assert (nproc < 96).mean() > 0.3  # termination actually happens
    out = render_rays(grid.to(DEV), o.to(DEV), d.to(DEV), grid_min=gmin.tolist(), t_stop=t_stop, **kw)
    err = np.abs(out.cpu().double().numpy() - rgba64).max(axis=1)
    # a ray whose transmittance lands within rounding of t_stop may keep one more sample
    # (contribution <= t_stop); all others agree to 1e-5
    assert (err <= 1e-5).mean() >= 0.99
    assert err.max() <= t_stop


def test_long_rays_and_saturation_vs_fp64():
    """S = 600 (the reference's own run, scripts/main.py) through the fwd/bwd pair, dense
    opacity near and at 1."""
    say("long rays S=600")
    grid, o, d, gmin, pd = small_scene(60, X=32, R=256, pd=0.08, alpha_scale=0.9, p_opaque=0.05)
    R = o.shape[0]
    kw = dict(num_samples=600, delta_step=0.0075, points_distance=pd, interp="nearest")
    gd = grid.to(DEV).requires_grad_(True)
    out = render_rays(gd, o.to(DEV), d.to(DEV), grid_min=gmin.tolist(), **kw)
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(61))
    ((out - tgt.to(DEV)) ** 2).mean().backward()
    up = (2 * (out.detach().cpu().double() - tgt.double()) / (4 * R)).numpy()
    rgba64, g64, _ = cref.render(grid.numpy(), o.numpy(), d.numpy(), grad_rgba=up, gmin=gmin.tolist(),
                                 div_mode="reciprocal", **kw)
    assert np.abs(out.detach().cpu().double().numpy() - rgba64).max() <= 1e-5
    ok, info = grad_close(gd.grad.cpu(), torch.from_numpy(g64))
    say("  grad", info)
    assert ok


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_transmittance_scan_fp64_on_gpu_near_saturation(interp):
    """The transmittance scan of long rays (S = 600) near opacity saturation, checked in FP64
    on the GPU: the reference chain's FP32 sample values (positions, lookup, clamp; the
    oracle's torch ops run on the device) composited in float64 on the device, against the
    fused FP32 kernel's RGBA and its transmittance 1 - A.  The only FP64 arithmetic on the GPU."""
    say("FP64 transmittance scan on the GPU", interp)
    grid, o, d, gmin, pd = small_scene(70, X=32, R=512, pd=0.08, alpha_scale=0.9, p_opaque=0.05)
    S, delta = 600, 0.0075
    kw = dict(num_samples=S, delta_step=delta, points_distance=pd, interp=interp)
    gdev, od, dd = grid.to(DEV), o.to(DEV), d.to(DEV)
    out = render_rays(gdev, od, dd, grid_min=gmin.tolist(), **kw).detach()
    with torch.no_grad():
        n = plain.normalize_samples(plain.sample_positions(od, dd, S, delta), gmin.to(DEV), pd, "reciprocal")
        lookup = plain.nearest_values if interp == "nearest" else plain.trilinear_values
        vals64 = lookup(n, gdev)[0].reshape(-1, S, 4).double()
        want = plain.compute_alpha_weighted_pixels(vals64)
        T64 = (1 - vals64[..., 3]).prod(-1)
    sat = (T64 < 1e-4).double().mean().item()
    say("  rays with final transmittance < 1e-4:", sat)
    assert sat > 0.3
    err = (out.double() - want).abs().max().item()
    terr = ((1 - out[:, 3].double()) - T64).abs().max().item()
    say("  rgba max abs err", err, "transmittance max abs err", terr)
    assert err <= 1e-5 and terr <= 1e-5


@pytest.mark.parametrize("norm_mode", ["gpu", "cpu"])
def test_camera_rays_bitwise_vs_plain(norm_mode):
    """pxl_camera_rays vs the reference's ray generation restated (oracle.plain.camera_rays) on
    the CPU with the same norm order, bitwise; in the default "gpu" mode also vs the reference's
    own ops run on the GPU (torch's GPU norm), which is what the reference computes on its
    default device."""
    say("camera rays", norm_mode)
    B, H, W, R = 5, 37, 37, 4096
    c2w = random_cameras(B, 70, y_scale=1.3)
    g = torch.Generator().manual_seed(71)
    imgs = torch.rand(B, H, W, 4, generator=g)
    cam = torch.randint(0, B, (R,), generator=g)
    uv = torch.rand(R, 2, generator=g)
    uv[:8] = torch.tensor([[0.0, 0.0], [1.0, 1.0], [0.5, 0.5], [0.5 / H, 0.5 / W],
                           [1.5 / H, 2.5 / W], [1.0, 0.0], [0.0, 1.0], [0.25, 0.75]])
    theta = 0.6911112070083618
    o_w, d_w, t_w = plain.camera_rays(imgs, c2w, theta, cam, uv[:, 0].contiguous(), uv[:, 1].contiguous(),
                                      norm_mode=norm_mode)
    cr = CameraRays(cam.to(torch.int32).to(DEV), uv.to(DEV), c2w.to(DEV), theta, imgs.to(DEV), norm_mode=norm_mode)
    o, d, t = camera_rays(cr)
    assert torch.equal(o.cpu(), o_w)
    assert torch.equal(t.cpu(), t_w)
    mism = int((d.cpu() != d_w).any(dim=1).sum())
    say("  direction mismatches vs the restatement (bitwise):", mism)
    assert mism == 0
    if norm_mode == "gpu":
        _, d_g, t_g = plain.camera_rays(imgs.to(DEV), c2w.to(DEV), theta, cam.to(DEV), uv[:, 0].contiguous().to(DEV),
                                        uv[:, 1].contiguous().to(DEV))
        mism = int((d.cpu() != d_g.cpu()).any(dim=1).sum())
        say("  direction mismatches vs torch's ops on the GPU (bitwise):", mism)
        assert mism == 0 and torch.equal(t.cpu(), t_g.cpu())


def test_torch_gpu_norm_order():
    """Pins oracle.plain.norm3's "gpu" order: torch's GPU x.norm(dim=-1) over a length-3 axis is
    sqrt((x*x + z*z) + y*y) with every op rounded to float32 -- for contiguous rows, for the
    reference's (B,R,3) direction tensor and for strided camera-matrix columns
    (src/ray_sampling.py:218,262); its CPU norm is the fma order, which differs on many rows."""
    say("torch GPU 3-norm order")
    g = torch.Generator().manual_seed(5)
    v = torch.randn(1 << 22, 3, generator=g) * torch.rand(1 << 22, 1, generator=g) * 10
    vd = v.to(DEV)
    got = vd.norm(dim=1).cpu()
    assert torch.equal(got, plain.norm3(v, "gpu"))
    assert torch.equal(plain.norm3(vd, "gpu").cpu(), got)
    diff_cpu = int((v.norm(dim=1) != got).sum())
    say("  rows where the CPU order differs:", diff_cpu)
    assert diff_cpu > 0  # the two orders are genuinely different
    m = torch.randn(257, 4, 4, generator=g)
    for col in range(3):  # strided (B,3) views, as camera_x_axis / camera_y_axis
        s = m.to(DEV)[:, :3, col]
        assert torch.equal(s.norm(dim=1).cpu(), plain.norm3(m[:, :3, col], "gpu"))
    d3 = v[: 100 * 4096].reshape(100, 4096, 3)
    assert torch.equal(d3.to(DEV).norm(dim=2).cpu(), plain.norm3(d3, "gpu"))


def test_pack_occupancy_matches_oracle_bits():
    say("pack occupancy")
    from plenoxels_amd import pack_occupancy
    occ = torch.rand(13, 17, 19, generator=torch.Generator().manual_seed(80)) < 0.3
    got = pack_occupancy(occ.to(DEV)).cpu()
    assert torch.equal(got, plain.occupancy_to_bits(occ))


@pytest.mark.parametrize("fused_camera", [False, True])
@pytest.mark.parametrize("t_stop", [0.0, 1e-4])
def test_sparse256_subset_train_step_vs_plain(t_stop, fused_camera):
    """The benchmark workload itself (256^3, ~5% occupancy, trilinear, camera-form rays,
    targets gathered from resident images) on a 2048-ray subset, fused train step vs the
    plain chain on the CPU with the same rays."""
    from plenoxels_amd import occupancy_dilate, pack_occupancy, scenes
    say("sparse256 subset t_stop", t_stop, "fused camera" if fused_camera else "separate ray pass")
    w = scenes.make_workload("sparse256", n_rays=1 << 18)
    n = 2048
    cam, uv = w.cam_ids[:n].contiguous(), w.uv[:n].contiguous()
    imgs = torch.rand(w.c2w.shape[0], 800, 800, 4, generator=torch.Generator().manual_seed(3))
    o, d, tgt = plain.camera_rays(imgs, w.c2w, w.camera_angle_x, cam.long(), uv[:, 0].contiguous(),
                                  uv[:, 1].contiguous(), norm_mode="gpu")
    occ_cells = plain.dilate_cells(w.occupancy)
    kw = dict(num_samples=w.num_samples, delta_step=w.delta_step, points_distance=w.points_distance,
              interp="trilinear", near=w.near)
    g = w.grid.clone().requires_grad_(True)
    want = plain.render(g, o, d, gmin=torch.tensor(w.grid_min), div_mode="reciprocal",
                        occupancy=occ_cells, t_stop=t_stop, **kw)
    ((want - tgt) ** 2).mean().backward()
    grid = w.grid.to(DEV)
    bits = occupancy_dilate(pack_occupancy(w.occupancy.to(DEV)), w.dims)
    assert torch.equal(bits.cpu(), plain.occupancy_to_bits(occ_cells))
    cr = CameraRays(cam.to(DEV), uv.to(DEV), w.c2w.to(DEV), w.camera_angle_x, imgs.to(DEV))
    cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance, "trilinear", w.near, t_stop)
    gg = torch.zeros_like(grid)
    out = torch.empty(n, 4, device=DEV)
    train_step(grid, gg, cfg=cfg, camera=cr, occupancy=bits, out=out, fused_camera=fused_camera)
    err = (out.cpu() - want.detach()).abs().max(dim=1).values
    say("  rgba: max err", err.max().item(), "rays > 1e-5:", int((err > 1e-5).sum()))
    ok, info = grad_close(gg.cpu(), g.grad)
    say("  grad", info)
    if t_stop == 0.0:
        assert err.max().item() <= 1e-5
        assert ok
    else:
        assert (err <= 1e-5).float().mean().item() >= 0.995


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_nan_cell_propagates_like_torch_clamp(interp):
    """torch.clamp keeps NaN, so a NaN cell read by a ray's in-grid sample makes that pixel
    NaN in the reference chain; the kernels must not hide it.  (The reference also gathers the
    wrapped cell of an out-of-grid sample and multiplies it by 0, so NaN * 0 can reach a pixel
    through a sample outside the grid; the kernels skip those samples.)"""
    say("nan propagation", interp)
    grid, o, d, gmin, pd = small_scene(100)
    grid[10:14, 10:14, :, 3] = float("nan")
    S, delta = 96, 0.04
    kw = dict(num_samples=S, delta_step=delta, points_distance=pd, interp=interp)
    want = plain.render(grid, o, d, gmin=gmin, div_mode="reciprocal", **kw)
    got = render_rays(grid.to(DEV), o.to(DEV), d.to(DEV), grid_min=gmin.tolist(), **kw).cpu()
    # in-grid reads only: zero the out-of-grid samples before compositing
    n = plain.normalize_samples(plain.sample_positions(o, d, S, delta), gmin, pd, "reciprocal")
    vals, mask = (plain.nearest_values if interp == "nearest" else plain.trilinear_values)(n, grid)
    vals = torch.where(mask.unsqueeze(-1), vals, torch.zeros_like(vals))
    want_in = plain.compute_alpha_weighted_pixels(vals.reshape(-1, S, 4))
    nan_in = torch.isnan(want_in).any(1)
    nan_got = torch.isnan(got).any(1)
    assert nan_in.any() and not nan_in.all()
    assert torch.equal(nan_got, nan_in)
    assert bool((nan_got <= torch.isnan(want).any(1)).all())
    ok = ~torch.isnan(want).any(1)
    assert (got[ok] - want[ok]).abs().max().item() <= 1e-5


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_empty_ragged_and_missing_batches(interp):
    """Batches of 0, 1, 63, 65 and 130 rays (the wave is 64 lanes) and a batch whose rays all
    miss the grid, through render_rays (autograd), train_step and, for trilinear, the corner
    blocks: RGBA and gradients vs the plain chain; an empty batch changes nothing."""
    from plenoxels_amd import corner_blocks
    say("empty / ragged / missing batches", interp)
    grid, o, d, gmin, pd = small_scene(30, R=130)
    tgt = torch.rand(130, 4, generator=torch.Generator().manual_seed(31))
    kw = dict(num_samples=96, delta_step=0.04, points_distance=pd, interp=interp)
    cfg = MarchConfig(96, 0.04, tuple(gmin.tolist()), pd, interp)
    gd = grid.to(DEV)
    cases = [(R, o[:R], d[:R], tgt[:R]) for R in (0, 1, 63, 65, 130)]
    cases.append(("miss", o[:70], -d[:70], tgt[:70]))  # from below the grid, pointing away
    for name, oc, dc, tc in cases:
        say("  batch", name)
        R = oc.shape[0]
        od, dd, td = oc.contiguous().to(DEV), dc.contiguous().to(DEV), tc.contiguous().to(DEV)
        gr = gd.clone().requires_grad_(True)
        out = render_rays(gr, od, dd, grid_min=gmin.tolist(), **kw)
        assert tuple(out.shape) == (R, 4)
        (out * td).sum().backward()
        gstep = torch.full(grid.shape, 7.0, device=DEV) if R == 0 else torch.zeros(grid.shape, device=DEV)
        sout = torch.empty(R, 4, device=DEV)
        sq = torch.empty(R, device=DEV)
        train_step(gd, gstep, cfg=cfg, origins=od, dirs=dd, targets=td, out=sout, ray_sq_err=sq)
        if R == 0:
            assert int(torch.count_nonzero(gr.grad)) == 0
            assert bool((gstep == 7.0).all()), "an empty batch wrote the gradient"
            continue
        want, gwant = _oracle(grid, oc, dc, tc, gmin=gmin, div_mode="reciprocal", **kw)
        assert (out.detach().cpu() - want).abs().max().item() <= 1e-5
        assert (sout.cpu() - want).abs().max().item() <= 1e-5
        ok, info = grad_close(gstep.cpu(), gwant)
        assert ok, info
        gp = grid.clone().requires_grad_(True)
        (plain.render(gp, oc, dc, gmin=gmin, div_mode="reciprocal", **kw) * tc).sum().backward()
        ok, info = grad_close(gr.grad.cpu(), gp.grad)
        assert ok, info
        if name == "miss":
            assert want.abs().max().item() == 0.0 and int(torch.count_nonzero(gwant)) == 0
            assert int(torch.count_nonzero(sout)) == 0 and int(torch.count_nonzero(gr.grad)) == 0
        if interp == "trilinear":
            blocks = corner_blocks(gd)
            gb = torch.full(grid.shape, 7.0, device=DEV)
            bout = torch.empty(R, 4, device=DEV)
            train_step(gd, gb, cfg=cfg, origins=od, dirs=dd, targets=td, out=bout, blocks=blocks,
                       overwrite=True)
            assert torch.equal(bout, sout)
            ok, info = grad_close(gb.cpu(), gwant)
            assert ok, info
            assert int(torch.count_nonzero(blocks)) == 0


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_early_termination_boundary_is_exact(interp):
    """Early termination keeps a sample whose transmittance equals t_stop and drops the first one
    below it.  Every cell has alpha exactly 0.5 and the rays run along +z through lattice points
    (delta = points_distance, all coordinates short binary fractions), so sample i inside the grid
    sees T = 2^-i exactly in any summation order: with t_stop = 2^-6 seven samples contribute
    (A = 1 - 2^-7 exactly), with the next float above 2^-6 six do (A = 1 - 2^-6); RGBA equals the
    oracle's and the gradient its autograd gradient."""
    say("early termination boundary", interp)
    X, pd = 12, 0.25
    g = torch.Generator().manual_seed(90)
    grid = torch.rand(X, X, X, 4, generator=g) * 0.8 + 0.1
    grid[..., 3] = 0.5
    gi, _, _, _ = plain.generate_grid(X, X, X, points_distance=pd)
    gmin = plain.grid_min(gi)
    R = 96
    ij = torch.randint(0, X, (R, 2), generator=g).float()
    o = torch.empty(R, 3)
    o[:, :2] = gmin[:2] + ij * pd
    o[:, 2] = gmin[2] - 3 * pd  # sample 3 is the first inside the grid
    d = torch.zeros(R, 3)
    d[:, 2] = 1.0
    tgt = torch.rand(R, 4, generator=g)
    S = 30
    for t_stop, n_kept in ((2.0 ** -6, 7), (float(np.nextafter(np.float32(2.0 ** -6), np.float32(1))), 6)):
        kw = dict(num_samples=S, delta_step=pd, points_distance=pd, interp=interp, t_stop=t_stop)
        want, gwant = _oracle(grid, o, d, tgt, gmin=gmin, div_mode="reciprocal", **kw)
        assert torch.equal(want[:, 3], torch.full((R,), 1 - 2.0 ** -n_kept))
        gd = grid.to(DEV).requires_grad_(True)
        out = render_rays(gd, o.to(DEV), d.to(DEV), grid_min=gmin.tolist(), **kw)
        ((out - tgt.to(DEV)) ** 2).mean().backward()
        out = out.detach().cpu()
        assert torch.equal(out[:, 3], want[:, 3]), (t_stop, out[:3, 3])
        assert (out - want).abs().max().item() <= 1e-6
        ok, info = grad_close(gd.grad.cpu(), gwant)
        assert ok, (t_stop, info)
        # the fused train step cuts at the same sample
        gg = torch.zeros(grid.shape, device=DEV)
        so = torch.empty(R, 4, device=DEV)
        train_step(grid.to(DEV), gg, cfg=MarchConfig(S, pd, tuple(gmin.tolist()), pd, interp, 0.0, t_stop),
                   origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=so)
        assert torch.equal(so.cpu()[:, 3], want[:, 3])
        ok, info = grad_close(gg.cpu(), gwant)
        assert ok, (t_stop, info)
