"""GPU parity of the per-function drop-ins (plenoxels_amd.ray_sampling / grid_functions /
optim / train) against the oracle on the same inputs.

Bit-exact where the reference's arithmetic is plain float32 elementwise work (sample
positions, normalisation, voxel choice, corner indices, gathers, lerp order, coordinates);
RGBA within 1e-5 absolute; gradients within 1e-4 relative (of the largest element, and per
element above 1% of it); Adam within a few float32 ulps of torch.optim.Adam.
"""
import math

import numpy as np
import pytest
import torch

from oracle import plain
from plenoxels_amd import grid_functions as gf
from plenoxels_amd import ray_sampling as rs
from plenoxels_amd.optim import Adam, tv_loss
from util import grad_close, random_cameras, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _cams(B=3, H=16, seed=5):
    c2w = random_cameras(B, seed, radius=2.2, y_scale=1.2)
    imgs = torch.rand(B, H, H, 4, generator=torch.Generator().manual_seed(seed + 1))
    return c2w, imgs


def test_normalize_samples_bitwise():
    say("normalize_samples_for_indecies")
    gi, _, _, _ = plain.generate_grid(20, 21, 22, points_distance=0.0125)
    p = torch.rand(5000, 3, generator=torch.Generator().manual_seed(1)) * 0.6 - 0.3
    want = plain.normalize_samples(p, plain.grid_min(gi), 0.0125, "reciprocal")
    got = rs.normalize_samples_for_indecies(gi.to(DEV), p.to(DEV), 0.0125)
    assert torch.equal(got.cpu(), want)


@pytest.mark.parametrize("even", [False, True])
def test_sample_camera_rays_batched_bitwise(even):
    say("sample_camera_rays_batched even_spread", even)
    c2w, imgs = _cams()
    R, S, delta = 49, 7, 0.21
    torch.manual_seed(123)
    uv = torch.rand(3, R, 2, device=DEV).cpu() if not even else plain.even_spread_uniforms(R, 3)
    torch.manual_seed(123)
    smp, pix, cpos, dirs = rs.sample_camera_rays_batched(c2w.to(DEV), 0.69, imgs.to(DEV), R, S, delta,
                                                         even_spread=even, camera_ray=False, device="cuda")
    d_w, pix_w = plain.generate_rays_batched(imgs, R, c2w, 0.69, uv, norm_mode="gpu")
    o_w = torch.repeat_interleave(c2w[:, :3, 3], R, 0)
    smp_w = plain.sample_positions(o_w, d_w, S, delta)
    assert torch.equal(dirs.cpu(), d_w)
    assert torch.equal(pix.cpu(), pix_w)
    assert torch.equal(smp.cpu(), smp_w)
    assert torch.equal(cpos.cpu(), c2w[:, :3, 3])
    # the reference's own ops on the GPU (its default device) with the same RNG draw
    d_g, pix_g = plain.generate_rays_batched(imgs.to(DEV), R, c2w.to(DEV), 0.69, uv.to(DEV))
    smp_g = plain.sample_positions(o_w.to(DEV), d_g, S, delta)
    assert torch.equal(dirs.cpu(), d_g.cpu()) and torch.equal(pix.cpu(), pix_g.cpu())
    assert torch.equal(smp.cpu(), smp_g.cpu())
    with pytest.raises(NotImplementedError):
        rs.sample_camera_rays_batched(c2w.to(DEV), 0.69, imgs.to(DEV), R, S, delta, False, True)


def _points(seed, X=9, N=3000):
    g = torch.Generator().manual_seed(seed)
    n = torch.rand(N, 3, generator=g) * (X + 3) - 1.5
    ties = torch.tensor([[0.5, 1.5, 2.5], [-0.5, 8.5, 7.5], [8.49, -0.51, 3.0], [9.0, 0.0, 4.5]])
    return torch.cat([ties, n], 0)


def test_bounds_wrap_bitwise():
    say("find_out_of_bound / fix_out_of_bounds")
    grid = torch.zeros(9, 10, 11, 4)
    n = _points(2, X=10)
    idx = torch.round(n).long()
    assert torch.equal(gf.find_out_of_bound(idx.to(DEV), grid).cpu(), plain.find_out_of_bound(idx, (9, 10, 11)))
    assert torch.equal(gf.find_out_of_bound(n.to(DEV), grid).cpu(), plain.find_out_of_bound(n, (9, 10, 11)))
    w = idx.to(DEV)
    a, b, c = gf.fix_out_of_bounds(w, grid)
    assert torch.equal(w.cpu(), plain.wrap_indices(idx, (9, 10, 11)))
    assert torch.equal(a.cpu(), w[:, 0].cpu())


def test_nearest_voxels_fwd_bwd():
    say("get_nearest_voxels")
    g = torch.Generator().manual_seed(3)
    grid = torch.rand(9, 9, 9, 4, generator=g)
    n = _points(4)
    gd = grid.to(DEV).requires_grad_(True)
    vals, mask = gf.get_nearest_voxels(n.to(DEV), gd)
    idx = torch.round(n).long()
    m_w = plain.find_out_of_bound(idx, (9, 9, 9))
    w = plain.wrap_indices(idx, (9, 9, 9))
    go = grid.clone().requires_grad_(True)
    v_w = go[w[:, 0], w[:, 1], w[:, 2]]
    assert torch.equal(mask.cpu(), m_w)
    assert torch.equal(vals.detach().cpu(), v_w.detach())
    up = torch.randn(vals.shape, generator=g)
    (vals * up.to(DEV)).sum().backward()
    (v_w * up).sum().backward()
    ok, info = grad_close(gd.grad.cpu(), go.grad)
    assert ok, info


def test_corners_collect_trilinear():
    say("corners / collect / trilinear")
    g = torch.Generator().manual_seed(6)
    X = 9
    grid = torch.rand(X, X, X, 4, generator=g) * 1.2 - 0.1
    n = _points(7, X=X)
    c = gf.get_grid_points_indices(n.to(DEV))
    assert torch.equal(c.cpu(), plain.get_grid_points_indices(n))
    gd = grid.to(DEV).requires_grad_(True)
    cells, mask = gf.collect_cell_information_via_indices(n.to(DEV), gd)
    go = grid.clone().requires_grad_(True)
    cells_w, mask_w = plain.collect_cell_information_via_indices(n, go)
    assert torch.equal(mask.cpu(), mask_w)
    assert torch.equal(cells.detach().cpu(), cells_w.detach())
    up = torch.randn(cells.shape, generator=g)
    (cells * up.to(DEV)).sum().backward()
    (cells_w * up).sum().backward()
    assert grad_close(gd.grad.cpu(), go.grad)[0]
    # trilinear on wrapped corners (what the reference's callers pass)
    cw = plain.wrap_indices(plain.get_grid_points_indices(n).reshape(-1, 3), (X, X, X)).reshape(-1, 8, 3)
    gd2 = grid.to(DEV).requires_grad_(True)
    tri = gf.trilinear_interpolation(n.to(DEV), cw.to(DEV), gd2)
    go2 = grid.clone().requires_grad_(True)
    tri_w = plain.trilinear_interpolation(n, cw, go2)
    assert torch.equal(tri.detach().cpu(), tri_w.detach())
    up = torch.randn(tri.shape, generator=g)
    (tri * up.to(DEV)).sum().backward()
    (tri_w * up).sum().backward()
    ok, info = grad_close(gd2.grad.cpu(), go2.grad)
    assert ok, info


@pytest.mark.parametrize("S", [1, 64, 100, 600])
def test_compute_alpha_weighted_pixels_fwd_bwd(S):
    say("compute_alpha_weighted_pixels S", S)
    g = torch.Generator().manual_seed(8 + S)
    x = torch.rand(2, 5, S, 4, generator=g)
    x[..., 3] *= 0.2
    x[0, 1, S // 2, 3] = 1.0
    xd = x.to(DEV).requires_grad_(True)
    out = rs.compute_alpha_weighted_pixels(xd)
    xo = x.double().requires_grad_(True)
    want = plain.compute_alpha_weighted_pixels(xo)
    assert out.shape == (2, 5, 4)
    assert (out.detach().cpu().double() - want.detach()).abs().max().item() <= 1e-5
    up = torch.randn(out.shape, generator=g)
    (out * up.to(DEV)).sum().backward()
    (want * up.double()).sum().backward()
    ok, info = grad_close(xd.grad.cpu(), xo.grad)
    assert ok, info


def test_compute_alpha_weighted_pixels_empty_batch():
    say("compute_alpha_weighted_pixels on zero rays")
    xd = torch.rand(0, 16, 4, device=DEV).requires_grad_(True)
    out = rs.compute_alpha_weighted_pixels(xd)
    want = plain.compute_alpha_weighted_pixels(torch.rand(0, 16, 4))
    assert out.shape == want.shape == (0, 4)
    out.sum().backward()
    assert xd.grad is not None and xd.grad.shape == (0, 16, 4)
    # zero samples per ray: the reference's sums over an empty sample axis are zero
    xz = torch.rand(3, 0, 4, device=DEV).requires_grad_(True)
    out = rs.compute_alpha_weighted_pixels(xz)
    want = plain.compute_alpha_weighted_pixels(torch.rand(3, 0, 4))
    assert out.shape == want.shape == (3, 4)
    assert torch.equal(out.detach().cpu(), want)
    out.sum().backward()
    assert xz.grad is not None and xz.grad.shape == (3, 0, 4)


@pytest.mark.parametrize("k,s", [(3, 1), (5, 2), (13, 3), (33, 8)])
def test_average_pool3d_grid_fwd_bwd(k, s):
    say("average_pool3d_grid", k, s)
    g = torch.Generator().manual_seed(k)
    # every pooled axis > 1: the reference's .squeeze() (src/grid_functions.py:180) breaks otherwise
    grid = torch.rand(40, 37, 35, 4, generator=g) if k < 20 else torch.rand(64, 60, 70, 4, generator=g)
    gd = grid.to(DEV).requires_grad_(True)
    out = gf.average_pool3d_grid(gd, receptive_field_size=k, stride=s)
    go = grid.double().requires_grad_(True)
    want = plain.average_pool3d_grid(go, k, s)
    assert out.shape == want.shape
    assert (out.detach().cpu().double() - want.detach()).abs().max().item() <= 1e-6
    up = torch.randn(out.shape, generator=g)
    (out * up.to(DEV)).sum().backward()
    (want * up.double()).sum().backward()
    ok, info = grad_close(gd.grad.cpu(), go.grad)
    assert ok, info


def test_pool_backward_overwrite_and_accumulate():
    """pxl_avg_pool3d_bwd writes every cell: overwriting a NaN-filled buffer gives the bits of
    accumulating into zeros (cells outside every window: 0, as the plane x = 39 of 40 is with
    k = 9, s = 2), and accumulating onto a filled buffer adds to it."""
    say("pool backward modes")
    g = torch.Generator().manual_seed(21)
    dims, k, s = (40, 37, 35), 9, 2
    shape = tuple((d - k) // s + 1 for d in dims) + (4,)
    gout = torch.randn(shape, generator=g).to(DEV)
    base = torch.randn(*dims, 4, generator=g).to(DEV)
    z = torch.zeros(*dims, 4, device=DEV)
    gf.pool_backward_(gout, dims, k, s, z)
    o = torch.full((*dims, 4), float("nan"), device=DEV)
    gf.pool_backward_(gout, dims, k, s, o, accumulate=False)
    assert torch.equal(o, z)
    a = base.clone()
    gf.pool_backward_(gout, dims, k, s, a)
    assert torch.equal(a, base + z)
    assert float(z[39].abs().max()) == 0.0  # windows along x cover 0..38


def test_generate_grid_bitwise():
    say("generate_grid")
    for dims, pd in (((5, 6, 7), 0.25), ((64, 63, 2), 3 / 63)):
        want = plain.generate_grid(*dims, points_distance=pd)
        got = gf.generate_grid(*dims, points_distance=pd, device="cuda")
        for a, b in zip(got, want):
            assert torch.equal(a.detach().cpu(), b.detach())
        assert got[1].requires_grad


def test_tv_loss_fwd_bwd():
    say("tv_loss")
    g = torch.Generator().manual_seed(12)
    grid = torch.rand(17, 19, 21, 4, generator=g)
    gd = grid.to(DEV).requires_grad_(True)
    v = tv_loss(gd)
    go = grid.double().requires_grad_(True)
    w = plain.tv_loss(go)
    assert abs(float(v.detach()) - float(w.detach())) <= 1e-6 * float(w.detach())
    (3.0 * v).backward()
    (3.0 * w).backward()
    ok, info = grad_close(gd.grad.cpu(), go.grad)
    assert ok, info


def test_adam_matches_torch():
    say("Adam")
    g = torch.Generator().manual_seed(13)
    p0 = torch.randn(6, 7, 8, 4, generator=g)
    grads = [torch.randn(6, 7, 8, 4, generator=g) * 10 ** (-i) for i in range(4)]
    pd = p0.clone().to(DEV).requires_grad_(True)
    gabs = torch.zeros_like(pd)
    half = torch.empty_like(pd, dtype=torch.bfloat16)
    opt = Adam([pd], lr=0.0075, grad_abs={pd: gabs}, half_copy={pd: half})
    pr = p0.clone().requires_grad_(True)
    ref = torch.optim.Adam([pr], lr=0.0075)
    for gr in grads:
        pd.grad = gr.to(DEV)
        opt.step()
        pr.grad = gr.clone()
        ref.step()
    assert torch.allclose(pd.detach().cpu(), pr.detach(), rtol=1e-6, atol=1e-6)
    st = opt.state[pd]
    for key in ("exp_avg", "exp_avg_sq"):  # within a rounding step of the largest moment
        a, b = st[key].cpu(), ref.state[pr][key]
        assert ((a - b).abs() <= 1e-6 * b.abs() + 1e-7 * b.abs().max()).all(), key
    assert torch.allclose(gabs.cpu(), sum(x.abs() for x in grads), rtol=1e-6)
    assert torch.equal(half.cpu(), pd.detach().cpu().to(torch.bfloat16))


@pytest.mark.parametrize("half_dtype", [None, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("zero_grad", [False, True])
def test_fused_tv_adam_equals_tv_backward_then_adam(half_dtype, zero_grad):
    """pxl_adam_tv_step: the bits of pxl_tv_bwd into the gradient followed by pxl_adam_step,
    over three steps on an odd-shaped grid (faces, edges and corners of the stencil), with the
    parameter ping-ponged between two buffers; the gradient buffer is left unchanged, or with
    zero_grad all zero afterwards, written only where it was nonzero (cells of -0.0 keep
    their sign)."""
    from plenoxels_amd.optim import adam_step_, adam_tv_step_, tv_backward_, tv_sumsq
    say("fused TV + Adam", half_dtype)
    g = torch.Generator().manual_seed(14)
    dims = (9, 6, 11)
    p0 = torch.randn(*dims, 4, generator=g).to(DEV)
    grads = []
    for i in range(3):  # built on the host, where -0.0 survives the fill, then copied bitwise
        gr = torch.randn(*dims, 4, generator=g) * 10 ** (-i)
        gr[:, 2] = 0.0  # cells the scatter never reached
        gr[0] = -0.0
        assert bool(torch.signbit(gr[0]).all())
        grads.append(gr.to(DEV))
    a_p, a_m, a_v, a_ga = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0), torch.zeros_like(p0)
    b_p, b_alt, b_m, b_v, b_ga = p0.clone(), torch.empty_like(p0), torch.zeros_like(p0), torch.zeros_like(p0), \
        torch.zeros_like(p0)
    a_h = None if half_dtype is None else torch.empty_like(p0, dtype=half_dtype)
    b_h = None if half_dtype is None else torch.empty_like(p0, dtype=half_dtype)
    for t, gr in enumerate(grads, start=1):
        # one TV sum for both paths (its FP64 atomics make repeated sums differ in the last bit)
        assert torch.equal(a_p, b_p)
        s_t = tv_sumsq(a_p)
        # separate passes on a copy of the gradient
        ga = gr.clone()
        tv_backward_(a_p, s_t, 0.37, ga)
        adam_step_(a_p, ga, a_m, a_v, t, lr=0.01, grad_abs=a_ga, half_out=a_h)
        # fused pass
        gb = gr.clone()
        adam_tv_step_(b_p, b_alt, gb, b_m, b_v, t, tv_sumsq=s_t, tv_weight=0.37, lr=0.01, grad_abs=b_ga,
                      half_out=b_h, zero_grad=zero_grad)
        if zero_grad:
            assert int(torch.count_nonzero(gb)) == 0
            assert bool(torch.signbit(gb[0]).all()), "a zero cell was written"
        else:
            assert torch.equal(gb, gr), "the fused step must not write the gradient"
        b_p, b_alt = b_alt, b_p
        for x, y in ((a_p, b_p), (a_m, b_m), (a_v, b_v), (a_ga, b_ga)):
            assert torch.equal(x, y)
        if half_dtype is not None:
            assert torch.equal(a_h, b_h)
    with pytest.raises(ValueError):
        adam_tv_step_(b_p, b_p, grads[0], b_m, b_v, 4, tv_sumsq=s_t, tv_weight=0.37, lr=0.01)


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_train_step_beta_term(interp):
    """The beta regulariser of scripts/train.py:170-177 fused into the train step."""
    from plenoxels_amd import MarchConfig, train_step
    from plenoxels_amd.render import beta_term_zero
    say("beta term", interp)
    grid, o, d, gmin, pd = small_scene(90)
    R, S, beta = o.shape[0], 64, 5e-3
    tgt = torch.rand(R, 4, generator=torch.Generator().manual_seed(91))
    kw = dict(num_samples=S, delta_step=0.06, points_distance=pd)
    go = grid.clone().requires_grad_(True)
    p = plain.sample_positions(o, d, S, 0.06)
    n = plain.normalize_samples(p, gmin, pd, "reciprocal")
    vals, _ = (plain.nearest_values if interp == "nearest" else plain.trilinear_values)(n, go)
    vals = vals.reshape(R, S, 4)
    out_w = plain.compute_alpha_weighted_pixels(vals)
    a = vals[..., -1]
    bl = beta * (torch.log(a + 1e-4) - torch.log(1 - a + 1e-4)).mean()
    (plain.mse_loss(out_w, tgt) + bl).backward()
    gg = torch.zeros(grid.shape, device=DEV)
    rb = torch.empty(R, device=DEV)
    train_step(grid.to(DEV), gg, cfg=MarchConfig(S, 0.06, tuple(gmin.tolist()), pd, interp), origins=o.to(DEV),
               dirs=d.to(DEV), targets=tgt.to(DEV), beta=beta, ray_beta=rb)
    bl_k = beta * (rb.sum().item() + R * S * beta_term_zero()) / (R * S)
    assert abs(bl_k - float(bl.detach())) <= 1e-5 * abs(float(bl.detach()))
    ok, info = grad_close(gg.cpu(), go.grad)
    say("  grad", info)
    assert ok, info


def _fit_inputs(B=3, H=24, R=64, steps=4, seed=20):
    c2w = random_cameras(B, seed, radius=2.0)
    imgs = torch.rand(B, H, H, 4, generator=torch.Generator().manual_seed(seed + 1))
    g = torch.Generator().manual_seed(seed + 2)
    uniforms = [torch.rand(B, R, 2, generator=g) for _ in range(steps)]
    return c2w, imgs, uniforms


@pytest.mark.parametrize("S,delta", [(48, 0.05), (600, 0.005)])
@pytest.mark.parametrize("start,dims,progressive", [(0, (24, 24, 24), False), (205, (48, 48, 48), True)])
def test_trainer_matches_reference_loop(start, dims, progressive, S, delta):
    """Trainer.step (fused) vs the reference loop restated op for op (oracle.plain.reference_fit),
    same uniforms: full-grid steps, and progressive-growing steps with k = 13..11 pooling; 48
    samples per ray, and the reference's default of 600 (scripts/train.py:218)."""
    from plenoxels_amd.train import Trainer
    say("trainer vs reference loop", start, dims, S)
    c2w, imgs, uniforms = _fit_inputs()
    # a non-constant start grid: from all zeros the first TV step (k < 19) divides by
    # sqrt(0) and both loops turn NaN, as the reference does
    init = torch.rand(*dims, 4, generator=torch.Generator().manual_seed(30)) * 0.6 - 0.1
    kw = dict(gridsize=dims, points_distance=0.06, num_samples=S, delta_step=delta, lr=0.01, tv=1e-3,
              beta=5e-3, init_grid=init)
    grid_w, gabs_w, hist_w = plain.reference_fit(imgs, c2w, 0.69, uniforms=uniforms, div_mode="reciprocal",
                                                 progressive_growing=progressive, start_step=start, **kw)
    kw["init_grid"] = init.to(DEV)
    tr = Trainer(imgs, c2w, 0.69, number_of_rays=64, device="cuda", progressive_growing=progressive,
                 uniforms_fn=lambda i: uniforms[i - start], **kw)
    for i in range(start, start + len(uniforms)):
        out = tr.step(i)
        rec = hist_w[i - start]
        assert set(out) - {"loss"} == set(rec), (i, sorted(out), sorted(rec))
        for key, v in rec.items():
            assert abs(float(out[key]) - v) <= 1e-4 * abs(v) + 1e-9, (i, key, float(out[key]), v)
    ok, info = grad_close(tr.grad_abs.cpu(), gabs_w, rel=1e-3)
    say("  |grad| sum", info)
    assert ok, info
    diff = (tr.grid.cpu() - grid_w).abs()
    say("  grid: max diff", diff.max().item(), "frac > 1e-5:", (diff > 1e-5).float().mean().item())
    assert (diff > 1e-5).float().mean().item() <= 1e-3
    assert diff.max().item() <= 2 * 0.01 * len(uniforms)


@pytest.mark.parametrize("n", [0, 1, 3, 5, 1000, 4099, 262144, (1 << 20) + 3])
def test_scaled_sum_matches_exact_sum(n):
    """pxl_scaled_sum (the loss value of a step, mse_loss's mean scripts/train.py:156, from the
    per-ray squared errors): float32(scale * sum) within one float32 ulp of the exactly rounded
    value (math.fsum), ragged lengths (tails of 1-3 floats, fewer float4s than threads), the
    same bits on every call, and the work buffer's ticket left ready for calls of other lengths."""
    import ctypes
    from plenoxels_amd import LossMean, _lib
    say("scaled sum n", n)
    L = _lib.lib()
    v = (torch.rand(n, generator=torch.Generator().manual_seed(n + 7), dtype=torch.float64) * 3).float()
    vd = v.to(DEV)
    work = torch.zeros(int(L.pxl_scaled_sum_work_bytes()), dtype=torch.uint8, device=DEV)
    small = torch.rand(7, generator=torch.Generator().manual_seed(1)).to(DEV)
    scale = 0.37
    outs = []
    for k in range(3):
        out = torch.full((), -1.0, device=DEV)
        _lib.check(L.pxl_scaled_sum(_lib.ptr(vd), n, scale, _lib.ptr(out), _lib.ptr(work), _lib.stream_handle(DEV)))
        outs.append(out.item())
        o2 = torch.empty((), device=DEV)  # a call of another length in between
        _lib.check(L.pxl_scaled_sum(_lib.ptr(small), 7, 1.0, _lib.ptr(o2), _lib.ptr(work), _lib.stream_handle(DEV)))
        assert o2.item() == float(np.float32(math.fsum(small.cpu().double().tolist())))
    exact = scale * math.fsum(v.double().tolist())
    want = float(np.float32(exact))
    assert outs[0] == outs[1] == outs[2]
    assert abs(outs[0] - want) <= 2.0 ** -23 * abs(want), (outs[0], want)
    if n > 0:
        lm = LossMean(DEV)(vd)
        assert abs(lm.item() - math.fsum(v.double().tolist()) / (4 * n)) <= 2.0 ** -23 * lm.item()
