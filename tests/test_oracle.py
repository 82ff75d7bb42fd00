"""Pins the oracle (oracle/plain.py, oracle/render_ref.c) before it is trusted.

Checks, per function: the reference's own outputs on tiny inputs (tests/golden), closed
forms and invariants the mathematics fixes, brute force, and agreement between the two
independent oracle formulations (vectorised torch vs scalar FP64 loop).
"""
import math

import numpy as np
import pytest
import torch

from oracle import cref
from oracle import plain

T = torch.from_numpy


# --------------------------------------------------------------------------- golden vectors

def test_golden_generate_rays_random(golden):
    imgs, c2w = T(golden["rays_imgs"]), T(golden["rays_c2w"])
    theta = float(golden["rays_theta"])
    d, pix = plain.generate_rays_batched(imgs, 7, c2w, theta, T(golden["rays_rand_uniforms"]))
    assert torch.equal(d, T(golden["rays_rand_dirs"]))
    assert torch.equal(pix, T(golden["rays_rand_pixels"]))


def test_golden_generate_rays_even(golden):
    imgs, c2w = T(golden["rays_imgs"]), T(golden["rays_c2w"])
    theta = float(golden["rays_theta"])
    uni = plain.even_spread_uniforms(9, imgs.shape[0])
    d, pix = plain.generate_rays_batched(imgs, 9, c2w, theta, uni)
    assert torch.equal(d, T(golden["rays_even_dirs"]))
    assert torch.equal(pix, T(golden["rays_even_pixels"]))


def test_golden_sample_camera_rays(golden):
    imgs, c2w = T(golden["rays_imgs"]), T(golden["rays_c2w"])
    theta = float(golden["rays_theta"])
    uni = T(golden["scr_uniforms"])
    B, R = uni.shape[:2]
    S, delta = int(golden["scr_S"]), float(golden["scr_delta"])
    d, pix = plain.generate_rays_batched(imgs, R, c2w, theta, uni)
    o = torch.repeat_interleave(c2w[:, :3, 3], R, 0)
    smp = plain.sample_positions(o, d, S, delta)
    assert torch.equal(smp, T(golden["scr_samples"]))
    assert torch.equal(pix, T(golden["scr_pixels"]))
    assert torch.equal(d, T(golden["scr_dirs"]))


def test_golden_generate_grid_and_normalize(golden):
    gi, gc, mg, gg = plain.generate_grid(5, 6, 7, points_distance=0.25)
    assert torch.equal(gi, T(golden["grid_coords"]))
    assert torch.equal(mg, T(golden["grid_meshgrid"]))
    assert torch.equal(gg, T(golden["grid_grid"]))
    n = plain.normalize_samples(T(golden["norm_in"]), plain.grid_min(gi), 0.25)
    assert torch.equal(n, T(golden["norm_out"]))


def test_golden_nearest(golden):
    grid, n = T(golden["near_grid"]), T(golden["near_n"])
    vals, mask = plain.nearest_values(n, grid)
    ref_vals = T(golden["near_vals"]) * T(golden["near_mask"]).unsqueeze(-1)
    assert torch.equal(mask, T(golden["near_mask"]))
    assert torch.equal(vals, ref_vals)


def test_golden_corners_and_trilinear(golden):
    grid, n = T(golden["near_grid"]), T(golden["near_n"])
    c = plain.get_grid_points_indices(n)
    assert torch.equal(c, T(golden["tri_corners"]))
    cells, mask = plain.collect_cell_information_via_indices(n, grid.clip(0, 1))
    assert torch.equal(cells, T(golden["tri_cells"]))
    assert torch.equal(mask, T(golden["tri_mask"]))
    vals, m2 = plain.trilinear_values(n, grid)
    assert torch.equal(m2, T(golden["tri_mask"]))
    assert torch.equal(vals, T(golden["tri_vals"]) * m2.unsqueeze(-1))


def test_golden_compositing(golden):
    x = T(golden["comp_in"]).clone().requires_grad_(True)
    res = plain.compute_alpha_weighted_pixels(x)
    assert torch.equal(res.detach(), T(golden["comp_out"]))
    (res * T(golden["comp_up"])).sum().backward()
    assert torch.allclose(x.grad, T(golden["comp_grad"]), rtol=0, atol=1e-6)


def _chain_inputs(golden):
    c2w, imgs = T(golden["chain_c2w"]), T(golden["chain_imgs"])
    uni = T(golden["chain_uniforms"])
    B, R = uni.shape[:2]
    theta = float(golden["chain_theta"])
    cam = torch.repeat_interleave(torch.arange(B), R)
    o, d, tgt = plain.camera_rays(imgs, c2w, theta, cam, uni[..., 0].reshape(-1), uni[..., 1].reshape(-1))
    return o, d, tgt


def test_golden_training_chain(golden):
    """scripts/train.py:142-156 composed from the reference functions vs oracle.render."""
    o, d, tgt = _chain_inputs(golden)
    assert torch.equal(d, T(golden["chain_dirs"]))
    grid = T(golden["chain_grid"]).clone().requires_grad_(True)
    X = grid.shape[0]
    pd = float(golden["chain_pd"])
    gi, _, _, _ = plain.generate_grid(X, X, X, points_distance=pd)
    out = plain.render(grid, o, d, num_samples=int(golden["chain_S"]),
                       delta_step=float(golden["chain_delta"]), gmin=plain.grid_min(gi),
                       points_distance=pd)
    loss = plain.mse_loss(out, tgt)
    loss.backward()
    assert torch.allclose(out.detach(), T(golden["chain_rgba"]), rtol=0, atol=1e-7)
    assert abs(float(loss.detach()) - float(golden["chain_loss"])) < 1e-8
    assert torch.allclose(grid.grad, T(golden["chain_grad"]), rtol=1e-5, atol=1e-9)


# --------------------------------------------------------------------------- closed forms

def test_generate_grid_closed_form():
    for s, pd in ((4, 0.5), (5, 0.25), (256, 0.0125)):
        gi, _, _, _ = plain.generate_grid(s, 2, 3, points_distance=pd)
        xs = gi[:, 0].reshape(s, 2, 3)[:, 0, 0].double()
        want = (torch.arange(s).double() - math.ceil(s / 2) + 1) * np.float32(pd)
        assert torch.allclose(xs, want, rtol=1e-6, atol=1e-7)


def test_round_half_even_and_wrap_brute_force():
    grid = torch.arange(3 * 4 * 5 * 4, dtype=torch.float32).reshape(3, 4, 5, 4) / 300.0
    pts = torch.tensor([[0.5, 1.5, 2.5], [-0.5, 0.49, 4.5], [-0.6, 1.0, 1.0],
                        [2.5, 3.5, 4.49], [2.49, 3.51, 0.0]])
    vals, mask = plain.nearest_values(pts, grid)
    # brute force: python's round() is half-to-even
    for r, p in enumerate(pts.tolist()):
        idx = [round(x) for x in p]
        inb = all(0 <= i < s for i, s in zip(idx, (3, 4, 5)))
        assert bool(mask[r]) == inb
        if inb:
            assert torch.equal(vals[r], grid[idx[0], idx[1], idx[2]])
        else:
            assert torch.equal(vals[r], torch.zeros(4))
    w = plain.wrap_indices(torch.tensor([[-1, -4, 7]]), (3, 4, 5))
    assert w.tolist() == [[2, 0, 2]]


def test_camera_rays_invariants():
    g = torch.Generator().manual_seed(3)
    B = 4
    c2w = torch.eye(4).repeat(B, 1, 1)
    for b in range(B):
        q, _ = torch.linalg.qr(torch.randn(3, 3, generator=g, dtype=torch.float64))
        c2w[b, :3, :3] = q.float()
        c2w[b, :3, 3] = torch.randn(3, generator=g)
    imgs = torch.rand(B, 10, 10, 4, generator=g)
    cam = torch.arange(B)
    half = torch.full((B,), 0.5)
    o, d, tgt = plain.camera_rays(imgs, c2w, 0.7, cam, half, half)
    # centre ray is the camera's -Z axis; origin is the translation column
    assert torch.allclose(d, -c2w[:, :3, 2], atol=1e-6)
    assert torch.equal(o, c2w[:, :3, 3])
    assert torch.equal(tgt, imgs[cam, 5, 5])
    u = torch.rand(B, generator=g)
    v = torch.rand(B, generator=g)
    _, d, tgt = plain.camera_rays(imgs, c2w, 0.7, cam, u, v)
    assert torch.allclose(d.norm(dim=1), torch.ones(B), atol=1e-6)
    # linear-angle camera: d ~ 0.7(u-.5) X - 0.7(v-.5) Y - Z (unit axes here)
    raw = (0.7 * (u - 0.5))[:, None] * c2w[:, :3, 0] - (0.7 * (v - 0.5))[:, None] * c2w[:, :3, 1] - c2w[:, :3, 2]
    assert torch.allclose(d, raw / raw.norm(dim=1, keepdim=True), atol=1e-6)
    row = torch.clamp(torch.round(10 * v), max=9).long()
    col = torch.clamp(torch.round(10 * u), max=9).long()
    assert torch.equal(tgt, imgs[cam, row, col])


def test_torch_cpu_norm_is_sequential_fma():
    """The three-element norm of torch on the CPU equals sqrt(fma(z,z,fma(y,y,x*x))); the
    HIP ray generator uses that order, so its directions match the CPU oracle bitwise."""
    g = torch.Generator().manual_seed(0)
    v = torch.randn(50000, 3, generator=g) * torch.rand(50000, 1, generator=g) * 5
    x, y, z = (v[:, a].double() for a in range(3))
    f32 = lambda t: t.float().double()
    want = torch.sqrt(f32(z * z + f32(y * y + f32(x * x)))).float()
    assert torch.equal(v.norm(dim=1), want)


def test_sample_positions_closed_form():
    g = torch.Generator().manual_seed(4)
    o = torch.randn(5, 3, generator=g, dtype=torch.float64)
    d = torch.nn.functional.normalize(torch.randn(5, 3, generator=g, dtype=torch.float64), dim=1)
    p = plain.sample_positions(o, d, 6, 0.25, near=1.0).reshape(5, 6, 3)
    dist = (p - o[:, None]).norm(dim=-1)
    want = 1.0 + 0.25 * torch.arange(1, 7, dtype=torch.float64)
    assert torch.allclose(dist, want.expand(5, 6), atol=1e-12)


def test_compositing_closed_forms():
    S = 17
    c = torch.rand(3, 2, S, 3, dtype=torch.float64)
    a = torch.full((3, 2, S, 1), 0.3, dtype=torch.float64)
    out = plain.compute_alpha_weighted_pixels(torch.cat([c, a], -1))
    assert torch.allclose(out[..., 3], torch.full((3, 2), 1 - 0.7 ** S, dtype=torch.float64))
    w = 0.3 * 0.7 ** torch.arange(S, dtype=torch.float64)
    assert torch.allclose(out[..., :3], (c * w[:, None]).sum(-2))
    # opaque sample: colour of the first alpha=1 sample, A = 1
    a = torch.zeros(1, 1, S, 1, dtype=torch.float64)
    a[..., 5, :] = 1.0
    a[..., 9, :] = 0.5
    out = plain.compute_alpha_weighted_pixels(torch.cat([c[:1, :1], a], -1))
    assert torch.allclose(out[0, 0, :3], c[0, 0, 5])
    assert float(out[0, 0, 3]) == 1.0
    out = plain.compute_alpha_weighted_pixels(torch.zeros(2, 2, S, 4, dtype=torch.float64))
    assert torch.equal(out, torch.zeros(2, 2, 4, dtype=torch.float64))


def test_compositing_gradient_closed_form_and_gradcheck():
    g = torch.Generator().manual_seed(9)
    S = 12
    x = torch.rand(1, 1, S, 4, generator=g, dtype=torch.float64)
    x[0, 0, 4, 3] = 1.0  # saturated sample: gradient still defined (no division)
    x.requires_grad_(True)
    up = torch.randn(1, 1, 4, generator=g, dtype=torch.float64)
    (plain.compute_alpha_weighted_pixels(x) * up).sum().backward()
    a = x.detach()[0, 0, :, 3]
    c = x.detach()[0, 0, :, :3]
    Tp = torch.cumprod(torch.cat([torch.ones(1, dtype=torch.float64), 1 - a[:-1]]), 0)
    gc = (a * Tp)[:, None] * up[0, 0, :3]
    # suffix S_i = sum_{j>i} a_j c_j prod_{i<m<j} (1-a_m)
    Sx = torch.zeros(S, 4, dtype=torch.float64)
    for i in range(S):
        acc = torch.zeros(4, dtype=torch.float64)
        tr = 1.0
        for j in range(i + 1, S):
            acc += tr * a[j] * torch.cat([c[j], torch.ones(1, dtype=torch.float64)])
            tr = tr * (1 - a[j])
        Sx[i] = acc
    ga = Tp * ((up[0, 0, :3] * (c - Sx[:, :3])).sum(-1) + up[0, 0, 3] * (1 - Sx[:, 3]))
    assert torch.allclose(x.grad[0, 0, :, :3], gc, atol=1e-12)
    assert torch.allclose(x.grad[0, 0, :, 3], ga, atol=1e-12)
    y = torch.rand(2, 1, 6, 4, generator=g, dtype=torch.float64).requires_grad_(True)
    assert torch.autograd.gradcheck(plain.compute_alpha_weighted_pixels, (y,))


def _affine_grid(dims, A, b):
    ii = torch.stack(torch.meshgrid(*[torch.arange(s, dtype=torch.float64) for s in dims],
                                    indexing="ij"), -1)
    return ii @ A.T + b


def test_trilinear_exact_on_affine_fields():
    g = torch.Generator().manual_seed(2)
    dims = (5, 6, 7)
    A = torch.rand(4, 3, generator=g, dtype=torch.float64) * 0.05
    b = torch.rand(4, generator=g, dtype=torch.float64) * 0.2
    grid = _affine_grid(dims, A, b)
    n = torch.rand(200, 3, generator=g, dtype=torch.float64) * torch.tensor([3.99, 4.99, 5.99], dtype=torch.float64)
    vals, mask = plain.trilinear_values(n, grid)
    assert bool(mask.all())
    assert torch.allclose(vals, n @ A.T + b, atol=1e-12)
    # integer coordinates return the cell itself
    ni = torch.tensor([[1.0, 2.0, 3.0], [4.0, 0.0, 6.0]], dtype=torch.float64)
    vals, _ = plain.trilinear_values(ni, grid)
    assert torch.allclose(vals, grid[[1, 4], [2, 0], [3, 6]].clip(0, 1))


def test_trilinear_matches_grid_sample_interior():
    g = torch.Generator().manual_seed(12)
    dims = (6, 7, 8)
    grid = torch.rand(*dims, 4, generator=g, dtype=torch.float64)
    n = torch.rand(100, 3, generator=g, dtype=torch.float64) * (torch.tensor(dims, dtype=torch.float64) - 1)
    vals, _ = plain.trilinear_values(n, grid)
    inp = grid.permute(3, 2, 1, 0).unsqueeze(0)  # (1,C,D=Z,H=Y,W=X)
    coords = (n / (torch.tensor(dims, dtype=torch.float64) - 1)) * 2 - 1
    gs = torch.nn.functional.grid_sample(inp, coords.view(1, 100, 1, 1, 3), mode="bilinear",
                                         align_corners=True)
    assert torch.allclose(vals, gs.view(4, 100).T, atol=1e-12)


def test_trilinear_boundary_quirks():
    grid = torch.rand(4, 4, 4, 4, dtype=torch.float64)
    # top slab 3 <= n < 4 interpolates toward wrapped cell 0; n = 4 is outside
    n = torch.tensor([[3.25, 1.0, 1.0], [4.0, 1.0, 1.0], [-0.25, 1.0, 1.0]], dtype=torch.float64)
    vals, mask = plain.trilinear_values(n, grid)
    assert mask.tolist() == [True, False, False]
    want = 0.25 * grid[0, 1, 1] + 0.75 * grid[3, 1, 1]
    assert torch.allclose(vals[0], want)


@pytest.mark.parametrize("dims", [(5, 6, 7), (9, 17, 10), (8, 8, 16)])
def test_occupancy_bits_roundtrip(dims):
    """The vectorised packing against the tile layout cell by cell (8x8x8 tiles of 16 words;
    cell (x,y,z): word tile*16 + (x&7)*2 + ((y>>2)&1), bit (y&3)*8 + (z&7)); every bit of the
    words is either a cell's or padding, and padding is 0."""
    g = torch.Generator().manual_seed(sum(dims))
    occ = torch.rand(*dims, generator=g) < 0.3
    words = plain.occupancy_to_bits(occ).numpy().view(np.uint32)
    X, Y, Z = dims
    TY, TZ = (Y + 7) // 8, (Z + 7) // 8
    assert words.size == ((X + 7) // 8) * TY * TZ * 16
    seen = np.zeros(words.size * 32, dtype=bool)
    o = occ.numpy()
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                t = ((x // 8) * TY + y // 8) * TZ + z // 8
                w = t * 16 + (x % 8) * 2 + (y % 8) // 4
                b = (y % 4) * 8 + z % 8
                assert not seen[w * 32 + b], "two cells share a bit"
                seen[w * 32 + b] = True
                assert bool((words[w] >> b) & 1) == bool(o[x, y, z])
    allbits = np.unpackbits(words.view(np.uint8), bitorder="little").astype(bool)
    assert not bool((allbits & ~seen).any()), "a padding bit is set"


# --------------------------------------------------------------------------- C oracle vs plain

def _scene(seed, dims=(9, 10, 11), R=24, S=40):
    g = torch.Generator().manual_seed(seed)
    grid = torch.rand(*dims, 4, generator=g) * 1.4 - 0.2
    pd = 0.21
    gi, _, _, _ = plain.generate_grid(*dims, points_distance=pd)
    gmin = plain.grid_min(gi)
    o = torch.randn(R, 3, generator=g) * 0.2 + torch.tensor([0.0, 0.0, -2.2])
    tgt = torch.randn(R, 3, generator=g) * 0.4
    d = torch.nn.functional.normalize(tgt - o, dim=1)
    return grid, o, d, gmin, pd


def _plain_fp64(grid, o, d, gmin, pd, interp, S, delta, occ=None, t_stop=0.0, div_mode="divide"):
    """Plain chain with float32 geometry and float64 values (what the C oracle computes)."""
    p = plain.sample_positions(o, d, S, delta)
    n = plain.normalize_samples(p, gmin, pd, div_mode).double()
    g64 = grid.double().clone().requires_grad_(True)
    if interp == "nearest":
        vals, _ = plain.nearest_values(n, g64, occ)
    else:
        vals, _ = plain.trilinear_values(n, g64, occ)
    out = plain.compute_alpha_weighted_pixels(vals.reshape(o.shape[0], S, 4), t_stop)
    return out, g64


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
@pytest.mark.parametrize("div_mode", ["divide", "reciprocal"])
def test_c_oracle_matches_plain_fp64(interp, div_mode):
    grid, o, d, gmin, pd = _scene(3)
    S, delta = 40, 0.09
    out, g64 = _plain_fp64(grid, o, d, gmin, pd, interp, S, delta, div_mode=div_mode)
    up = torch.randn(out.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(4))
    (out * up).sum().backward()
    rgba, gg, _ = cref.render(grid.numpy(), o.numpy(), d.numpy(), grad_rgba=up.numpy(),
                              num_samples=S, delta_step=delta, gmin=gmin.tolist(),
                              points_distance=pd, interp=interp, div_mode=div_mode)
    assert np.abs(rgba - out.detach().numpy()).max() < 1e-12
    assert np.abs(gg - g64.grad.numpy()).max() < 1e-12
    assert np.abs(g64.grad.numpy()).max() > 1e-3  # non-trivial


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_c_oracle_occupancy_and_tstop_match_plain(interp):
    grid, o, d, gmin, pd = _scene(5)
    S, delta = 40, 0.09
    occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(6)) < 0.5
    if interp == "trilinear":
        occ = plain.dilate_cells(occ) & (torch.rand(grid.shape[:3]) < 0.7)
    bits = plain.occupancy_to_bits(occ).numpy()
    for t_stop in (0.0, 0.05):
        out, g64 = _plain_fp64(grid, o, d, gmin, pd, interp, S, delta, occ=occ, t_stop=t_stop)
        up = torch.randn(out.shape, dtype=torch.float64)
        (out * up).sum().backward()
        rgba, gg, nproc = cref.render(grid.numpy(), o.numpy(), d.numpy(), grad_rgba=up.numpy(),
                                      occupancy_bits=bits, num_samples=S, delta_step=delta,
                                      gmin=gmin.tolist(), points_distance=pd, interp=interp,
                                      t_stop=t_stop)
        assert np.abs(rgba - out.detach().numpy()).max() < 1e-12
        assert np.abs(gg - g64.grad.numpy()).max() < 1e-12
        if t_stop:
            assert (nproc < S).any()


def test_c_oracle_geometry_bitwise_equals_plain_fp32():
    grid, o, d, gmin, pd = _scene(7)
    S, delta = 40, 0.09
    p = plain.sample_positions(o, d, S, delta, near=0.13)
    for mode in ("divide", "reciprocal"):
        n = plain.normalize_samples(p, gmin, pd, mode).reshape(o.shape[0], S, 3)
        for r in (0, 5, 17):
            for k in (1, 13, 40):
                nc = cref.sample_coord(o[r].numpy(), d[r].numpy(), k, grid.shape[:3],
                                       delta_step=delta, near=0.13, gmin=gmin.tolist(),
                                       points_distance=pd, div_mode=mode)
                assert np.array_equal(nc, n[r, k - 1].numpy())


def test_c_oracle_central_differences():
    grid, o, d, gmin, pd = _scene(8, dims=(6, 6, 6), R=12, S=30)
    kw = dict(num_samples=30, delta_step=0.1, gmin=gmin.tolist(), points_distance=pd, interp="trilinear")
    g = grid.double().numpy().clip(0.05, 0.95)  # away from the clamp kinks
    tgt = np.random.default_rng(0).random((12, 4))
    rgba, _, _ = cref.render(g, o.numpy(), d.numpy(), **kw)
    up = 2 * (rgba - tgt) / rgba.size
    _, gg, _ = cref.render(g, o.numpy(), d.numpy(), grad_rgba=up, **kw)
    idx = np.argsort(-np.abs(gg).reshape(-1))[:20]
    eps = 1e-6
    for L in idx:
        gp = g.copy().reshape(-1)
        gm = g.copy().reshape(-1)
        gp[L] += eps
        gm[L] -= eps
        lp = ((cref.render(gp.reshape(g.shape), o.numpy(), d.numpy(), **kw)[0] - tgt) ** 2).mean()
        lm = ((cref.render(gm.reshape(g.shape), o.numpy(), d.numpy(), **kw)[0] - tgt) ** 2).mean()
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - gg.reshape(-1)[L]) <= 1e-6 * max(1e-3, abs(fd))


def test_golden_pool_and_tv(golden):
    x = T(golden["pool_in"])
    assert torch.allclose(plain.average_pool3d_grid(x, 5, 1), T(golden["pool_k5_s1"]), atol=1e-6)
    assert torch.allclose(plain.average_pool3d_grid(x, 7, 2), T(golden["pool_k7_s2"]), atol=1e-6)
    t = T(golden["tv_in"]).clone().requires_grad_(True)
    v = plain.tv_loss(t)
    v.backward()
    assert abs(float(v.detach()) - float(golden["tv_out"])) < 1e-5
    assert torch.allclose(t.grad, T(golden["tv_grad"]), atol=1e-7)


def test_tv_loss_closed_form():
    g = torch.zeros(3, 4, 5, 4, dtype=torch.float64)
    g[1, 2, 3, 0] = 2.0  # one spike: 6 neighbours (all inside) each contribute 2^2
    assert float(plain.tv_loss(g)) == pytest.approx(math.sqrt(6 * 4.0))
    g = torch.arange(3 * 4 * 5 * 4, dtype=torch.float64).reshape(3, 4, 5, 4)
    # linear ramp: differences are the strides 4 (z), 20 (y), 80 (x)
    want = math.sqrt(4 * (3 * 4 * 4 * 16 + 3 * 3 * 5 * 400 + 2 * 4 * 5 * 6400))
    assert float(plain.tv_loss(g)) == pytest.approx(want)


def test_norm3_orders_are_within_an_ulp_and_distinct():
    """Both restated 3-norm orders are the float32 norm to within two ulps of the FP64 value (three
    products and sums rounded, then a correctly rounded root), agree with torch's CPU norm in
    the "cpu" order, and genuinely differ from each other on some vectors -- a swapped or dropped
    term fails one of these."""
    g = torch.Generator().manual_seed(12)
    v = torch.randn(1 << 18, 3, generator=g) * torch.rand(1 << 18, 1, generator=g) * 5
    exact = v.double().norm(dim=1)
    for mode in ("cpu", "gpu"):
        got = plain.norm3(v, mode).double()
        assert ((got - exact).abs() <= 2.0 ** -22 * exact).all(), mode
    assert torch.equal(plain.norm3(v, "cpu"), v.norm(dim=1))
    assert int((plain.norm3(v, "cpu") != plain.norm3(v, "gpu")).sum()) > 100
    # "gpu" is the x, z pair summed first: swapping y and z changes it, swapping x and z does not
    assert torch.equal(plain.norm3(v[:, [2, 1, 0]], "gpu"), plain.norm3(v, "gpu"))
    assert not torch.equal(plain.norm3(v[:, [0, 2, 1]], "gpu"), plain.norm3(v, "gpu"))


@pytest.mark.parametrize("start,progressive", [(0, False), (185, True), (229, True)])
def test_reference_fit_chunked_equals_one_graph(start, progressive):
    """reference_fit's ray-chunked step (for batches too large for one autograd graph) gives the
    one-graph step's losses, |grad| and updated grid up to summation order: full-grid steps,
    pooled steps, the TV and beta terms."""
    from util import random_cameras
    B, H, R = 3, 16, 40
    c2w = random_cameras(B, 5, radius=2.0)
    imgs = torch.rand(B, H, H, 4, generator=torch.Generator().manual_seed(1))
    g = torch.Generator().manual_seed(2)
    uni = [torch.rand(B, R, 2, generator=g) for _ in range(3)]
    init = torch.rand(24, 24, 24, 4, generator=torch.Generator().manual_seed(3)) * 0.6 - 0.1
    kw = dict(gridsize=(24, 24, 24), points_distance=0.06, num_samples=40, delta_step=0.05, lr=0.01, tv=1e-3,
              beta=5e-3, uniforms=uni, start_step=start, init_grid=init, progressive_growing=progressive)
    ga, aa, ha = plain.reference_fit(imgs, c2w, 0.69, **kw)
    gb, ab, hb = plain.reference_fit(imgs, c2w, 0.69, chunk_rays=17, **kw)
    assert (ga - gb).abs().max().item() <= 1e-6
    assert (aa - ab).abs().max().item() <= 1e-6 * aa.abs().max().item()
    for x, y in zip(ha, hb):
        assert set(x) == set(y)
        for k in x:
            assert abs(x[k] - y[k]) <= 1e-6 * abs(y[k])
