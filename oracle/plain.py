"""Plain-PyTorch restatement of the reference render-and-optimise chain (oracle; test infrastructure).

Every function follows the reference op for op -- same torch ops, same order, same dtypes --
so that, run on the CPU in float32, it reproduces the reference bit for bit where the
reference is deterministic.  It adds explicit inputs where the reference draws hidden
randomness (per-ray camera index and uniforms instead of ``torch.rand``) and the opt-in
extensions the benchmark scenes need (``near`` offset, occupancy skip, early termination),
whose defaults reproduce the reference exactly.

Division semantics.  ``normalize_samples_for_indecies`` divides by a Python float
(src/ray_sampling.py:13).  ATen's CPU kernel divides; its GPU kernel multiplies by the
float32 reciprocal of the divisor.  ``div_mode="divide"`` restates the former,
``div_mode="reciprocal"`` the latter; both are pinned in tests (the second on the GPU).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
baseline legs, never by the product package.
"""
import math

import numpy as np
import torch
import torch.nn.functional as F


# --------------------------------------------------------------------------- grid layout

def generate_grid(sx, sy, sz, points_distance=0.5, info_size=4, device="cpu"):
    """src/grid_functions.py:184-217.

    Returns (grid_coords (XYZ,3) f32, grid_cells (X,Y,Z,C) f32 zeros requiring grad,
    meshgrid (X,Y,Z,3) int64, grid_grid (X,Y,Z,3) f32).  Cell (i,j,k) sits at world
    ((i,j,k) - ceil(S/2) + 1) * pd.
    Pinned by: closed form (tests/test_oracle.py::test_generate_grid_closed_form) and the
    reference golden vector (tests/golden).
    """
    gx, gy, gz = (torch.arange(s, device=device) for s in (sx, sy, sz))
    cx, cy, cz = torch.meshgrid(gx, gy, gz, indexing="ij")
    meshgrid = torch.stack([cx, cy, cz], dim=-1)
    cx = cx - float(np.ceil(sx / 2)) + 1
    cy = cy - float(np.ceil(sy / 2)) + 1
    cz = cz - float(np.ceil(sz / 2)) + 1
    cx, cy, cz = cx * points_distance, cy * points_distance, cz * points_distance
    grid_grid = torch.stack([cx, cy, cz], dim=-1)
    grid_coords = grid_grid.reshape(sx * sy * sz, 3)
    grid_cells = torch.zeros([sx, sy, sz, info_size], requires_grad=True, device=device)
    return grid_coords, grid_cells, meshgrid, grid_grid


def grid_min(grid_indices):
    """The ``grid_indices.min(0)[0]`` term of src/ray_sampling.py:13 (min is order-exact)."""
    return grid_indices.min(0)[0]


def reciprocal_f32(points_distance):
    """float32(1 / float32(pd)): the multiplier ATen's GPU true-division uses for a scalar
    divisor (ATen BinaryDivTrueKernel.cu, outside the reference).  Pinned on the GPU by
    tests/test_gpu_render.py::test_torch_gpu_scalar_division_is_reciprocal_multiply."""
    return float(np.float32(1.0) / np.float32(points_distance))


# --------------------------------------------------------------------------- rays

def norm3(x, norm_mode=None):
    """x.norm(dim=-1) of (..., 3) float32 vectors (src/ray_sampling.py:218,262).

    norm_mode None: torch's own norm on x's device (what the reference computes there).
    "cpu": ATen's CPU reduction order, sqrt(fma(z, z, fma(y, y, x*x))) -- on a CPU tensor this
    is torch's norm itself.  "gpu": ATen's GPU reduction order, sqrt((x*x + z*z) + y*y), each
    op rounded to float32 (its reduce kernel gives a length-3 row two threads: elements 0 and
    2 on one, 1 on the other) and a correctly rounded square root.  The root is taken in
    float64 and rounded once to float32, which is the correctly rounded float32 root (53 >=
    2*24 + 2 bits); torch's elementwise float32 sqrt on the CPU is not correctly rounded (it
    differs from it on a few percent of inputs), while both of torch's GPU roots are.
    Pinned by tests/test_oracle.py (CPU order vs torch on the CPU) and
    tests/test_gpu_render.py::test_torch_gpu_norm_order (GPU order vs torch on the GPU,
    bitwise over millions of vectors).
    """
    if norm_mode is None:
        return x.norm(dim=-1)
    if norm_mode == "cpu":
        return x.cpu().norm(dim=-1).to(x.device)
    if norm_mode == "gpu":
        a, b, c = x[..., 0], x[..., 1], x[..., 2]
        return torch.sqrt(((a * a + c * c) + b * b).double()).float()
    raise ValueError(norm_mode)


def camera_rays(imgs, c2w, camera_angle_x, cam_ids, u, v, norm_mode=None):
    """Per-ray form of src/ray_sampling.py:212-264 (generate_rays_batched).

    imgs (B,H,W,4), c2w (B,4,4), cam_ids (R,) int64, u/v (R,) in [0,1].  Returns
    origins (R,3) = c2w[cam,:3,3] (src/ray_sampling.py:159), unit directions (R,3) and
    the target pixels (R,4) imgs[cam, min(round(W*v),W-1), min(round(H*u),H-1)].
    Elementwise ops are applied per ray in the reference's order, so the values equal
    the reference's (B,R)-shaped computation for the same (cam, u, v).  norm_mode: the
    summation order of the two 3-norms (see norm3; None = torch's on this device).
    Pinned by: reference golden vectors (generate_rays_batched), closed forms
    (unit norm, centre ray = -Z axis, pixel round/clamp).
    """
    cam_x = c2w[:, :3, 0]
    cam_y = c2w[:, :3, 1]
    cam_z = -c2w[:, :3, 2]
    aspect = norm3(cam_x, norm_mode) / norm3(cam_y, norm_mode)
    H, W = imgs.shape[1], imgs.shape[2]
    col = (H * u).round().clamp(max=H - 1).to(torch.long)
    row = (W * v).round().clamp(max=W - 1).to(torch.long)
    targets = imgs[cam_ids, row, col]
    us = camera_angle_x * (u - 0.5)
    vs = -((camera_angle_x * (1 / aspect))[cam_ids] * (v - 0.5))
    d = us.unsqueeze(-1) * cam_x[cam_ids] + vs.unsqueeze(-1) * cam_y[cam_ids] + cam_z[cam_ids]
    d = d / norm3(d, norm_mode).unsqueeze(-1)
    origins = c2w[cam_ids, :3, 3]
    return origins, d, targets


def even_spread_uniforms(number_of_rays, num_cameras, device="cpu"):
    """src/ray_sampling.py:220-223: cartesian_prod(linspace(0,1,sqrt R)) repeated per camera."""
    n = int(np.round(np.sqrt(number_of_rays)))
    ls = torch.linspace(0, end=1, steps=n, device=device)
    return torch.cartesian_prod(ls, ls).repeat([num_cameras, 1, 1])


def generate_rays_batched(imgs, number_of_rays, c2w, camera_angle_x, uniforms, norm_mode=None):
    """src/ray_sampling.py:195-264 with the uniforms (B,R,2) passed in instead of drawn.

    Returns (dirs (B*R,3), pixels_to_rays (B*R,4)) in camera-major order.
    """
    B, R = uniforms.shape[0], uniforms.shape[1]
    cam_ids = torch.repeat_interleave(torch.arange(B, device=imgs.device), R, 0)
    u = uniforms[:, :, 0].reshape(-1)
    v = uniforms[:, :, 1].reshape(-1)
    _, d, tgt = camera_rays(imgs, c2w, camera_angle_x, cam_ids, u, v, norm_mode)
    return d, tgt


def device_norm_mode(div_mode):
    """The norm order that goes with a division mode: both are the semantics of one device
    (ATen CPU: exact division, CPU norm; ATen GPU: reciprocal multiply, GPU norm)."""
    return {"divide": "cpu", "reciprocal": "gpu"}[div_mode]


# --------------------------------------------------------------------------- samples

def sample_positions(origins, dirs, num_samples, delta_step, near=0.0):
    """src/ray_sampling.py:161-167: p = o + d * t_k, t_k = delta*k (+ near), k = 1..S.

    Layout (R*S,3), sample-fastest.  near=0 reproduces the reference exactly
    (t + 0.0 == t).  Pinned by closed form ||p - o|| = t_k and golden vectors.
    """
    R = origins.shape[0]
    k = torch.arange(num_samples + 1, device=origins.device, dtype=origins.dtype)[1:]
    t = delta_step * k
    if near:
        t = t + near
    t = t.repeat(R).unsqueeze(1)
    o = torch.repeat_interleave(origins, num_samples, 0)
    d = torch.repeat_interleave(dirs, num_samples, 0)
    return o + d * t


def normalize_samples(samples, gmin, points_distance, div_mode="divide"):
    """src/ray_sampling.py:12-13 with grid_indices.min hoisted (gmin)."""
    diff = samples - gmin
    if div_mode == "divide":
        return diff / points_distance
    if div_mode == "reciprocal":
        return diff * reciprocal_f32(points_distance)
    raise ValueError(div_mode)


# --------------------------------------------------------------------------- gathers

def find_out_of_bound(indices, dims):
    """src/grid_functions.py:47-63 (True = inside)."""
    X, Y, Z = dims
    i0, i1, i2 = indices[:, 0], indices[:, 1], indices[:, 2]
    return (i0 < X) & (i0 >= 0) & (i1 < Y) & (i1 >= 0) & (i2 < Z) & (i2 >= 0)


def wrap_indices(indices, dims):
    """src/grid_functions.py:66-79: periodic wrap, sign of the divisor (-1 -> X-1)."""
    out = indices.clone()
    for a in range(3):
        out[:, a] %= dims[a]
    return out


def nearest_values(n, grid, occupancy=None):
    """src/grid_functions.py:103-114 composed as scripts/train.py:146-147.

    idx = round-half-even(n) -> int64; in-bounds mask on the rounded index; wrap; gather
    from grid.clip(0,1); multiply by the mask.  occupancy (X,Y,Z) bool, optional: a sample
    whose nearest cell is unoccupied is skipped (value 0, no gradient).
    Returns (values (N,C), keep mask (N,)).
    """
    dims = grid.shape[:3]
    idx = torch.round(n).to(torch.long)
    mask = find_out_of_bound(idx, dims)
    w = wrap_indices(idx, dims)
    vals = grid.clip(0, 1)[w[:, 0], w[:, 1], w[:, 2]]
    if occupancy is not None:
        mask = mask & occupancy[w[:, 0], w[:, 1], w[:, 2]]
    return vals * mask.unsqueeze(-1), mask


def get_grid_points_indices(n):
    """src/grid_functions.py:220-246: (N,8,3) int64 corners, order x{ceil,floor} x y{..} x z{..}."""
    cx, cy, cz = (torch.ceil(n[:, a]).unsqueeze(1) for a in range(3))
    fx, fy, fz = (torch.floor(n[:, a]).unsqueeze(1) for a in range(3))
    pts = [torch.cat([x, y, z], 1).unsqueeze(1)
           for x in (cx, fx) for y in (cy, fy) for z in (cz, fz)]
    return torch.cat(pts, 1).type(torch.long)


def trilinear_interpolation(n, corners, grid):
    """src/grid_functions.py:7-44 (signed frac weights; corners must be pre-wrapped)."""
    f = torch.frac(n)
    s0 = grid[corners[:, 4:, 0], corners[:, 4:, 1], corners[:, 4:, 2]]
    s1 = grid[corners[:, :4, 0], corners[:, :4, 1], corners[:, :4, 2]]
    fs = f[:, 0].unsqueeze(-1).unsqueeze(-1)
    v = (s1 * fs) + (s0 * (1 - fs))
    fs = f[:, 1].unsqueeze(-1).unsqueeze(-1)
    v = (v[:, :2] * fs) + (v[:, 2:] * (1 - fs))
    fs = f[:, 2].unsqueeze(-1)
    return (v[:, 0] * fs) + (v[:, 1] * (1 - fs))


def trilinear_values(n, grid, cell_occupancy=None):
    """Trilinear composition of src/grid_functions.py:154-170 and :7-44 in the training chain.

    Float in-bounds mask on the unrounded coordinate (0 <= n < X per axis, :162), corners
    wrapped (:168), interpolation of grid.clip(0,1), times the mask.  cell_occupancy
    (X,Y,Z) bool, optional: a sample whose floor cell (corner 7, wrapped) is unoccupied is
    skipped.  Returns (values (N,C), keep mask (N,)).
    """
    dims = grid.shape[:3]
    mask = find_out_of_bound(n, dims)
    c = get_grid_points_indices(n)
    N = c.shape[0]
    c = wrap_indices(c.reshape(N * 8, 3), dims).reshape(N, 8, 3)
    vals = trilinear_interpolation(n, c, grid.clip(0, 1))
    if cell_occupancy is not None:
        mask = mask & cell_occupancy[c[:, 7, 0], c[:, 7, 1], c[:, 7, 2]]
    return vals * mask.unsqueeze(-1), mask


def collect_cell_information_via_indices(n, grid):
    """src/grid_functions.py:154-170: (N*8,C) gathered corners and the float in-bounds mask."""
    dims = grid.shape[:3]
    mask = find_out_of_bound(n, dims)
    c = get_grid_points_indices(n)
    c = wrap_indices(c.reshape(c.shape[0] * c.shape[1], 3), dims)
    return grid[c[:, 0], c[:, 1], c[:, 2]], mask


# --------------------------------------------------------------------------- compositing

def compute_alpha_weighted_pixels(samples, t_stop=0.0):
    """src/ray_sampling.py:172-192 on (..., S, 4): T = cumprod(1 - [0, a_0..a_{S-2}]),
    w = a*T, RGB = sum w*c, A = sum w.  t_stop > 0 (extension) drops every sample whose
    transmittance T_i is below t_stop (value and gradient zero).
    """
    if t_stop > 0:
        with torch.no_grad():
            a = samples[..., -1]
            pad = torch.cat([torch.ones_like(a[..., :1]), 1 - a[..., :-1]], dim=-1)
            keep = pad.cumprod(dim=-1) >= t_stop
        samples = samples * keep.unsqueeze(-1)
    alpha = samples[..., -1]
    padded = torch.cat([torch.zeros_like(alpha[..., :1]), alpha], dim=-1)
    cumprod = (1 - padded[..., :-1]).cumprod(dim=-1)
    weights = (alpha * cumprod).unsqueeze(-1)
    rgb = (samples[..., :-1] * weights).sum(-2)
    final_alpha = weights.sum(dim=-2)
    return torch.cat([rgb, final_alpha], dim=-1)


# --------------------------------------------------------------------------- full chain

def render(grid, origins, dirs, *, num_samples, delta_step, gmin, points_distance,
           interp="nearest", near=0.0, div_mode="divide", occupancy=None, t_stop=0.0):
    """The training chain scripts/train.py:130-153 on given rays: (R,4) RGBA.

    grid is the raw (X,Y,Z,4) parameter (clipped inside).  For interp="trilinear" the
    occupancy is the per-cell mask indexed by the floor cell.
    """
    R = origins.shape[0]
    p = sample_positions(origins, dirs, num_samples, delta_step, near)
    n = normalize_samples(p, gmin, points_distance, div_mode)
    if interp == "nearest":
        vals, _ = nearest_values(n, grid, occupancy)
    elif interp == "trilinear":
        vals, _ = trilinear_values(n, grid, occupancy)
    else:
        raise ValueError(interp)
    vals = vals.reshape(R, num_samples, grid.shape[-1])
    return compute_alpha_weighted_pixels(vals, t_stop)


def render_chunked_backward(grid, origins, dirs, targets, *, chunk_rays, **kw):
    """Forward + mean-MSE (scripts/train.py:156) + autograd backward over ray chunks.

    The loss is normalised by the full batch (4 * R), exactly as one mse_loss over all rays,
    so grid.grad equals the unchunked gradient up to summation order.  Returns (rgba, loss).
    """
    R = origins.shape[0]
    outs = []
    loss = 0.0
    for s in range(0, R, chunk_rays):
        e = min(R, s + chunk_rays)
        out = render(grid, origins[s:e], dirs[s:e], **kw)
        l = ((out - targets[s:e]) ** 2).sum() / (4 * R)
        l.backward()
        loss += float(l.detach())
        outs.append(out.detach())
    return torch.cat(outs, 0), loss


def mse_loss(pred, target):
    """scripts/train.py:156 (reduction = mean)."""
    return F.mse_loss(pred, target)


# --------------------------------------------------------------------------- occupancy

def occupancy_to_bits(occ):
    """(X,Y,Z) bool -> int32 words in the kernels' occupancy layout: 8x8x8 tiles of 16 words,
    tile ((x>>3)*TY + (y>>3))*TZ + (z>>3) with TY = ceil(Y/8), TZ = ceil(Z/8); cell (x,y,z) is
    bit (y&3)*8 + (z&7) of word (x&7)*2 + ((y>>2)&1) of its tile (padding cells 0)."""
    o = occ.to(torch.bool).cpu().numpy()
    X, Y, Z = o.shape
    TX, TY, TZ = (X + 7) // 8, (Y + 7) // 8, (Z + 7) // 8
    pad = np.zeros((TX * 8, TY * 8, TZ * 8), dtype=bool)
    pad[:X, :Y, :Z] = o
    # axes (tx, xi, ty, yhi, ylo, tz, zi) -> word (tx, ty, tz, xi, yhi), bit (ylo, zi)
    t = pad.reshape(TX, 8, TY, 2, 4, TZ, 8).transpose(0, 2, 5, 1, 3, 4, 6)
    bits = t.reshape(-1, 32).astype(np.uint64)
    words = (bits << np.arange(32, dtype=np.uint64)).sum(axis=1).astype(np.uint32)
    return torch.from_numpy(words.view(np.int32).copy())


def dilate_cells(occ):
    """Cell occupancy for trilinear skipping: cell (i,j,k) is occupied iff any of its 8
    corners (i+a, j+b, k+c), a,b,c in {0,1}, wrapped, is occupied."""
    out = occ.clone()
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                out |= torch.roll(occ, shifts=(-a, -b, -c), dims=(0, 1, 2))
    return out


# --------------------------------------------------------------------------- training loop

RECEPTIVE_FIELDS = [i for i in range(93, 2, -2) if i % 2 != 0]  # scripts/train.py:91


def average_pool3d_grid(tensor, receptive_field_size=3, stride=None):
    """src/grid_functions.py:173-181."""
    inp = tensor.permute(3, 0, 1, 2).unsqueeze(0)
    k = (receptive_field_size,) * 3
    out = F.avg_pool3d(inp, k, stride=stride)
    return out.squeeze().permute(1, 2, 3, 0)


def tv_loss(input_tensor):
    """scripts/train.py:44-65."""
    gx = (input_tensor[:, :-1, :, :] - input_tensor[:, 1:, :, :]).pow(2)
    gy = (input_tensor[:, :, :-1, :] - input_tensor[:, :, 1:, :]).pow(2)
    gz = (input_tensor[:-1, :, :, :] - input_tensor[1:, :, :, :]).pow(2)
    return torch.sqrt(gx.sum() + gy.sum() + gz.sum())


def _chunked_step(cells, o, d, pix, N, S, delta_step, near, gmin, pd, div_mode, tv, beta, chunk, opt):
    """One step of reference_fit over ray chunks (see its chunk_rays): mse_loss's mean and the
    beta term's mean as sums over chunks divided by the full batch's counts."""
    leaf = cells.detach().requires_grad_(True)
    pix = pix.to(leaf.dtype)
    closs = 0.0
    bl = 0.0
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        p = sample_positions(o[s:e], d[s:e], S, delta_step, near)
        n = normalize_samples(p, gmin, pd, div_mode)
        nearest, _ = nearest_values(n, leaf)
        nearest = nearest.reshape(1, e - s, S, 4)
        col = compute_alpha_weighted_pixels(nearest).reshape(e - s, 4)
        part = ((col - pix[s:e]) ** 2).sum() / (4 * N)
        total = part
        closs += float(part.detach())
        if beta > 0:
            a = nearest[..., -1]
            bpart = beta * (torch.log(a + 1e-4) - torch.log(1 - a + 1e-4)).sum() / (N * S)
            bl += float(bpart.detach())
            total = total + bpart
        total.backward()
    rec = {"closs": closs}
    if tv > 0:
        t = tv * tv_loss(leaf)
        t.backward()
        rec["tvloss"] = float(t.detach())
    if beta > 0:
        rec["betaloss"] = bl
    opt.zero_grad()
    cells.backward(leaf.grad)  # through the pooling (or straight into the leaf grid)
    opt.step()
    return rec


def reference_fit(imgs, c2w, camera_angle_x, *, gridsize, points_distance, num_samples, delta_step, lr, tv,
                  beta, uniforms, progressive_growing=True, div_mode="divide", start_step=0, init_grid=None,
                  norm_mode="match", near=0.0, chunk_rays=None, accum_dtype=None):
    """scripts/train.py:79-184 op for op on given per-step uniforms (list of (B,R,2)).

    Returns (grid, grid_grad_abs_sum, per-step loss dicts).  Rays: camera-major, R per camera;
    or, for an entry (cam_ids (N,), uv (N,2)), N rays from the given cameras (an extension for
    batch sizes the camera count does not divide: the same chain on one (1, N) ray batch).
    div_mode / norm_mode: the division and 3-norm semantics ("match": the norm of the device
    whose division div_mode names, see device_norm_mode; None: torch's on imgs' device).
    near: first sample at near + delta (extension; 0 is the reference).
    chunk_rays: for batches too large for one autograd graph (torch's GPU kernels refuse launch
    shapes beyond a few hundred million samples), the step's loss terms are formed and
    back-propagated ray chunk by ray chunk into the rendered grid cut from the graph, whose
    gradient then flows once through the pooling: the same gradient up to summation order.
    accum_dtype=torch.float64 (with chunk_rays): the FP64 evaluation of the same step -- sample
    placement and voxel choice stay float32 as in the reference, while the pooled grid, the
    gathered values, the compositing, the losses and every gradient sum are float64; the FP32
    parameter's gradient is that sum rounded once.
    """
    if norm_mode == "match":
        norm_mode = device_norm_mode(div_mode)
    dev = imgs.device
    grid_indices, grid_full, _, grid_grid_original = generate_grid(*gridsize, points_distance=points_distance,
                                                                   device=dev)
    if init_grid is not None:  # (the reference starts from zeros, src/grid_functions.py:214)
        with torch.no_grad():
            grid_full.copy_(init_grid)
    grad_abs = torch.zeros_like(grid_full)
    opt = torch.optim.Adam([grid_full], lr=lr)
    hist = []
    for i, uv in enumerate(uniforms, start=start_step):
        k = RECEPTIVE_FIELDS[i // 5] if i // 5 < len(RECEPTIVE_FIELDS) else 0
        if k > 1 and progressive_growing:
            stride = max(1, k // 4)
            start = int(k / 2)
            gg = grid_grid_original[start::stride, start::stride, start::stride]
            gi = gg.reshape(-1, 3)
            pd = points_distance * stride
            src = grid_full if accum_dtype is None else grid_full.to(accum_dtype)
            cells = average_pool3d_grid(src, receptive_field_size=k, stride=stride)
        else:
            cells = grid_full if accum_dtype is None else grid_full.to(accum_dtype)
            gi = grid_grid_original.reshape(-1, 3)
            pd = points_distance
        if isinstance(uv, tuple):  # explicit (cam_ids, uv) rays
            cam, uvr = uv
            cam = cam.to(imgs.device).long()
            B, R = 1, cam.shape[0]
            o, d, pix = camera_rays(imgs, c2w, camera_angle_x, cam, uvr[:, 0].contiguous(), uvr[:, 1].contiguous(),
                                    norm_mode)
        else:
            B, R = uv.shape[0], uv.shape[1]
            d, pix = generate_rays_batched(imgs, R, c2w, camera_angle_x, uv, norm_mode)
            o = torch.repeat_interleave(c2w[:, :3, 3], R, 0)
        N = B * R
        if chunk_rays is not None and (N > chunk_rays or accum_dtype is not None):
            rec = _chunked_step(cells, o, d, pix, N, num_samples, delta_step, near, grid_min(gi), pd, div_mode,
                                tv if k < 19 else 0.0, beta, chunk_rays, opt)
            grad_abs += torch.abs(grid_full.grad)
            hist.append(rec)
            continue
        p = sample_positions(o, d, num_samples, delta_step, near)
        n = normalize_samples(p, grid_min(gi), pd, div_mode)
        nearest, mask = nearest_values(n, cells)
        nearest = nearest.reshape(B, R, num_samples, 4)
        col = compute_alpha_weighted_pixels(nearest).reshape(B * R, 4)
        closs = F.mse_loss(col, pix)
        loss = closs
        rec = {"closs": float(closs.detach())}
        if tv > 0 and k < 19:
            t = tv * tv_loss(cells)
            loss = loss + t
            rec["tvloss"] = float(t.detach())
        if beta > 0:
            a = nearest[..., -1]
            b = beta * (torch.log(a + 1e-4) - torch.log(1 - a + 1e-4)).mean()
            loss = loss + b
            rec["betaloss"] = float(b.detach())
        opt.zero_grad()
        loss.backward()
        opt.step()
        grad_abs += torch.abs(grid_full.grad)
        hist.append(rec)
    return grid_full.detach(), grad_abs, hist
