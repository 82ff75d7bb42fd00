"""Fused render path: torch.autograd.Function over the libpxl kernels.

``render_rays`` returns, for a batch of rays, the (R,4) RGBA that the reference training
chain produces (scripts/train.py:130-153 -- sample placement, normalisation, nearest or
trilinear gather of grid.clip(0,1) with the out-of-grid mask, compute_alpha_weighted_pixels)
and differentiates it back to the grid with a hand-written backward kernel instead of
autograd's per-sample graph.  ``train_step`` fuses render, the mean-MSE gradient
(scripts/train.py:156) and the backward into one pass.

Extensions (keyword arguments; defaults reproduce the reference): ``near`` offset of the
sample positions, ``t_stop`` early termination, ``occupancy`` skipping, ``div_mode``.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

_DTYPES = {torch.float32: _lib.DTYPE_F32, torch.bfloat16: _lib.DTYPE_BF16,
           torch.float16: _lib.DTYPE_F16}
_INTERP = {"nearest": _lib.INTERP_NEAREST, "trilinear": _lib.INTERP_TRILINEAR}


@dataclass(frozen=True)
class MarchConfig:
    """How samples are placed and looked up (src/ray_sampling.py:161-167, :12-13)."""
    num_samples: int
    delta_step: float
    grid_min: tuple
    points_distance: float
    interp: str = "nearest"
    near: float = 0.0
    t_stop: float = 0.0
    div_mode: str = "reciprocal"  # "reciprocal": ATen GPU division; "divide": ATen CPU

    def to_struct(self):
        m = _lib.March()
        m.n_samples = int(self.num_samples)
        if self.interp not in _INTERP:
            raise ValueError(f"interp must be one of {sorted(_INTERP)}, got {self.interp!r}")
        m.interp = _INTERP[self.interp]
        if self.div_mode not in ("reciprocal", "divide"):
            raise ValueError(f"div_mode must be 'reciprocal' or 'divide', got {self.div_mode!r}")
        m.flags = _lib.FLAG_EXACT_DIV if self.div_mode == "divide" else 0
        m.delta = float(np.float32(self.delta_step))
        m.near = float(np.float32(self.near))
        m.t_stop = float(self.t_stop)
        m.pd = float(np.float32(self.points_distance))
        m.inv_pd = float(np.float32(1.0) / np.float32(self.points_distance))
        for a in range(3):
            m.grid_min[a] = float(self.grid_min[a])
        return m


def grid_min_of(grid_indices):
    """grid_indices.min(0)[0] (src/ray_sampling.py:13) as three float32 values, computed once
    per grid instead of once per step."""
    if torch.is_tensor(grid_indices):
        return tuple(float(x) for x in grid_indices.min(0)[0].float().cpu().tolist())
    return tuple(float(np.float32(x)) for x in grid_indices)


def _grid_struct(grid, occupancy=None):
    if grid.dim() != 4 or grid.shape[-1] != 4:
        raise ValueError(f"grid must be (X,Y,Z,4), got {tuple(grid.shape)}")
    if grid.dtype not in _DTYPES:
        raise TypeError(f"grid dtype must be float32, bfloat16 or float16, got {grid.dtype}")
    if not grid.is_cuda:
        raise ValueError("grid must be a HIP device tensor")
    if not grid.is_contiguous():
        raise ValueError("grid must be contiguous")
    g = _lib.Grid()
    g.data = grid.data_ptr()
    for a in range(3):
        g.dims[a] = grid.shape[a]
    g.channels = 4
    g.dtype = _DTYPES[grid.dtype]
    if occupancy is not None:
        if occupancy.dtype != torch.int32 or not occupancy.is_cuda or not occupancy.is_contiguous():
            raise ValueError("occupancy must be a contiguous int32 device tensor of packed bits")
        need = occupancy_words(grid.shape[:3])
        if occupancy.numel() < need:
            raise ValueError(f"occupancy needs {need} words, got {occupancy.numel()}")
        g.occupancy = occupancy.data_ptr()
    return g


def _rays_struct(origins=None, dirs=None, camera=None):
    r = _lib.Rays()
    if origins is not None:
        for name, t in (("origins", origins), ("dirs", dirs)):
            if t is None or t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 3:
                raise ValueError(f"{name} must be (R,3) float32")
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous device tensor")
        if origins.shape[0] != dirs.shape[0]:
            raise ValueError("origins and dirs must have the same number of rays")
        r.n_rays = origins.shape[0]
        r.origins = origins.data_ptr()
        r.dirs = dirs.data_ptr()
        return r
    cam = camera
    r.n_rays = cam.cam_ids.shape[0]
    r.cam_ids = cam.cam_ids.data_ptr()
    r.uv = cam.uv.data_ptr()
    r.c2w = cam.c2w.data_ptr()
    r.n_cams = cam.c2w.shape[0]
    r.camera_angle_x = float(np.float32(cam.camera_angle_x))
    r.flags = _lib.RAYS_CPU_NORM if cam.norm_mode == "cpu" else 0
    if cam.imgs is not None:
        r.imgs = cam.imgs.data_ptr()
        r.img_h = cam.imgs.shape[1]
        r.img_w = cam.imgs.shape[2]
    return r


@dataclass
class CameraRays:
    """Rays in camera form: camera index and uniforms (u,v) per ray, plus the camera stack
    (src/ray_sampling.py:195-264 with the uniforms drawn by the caller).

    norm_mode: summation order of the direction and aspect 3-norms -- "gpu" (default), torch's
    GPU norm, sqrt((x*x + z*z) + y*y), what the reference computes on its default device
    'cuda'; "cpu", torch's CPU norm, what it computes with device='cpu'."""
    cam_ids: torch.Tensor          # (R,) int32
    uv: torch.Tensor               # (R,2) float32 in [0,1]
    c2w: torch.Tensor              # (B,4,4) float32
    camera_angle_x: float
    imgs: torch.Tensor = None      # (B,H,W,4) float32 targets
    norm_mode: str = "gpu"

    def validate(self):
        if self.norm_mode not in ("gpu", "cpu"):
            raise ValueError(f"norm_mode must be 'gpu' or 'cpu', got {self.norm_mode!r}")
        if self.cam_ids.dtype != torch.int32 or self.cam_ids.dim() != 1:
            raise ValueError("cam_ids must be (R,) int32")
        if self.uv.dtype != torch.float32 or tuple(self.uv.shape) != (self.cam_ids.shape[0], 2):
            raise ValueError("uv must be (R,2) float32")
        if self.c2w.dtype != torch.float32 or self.c2w.dim() != 3 or tuple(self.c2w.shape[1:]) != (4, 4):
            raise ValueError("c2w must be (B,4,4) float32")
        ts = [self.cam_ids, self.uv, self.c2w] + ([self.imgs] if self.imgs is not None else [])
        for t in ts:
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError("camera-form tensors must be contiguous device tensors")
        if self.imgs is not None:
            if self.imgs.dtype != torch.float32 or self.imgs.dim() != 4 or self.imgs.shape[-1] != 4:
                raise ValueError("imgs must be (B,H,W,4) float32")
            if self.imgs.shape[0] != self.c2w.shape[0]:
                raise ValueError("imgs and c2w must cover the same cameras")


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, origins, dirs, cfg, occupancy, render_copy):
        L = _lib.lib()
        src = grid if render_copy is None else render_copy  # the cells the kernels gather
        R = origins.shape[0]
        out = torch.empty((R, 4), dtype=torch.float32, device=grid.device)
        need_aux = ctx.needs_input_grad[0]
        aux = (torch.empty((R, int(L.pxl_aux_floats(int(cfg.num_samples)))), dtype=torch.float32,
                           device=grid.device) if need_aux else None)
        g = _grid_struct(src, occupancy)
        m = cfg.to_struct()
        r = _rays_struct(origins, dirs)
        _lib.check(L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r),
                                    _lib.ptr(out), _lib.ptr(aux), _lib.stream_handle(grid.device)))
        ctx.cfg = cfg
        if need_aux:
            ctx.save_for_backward(src, origins, dirs, aux, occupancy)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        src, origins, dirs, aux, occupancy = ctx.saved_tensors
        L = _lib.lib()
        grad_out = grad_out.contiguous().float()
        # the gradient is accumulated in FP32 and returned in FP32: the differentiable grid is
        # always an FP32 tensor (render_rays rejects half-precision leaves)
        grad = torch.zeros(src.shape, dtype=torch.float32, device=src.device)
        g = _grid_struct(src, occupancy)
        m = ctx.cfg.to_struct()
        r = _rays_struct(origins, dirs)
        _lib.check(L.pxl_render_bwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r),
                                    _lib.ptr(grad_out), _lib.ptr(aux), _lib.ptr(grad),
                                    _lib.stream_handle(src.device)))
        return grad, None, None, None, None, None


def render_rays(grid, origins, dirs, *, num_samples, delta_step, grid_min, points_distance,
                interp="nearest", near=0.0, t_stop=0.0, occupancy=None, div_mode="reciprocal",
                render_copy=None):
    """(R,4) RGBA of the reference chain for explicit rays, differentiable w.r.t. grid.

    grid: raw (X,Y,Z,4) parameter (clipped inside); origins/dirs: (R,3) float32.
    grid_min: grid_indices.min(0)[0] (see grid_min_of).  occupancy: packed int32 bit words
    (see pxl.h); trilinear reads the floor cell's bit.
    Precision: gradients are FP32.  A grid that requires grad must be float32; to render from
    reduced-precision storage, pass the FP32 master as `grid` and a bfloat16/float16 copy of it
    as `render_copy` -- the kernels gather from the copy (half the bytes) and the FP32 gradient
    goes to the master.  A bfloat16/float16 `grid` is accepted for forward-only use.
    """
    if grid.dtype != torch.float32 and grid.requires_grad and torch.is_grad_enabled():
        raise TypeError(f"render_rays differentiates float32 grids (gradients stay FP32); got a {grid.dtype} "
                        "leaf -- keep a float32 master as `grid` and pass the half copy as render_copy=")
    if render_copy is not None:
        if grid.dtype != torch.float32:
            raise TypeError("render_copy goes with a float32 master grid")
        if tuple(render_copy.shape) != tuple(grid.shape) or render_copy.device != grid.device:
            raise ValueError("render_copy must have the grid's shape and device")
        render_copy = render_copy.detach()
    cfg = MarchConfig(int(num_samples), float(delta_step), tuple(grid_min), float(points_distance),
                      interp, float(near), float(t_stop), div_mode)
    return _RenderFn.apply(grid, origins.contiguous(), dirs.contiguous(), cfg, occupancy, render_copy)


def camera_rays(camera, want_targets=True):
    """(origins, dirs, targets) for camera-form rays: src/ray_sampling.py:212-264 per ray."""
    camera.validate()
    L = _lib.lib()
    R = camera.cam_ids.shape[0]
    dev = camera.c2w.device
    o = torch.empty((R, 3), dtype=torch.float32, device=dev)
    d = torch.empty((R, 3), dtype=torch.float32, device=dev)
    t = torch.empty((R, 4), dtype=torch.float32, device=dev) if want_targets else None
    r = _rays_struct(camera=camera)
    _lib.check(L.pxl_camera_rays(ctypes.byref(r), _lib.ptr(o), _lib.ptr(d), _lib.ptr(t),
                                 _lib.stream_handle(dev)))
    return o, d, t


def corner_blocks(grid):
    """Zeroed corner blocks for `grid` ((X*Y*Z, 16) float32 on its device, one 64-byte entry
    per cell): the accumulation buffer of train_step(..., blocks=...), kept zero between
    steps by blocks_reduce."""
    X, Y, Z = (int(v) for v in grid.shape[:3])
    return torch.zeros((X * Y * Z, 16), dtype=torch.float32, device=grid.device)


class CornerBlockPair:
    """Two corner-block buffers used alternately: each reduce sums the active buffer and zeroes
    the other one in the same pass, so no separate clearing pass is needed."""

    def __init__(self, grid):
        self.buffers = [corner_blocks(grid), corner_blocks(grid)]
        self.active = 0

    @property
    def blocks(self):
        return self.buffers[self.active]

    @property
    def other(self):
        return self.buffers[1 - self.active]

    def flip(self):
        self.active = 1 - self.active


def blocks_reduce(grid, blocks, grad_grid, occupancy=None, accumulate=True, clear=None, reachable_only=False):
    """grad_grid (+)= the gradient held in corner blocks (pxl_blocks_reduce).  clear: the
    blocks to zero for the next step -- default `blocks` itself (a second pass); another
    buffer is zeroed within the same pass; False: none (clear them with blocks_clear).
    occupancy: the one the training steps used.  reachable_only (with accumulate=False): write
    only the cells a step with this occupancy can make nonzero, for a gradient buffer that holds
    zeros everywhere else (zeroed once, used with one occupancy)."""
    L = _lib.lib()
    _check_blocks(grid, blocks, grad_grid)
    if clear is None:
        clear = blocks
    elif clear is False:
        clear = None
    elif clear is not blocks:
        _check_blocks(grid, clear, None)
    g = _grid_struct(grid, occupancy)
    if reachable_only and accumulate:
        raise ValueError("reachable_only overwrites: use it with accumulate=False")
    mode = (_lib.REDUCE_ACCUMULATE if accumulate else
            _lib.REDUCE_OVERWRITE_REACHABLE if reachable_only else _lib.REDUCE_OVERWRITE)
    _lib.check(L.pxl_blocks_reduce(ctypes.byref(g), _lib.ptr(blocks), _lib.ptr(grad_grid), mode,
                                   _lib.ptr(clear), _lib.stream_handle(grid.device)))


def blocks_clear(grid, blocks, occupancy=None):
    """Zero the corner-block entries a step with `occupancy` can write (pxl_blocks_clear), on the
    current stream."""
    _check_blocks(grid, blocks, None)
    g = _grid_struct(grid, occupancy)
    _lib.check(_lib.lib().pxl_blocks_clear(ctypes.byref(g), _lib.ptr(blocks), _lib.stream_handle(grid.device)))


def _check_blocks(grid, blocks, grad_grid):
    X, Y, Z = (int(v) for v in grid.shape[:3])
    if (blocks.dtype != torch.float32 or not blocks.is_contiguous() or blocks.numel() != X * Y * Z * 16
            or blocks.device != grid.device):
        raise ValueError("blocks must be contiguous float32 with X*Y*Z*16 elements on the grid's device")
    if grad_grid is not None and (grad_grid.dtype != torch.float32 or tuple(grad_grid.shape) != (X, Y, Z, 4)
                                  or not grad_grid.is_contiguous()):
        raise ValueError("grad_grid must be a contiguous (X,Y,Z,4) float32 tensor")


def train_step(grid, grad_grid, *, cfg, origins=None, dirs=None, camera=None, targets=None,
               n_total_rays=None, occupancy=None, out=None, ray_sq_err=None, beta=0.0,
               n_total_samples=None, ray_beta=None, blocks=None, overwrite=False, clear=None,
               fused_camera=False):
    """Render + loss gradient + backward in one kernel pass (scripts/train.py:130-181 minus
    the optimiser and the TV term).

    Accumulates d loss / d grid into grad_grid ((X,Y,Z,4) float32; zero it first for a plain
    gradient).  loss = mse_loss(out, targets) + beta * mean_samples(log(a+1e-4) -
    log(1-a+1e-4)).  Rays are explicit (origins, dirs) or a CameraRays; targets (R,4), or
    None to gather them from camera.imgs.  n_total_rays / n_total_samples: the normalisers
    of the full batch this call is part of (default: this call's).  ray_sq_err (R,) and
    ray_beta (R,) receive per-ray loss partials when given (see pxl_loss in pxl.h).

    Camera-form rays are generated by pxl_camera_rays into (R,3)+(R,3)+(R,4) buffers first and
    the step runs on explicit rays (the same bits; the separate generation pass takes the
    dependent camera loads and the scattered target gather off the step's critical path, 16%
    faster on sparse256); fused_camera=True generates them inside the step kernel instead.

    blocks (trilinear only): zeroed corner blocks (corner_blocks(grid)); the gradient is
    accumulated there and reduced into grad_grid before returning -- grad_grid is then
    overwritten when overwrite=True (no zero-fill needed), else added to; overwrite="reachable"
    writes only the cells a step with this occupancy can make nonzero (blocks_reduce's
    reachable_only); `clear` as in blocks_reduce.  grad_grid=None leaves it in the blocks for a later blocks_reduce
    (several calls, one reduce).
    """
    L = _lib.lib()
    if blocks is not None:
        _check_blocks(grid, blocks, grad_grid)
        if cfg.interp != "trilinear":
            raise ValueError("corner blocks hold trilinear gradients")
    elif grad_grid is None or grad_grid.dtype != torch.float32 or tuple(grad_grid.shape) != tuple(grid.shape) \
            or not grad_grid.is_contiguous():
        raise ValueError("grad_grid must be a contiguous float32 tensor shaped like grid")
    if camera is not None:
        camera.validate()
        if targets is None and camera.imgs is None:
            raise ValueError("targets or camera.imgs is required")
        if fused_camera:
            r = _rays_struct(camera=camera)
        else:
            origins, dirs, gathered = camera_rays(camera, want_targets=targets is None)
            targets = gathered if targets is None else targets
            r = _rays_struct(origins, dirs)
    else:
        r = _rays_struct(origins.contiguous(), dirs.contiguous())
    R = r.n_rays
    if targets is not None:
        if targets.dtype != torch.float32 or tuple(targets.shape) != (R, 4) or not targets.is_contiguous():
            raise ValueError("targets must be contiguous (R,4) float32")
    for name, t in (("out", out), ("ray_sq_err", ray_sq_err), ("ray_beta", ray_beta)):
        if t is not None and (t.dtype != torch.float32 or not t.is_contiguous() or
                              t.numel() != (4 * R if name == "out" else R)):
            raise ValueError(f"{name} has the wrong shape or dtype")
    n_rays_total = R if n_total_rays is None else int(n_total_rays)
    n_samples_total = (R * cfg.num_samples) if n_total_samples is None else int(n_total_samples)
    loss = _lib.Loss()
    loss.targets = None if targets is None else targets.data_ptr()
    # an empty batch (R = 0) launches nothing; its scales are never used
    loss.grad_scale = 2.0 / (4.0 * n_rays_total) if n_rays_total > 0 else 0.0
    loss.beta_scale = float(beta) / n_samples_total if beta and n_samples_total > 0 else 0.0
    loss.ray_sq_err = None if ray_sq_err is None else ray_sq_err.data_ptr()
    loss.ray_beta = None if ray_beta is None else ray_beta.data_ptr()
    g = _grid_struct(grid, occupancy)
    m = cfg.to_struct()
    stream = _lib.stream_handle(grid.device)
    if blocks is None:
        _lib.check(L.pxl_train_step(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), ctypes.byref(loss),
                                    _lib.ptr(out), _lib.ptr(grad_grid), stream))
    else:
        _lib.check(L.pxl_train_step_blocks(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), ctypes.byref(loss),
                                           _lib.ptr(out), _lib.ptr(blocks), stream))
        if grad_grid is not None:
            if clear is None:
                clear = blocks
            elif clear is not blocks:
                _check_blocks(grid, clear, None)
            mode = (_lib.REDUCE_ACCUMULATE if not overwrite else
                    _lib.REDUCE_OVERWRITE_REACHABLE if overwrite == "reachable" else _lib.REDUCE_OVERWRITE)
            _lib.check(L.pxl_blocks_reduce(ctypes.byref(g), _lib.ptr(blocks), _lib.ptr(grad_grid),
                                           mode, _lib.ptr(clear), stream))
    return ray_sq_err


BETA_EPS = 1e-4


def beta_term_zero():
    """f(0) of the per-sample beta term f(a) = log(a + 1e-4) - log(1 - a + 1e-4)."""
    return float(np.log(np.float32(BETA_EPS)) - np.log(np.float32(1.0 + BETA_EPS)))


class LossMean:
    """mse_loss's mean (scripts/train.py:156) from per-ray squared errors (train_step's
    ray_sq_err) on the device: scale * sum in FP64, one launch, the same bits on every call
    (pxl_scaled_sum).  Holds the call's small work buffer: use one instance per stream."""

    def __init__(self, device):
        n = int(_lib.lib().pxl_scaled_sum_work_bytes())
        self.work = torch.zeros(n, dtype=torch.uint8, device=device)

    def __call__(self, ray_sq_err, n_rays=None):
        """0-dim float32 tensor: sum(ray_sq_err) / (4 * n_rays) (n_rays defaults to its length)."""
        v = ray_sq_err
        if v.dtype != torch.float32 or v.dim() != 1 or not v.is_contiguous() or v.device != self.work.device:
            raise ValueError("ray_sq_err must be a contiguous 1-D float32 tensor on the work buffer's device")
        n_rays = v.shape[0] if n_rays is None else int(n_rays)
        if n_rays <= 0:
            raise ValueError("n_rays must be positive")
        out = torch.empty((), dtype=torch.float32, device=v.device)
        _lib.check(_lib.lib().pxl_scaled_sum(_lib.ptr(v), v.shape[0], 1.0 / (4.0 * n_rays), _lib.ptr(out),
                                             _lib.ptr(self.work), _lib.stream_handle(v.device)))
        return out


def fill_(t, value=0.0):
    """In-place fill of a contiguous float32 device tensor with the library's fill kernel."""
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("fill_ needs a contiguous float32 device tensor")
    _lib.check(_lib.lib().pxl_fill_f32(_lib.ptr(t), t.numel(), ctypes.c_float(value),
                                       _lib.stream_handle(t.device)))
    return t


def fill_reachable_(grad, occupancy, interp, value=0.0):
    """The zero-fill of a training step with `occupancy` (pxl_fill_reachable_f32): grad (X,Y,Z,4)
    float32 = value at every cell the step's scatter can write (nearest: the occupied cells;
    trilinear: every cell of the 8x8x8 tiles their corners reach), other cells untouched -- for
    a gradient buffer zeroed once and used with one occupancy."""
    if grad.dtype != torch.float32 or grad.dim() != 4 or grad.shape[-1] != 4 or not grad.is_contiguous():
        raise ValueError("grad must be a contiguous (X,Y,Z,4) float32 tensor")
    if interp not in _INTERP:
        raise ValueError(f"interp must be one of {sorted(_INTERP)}")
    g = _grid_struct(grad, occupancy)
    _lib.check(_lib.lib().pxl_fill_reachable_f32(ctypes.byref(g), _INTERP[interp], _lib.ptr(grad),
                                                 ctypes.c_float(value), _lib.stream_handle(grad.device)))
    return grad


def occupancy_words(dims):
    """int32 words of the occupancy bits of an (X,Y,Z) grid: 16 per 8x8x8 tile (pxl.h)."""
    X, Y, Z = (int(v) for v in dims)
    n = int(_lib.lib().pxl_occupancy_words(X, Y, Z))
    if n < 0:
        raise ValueError(f"invalid grid dims {(X, Y, Z)}")
    return n


def occupancy_dilate(occ_bits, dims):
    """Cell occupancy for trilinear skipping from point occupancy bits (see pxl.h)."""
    X, Y, Z = dims
    if occ_bits.numel() < occupancy_words(dims):
        raise ValueError(f"occupancy needs {occupancy_words(dims)} words, got {occ_bits.numel()}")
    out = torch.empty_like(occ_bits)
    _lib.check(_lib.lib().pxl_occupancy_dilate(_lib.ptr(occ_bits), X, Y, Z, _lib.ptr(out),
                                               _lib.stream_handle(occ_bits.device)))
    return out


def alpha_occupancy(grid, interp="nearest"):
    """Occupancy bits of the cells whose raw alpha is >= 0 (trilinear: floor cells any of whose
    8 corners is), from the grid itself (pxl_alpha_occupancy).  Skipping the other samples
    changes no finite result of the render and training kernels: their clamped alpha is 0 and
    the clamp's gradient mask is 0."""
    X, Y, Z = (int(v) for v in grid.shape[:3])
    bits = torch.empty(occupancy_words((X, Y, Z)), dtype=torch.int32, device=grid.device)
    g = _grid_struct(grid, None)
    _lib.check(_lib.lib().pxl_alpha_occupancy(ctypes.byref(g), 1 if interp == "trilinear" else 0, _lib.ptr(bits),
                                              _lib.stream_handle(grid.device)))
    return bits


def pack_occupancy(occ):
    """(X,Y,Z) bool/uint8 device tensor -> int32 bit words for the kernels' occupancy argument
    (8x8x8 tiles of 16 words; cell (x,y,z) is bit (y&3)*8 + (z&7) of word (x&7)*2 + ((y>>2)&1)
    of tile ((x>>3)*TY + (y>>3))*TZ + (z>>3), TY = ceil(Y/8), TZ = ceil(Z/8))."""
    if not occ.is_cuda or occ.dim() != 3:
        raise ValueError("occupancy must be an (X,Y,Z) device tensor")
    src = occ.to(torch.uint8).contiguous()
    X, Y, Z = (int(v) for v in src.shape)
    out = torch.empty(occupancy_words((X, Y, Z)), dtype=torch.int32, device=occ.device)
    _lib.check(_lib.lib().pxl_pack_occupancy(_lib.ptr(src), X, Y, Z, _lib.ptr(out),
                                             _lib.stream_handle(occ.device)))
    return out


def render_camera(grid, camera, cfg, occupancy=None):
    """Forward-only render of camera-form rays: (R,4) RGBA (the chain of
    src/visualization.py:125-146 in one kernel)."""
    camera.validate()
    L = _lib.lib()
    r = _rays_struct(camera=camera)
    out = torch.empty((r.n_rays, 4), dtype=torch.float32, device=grid.device)
    g = _grid_struct(grid, occupancy)
    m = cfg.to_struct()
    _lib.check(L.pxl_render_fwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), _lib.ptr(out), None,
                                _lib.stream_handle(grid.device)))
    return out
