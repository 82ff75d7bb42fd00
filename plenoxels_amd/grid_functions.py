"""Drop-in counterparts of the reference's src/grid_functions.py hot-path functions.

Same names, arguments, shapes and dtypes as the reference; the arithmetic runs in libpxl's
HIP kernels (include/pxl.h).  Grids are float32 (X,Y,Z,4).  The gathers are differentiable
w.r.t. the grid (the gradient the reference's training chain takes), not w.r.t. the sample
coordinates.  trilinear_interpolation wraps out-of-range corners where the reference raises
IndexError (its caller always passes wrapped corners).
"""
import ctypes

import numpy as np
import torch
import torch.nn.functional as F  # noqa: F401  (reference module namespace)

from . import _lib
from .render import _grid_struct


def _dev(t):
    if not t.is_cuda:
        raise ValueError("plenoxels_amd runs on a HIP device; move the tensors there first")
    return t.device


def _f32_grid(grid):
    if grid.dtype != torch.float32:
        raise TypeError(f"grid must be float32, got {grid.dtype}")
    if grid.dim() != 4 or grid.shape[-1] != 4:
        raise ValueError(f"grid must be (X,Y,Z,4), got {tuple(grid.shape)}")
    return grid.contiguous()


def _points(n):
    n = n.float().contiguous()
    if n.dim() != 2 or n.shape[1] != 3:
        raise ValueError("points must be (N,3)")
    return n


# --------------------------------------------------------------------------- bounds and wrap

def find_out_of_bound(indecies, grid):
    """src/grid_functions.py:47-63: (N,) bool, True where the (int or float) index is inside."""
    dev = _dev(indecies)
    if indecies.dtype.is_floating_point:
        idx, is_float = indecies.float().contiguous(), 1
    else:
        idx, is_float = indecies.to(torch.int64).contiguous(), 0
    N = idx.shape[0]
    mask = torch.empty(N, dtype=torch.bool, device=dev)
    _lib.check(_lib.lib().pxl_in_bounds(_lib.ptr(idx), is_float, N, _lib.dims_arg(grid.shape[:3]),
                                        _lib.ptr(mask), _lib.stream_handle(dev)))
    return mask


def fix_out_of_bounds(indices, grid):
    """src/grid_functions.py:66-79: wraps the int64 (N,3) indices in place (Python modulo) and
    returns the three column views."""
    dev = _dev(indices)
    if indices.dtype != torch.int64:
        raise TypeError("indices must be int64")
    work = indices if indices.is_contiguous() else indices.contiguous()
    _lib.check(_lib.lib().pxl_wrap_indices(_lib.ptr(work), work.shape[0], _lib.dims_arg(grid.shape[:3]),
                                           _lib.stream_handle(dev)))
    if work is not indices:
        indices.copy_(work)
    return indices[:, 0], indices[:, 1], indices[:, 2]


# --------------------------------------------------------------------------- nearest

class _NearestFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n, grid):
        g = _f32_grid(grid)
        N = n.shape[0]
        vals = torch.empty((N, 4), dtype=torch.float32, device=n.device)
        mask = torch.empty(N, dtype=torch.bool, device=n.device)
        gs = _grid_struct(g)
        _lib.check(_lib.lib().pxl_nearest_gather(_lib.ptr(n), N, ctypes.byref(gs), _lib.ptr(vals),
                                                 _lib.ptr(mask), _lib.stream_handle(n.device)))
        ctx.save_for_backward(n, g)
        ctx.mark_non_differentiable(mask)
        return vals, mask

    @staticmethod
    def backward(ctx, gvals, gmask):
        n, g = ctx.saved_tensors
        grad = torch.zeros_like(g)
        gv = gvals.float().contiguous()
        gs = _grid_struct(g)
        _lib.check(_lib.lib().pxl_nearest_scatter(_lib.ptr(n), n.shape[0], ctypes.byref(gs), _lib.ptr(gv),
                                                  _lib.ptr(grad), _lib.stream_handle(n.device)))
        return None, grad


def get_nearest_voxels(normalized_samples_for_indices, grid, receptive_field=1):
    """src/grid_functions.py:103-114: (grid[round-half-even(n) wrapped] (N,4), in-bounds mask (N,)).

    Every point is gathered (out-of-grid ones at the wrapped index), as the reference does
    before its caller multiplies by the mask.
    """
    _dev(normalized_samples_for_indices)
    return _NearestFn.apply(_points(normalized_samples_for_indices), grid)


# --------------------------------------------------------------------------- trilinear

def get_grid_points_indices(normalized_samples_for_indecies):
    """src/grid_functions.py:220-246: (N,8,3) int64 corners, x{ceil,floor} x y{..} x z{..}."""
    dev = _dev(normalized_samples_for_indecies)
    n = _points(normalized_samples_for_indecies)
    out = torch.empty((n.shape[0], 8, 3), dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().pxl_corner_indices(_lib.ptr(n), n.shape[0], _lib.ptr(out), _lib.stream_handle(dev)))
    return out


class _CollectFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n, grid):
        g = _f32_grid(grid)
        N = n.shape[0]
        out = torch.empty((N * 8, 4), dtype=torch.float32, device=n.device)
        mask = torch.empty(N, dtype=torch.bool, device=n.device)
        gs = _grid_struct(g)
        _lib.check(_lib.lib().pxl_collect_cells(_lib.ptr(n), N, ctypes.byref(gs), _lib.ptr(out),
                                                _lib.ptr(mask), _lib.stream_handle(n.device)))
        ctx.save_for_backward(n, g)
        ctx.mark_non_differentiable(mask)
        return out, mask

    @staticmethod
    def backward(ctx, gout, gmask):
        n, g = ctx.saved_tensors
        grad = torch.zeros_like(g)
        go = gout.float().contiguous()
        gs = _grid_struct(g)
        _lib.check(_lib.lib().pxl_collect_cells_bwd(_lib.ptr(n), n.shape[0], ctypes.byref(gs), _lib.ptr(go),
                                                    _lib.ptr(grad), _lib.stream_handle(n.device)))
        return None, grad


def collect_cell_information_via_indices(normalized_samples_for_indices, B):
    """src/grid_functions.py:154-170: (values at the 8 wrapped corners (N*8,4), float
    in-bounds mask (N,))."""
    _dev(normalized_samples_for_indices)
    return _CollectFn.apply(_points(normalized_samples_for_indices), B)


class _TrilinearFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n, corners, grid):
        g = _f32_grid(grid)
        N = n.shape[0]
        out = torch.empty((N, 4), dtype=torch.float32, device=n.device)
        gs = _grid_struct(g)
        _lib.check(_lib.lib().pxl_trilinear(_lib.ptr(n), _lib.ptr(corners), N, ctypes.byref(gs),
                                            _lib.ptr(out), _lib.stream_handle(n.device)))
        ctx.save_for_backward(n, corners, g)
        return out

    @staticmethod
    def backward(ctx, gout):
        n, corners, g = ctx.saved_tensors
        grad = torch.zeros_like(g)
        go = gout.float().contiguous()
        gs = _grid_struct(g)
        _lib.check(_lib.lib().pxl_trilinear_bwd(_lib.ptr(n), _lib.ptr(corners), n.shape[0], ctypes.byref(gs),
                                                _lib.ptr(go), _lib.ptr(grad), _lib.stream_handle(n.device)))
        return None, None, grad


def trilinear_interpolation(normalized_samples_for_indecies, selected_points, grid_cells):
    """src/grid_functions.py:7-44: (N,4) values interpolated from the (pre-wrapped) corners
    (N,8,3) with signed frac weights, in the reference's lerp order."""
    _dev(normalized_samples_for_indecies)
    c = selected_points.to(torch.int64).contiguous()
    if c.dim() != 3 or tuple(c.shape[1:]) != (8, 3):
        raise ValueError("selected_points must be (N,8,3)")
    return _TrilinearFn.apply(_points(normalized_samples_for_indecies), c, grid_cells)


# --------------------------------------------------------------------------- grid layout, pooling

def _pool_workspace(dims, k, s, device):
    n = _lib.lib().pxl_avg_pool3d_workspace_floats(*dims, k, s)
    if n < 0:
        raise ValueError(f"average pooling needs 1 <= k <= dims and s >= 1 (k={k}, s={s}, dims={dims})")
    return torch.empty(n, dtype=torch.float32, device=device)


def pool_backward_(grad_out, dims, k, s, grad_grid, accumulate=True):
    """grad_grid (X,Y,Z,4) += (accumulate) or = the average-pooling backward of grad_out
    (pxl_avg_pool3d_bwd; every cell is written, 0 outside all windows)."""
    go = grad_out.float().contiguous()
    ws = _pool_workspace(dims, k, s, go.device)
    _lib.check(_lib.lib().pxl_avg_pool3d_bwd(_lib.ptr(go), *dims, k, s, _lib.ptr(grad_grid),
                                             1 if accumulate else 0, _lib.ptr(ws),
                                             _lib.stream_handle(go.device)))


class _PoolFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, k, s):
        g = _f32_grid(grid)
        X, Y, Z = g.shape[:3]
        ws = _pool_workspace((X, Y, Z), k, s, g.device)
        out = torch.empty(((X - k) // s + 1, (Y - k) // s + 1, (Z - k) // s + 1, 4), dtype=torch.float32,
                          device=g.device)
        _lib.check(_lib.lib().pxl_avg_pool3d(_lib.ptr(g), X, Y, Z, k, s, _lib.ptr(out), _lib.ptr(ws),
                                             _lib.stream_handle(g.device)))
        ctx.k, ctx.s, ctx.dims = k, s, (X, Y, Z)
        return out

    @staticmethod
    def backward(ctx, gout):
        grad = torch.empty((*ctx.dims, 4), dtype=torch.float32, device=gout.device)
        pool_backward_(gout, ctx.dims, ctx.k, ctx.s, grad, accumulate=False)
        return grad, None, None


def average_pool3d_grid(tensor, receptive_field_size=3, stride=None):
    """src/grid_functions.py:173-181: avg_pool3d of the (X,Y,Z,4) grid with a cubic window,
    stride defaulting to the window; returns (Xo,Yo,Zo,4) (contiguous; the reference returns
    a permuted view with the same values)."""
    _dev(tensor)
    k = int(receptive_field_size)
    s = k if stride is None else int(stride)
    return _PoolFn.apply(tensor, k, s)


def generate_grid(sx, sy, sz, points_distance=0.5, info_size=4, device="cuda"):
    """src/grid_functions.py:184-217: (grid_coords (XYZ,3) f32, grid_cells (X,Y,Z,C) zeros
    requiring grad, meshgrid (X,Y,Z,3) int64, grid_grid (X,Y,Z,3) f32)."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("plenoxels_amd runs on a HIP device")
    coords = torch.empty((sx, sy, sz, 3), dtype=torch.float32, device=dev)
    mesh = torch.empty((sx, sy, sz, 3), dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().pxl_grid_coords(sx, sy, sz, float(np.float32(points_distance)), _lib.ptr(coords),
                                          _lib.ptr(mesh), _lib.stream_handle(dev)))
    grid_cells = torch.zeros([sx, sy, sz, info_size], requires_grad=True, device=dev)
    return coords.reshape(sx * sy * sz, 3), grid_cells, mesh, coords
