"""Optimiser and regularisers of the reference training step (scripts/train.py:44-65, 89,
163-184), on libpxl kernels.

Adam: torch.optim.Adam's update (defaults betas (0.9, 0.999), eps 1e-8, no weight decay,
no amsgrad) for FP32 parameters in one fused pass that also accumulates |grad| (the
grid_cells_full_grad of scripts/train.py:184) and can write a BF16/FP16 render copy of the
updated parameter.  State keys match torch's ("step", "exp_avg", "exp_avg_sq").

tv_loss: the reference's L2 total variation sqrt(sum of squared neighbour differences over
the three axes), summed on the device in FP64.
"""
import torch

from . import _lib

_HALF = {torch.bfloat16: _lib.DTYPE_BF16, torch.float16: _lib.DTYPE_F16}


class Adam(torch.optim.Optimizer):
    """Drop-in for torch.optim.Adam([grid], lr) as scripts/train.py:89 builds it.

    grad_abs: optional {param: tensor} of |grad| accumulators updated in the same pass;
    half_copy: optional {param: bf16/fp16 tensor} receiving the updated parameter.
    """

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0, amsgrad=False,
                 grad_abs=None, half_copy=None):
        if weight_decay != 0 or amsgrad:
            raise NotImplementedError("weight_decay and amsgrad are not used by the reference")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps))
        self.grad_abs = grad_abs or {}
        self.half_copy = half_copy or {}

    @torch.no_grad()
    def step(self, closure=None):
        loss = closure() if closure is not None else None
        L = _lib.lib()
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous():
                    raise ValueError("Adam here updates contiguous float32 device parameters")
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                g = p.grad.float().contiguous()
                ga = self.grad_abs.get(p)
                hc = self.half_copy.get(p)
                _lib.check(L.pxl_adam_step(
                    _lib.ptr(p), _lib.ptr(g), _lib.ptr(st["exp_avg"]), _lib.ptr(st["exp_avg_sq"]),
                    _lib.ptr(ga), p.numel(), float(group["lr"]), float(b1), float(b2), float(group["eps"]),
                    int(st["step"]), _lib.ptr(hc), _HALF[hc.dtype] if hc is not None else 0,
                    _lib.stream_handle(p.device)))
        return loss


def adam_step_(param, grad, exp_avg, exp_avg_sq, step, *, lr, betas=(0.9, 0.999), eps=1e-8,
               grad_abs=None, half_out=None):
    """One fused Adam update on raw tensors (no Optimizer bookkeeping)."""
    _lib.check(_lib.lib().pxl_adam_step(
        _lib.ptr(param), _lib.ptr(grad), _lib.ptr(exp_avg), _lib.ptr(exp_avg_sq), _lib.ptr(grad_abs),
        param.numel(), float(lr), float(betas[0]), float(betas[1]), float(eps), int(step),
        _lib.ptr(half_out), _HALF[half_out.dtype] if half_out is not None else 0,
        _lib.stream_handle(param.device)))


def adam_tv_step_(param_in, param_out, grad, exp_avg, exp_avg_sq, step, *, tv_sumsq, tv_weight, lr,
                  betas=(0.9, 0.999), eps=1e-8, grad_abs=None, half_out=None, zero_grad=False):
    """tv_backward_(param_in, tv_sumsq, tv_weight, grad) then adam_step_ in one pass
    (pxl_adam_tv_step): the same bits, with the updated parameter written to param_out (a
    different (X,Y,Z,4) float32 buffer).  grad is left unchanged, or with zero_grad set to zero
    after use (only its nonzero elements are written), ready for the next step's scatter."""
    if param_in.data_ptr() == param_out.data_ptr():
        raise ValueError("param_out must be a different buffer from param_in")
    X, Y, Z = param_in.shape[:3]
    _lib.check(_lib.lib().pxl_adam_tv_step(
        _lib.ptr(param_in), _lib.ptr(param_out), _lib.ptr(grad), _lib.ptr(exp_avg), _lib.ptr(exp_avg_sq),
        _lib.ptr(grad_abs), X, Y, Z, _lib.ptr(tv_sumsq), float(tv_weight), float(lr), float(betas[0]),
        float(betas[1]), float(eps), int(step), _lib.ptr(half_out),
        _HALF[half_out.dtype] if half_out is not None else 0, 1 if zero_grad else 0,
        _lib.stream_handle(param_in.device)))


def tv_sumsq(grid, out=None):
    """Device FP64 scalar: sum over the three axes of squared neighbour differences."""
    if grid.dtype != torch.float32 or grid.dim() != 4 or grid.shape[-1] != 4:
        raise ValueError("tv needs a float32 (X,Y,Z,4) grid")
    g = grid.contiguous()
    s = torch.zeros(1, dtype=torch.float64, device=g.device) if out is None else out
    X, Y, Z = g.shape[:3]
    _lib.check(_lib.lib().pxl_tv_sumsq(_lib.ptr(g), X, Y, Z, _lib.ptr(s), _lib.stream_handle(g.device)))
    return s


def tv_backward_(grid, sumsq, weight, grad):
    """grad += d(weight * sqrt(sumsq))/dgrid (sumsq from tv_sumsq, read on the device)."""
    g = grid.contiguous()
    X, Y, Z = g.shape[:3]
    _lib.check(_lib.lib().pxl_tv_bwd(_lib.ptr(g), X, Y, Z, _lib.ptr(sumsq), float(weight), _lib.ptr(grad),
                                     _lib.stream_handle(g.device)))


class _TVFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid):
        s = tv_sumsq(grid)
        ctx.save_for_backward(grid, s)
        return torch.sqrt(s).to(torch.float32).reshape(())

    @staticmethod
    def backward(ctx, g):
        grid, s = ctx.saved_tensors
        grad = torch.zeros_like(grid, dtype=torch.float32)
        tv_backward_(grid, s, float(g), grad)
        return grad


def tv_loss(input_tensor):
    """scripts/train.py:44-65: sqrt(sum of squared differences along the three axes) of a
    (X,Y,Z,4) grid; differentiable."""
    return _TVFn.apply(input_tensor)
