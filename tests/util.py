"""Shared scene builders for the tests (small seeded inputs on the CPU)."""
import sys

import numpy as np
import torch

from oracle import plain


def say(*a):
    """Print and flush, so a GPU run that dies still shows which case it was in."""
    print(*a, flush=True)
    sys.stdout.flush()


def small_scene(seed, X=24, R=512, pd=0.1, alpha_scale=0.12, p_opaque=0.02, p_out=0.05):
    """Random raw grid (values spill outside [0,1] to exercise the clamp; a few exactly opaque
    cells), rays from z ~ -2.5 through the grid cube."""
    g = torch.Generator().manual_seed(seed)
    grid = torch.rand(X, X, X, 4, generator=g) * 1.2 - 0.1
    a = torch.rand(X, X, X, generator=g) * alpha_scale
    opaque = torch.rand(X, X, X, generator=g) < p_opaque
    a[opaque] = 1.0
    neg = torch.rand(X, X, X, generator=g) < p_out
    a[neg] = -0.05
    grid[..., 3] = a
    gi, _, _, _ = plain.generate_grid(X, X, X, points_distance=pd)
    gmin = plain.grid_min(gi)
    lo, hi = float(gmin[0]), float(gmin[0]) + (X - 1) * pd
    o = torch.empty(R, 3)
    o[:, :2] = torch.rand(R, 2, generator=g) * (hi - lo) * 0.6 + lo + 0.2 * (hi - lo)
    o[:, 2] = lo - 1.4
    tgt = torch.rand(R, 3, generator=g) * (hi - lo) * 1.1 + lo - 0.05 * (hi - lo)
    d = torch.nn.functional.normalize(tgt - o, dim=1)
    return grid.contiguous(), o.contiguous(), d.contiguous(), gmin, pd


def random_cameras(B, seed, radius=2.5, y_scale=1.0):
    g = torch.Generator().manual_seed(seed)
    c2w = torch.zeros(B, 4, 4)
    for b in range(B):
        q, _ = torch.linalg.qr(torch.randn(3, 3, generator=g, dtype=torch.float64))
        q = q.float()
        q[:, 1] *= y_scale
        c2w[b, :3, :3] = q
        c2w[b, :3, 3] = q[:, 2] * radius
        c2w[b, 3, 3] = 1
    return c2w


def grad_close(got, want, rel=1e-4):
    """max |got - want| <= rel * max |want|, and per element within rel for every element
    above 1% of the largest."""
    got = got.double()
    want = want.double()
    scale = want.abs().max().item()
    err = (got - want).abs()
    big = want.abs() > 1e-2 * scale
    per = (err[big] / want.abs()[big]).max().item() if bool(big.any()) else 0.0
    return err.max().item() <= rel * max(scale, 1e-30) and per <= rel, (err.max().item(), scale, per)
