"""Property tests of the oracle (oracle/plain.py) over generated inputs, on the CPU.

Each property is fixed by the mathematics or by Python/NumPy semantics, not by the oracle's
own formula: round-half-even (src/grid_functions.py:103-114 rounds with torch.round), the
sign-of-divisor wrap of src/grid_functions.py:66-79, the in-bounds test of :47-63, trilinear
interpolation (:7-44) reproducing affine fields exactly, and, for the composed chain of
scripts/train.py:142-156, torch.autograd.gradcheck in FP64 of the rendered RGBA with
respect to the grid.
"""
import math

import numpy as np
import pytest
import torch
from hypothesis import given, settings
from hypothesis import strategies as st

from oracle import plain

SETTINGS = settings(max_examples=60, deadline=None, derandomize=True)


@SETTINGS
@given(st.integers(-10_000, 10_000), st.sampled_from([-0.5, 0.5]))
def test_ties_round_to_even(k, half):
    x = torch.tensor([[k + half, k + half, k + half]], dtype=torch.float32)
    idx = torch.round(x).to(torch.long)
    want = int(np.rint(np.float64(k + half)))       # IEEE round-half-even
    assert want % 2 == 0
    assert idx.tolist() == [[want] * 3]


@SETTINGS
@given(st.lists(st.floats(-1e4, 1e4, allow_nan=False, width=32), min_size=3, max_size=3))
def test_rounding_is_nearest(v):
    x = torch.tensor([v], dtype=torch.float32)
    r = torch.round(x)
    assert bool(((r - x).abs() <= 0.5).all())
    assert r.tolist() == [[float(np.rint(np.float32(a))) for a in v]]


@SETTINGS
@given(st.tuples(st.integers(1, 9), st.integers(1, 9), st.integers(1, 9)),
       st.lists(st.tuples(st.integers(-40, 40), st.integers(-40, 40), st.integers(-40, 40)),
                min_size=1, max_size=20))
def test_wrap_is_python_modulo_and_bounds_are_half_open(dims, pts):
    idx = torch.tensor(pts, dtype=torch.long)
    w = plain.wrap_indices(idx, dims)
    assert w.tolist() == [[a % n for a, n in zip(p, dims)] for p in pts]
    inside = plain.find_out_of_bound(idx, dims)
    assert inside.tolist() == [all(0 <= a < n for a, n in zip(p, dims)) for p in pts]
    # the float test of the trilinear path on the unrounded coordinate
    f = idx.float() + 0.25
    inside_f = plain.find_out_of_bound(f, dims)
    assert inside_f.tolist() == [all(0 <= a + 0.25 < n for a, n in zip(p, dims)) for p in pts]


@SETTINGS
@given(st.tuples(st.integers(2, 7), st.integers(2, 7), st.integers(2, 7)),
       st.floats(-0.05, 0.05), st.lists(st.floats(-0.02, 0.02), min_size=3, max_size=3),
       st.integers(0, 2 ** 31 - 1))
def test_trilinear_reproduces_affine_fields(dims, a0, coef, seed):
    """Every channel an affine function of the cell index, inside [0, 1] so the clamp is
    inactive: interpolation at any point of [0, X-1) x [0, Y-1) x [0, Z-1) (where no corner
    wraps) equals the function there."""
    X, Y, Z = dims
    ax = [torch.arange(m, dtype=torch.float64) for m in (X, Y, Z)]
    i, j, k = torch.meshgrid(*ax, indexing="ij")
    a1, a2, a3 = coef
    f = 0.5 + a1 * i + a2 * j + a3 * k + a0
    grid = torch.stack([f, f * 0.5 + 0.1, 0.9 - f * 0.5, f], -1)
    g = torch.Generator().manual_seed(seed)
    n = torch.rand(32, 3, generator=g, dtype=torch.float64) * torch.tensor([X - 1, Y - 1, Z - 1]).double()
    vals, keep = plain.trilinear_values(n, grid)
    assert bool(keep.all())
    fn = 0.5 + a1 * n[:, 0] + a2 * n[:, 1] + a3 * n[:, 2] + a0
    want = torch.stack([fn, fn * 0.5 + 0.1, 0.9 - fn * 0.5, fn], -1)
    assert torch.allclose(vals, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_render_chain_gradcheck_fp64(interp):
    """d RGBA / d grid of the whole chain (placement, lookup, clamp, compositing) in FP64,
    against central differences (torch.autograd.gradcheck), on a 4^3 grid whose values stay
    away from the clamp kinks at 0 and 1."""
    g = torch.Generator().manual_seed(7)
    X = 4
    grid = (torch.rand(X, X, X, 4, generator=g, dtype=torch.float64) * 0.8 + 0.1)
    grid[..., 3] *= 0.5
    gi, _, _, _ = plain.generate_grid(X, X, X, points_distance=0.5)
    gmin = plain.grid_min(gi).double()
    R = 6
    o = torch.tensor([[0.1, 0.2, -2.5]], dtype=torch.float64).repeat(R, 1)
    o[:, :2] += torch.rand(R, 2, generator=g, dtype=torch.float64) * 0.4 - 0.2
    d = torch.nn.functional.normalize(torch.tensor([[0.05, -0.03, 1.0]], dtype=torch.float64)
                                      + torch.rand(R, 3, generator=g, dtype=torch.float64) * 0.1, dim=1)
    kw = dict(num_samples=20, delta_step=0.22, gmin=gmin, points_distance=0.5, interp=interp, div_mode="divide")
    out = plain.render(grid, o, d, **kw)
    assert float(out[:, 3].min()) > 0.05  # the rays cross the grid
    assert torch.autograd.gradcheck(lambda t: plain.render(t, o, d, **kw),
                                    (grid.clone().requires_grad_(True),), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_composite_of_constant_alpha_matches_geometric_series():
    """S samples of constant alpha a and colour c: A = 1 - (1-a)^S, RGB = c A, for a sweep of
    a including 0 and 1."""
    for a in (0.0, 1e-6, 0.01, 0.3, 0.999, 1.0):
        for S in (1, 7, 64, 600):
            x = torch.empty(1, S, 4, dtype=torch.float64)
            x[..., :3] = torch.tensor([0.2, 0.5, 0.9], dtype=torch.float64)
            x[..., 3] = a
            out = plain.compute_alpha_weighted_pixels(x)[0]
            A = 1.0 if a == 1.0 else -math.expm1(S * math.log1p(-a))  # 1 - (1-a)^S without cancellation
            assert math.isclose(float(out[3]), A, rel_tol=1e-12, abs_tol=1e-15)
            assert torch.allclose(out[:3], torch.tensor([0.2, 0.5, 0.9], dtype=torch.float64) * A,
                                  rtol=1e-12, atol=1e-15)
