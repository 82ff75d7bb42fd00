"""pxl_count_cells_host (roofline accounting, host code) vs a brute-force count over the
oracle's float32 geometry."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import plain
from plenoxels_amd import MarchConfig, _lib
from util import small_scene


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_count_cells_host_brute_force(interp):
    grid, o, d, gmin, pd = small_scene(90, X=12, R=64, pd=0.2)
    S, delta = 40, 0.09
    occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(91)) < 0.6
    occ_used = plain.dilate_cells(occ) if interp == "trilinear" else occ
    bits = plain.occupancy_to_bits(occ_used).numpy().view(np.uint32).copy()
    nproc = np.random.default_rng(0).integers(0, S + 1, size=64).astype(np.int32)
    cfg = MarchConfig(S, delta, tuple(gmin.tolist()), pd, interp)
    m = cfg.to_struct()
    g = _lib.Grid()
    for a in range(3):
        g.dims[a] = grid.shape[a]
    g.channels = 4
    g.occupancy = bits.ctypes.data
    oh, dh = o.numpy().copy(), d.numpy().copy()
    counts = np.zeros(64, np.int64)
    _lib.check(_lib.lib().pxl_count_cells_host(ctypes.byref(g), ctypes.byref(m), oh.ctypes.data,
                                               dh.ctypes.data, 64, nproc.ctypes.data, counts.ctypes.data, 3))
    n = plain.normalize_samples(plain.sample_positions(o, d, S, delta), gmin, pd, "reciprocal")
    n = n.reshape(64, S, 3)
    X = grid.shape[0]
    for r in range(64):
        cells = set()
        for k in range(int(nproc[r])):
            p = n[r, k]
            if interp == "nearest":
                idx = torch.round(p).long()
                if bool(((idx >= 0) & (idx < X)).all()) and bool(occ_used[idx[0], idx[1], idx[2]]):
                    cells.add(int((idx[0] * X + idx[1]) * X + idx[2]))
            else:
                if not bool(((p >= 0) & (p < X)).all()):
                    continue
                c = plain.wrap_indices(plain.get_grid_points_indices(p[None])[0], (X, X, X))
                if not bool(occ_used[c[7, 0], c[7, 1], c[7, 2]]):
                    continue
                for j in range(8):
                    cells.add(int((c[j, 0] * X + c[j, 1]) * X + c[j, 2]))
        assert counts[r] == len(cells), r
    assert counts.sum() > 0
