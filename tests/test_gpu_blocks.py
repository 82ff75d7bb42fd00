"""GPU parity of the corner-block gradient path (pxl_train_step_blocks + pxl_blocks_reduce).

The blocks path must give the gradient of the direct path (pxl_train_step's atomics into the
grid) up to FP32 summation order, and the plain chain's autograd gradient within the usual
1e-4 relative bar: with and without occupancy, for every storage dtype, on rays that cross the
grid's top faces (ceil index wrapping to 0) and on samples at exact integer coordinates (ceil ==
floor, the zero-weight corner), over several accumulating calls, and the blocks must be zero again
after each reduce.
"""
import pytest
import torch

from oracle import plain
from plenoxels_amd import (CameraRays, CornerBlockPair, MarchConfig, blocks_reduce, corner_blocks,
                           occupancy_dilate, pack_occupancy, train_step)
from util import grad_close, say, small_scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _direct_and_blocks(grid, o, d, tgt, cfg, occupancy=None, storage=torch.float32):
    gd = grid.to(storage).to(DEV)
    g1 = torch.zeros(grid.shape, device=DEV)
    out1 = torch.empty(o.shape[0], 4, device=DEV)
    train_step(gd, g1, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out1, occupancy=occupancy)
    blocks = corner_blocks(gd)
    g2 = torch.full(grid.shape, 7.0, device=DEV)  # overwritten, not added to
    out2 = torch.empty(o.shape[0], 4, device=DEV)
    train_step(gd, g2, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), out=out2, occupancy=occupancy,
               blocks=blocks, overwrite=True)
    assert torch.equal(out1, out2)
    assert int(torch.count_nonzero(blocks)) == 0, "blocks not re-zeroed by the reduce"
    return g1.cpu(), g2.cpu(), out2.cpu()


@pytest.mark.parametrize("storage", [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("occupied", [False, True])
def test_blocks_equal_direct_and_plain(storage, occupied):
    say("blocks vs direct vs plain", storage, "occupancy" if occupied else "dense")
    grid, o, d, gmin, pd = small_scene(60, X=20, R=768)
    X = grid.shape[0]
    tgt = torch.rand(o.shape[0], 4, generator=torch.Generator().manual_seed(61))
    occ_cells = None
    bits = None
    if occupied:
        occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(62)) < 0.35
        occ_cells = plain.dilate_cells(occ)
        bits = occupancy_dilate(pack_occupancy(occ.to(DEV)), grid.shape[:3])
    cfg = MarchConfig(120, 0.035, tuple(gmin.tolist()), pd, "trilinear")
    g1, g2, out = _direct_and_blocks(grid, o, d, tgt, cfg, bits, storage)
    ok, info = grad_close(g2, g1, rel=1e-5)
    say("  blocks vs direct", info)
    assert ok, info
    # the rays reach the top faces: some samples interpolate across the wrap to index 0
    n = plain.normalize_samples(plain.sample_positions(o, d, 120, 0.035), gmin, pd, "reciprocal")
    assert bool(((n > X - 1) & (n < X)).any()), "no sample in the wrapping top slab"
    rounded = grid.to(storage).float()
    g = rounded.clone().requires_grad_(True)
    want = plain.render(g, o, d, num_samples=120, delta_step=0.035, gmin=gmin, points_distance=pd,
                        interp="trilinear", div_mode="reciprocal", occupancy=occ_cells)
    ((want - tgt) ** 2).mean().backward()
    assert (out - want.detach()).abs().max().item() <= 1e-5
    ok, info = grad_close(g2, g.grad)
    say("  blocks vs plain autograd", info)
    assert ok, info


def test_blocks_exact_integer_coordinates():
    """Axis-aligned rays through cell centres with delta = pd: every coordinate is an exact
    integer, so ceil == floor and the ceil corners carry weight 0 (the reference adds them to
    the floor cell; the blocks attribute the x and y ceil corners to the next cell, and the z
    ones, resolved at the scatter, to the floor cell -- the same finite sum)."""
    say("blocks at integer coordinates")
    X, pd = 16, 0.25
    g = torch.Generator().manual_seed(70)
    grid = torch.rand(X, X, X, 4, generator=g) * 0.3
    gi, _, _, _ = plain.generate_grid(X, X, X, points_distance=pd)
    gmin = plain.grid_min(gi)
    R = 64
    o = torch.zeros(R, 3)
    ij = torch.randint(0, X, (R, 2), generator=g).float()
    o[:, 0] = gmin[0] - 3 * pd  # enters along +x at exact multiples of pd
    o[:, 1] = gmin[1] + ij[:, 0] * pd
    o[:, 2] = gmin[2] + ij[:, 1] * pd
    d = torch.zeros(R, 3)
    d[:, 0] = 1.0
    tgt = torch.rand(R, 4, generator=g)
    n = plain.normalize_samples(plain.sample_positions(o, d, 24, pd), gmin, pd, "reciprocal")
    inb = ((n >= 0) & (n < X)).all(1)
    assert bool(inb.any()) and torch.equal(n[inb], torch.floor(n[inb])), "coordinates are not exact integers"
    cfg = MarchConfig(24, pd, tuple(gmin.tolist()), pd, "trilinear")
    g1, g2, _ = _direct_and_blocks(grid, o, d, tgt, cfg)
    ok, info = grad_close(g2, g1, rel=1e-6)
    say("  blocks vs direct", info)
    assert ok, info


def test_blocks_accumulate_over_calls_then_reduce():
    """Three ray chunks into the same blocks, one reduce adding into a non-zero gradient; then a
    reduce of the (re-zeroed) blocks changes nothing."""
    say("blocks: chunks + one reduce")
    from plenoxels_amd import scenes
    w = scenes.make_workload("sparse256", n_rays=3 * 2048)
    grid = w.grid.to(DEV)
    bits = occupancy_dilate(pack_occupancy(w.occupancy.to(DEV)), w.dims)
    imgs = torch.rand(w.c2w.shape[0], 800, 800, 4, generator=torch.Generator().manual_seed(80)).to(DEV)
    cfg = MarchConfig(w.num_samples, w.delta_step, w.grid_min, w.points_distance, "trilinear", w.near)
    base = torch.randn(grid.shape, generator=torch.Generator().manual_seed(81)).to(DEV) * 1e-3
    direct = base.clone()
    blocks = corner_blocks(grid)
    R = w.cam_ids.shape[0]
    for s in range(0, R, 2048):
        cr = CameraRays(w.cam_ids[s:s + 2048].to(DEV), w.uv[s:s + 2048].to(DEV), w.c2w.to(DEV), w.camera_angle_x, imgs)
        train_step(grid, direct, cfg=cfg, camera=cr, occupancy=bits, n_total_rays=R)
        train_step(grid, None, cfg=cfg, camera=cr, occupancy=bits, n_total_rays=R, blocks=blocks)
    assert int(torch.count_nonzero(blocks)) > 0
    got = base.clone()
    blocks_reduce(grid, blocks, got, occupancy=bits, accumulate=True)
    assert int(torch.count_nonzero(blocks)) == 0
    # the usual 1e-4 bar: subtracting the 1e-3-sized base rounds at ~6e-11, which is 2e-5 of an
    # element at 1% of this gradient's largest (2.5e-4)
    ok, info = grad_close((got - base).cpu(), (direct - base).cpu(), rel=1e-4)
    say("  chunks: blocks vs direct", info)
    assert ok, info
    again = got.clone()
    blocks_reduce(grid, blocks, again, occupancy=bits, accumulate=True)
    assert torch.equal(again, got)


@pytest.mark.parametrize("start,dims,progressive", [(0, (24, 24, 24), False), (205, (48, 48, 48), True)])
def test_trainer_corner_blocks_match_direct(start, dims, progressive):
    """Trainer (trilinear, TV and beta on) with corner blocks vs with direct atomics, same
    uniforms: per-step losses, |grad| sums and grids agree (pooled and full-grid steps)."""
    from plenoxels_amd.train import Trainer
    from util import random_cameras
    say("trainer corner blocks vs direct", start, dims)
    B, H, R, steps = 3, 24, 64, 4
    c2w = random_cameras(B, 20, radius=2.0)
    imgs = torch.rand(B, H, H, 4, generator=torch.Generator().manual_seed(21))
    g = torch.Generator().manual_seed(22)
    uniforms = [torch.rand(B, R, 2, generator=g) for _ in range(steps)]
    init = (torch.rand(*dims, 4, generator=torch.Generator().manual_seed(30)) * 0.6 - 0.1).to(DEV)
    runs = []
    for blocks in (False, True):
        tr = Trainer(imgs, c2w, 0.69, gridsize=dims, points_distance=0.06, number_of_rays=R, num_samples=48,
                     delta_step=0.05, lr=0.01, tv=1e-3, beta=5e-3, device="cuda", progressive_growing=progressive,
                     interp="trilinear", uniforms_fn=lambda i: uniforms[i - start], init_grid=init,
                     corner_blocks=blocks, skip_empty=False)
        hist = [{k: float(v) for k, v in tr.step(i).items()} for i in range(start, start + steps)]
        runs.append((tr, hist))
    (ta, ha), (tb, hb) = runs
    for a, b in zip(ha, hb):
        assert set(a) == set(b)
        for key in a:
            assert abs(a[key] - b[key]) <= 1e-5 * abs(a[key]) + 1e-9, (key, a[key], b[key])
    ok, info = grad_close(tb.grad_abs.cpu(), ta.grad_abs.cpu(), rel=1e-3)
    say("  |grad| sum", info)
    assert ok, info
    diff = (tb.grid - ta.grid).abs()
    say("  grid: max diff", diff.max().item(), "frac > 1e-5:", (diff > 1e-5).float().mean().item())
    # Adam's first steps move an element by ~lr * sign(g): where a cell's gradient sums to ~0,
    # the summation order of the two paths may flip that sign, as between any two runs
    assert (diff > 1e-5).float().mean().item() <= 1e-3


@pytest.mark.parametrize("occupied", [False, True])
def test_reduce_without_clear_then_blocks_clear(occupied):
    """reduce(clear=False) leaves the blocks as they were; blocks_clear then zeroes exactly the
    entries a step with that occupancy can write (entry (x, y, z): bit of (x, y, z) or of
    (x, y, z - 1 wrapped)), every entry without occupancy."""
    from plenoxels_amd import blocks_clear
    say("reduce without clear + blocks_clear", "occupancy" if occupied else "dense")
    grid, o, d, gmin, pd = small_scene(95, X=20, R=256)
    tgt = torch.rand(o.shape[0], 4, generator=torch.Generator().manual_seed(96))
    occ_dev, written = None, None
    if occupied:
        occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(97)) < 0.35
        cells = plain.dilate_cells(occ)
        written = (cells | torch.roll(cells, shifts=1, dims=2)).reshape(-1)
        occ_dev = occupancy_dilate(pack_occupancy(occ.to(DEV)), grid.shape[:3])
    cfg = MarchConfig(120, 0.035, tuple(gmin.tolist()), pd, "trilinear")
    gd = grid.to(DEV)
    blocks = corner_blocks(gd)
    train_step(gd, None, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=occ_dev,
               blocks=blocks)
    before = blocks.clone()
    g = torch.empty(grid.shape, device=DEV)
    blocks_reduce(gd, blocks, g, occupancy=occ_dev, accumulate=False, clear=False)
    assert torch.equal(blocks, before)
    g_direct = torch.zeros(grid.shape, device=DEV)
    train_step(gd, g_direct, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=occ_dev)
    ok, info = grad_close(g.cpu(), g_direct.cpu(), rel=1e-5)
    assert ok, info
    blocks.fill_(1.0)
    blocks_clear(gd, blocks, occupancy=occ_dev)
    b = blocks.cpu()
    if occupied:
        assert int(torch.count_nonzero(b[written])) == 0
        assert bool((b[~written] == 1.0).all())
    else:
        assert int(torch.count_nonzero(b)) == 0


@pytest.mark.parametrize("occupied", [False, True])
def test_double_buffer_reduce_clears_the_other_buffer(occupied):
    """reduce(blocks=A, clear=B): the gradient equals the single-buffer reduce; B's entries that
    a step can write (all entries without occupancy) are zeroed in the same pass and A is left
    for its own clearing; two steps alternating the buffers give the direct path's gradients."""
    say("double-buffered reduce", "occupancy" if occupied else "dense")
    grid, o, d, gmin, pd = small_scene(90, X=20, R=512)
    tgt = torch.rand(o.shape[0], 4, generator=torch.Generator().manual_seed(91))
    occ_dev = None
    occ_cells = None
    if occupied:
        occ = torch.rand(grid.shape[:3], generator=torch.Generator().manual_seed(92)) < 0.35
        occ_cells = plain.dilate_cells(occ).reshape(-1)
        occ_dev = occupancy_dilate(pack_occupancy(occ.to(DEV)), grid.shape[:3])
    cfg = MarchConfig(120, 0.035, tuple(gmin.tolist()), pd, "trilinear")
    gd = grid.to(DEV)
    pair = CornerBlockPair(gd)
    pair.other.fill_(1.0)  # stale contents the pass must clear
    rays = dict(cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=occ_dev)
    g_direct = torch.zeros(grid.shape, device=DEV)
    train_step(gd, g_direct, **rays)
    g = torch.empty(grid.shape, device=DEV)
    train_step(gd, g, blocks=pair.blocks, overwrite=True, clear=pair.other, **rays)
    ok, info = grad_close(g.cpu(), g_direct.cpu(), rel=1e-5)
    say("  step 1 vs direct", info)
    assert ok, info
    other = pair.other.cpu().reshape(-1, 16)
    if occupied:
        # entry (x, y, z) holds corners of floor cells (x, y, z) and (x, y, z - 1 wrapped)
        cells = occ_cells.reshape(grid.shape[:3])
        written = (cells | torch.roll(cells, shifts=1, dims=2)).reshape(-1)
        assert int(torch.count_nonzero(other[written])) == 0
        assert bool((other[~written] == 1.0).all())
    else:
        assert int(torch.count_nonzero(other)) == 0
    assert int(torch.count_nonzero(pair.blocks)) > 0  # not cleared by its own reduce
    pair.flip()
    pair.other.zero_()  # (a third step would find it zeroed by the second reduce)
    g2 = torch.empty(grid.shape, device=DEV)
    train_step(gd, g2, blocks=pair.blocks, overwrite=True, clear=pair.other, **rays)
    ok, info = grad_close(g2.cpu(), g_direct.cpu(), rel=1e-5)
    say("  step 2 vs direct", info)
    assert ok, info


@pytest.mark.parametrize("occupied", [False, True])
def test_blocks_odd_shape_equal_direct(occupied):
    """A 6 x 5 x 300 grid (the reduce splits z into chunks of 256; x and y wrap within a few
    rows) with rays crossing it along z: the blocks path equals the direct path."""
    say("blocks odd shape", "occupancy" if occupied else "dense")
    X, Y, Z, pd = 6, 5, 300, 0.02
    g = torch.Generator().manual_seed(100)
    grid = torch.rand(X, Y, Z, 4, generator=g) * 0.2
    gi, _, _, _ = plain.generate_grid(X, Y, Z, points_distance=pd)
    gmin = plain.grid_min(gi)
    R = 256
    o = torch.empty(R, 3)
    o[:, 0] = gmin[0] + torch.rand(R, generator=g) * (X - 1) * pd
    o[:, 1] = gmin[1] + torch.rand(R, generator=g) * (Y - 1) * pd
    o[:, 2] = gmin[2] - 0.3
    d = torch.nn.functional.normalize(torch.stack([(torch.rand(R, generator=g) - 0.5) * 0.04,
                                                   (torch.rand(R, generator=g) - 0.5) * 0.04,
                                                   torch.ones(R)], 1), dim=1)
    tgt = torch.rand(R, 4, generator=g)
    bits = None
    if occupied:
        occ = torch.rand(X, Y, Z, generator=torch.Generator().manual_seed(101)) < 0.3
        bits = occupancy_dilate(pack_occupancy(occ.to(DEV)), (X, Y, Z))
    cfg = MarchConfig(400, 0.017, tuple(gmin.tolist()), pd, "trilinear")
    g1, g2, _ = _direct_and_blocks(grid, o, d, tgt, cfg, bits)
    assert float(g1.abs().max()) > 0
    ok, info = grad_close(g2, g1, rel=1e-5)
    say("  blocks vs direct", info)
    assert ok, info


def _tiles_of(cells):
    """Per 8x8x8 tile of a (X,Y,Z) bool grid: any cell set (padded to whole tiles)."""
    X, Y, Z = cells.shape
    TX, TY, TZ = (X + 7) // 8, (Y + 7) // 8, (Z + 7) // 8
    pad = torch.zeros(TX * 8, TY * 8, TZ * 8, dtype=torch.bool)
    pad[:X, :Y, :Z] = cells
    return pad.reshape(TX, 8, TY, 8, TZ, 8).any(5).any(3).any(1)


def _cells_of_tiles(tiles, dims):
    X, Y, Z = dims
    full = tiles.repeat_interleave(8, 0).repeat_interleave(8, 1).repeat_interleave(8, 2)
    return full[:X, :Y, :Z]


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_fill_reachable_covers_every_cell_a_step_writes(interp):
    """pxl_fill_reachable_f32 fills the occupied cells themselves (nearest), or whole 8^3 tiles
    with an occupied cell in the tile or a lower neighbour tile, wrapped (trilinear: corners reach
    one cell up); that covers every cell the step's scatter can write, and nothing else is
    touched."""
    from plenoxels_amd import fill_reachable_
    say("fill reachable", interp)
    dims = (37, 34, 45)  # partial tiles on every axis
    occ = torch.zeros(dims, dtype=torch.bool)
    occ[3:14, 9:20, 5:17] = torch.rand(11, 11, 12, generator=torch.Generator().manual_seed(70)) < 0.02
    occ[36, 33, 44] = True  # a floor cell at the top corner: its ceil corners wrap to index 0
    pts = pack_occupancy(occ.to(DEV))
    bits = occupancy_dilate(pts, dims) if interp == "trilinear" else pts
    g = torch.full(dims + (4,), 7.0, device=DEV)
    fill_reachable_(g, bits, interp)
    filled = (g.cpu() == 0.0).all(-1)
    assert bool(((g.cpu() == 0.0) | (g.cpu() == 7.0)).all())
    cells = plain.dilate_cells(occ) if interp == "trilinear" else occ  # floor cells a step keeps
    if interp == "trilinear":
        tiles = _tiles_of(cells)
        t = tiles.clone()
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    t |= torch.roll(tiles, shifts=(a, b, c), dims=(0, 1, 2))
        assert torch.equal(filled, _cells_of_tiles(t, dims))
    else:
        assert torch.equal(filled, occ)
    # every cell the scatter can write: the kept cell itself, or the 8 corners of a kept floor cell
    writes = cells.clone()
    if interp == "trilinear":
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    writes |= torch.roll(cells, shifts=(a, b, c), dims=(0, 1, 2))
    assert bool(filled[writes].all())
    assert int(filled.sum()) < filled.numel()  # something was left alone


def test_reduce_reachable_only_equals_the_dense_reduce_where_it_writes():
    """PXL_REDUCE_OVERWRITE_REACHABLE writes the dense reduce's values at every cell a source entry
    of a set occupancy bit reaches, leaves every other cell untouched, and the dense result is zero
    at those other cells -- so on a buffer zeroed once the two modes give the same gradient."""
    say("reduce reachable only")
    grid, o, d, gmin, pd = small_scene(72, X=22, R=900)
    dims = tuple(grid.shape[:3])
    occ = torch.rand(dims, generator=torch.Generator().manual_seed(73)) < 0.05
    bits = occupancy_dilate(pack_occupancy(occ.to(DEV)), dims)
    cells = plain.dilate_cells(occ)
    tgt = torch.rand(o.shape[0], 4, generator=torch.Generator().manual_seed(74))
    cfg = MarchConfig(140, 0.03, tuple(gmin.tolist()), pd, "trilinear")
    gd = grid.to(DEV)
    blocks = corner_blocks(gd)
    train_step(gd, None, cfg=cfg, origins=o.to(DEV), dirs=d.to(DEV), targets=tgt.to(DEV), occupancy=bits,
               blocks=blocks)
    dense = torch.full(grid.shape, 5.0, device=DEV)
    blocks_reduce(gd, blocks, dense, occupancy=bits, accumulate=False, clear=False)
    sparse = torch.full(grid.shape, 7.0, device=DEV)
    blocks_reduce(gd, blocks, sparse, occupancy=bits, accumulate=False, clear=blocks, reachable_only=True)
    assert int(torch.count_nonzero(blocks)) == 0
    # entry (x, y, z) may hold sums when the floor cell (x, y, z) or (x, y, z - 1) is kept; cell
    # (x, y, z) reads entries (x - 1 + bx, y - 1 + by, z)
    written = cells | torch.roll(cells, 1, 2)
    reach = written | torch.roll(written, 1, 0) | torch.roll(written, 1, 1) | torch.roll(written, (1, 1), (0, 1))
    dense, sparse = dense.cpu(), sparse.cpu()
    assert torch.equal(sparse[reach], dense[reach])
    assert bool((sparse[~reach] == 7.0).all()) and bool((dense[~reach] == 0.0).all())
    assert bool((dense[reach] != 0).any())
