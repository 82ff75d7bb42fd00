"""The grid_cells_trained.pth format (scripts/train.py:197-210) written by plenoxels_amd.checkpoint:
key set, tensor dtypes/shapes/devices, the "param" value types, the float32 upcast of
reduced-precision grids, and loading with torch.load(weights_only=True) -- both our files and a
dict built the way the reference builds it, read the way its consumers read it
(src/visualization.py:114-122, scripts/visulize_grid.py:23-30,
scripts/compare_inference_to_image.py:41-50)."""
import torch

from plenoxels_amd.checkpoint import PARAM_KEYS, load_grid, save_grid

# the reference's dict, key by key, with the Python types its fit() passes
# (scripts/train.py:197-210 with the CLI defaults of :215-233 and gridsize a list, :24-27)
REF_PARAM_TYPES = {"device": str, "number_of_rays": int, "num_samples": int, "delta_step": float,
                   "even_spread": bool, "camera_ray": bool, "points_distance": float, "gridsize": list}


def _reference_style_save(path, grid, grad):
    """A .pth exactly as scripts/train.py:194-210 writes it (a leaf that required grad, detached
    and moved to the CPU)."""
    cells = grid.clone().requires_grad_(True)
    torch.save({
        "grid": cells.detach().cpu(),
        "grid_grad": grad.detach().cpu(),
        "param": {"device": "cuda", "number_of_rays": 128, "num_samples": 600, "delta_step": 0.0125,
                  "even_spread": False, "camera_ray": False, "points_distance": 0.0125,
                  "gridsize": [256, 256, 256]},
    }, path)


def test_save_grid_round_trip_keys_and_types(tmp_path):
    g = torch.Generator().manual_seed(0)
    grid = torch.rand(6, 5, 7, 4, generator=g) * 1.4 - 0.2
    grad = torch.rand(6, 5, 7, 4, generator=g)
    p = tmp_path / "grid_cells_trained.pth"
    save_grid(p, grid.requires_grad_(True), grad, device="cuda:0", number_of_rays=75, num_samples=600,
              delta_step=0.0125, even_spread=False, camera_ray=False, points_distance=0.0125,
              gridsize=(6, 5, 7))
    d = torch.load(p, weights_only=True)
    assert set(d) == {"grid", "grid_grad", "param"}
    assert tuple(d["param"]) == PARAM_KEYS and set(PARAM_KEYS) == set(REF_PARAM_TYPES)
    for k, t in REF_PARAM_TYPES.items():
        assert type(d["param"][k]) is t, (k, type(d["param"][k]))
    assert d["param"]["gridsize"] == [6, 5, 7] and all(type(v) is int for v in d["param"]["gridsize"])
    assert d["param"]["number_of_rays"] == 75 and d["param"]["num_samples"] == 600
    for k, want in (("grid", grid), ("grid_grad", grad)):
        t = d[k]
        assert t.dtype == torch.float32 and t.device.type == "cpu" and not t.requires_grad
        assert tuple(t.shape) == (6, 5, 7, 4) and t.is_contiguous()
        assert torch.equal(t, want.detach())
    back = load_grid(p)
    assert torch.equal(back["grid"], grid.detach()) and back["param"] == d["param"]


def test_half_grid_is_saved_as_exact_float32(tmp_path):
    g = torch.Generator().manual_seed(1)
    for dt in (torch.bfloat16, torch.float16):
        grid = (torch.rand(4, 4, 4, 4, generator=g) * 2 - 0.5).to(dt)
        p = tmp_path / f"g_{dt}.pth"
        save_grid(p, grid, torch.zeros(4, 4, 4, 4, dtype=dt), device="cuda", number_of_rays=16, num_samples=8,
                  delta_step=0.1, even_spread=True, camera_ray=False, points_distance=0.1, gridsize=[4, 4, 4])
        d = load_grid(p)
        assert d["grid"].dtype == torch.float32 and d["grid_grad"].dtype == torch.float32
        assert torch.equal(d["grid"], grid.float())  # the upcast is exact


def test_reads_a_reference_written_file(tmp_path):
    """A dict written as the reference writes it loads with weights_only=True and answers the
    reference consumers' accesses."""
    g = torch.Generator().manual_seed(2)
    grid = torch.rand(8, 8, 8, 4, generator=g)
    grad = torch.rand(8, 8, 8, 4, generator=g)
    p = tmp_path / "ref.pth"
    _reference_style_save(p, grid, grad)
    d = load_grid(p)
    assert torch.equal(d["grid"], grid) and torch.equal(d["grid_grad"], grad)
    # src/visualization.py:114-122 and scripts/compare_inference_to_image.py:41-50
    assert d["grid"].clip(0, 1).shape == (8, 8, 8, 4)
    assert d["param"]["points_distance"] == 0.0125 and d["param"]["delta_step"] == 0.0125
    for k, t in REF_PARAM_TYPES.items():
        assert type(d["param"][k]) is t


def test_load_grid_rejects_a_file_without_the_grid(tmp_path):
    p = tmp_path / "bad.pth"
    torch.save({"param": {}}, p)
    try:
        load_grid(p)
    except KeyError as e:
        assert "grid" in str(e)
    else:
        raise AssertionError("load_grid accepted a dict without 'grid'")
