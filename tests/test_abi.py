"""C-ABI checks that need no GPU: libpxl loads, exports every symbol include/pxl.h declares,
its struct layouts match the ctypes mirrors, and argument validation reports errors before
any HIP call."""
import ctypes
import os
import re
import subprocess

import pytest

from plenoxels_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pxl.h")


def header_functions(header=HEADER):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pxl_[a-z0-9_]+)\s*\(", src)))


def test_probe_library_exports_its_header_and_nothing_of_the_render_path():
    """libpxl_probe (measurement and test kernels) exports what include/pxl_probe.h declares; the
    product library exports no probe or test hook."""
    P = _lib.probe_lib()
    names = header_functions(os.path.join(ROOT, "include", "pxl_probe.h"))
    assert names == sorted(_lib.PROBE_SIGNATURES)
    for n in names:
        assert hasattr(P, n), n
    assert not any("probe" in n or "debug" in n for n in header_functions())
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("pxl_")}
    assert exported == set(header_functions()), exported ^ set(header_functions())
    assert P.pxl_probe_memory_rate(None, 16, 0, 4, 1, 1, 0, None, None) == _lib.PXL_E_INVALID
    assert P.pxl_probe_memory_rate(ctypes.c_void_p(0x1000), 16, 0, 3, 1, 1, 0, None, None) == _lib.PXL_E_INVALID
    assert P.pxl_probe_memory_rate(ctypes.c_void_p(0x1000), 4, 0, 4, 1, 1, 1, None, None) == _lib.PXL_E_INVALID
    assert P.pxl_probe_memory_rate(ctypes.c_void_p(0x1000), 16, 2, 4, 1, 1, 0, None, None) == _lib.PXL_E_INVALID
    assert P.pxl_probe_memory_rate(ctypes.c_void_p(0x1000), 16, 0, 4, 0, 1, 0, None, None) == _lib.PXL_OK


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert L.pxl_abi_version() == _lib.ABI_VERSION == 3


def test_struct_layout_matches_header(tmp_path):
    probe = tmp_path / "probe.c"
    fields = {
        "pxl_grid": ["data", "dims", "channels", "dtype", "occupancy"],
        "pxl_march": ["n_samples", "interp", "flags", "reserved", "delta", "near", "t_stop", "pd",
                      "inv_pd", "grid_min"],
        "pxl_rays": ["n_rays", "origins", "dirs", "cam_ids", "uv", "c2w", "n_cams",
                     "camera_angle_x", "flags", "imgs", "img_h", "img_w"],
        "pxl_loss": ["targets", "grad_scale", "beta_scale", "ray_sq_err", "ray_beta"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pxl.h"', "int main(void){"]
    for s, fs in fields.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(probe), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    mirrors = {"pxl_grid": _lib.Grid, "pxl_march": _lib.March, "pxl_rays": _lib.Rays, "pxl_loss": _lib.Loss}
    for s, cls in mirrors.items():
        assert int(got[s]) == ctypes.sizeof(cls), s
        for f in fields[s]:
            assert int(got[f"{s}.{f}"]) == getattr(cls, f).offset, f"{s}.{f}"


def _valid():
    g = _lib.Grid()
    g.data = 0x1000
    g.dims[0] = g.dims[1] = g.dims[2] = 8
    g.channels = 4
    g.dtype = _lib.DTYPE_F32
    m = _lib.March()
    m.n_samples = 16
    m.interp = _lib.INTERP_NEAREST
    m.delta = 0.1
    m.pd = 0.2
    m.inv_pd = 5.0
    r = _lib.Rays()
    r.n_rays = 0
    r.origins = 0x2000
    r.dirs = 0x3000
    return g, m, r


def _fwd(g, m, r, out=0x4000):
    return _lib.lib().pxl_render_fwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r),
                                     ctypes.c_void_p(out), None, None)


def test_validation_errors_before_any_hip_call():
    L = _lib.lib()
    g, m, r = _valid()
    assert _fwd(g, m, r) == _lib.PXL_OK  # zero rays: validated, nothing launched
    assert L.pxl_render_fwd(None, ctypes.byref(m), ctypes.byref(r), None, None, None) == _lib.PXL_E_INVALID
    assert "grid" in _lib.last_error()
    g2, m2, r2 = _valid()
    g2.channels = 3
    assert _fwd(g2, m2, r2) == _lib.PXL_E_DTYPE
    g2, m2, r2 = _valid()
    g2.dtype = 7
    assert _fwd(g2, m2, r2) == _lib.PXL_E_DTYPE
    g2, m2, r2 = _valid()
    g2.dims[0] = g2.dims[1] = g2.dims[2] = 2048  # 2^33 cells
    assert _fwd(g2, m2, r2) == _lib.PXL_E_SIZE
    g2, m2, r2 = _valid()
    g2.dims[2] = 0
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    m2.n_samples = -1
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    m2.interp = 5
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    m2.pd = 0.0
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    m2.flags = 8
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    r2.n_rays = -3
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    r2.n_rays = 4
    r2.origins = None  # camera form without cameras
    assert _fwd(g2, m2, r2) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    r2.n_rays = 4
    assert _fwd(g2, m2, r2, out=0) == _lib.PXL_E_INVALID
    g2, m2, r2 = _valid()
    m2.n_samples = _lib.TRAIN_MAX_SAMPLES + 1
    r2.n_rays = 4
    loss = _lib.Loss()
    loss.grad_scale = 1.0
    rc = L.pxl_train_step(ctypes.byref(g2), ctypes.byref(m2), ctypes.byref(r2), ctypes.byref(loss),
                          None, ctypes.c_void_p(0x5000), None)
    assert rc == _lib.PXL_E_INVALID  # targets missing for explicit rays
    loss.targets = 0x6000
    rc = L.pxl_train_step(ctypes.byref(g2), ctypes.byref(m2), ctypes.byref(r2), ctypes.byref(loss),
                          None, ctypes.c_void_p(0x5000), None)
    assert rc == _lib.PXL_E_INVALID and str(_lib.TRAIN_MAX_SAMPLES) in _lib.last_error()
    r3 = _lib.Rays()
    r3.n_rays = 4
    r3.origins, r3.dirs = 0x2000, 0x3000
    r3.flags = 2  # unknown ray flag
    assert _fwd(g2, m2, r3) == _lib.PXL_E_INVALID and "flags" in _lib.last_error()
    assert L.pxl_train_step(ctypes.byref(g2), ctypes.byref(m2), ctypes.byref(r2), None, None,
                            ctypes.c_void_p(0x5000), None) == _lib.PXL_E_INVALID
    d3 = _lib.dims_arg((4, 4, 0))
    assert L.pxl_wrap_indices(ctypes.c_void_p(0x10), 4, d3, None) == _lib.PXL_E_INVALID
    assert L.pxl_avg_pool3d(ctypes.c_void_p(0x10), 4, 4, 4, 5, 1, ctypes.c_void_p(0x20), ctypes.c_void_p(0x30), None) == _lib.PXL_E_INVALID
    assert L.pxl_avg_pool3d_workspace_floats(4, 4, 4, 5, 1) == -1
    assert L.pxl_avg_pool3d_workspace_floats(8, 6, 10, 3, 2) == (8 * 6 * 4 + 8 * 2 * 4) * 4
    assert L.pxl_adam_step(ctypes.c_void_p(0x10), ctypes.c_void_p(0x10), ctypes.c_void_p(0x10),
                           ctypes.c_void_p(0x10), None, 4, 1e-3, 0.9, 0.999, 1e-8, 0, None, 0,
                           None) == _lib.PXL_E_INVALID  # step counts from 1
    g3, _, _ = _valid()
    g3.dtype = _lib.DTYPE_BF16
    assert L.pxl_nearest_gather(ctypes.c_void_p(0x10), 4, ctypes.byref(g3), ctypes.c_void_p(0x20), None,
                                None) == _lib.PXL_E_DTYPE
    assert L.pxl_fill_f32(None, 4, ctypes.c_float(0.0), None) == _lib.PXL_E_INVALID
    assert L.pxl_fill_f32(ctypes.c_void_p(0x1004), 8, ctypes.c_float(0.0), None) == _lib.PXL_E_INVALID
    assert L.pxl_occupancy_dilate(ctypes.c_void_p(0x10), 4, 4, 4, ctypes.c_void_p(0x10), None) == _lib.PXL_E_INVALID
    # occupancy tiles: 16 words per 8x8x8 tile, dims rounded up to 8
    assert L.pxl_occupancy_words(5, 6, 7) == 16
    assert L.pxl_occupancy_words(9, 17, 10) == 2 * 3 * 2 * 16
    assert L.pxl_occupancy_words(256, 256, 256) == 256 ** 3 // 32
    assert L.pxl_occupancy_words(0, 4, 4) == -1
    assert L.pxl_pack_occupancy(None, 4, 4, 4, ctypes.c_void_p(0x10), None) == _lib.PXL_E_INVALID
    assert L.pxl_pack_occupancy(ctypes.c_void_p(0x10), 4, 0, 4, ctypes.c_void_p(0x20), None) == _lib.PXL_E_INVALID
    # 1 x 1 x (2^31 - 1) pads to 2^28 tiles = 2^32 words, past the kernels' 32-bit word index
    assert L.pxl_occupancy_dilate(ctypes.c_void_p(0x10), 1, 1, 2 ** 31 - 1, ctypes.c_void_p(0x20), None) == _lib.PXL_E_SIZE
    g4, m4, r4 = _valid()
    g4.dims[0] = g4.dims[1] = 1
    g4.dims[2] = 2 ** 31 - 1
    g4.occupancy = ctypes.c_void_p(0x40)
    assert L.pxl_render_fwd(ctypes.byref(g4), ctypes.byref(m4), ctypes.byref(r4), ctypes.c_void_p(0x50), None,
                            None) == _lib.PXL_E_SIZE
    assert "occupancy" in _lib.last_error()
    assert L.pxl_adam_tv_step(ctypes.c_void_p(0x100), ctypes.c_void_p(0x100), ctypes.c_void_p(0x200),
                              ctypes.c_void_p(0x300), ctypes.c_void_p(0x400), None, 4, 4, 4,
                              ctypes.c_void_p(0x500), ctypes.c_float(0.1), 1e-2, 0.9, 0.999, 1e-8, 1, None, 0, 0,
                              None) == _lib.PXL_E_INVALID  # param_out == param_in
    assert "param_in" in _lib.last_error()
    assert L.pxl_adam_tv_step(ctypes.c_void_p(0x100), ctypes.c_void_p(0x108), ctypes.c_void_p(0x200),
                              ctypes.c_void_p(0x300), ctypes.c_void_p(0x400), None, 4, 4, 4,
                              ctypes.c_void_p(0x500), ctypes.c_float(0.1), 1e-2, 0.9, 0.999, 1e-8, 1, None, 0, 0,
                              None) == _lib.PXL_E_INVALID  # misaligned param_out
    assert L.pxl_aux_floats(256) == 10
    assert L.pxl_aux_floats(0) == 2
    assert L.pxl_aux_floats(65) == 6


def test_empty_inputs_are_no_ops_whatever_their_pointers():
    """torch hands an empty tensor a NULL data pointer: every zero-size call succeeds without
    reading its buffers (and launches nothing), while negative sizes stay errors."""
    L = _lib.lib()
    g, m, r = _valid()
    r.origins = r.dirs = None
    assert _fwd(g, m, r, out=0) == _lib.PXL_OK
    assert L.pxl_render_bwd(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), None, None, None,
                            None) == _lib.PXL_OK
    loss = _lib.Loss()
    loss.grad_scale = 0.0
    assert L.pxl_train_step(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), ctypes.byref(loss),
                            None, ctypes.c_void_p(0x5000), None) == _lib.PXL_OK
    assert L.pxl_train_step(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), ctypes.byref(loss),
                            None, None, None) == _lib.PXL_E_INVALID  # the gradient grid is never empty
    assert L.pxl_camera_rays(ctypes.byref(r), None, None, None, None) == _lib.PXL_OK
    assert L.pxl_sample_points(None, 1, None, 0, 16, ctypes.c_float(0.1), ctypes.c_float(0.0), None,
                               None) == _lib.PXL_OK
    assert L.pxl_composite_fwd(None, 0, 16, None, None) == _lib.PXL_OK
    assert L.pxl_composite_bwd(None, 0, 16, None, None, None) == _lib.PXL_OK
    assert L.pxl_composite_fwd(None, -1, 16, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_nearest_gather(None, 0, ctypes.byref(g), None, None, None) == _lib.PXL_OK
    assert L.pxl_nearest_gather(None, -1, ctypes.byref(g), None, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_trilinear_bwd(None, None, 0, ctypes.byref(g), None, None, None) == _lib.PXL_OK
    assert L.pxl_corner_indices(None, 0, None, None) == _lib.PXL_OK
    assert L.pxl_adam_step(None, None, None, None, None, 0, 1e-3, 0.9, 0.999, 1e-8, 1, None, 0,
                           None) == _lib.PXL_OK
    assert L.pxl_fill_f32(None, 0, ctypes.c_float(0.0), None) == _lib.PXL_OK


def test_corner_block_validation_before_any_hip_call():
    L = _lib.lib()
    g, m, r = _valid()
    r.n_rays = 4
    loss = _lib.Loss()
    loss.grad_scale = 1.0
    loss.targets = 0x6000
    args = (ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), ctypes.byref(loss), None, ctypes.c_void_p(0x5000), None)
    assert L.pxl_train_step_blocks(*args) == _lib.PXL_E_INVALID  # nearest scatters one cell per sample
    assert "trilinear" in _lib.last_error()
    m.interp = _lib.INTERP_TRILINEAR
    g.dims[0] = g.dims[1] = g.dims[2] = 1024  # 2^30 cells > 2^30 - 1
    assert L.pxl_train_step_blocks(*args) == _lib.PXL_E_SIZE
    g.dims[0] = g.dims[1] = g.dims[2] = 8
    assert L.pxl_train_step_blocks(ctypes.byref(g), ctypes.byref(m), ctypes.byref(r), ctypes.byref(loss), None,
                                   None, None) == _lib.PXL_E_INVALID
    assert L.pxl_blocks_floats(ctypes.byref(g)) == 8 * 8 * 8 * 16
    assert L.pxl_blocks_clear(ctypes.byref(g), None, None) == _lib.PXL_E_INVALID
    assert L.pxl_blocks_clear(None, ctypes.c_void_p(0x10), None) == _lib.PXL_E_INVALID
    g.dims[1] = 0
    assert L.pxl_blocks_floats(ctypes.byref(g)) == -1
    assert L.pxl_blocks_reduce(ctypes.byref(g), ctypes.c_void_p(0x10), ctypes.c_void_p(0x20), 1, None, None) == _lib.PXL_E_INVALID
    g.dims[1] = 8
    assert L.pxl_blocks_reduce(ctypes.byref(g), None, ctypes.c_void_p(0x20), 1, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_blocks_reduce(None, ctypes.c_void_p(0x10), ctypes.c_void_p(0x20), 1, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_alpha_occupancy(ctypes.byref(g), 1, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_alpha_occupancy(None, 1, ctypes.c_void_p(0x10), None) == _lib.PXL_E_INVALID


def test_scaled_sum_validation():
    L = _lib.lib()
    assert L.pxl_scaled_sum_work_bytes() == 1024
    work, out = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    assert L.pxl_scaled_sum(ctypes.c_void_p(0x3000), -1, 1.0, out, work, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum(None, 5, 1.0, out, work, None) == _lib.PXL_E_INVALID  # n > 0 needs v
    assert L.pxl_scaled_sum(ctypes.c_void_p(0x3000), 5, 1.0, None, work, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum(ctypes.c_void_p(0x3000), 5, 1.0, out, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum(ctypes.c_void_p(0x3004), 5, 1.0, out, work, None) == _lib.PXL_E_INVALID
    assert "aligned" in _lib.last_error()
    assert L.pxl_scaled_sum(ctypes.c_void_p(0x3000), 5, 1.0, out, ctypes.c_void_p(0x1008), None) == _lib.PXL_E_INVALID
    # reachable-only fill and reduce modes: occupancy required, unknown modes rejected
    g, m, r = _valid()
    assert L.pxl_fill_reachable_f32(ctypes.byref(g), 0, ctypes.c_void_p(0x3000), ctypes.c_float(0.0),
                                    None) == _lib.PXL_E_INVALID  # no occupancy
    g.occupancy = 0x7000
    assert L.pxl_fill_reachable_f32(ctypes.byref(g), 5, ctypes.c_void_p(0x3000), ctypes.c_float(0.0),
                                    None) == _lib.PXL_E_INVALID  # unknown interp
    assert L.pxl_fill_reachable_f32(ctypes.byref(g), 0, ctypes.c_void_p(0x3004), ctypes.c_float(0.0),
                                    None) == _lib.PXL_E_INVALID  # misaligned
    assert L.pxl_blocks_reduce(ctypes.byref(g), ctypes.c_void_p(0x3000), ctypes.c_void_p(0x4000), 3, None,
                               None) == _lib.PXL_E_INVALID and "mode" in _lib.last_error()
    # the allocation-free variant validates the same way, before it looks up its scratch
    assert L.pxl_scaled_sum_noalloc(ctypes.c_void_p(0x3000), -1, 1.0, out, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum_noalloc(None, 5, 1.0, out, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum_noalloc(ctypes.c_void_p(0x3000), 5, 1.0, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum_noalloc(ctypes.c_void_p(0x3004), 5, 1.0, out, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum_noalloc_scratch_bytes() == 1024


def test_no_device_is_reported_not_aborted():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    g, m, r = _valid()
    r.n_rays = 4
    rc = _fwd(g, m, r)
    assert rc in (_lib.PXL_E_NO_DEVICE, _lib.PXL_E_HIP)
    assert _lib.last_error()
    rc = _lib.lib().pxl_scaled_sum_noalloc(ctypes.c_void_p(0x3000), 5, 1.0, ctypes.c_void_p(0x4000), None)
    assert rc in (_lib.PXL_E_NO_DEVICE, _lib.PXL_E_HIP) and "noalloc" in _lib.last_error()


def test_python_wrappers_reject_cpu_tensors():
    import torch
    from plenoxels_amd import render_rays
    grid = torch.zeros(4, 4, 4, 4)
    o = torch.zeros(2, 3)
    with pytest.raises(ValueError):
        render_rays(grid, o, o, num_samples=4, delta_step=0.1, grid_min=(0, 0, 0), points_distance=0.1)


def test_missing_library_fails_loudly(monkeypatch):
    """No CPU fallback: without libpxl.so every product entry raises instead of computing."""
    from plenoxels_amd import LossMean
    monkeypatch.setattr(_lib, "LIB_PATH", os.path.join(ROOT, "no-such-dir", "libpxl.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError, match="not built"):
        _lib.lib()
    with pytest.raises(ImportError, match="not built"):
        LossMean("cpu")
