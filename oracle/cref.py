"""ctypes wrapper of the FP64 C oracle oracle/render_ref.c (test infrastructure only)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpxl_oracle.so")


class RefParams(ctypes.Structure):
    _fields_ = [
        ("X", ctypes.c_int64), ("Y", ctypes.c_int64), ("Z", ctypes.c_int64),
        ("n_samples", ctypes.c_int32), ("interp", ctypes.c_int32),
        ("div_mode", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("delta", ctypes.c_float), ("near", ctypes.c_float),
        ("pd", ctypes.c_float), ("inv_pd", ctypes.c_float),
        ("gmin", ctypes.c_float * 3), ("t_stop", ctypes.c_float),
    ]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.ref_render.argtypes = [ctypes.POINTER(RefParams), P, P, P, P, ctypes.c_int64, P, P, P, P]
        L.ref_render.restype = ctypes.c_int
        L.ref_rays_touching.argtypes = [ctypes.POINTER(RefParams), P, P, P, P, ctypes.c_int64,
                                        P, ctypes.c_int64, P]
        L.ref_rays_touching.restype = ctypes.c_int
        L.ref_sample_coord.argtypes = [ctypes.POINTER(RefParams), P, P, ctypes.c_int, P]
        L.ref_sample_coord.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_params(dims, *, num_samples, delta_step, gmin, points_distance, interp="nearest",
                near=0.0, div_mode="divide", t_stop=0.0):
    p = RefParams()
    p.X, p.Y, p.Z = (int(x) for x in dims)
    p.n_samples = int(num_samples)
    p.interp = {"nearest": 0, "trilinear": 1}[interp]
    p.div_mode = {"divide": 0, "reciprocal": 1}[div_mode]
    p.delta = float(np.float32(delta_step))
    p.near = float(np.float32(near))
    p.pd = float(np.float32(points_distance))
    p.inv_pd = float(np.float32(1.0) / np.float32(points_distance))
    for a in range(3):
        p.gmin[a] = float(gmin[a])
    p.t_stop = float(t_stop)
    return p


def render(grid, origins, dirs, *, grad_rgba=None, occupancy_bits=None, **kw):
    """grid: (X,Y,Z,4) array-like (converted to float64); origins/dirs (R,3) float32.

    Returns (rgba (R,4) f64, grad_grid (X,Y,Z,4) f64 or None, n_proc (R,) int32).
    """
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64))
    o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32))
    d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float32))
    R = o.shape[0]
    p = make_params(g.shape[:3], **kw)
    out = np.zeros((R, 4), np.float64)
    nproc = np.zeros(R, np.int32)
    occ = None if occupancy_bits is None else np.ascontiguousarray(
        np.asarray(occupancy_bits).view(np.uint32))
    gg = None
    gr = None
    if grad_rgba is not None:
        gr = np.ascontiguousarray(np.asarray(grad_rgba, dtype=np.float64))
        gg = np.zeros(g.shape, np.float64)
    rc = lib().ref_render(ctypes.byref(p), _ptr(g), _ptr(occ), _ptr(o), _ptr(d), R, _ptr(out),
                          _ptr(gr), _ptr(gg), _ptr(nproc))
    if rc != 0:
        raise MemoryError("ref_render failed")
    return out, gg, nproc


def render_parallel(grid, origins, dirs, *, threads=None, chunk=16384, occupancy_bits=None, progress=None, **kw):
    """Forward-only render of many rays on `threads` host threads (ref_render keeps no state
    and ctypes releases the GIL): (R,4) float64 RGBA.  progress(done, total) is called after
    each chunk."""
    from concurrent.futures import ThreadPoolExecutor
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64))
    o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32))
    d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float32))
    R = o.shape[0]
    p = make_params(g.shape[:3], **kw)
    out = np.zeros((R, 4), np.float64)
    nproc = np.zeros(R, np.int32)
    occ = None if occupancy_bits is None else np.ascontiguousarray(
        np.asarray(occupancy_bits).view(np.uint32))
    L = lib()

    def work(s):
        e = min(R, s + chunk)
        rc = L.ref_render(ctypes.byref(p), _ptr(g), _ptr(occ), _ptr(o[s:e]), _ptr(d[s:e]), e - s,
                          _ptr(out[s:e]), None, None, _ptr(nproc[s:e]))
        if rc != 0:
            raise MemoryError("ref_render failed")
        return e - s

    done = 0
    with ThreadPoolExecutor(max_workers=threads or min(16, os.cpu_count() or 1)) as ex:
        for n in ex.map(work, range(0, R, chunk)):
            done += n
            if progress is not None:
                progress(done, R)
    return out


def rays_touching(grid, origins, dirs, query_cells, *, occupancy_bits=None, **kw):
    """(R, Q) uint8: 1 where ray r reads linear cell query_cells[q] (query sorted here)."""
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64))
    o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32))
    d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float32))
    q = np.asarray(query_cells, dtype=np.int64)
    order = np.argsort(q)
    qs = np.ascontiguousarray(q[order])
    R = o.shape[0]
    hits = np.zeros((R, qs.size), np.uint8)
    occ = None if occupancy_bits is None else np.ascontiguousarray(
        np.asarray(occupancy_bits).view(np.uint32))
    p = make_params(g.shape[:3], **kw)
    lib().ref_rays_touching(ctypes.byref(p), _ptr(g), _ptr(occ), _ptr(o), _ptr(d), R,
                            _ptr(qs), qs.size, _ptr(hits))
    out = np.zeros_like(hits)
    out[:, order] = hits
    return out


def sample_coord(origin, direction, k, dims, **kw):
    p = make_params(dims, num_samples=1, **kw)
    o = np.ascontiguousarray(np.asarray(origin, np.float32))
    d = np.ascontiguousarray(np.asarray(direction, np.float32))
    n = np.zeros(3, np.float32)
    lib().ref_sample_coord(ctypes.byref(p), _ptr(o), _ptr(d), int(k), _ptr(n))
    return n
