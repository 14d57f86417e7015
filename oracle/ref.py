"""ctypes wrapper of oracle/_build/libfe_oracle.so -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests, the smoke check and the benchmark's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfe_oracle.so")
LOG2E = 1.4426950408889634

_lib = None

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.fe_ref_masked_bound_check.restype = ctypes.c_int64
    lib.fe_ref_masked_bound_check.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                              ctypes.c_void_p]
    lib.fe_ref_render_f64.restype = None
    lib.fe_ref_render_f64.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                      ctypes.c_int64, _f64p, _f64p, ctypes.c_int]
    lib.fe_ref_render_f32.restype = None
    lib.fe_ref_render_f32.argtypes = [ctypes.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_int64, _f32p,
                                      _f32p, ctypes.c_int]
    lib.fe_ref_records.restype = None
    lib.fe_ref_records.argtypes = [ctypes.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_float,
                                   ctypes.c_float, _f32p, ctypes.POINTER(ctypes.c_int)]
    lib.fe_ref_tiles.restype = ctypes.c_int64
    lib.fe_ref_tiles.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, _u32p, _u32p]
    lib.fe_ref_ray_rects.restype = None
    lib.fe_ref_ray_rects.argtypes = [ctypes.c_int64, ctypes.c_int, _f32p, _f32p]
    lib.fe_ref_tiles_rects.restype = ctypes.c_int64
    lib.fe_ref_tiles_rects.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int64, _f32p, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_void_p, ctypes.c_void_p, _u32p, _u32p]
    lib.fe_ref_order.restype = None
    lib.fe_ref_order.argtypes = [ctypes.c_int, _f32p, _f32p, _f32p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
    _lib = lib
    return lib


def _pix(pixels):
    if pixels is None:
        return None, 0, None
    p = np.ascontiguousarray(pixels, dtype=np.int64)
    return p.ctypes.data_as(ctypes.c_void_p), int(p.size), p


def render_f64(scene, cam, pixels=None, nthreads=0):
    """Dense float64 render of one pinhole view -> (depth, alpha) over `pixels` (or all)."""
    lib = load()
    pp, npix, keep = _pix(pixels)
    m = npix if pixels is not None else cam.H * cam.W
    d = np.empty(m)
    a = np.empty(m)
    lib.fe_ref_render_f64(scene.n, np.ascontiguousarray(scene.means, np.float64),
                          np.ascontiguousarray(scene.prec, np.float64).reshape(-1),
                          np.ascontiguousarray(scene.weights_log, np.float64),
                          np.ascontiguousarray(cam.R, np.float64).reshape(-1), np.ascontiguousarray(cam.t, np.float64),
                          cam.fx, cam.fy, cam.cx, cam.cy, cam.H, cam.W, scene.beta2, scene.beta3, pp, npix, d, a,
                          int(nthreads))
    return d, a


def render_f32(scene, cam, pixels=None, nthreads=0):
    """Dense float32 CPU render (the timed CPU baseline)."""
    lib = load()
    pp, npix, keep = _pix(pixels)
    m = npix if pixels is not None else cam.H * cam.W
    d = np.empty(m, np.float32)
    a = np.empty(m, np.float32)
    lib.fe_ref_render_f32(scene.n, np.ascontiguousarray(scene.means, np.float32),
                          np.ascontiguousarray(scene.prec, np.float32).reshape(-1),
                          np.ascontiguousarray(scene.weights_log, np.float32),
                          np.ascontiguousarray(cam.R, np.float32).reshape(-1), np.ascontiguousarray(cam.t, np.float32),
                          cam.fx, cam.fy, cam.cx, cam.cy, cam.H, cam.W, scene.beta2, scene.beta3, pp, npix, d, a,
                          int(nthreads))
    return d, a


def records(scene, cam):
    """(N,16) float32 per-Gaussian records of one view and the any-degenerate flag."""
    lib = load()
    n = scene.n
    rec = np.zeros((max(n, 1), 16), np.float32)
    deg = ctypes.c_int(0)
    lib.fe_ref_records(n, np.ascontiguousarray(scene.means, np.float32),
                       np.ascontiguousarray(scene.prec, np.float32).reshape(-1),
                       np.ascontiguousarray(scene.weights_log, np.float32),
                       np.ascontiguousarray(cam.R, np.float32).reshape(-1), np.ascontiguousarray(cam.t, np.float32),
                       np.float32(cam.fx), np.float32(cam.fy), rec.reshape(-1), ctypes.byref(deg))
    return rec[:n], int(deg.value)


def cull_constants(n, alpha_eps=2e-5, depth_eps=1e-6):
    """(c_alpha, delta2, floor_thr) exactly as the planner computes them (float32 of float64
    math, with the budgets rounded to float32 as the C ABI stores them)."""
    ln_n = np.log(max(n, 1))
    ae = float(np.float32(alpha_eps))
    de = float(np.float32(depth_eps))
    return (float(np.float32(ln_n - np.log(ae))), float(np.float32((ln_n - np.log(de)) * LOG2E)),
            float(np.float32(13.0 * 3.321928094887362 - ln_n * LOG2E)))


def _fptr(a):
    """(array kept alive, void pointer) of an optional float32 vector."""
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, np.float32)
    return arr, arr.ctypes.data_as(ctypes.c_void_p)


def tiles(scene, cam, ts, tile_lmin=None, tile_seed=None):
    """Per-tile sorted Gaussian lists: (counts (T,), concatenated ids) for phase 1, or for
    phase 2 when the per-tile thresholds `tile_lmin` (log2 units) are given, with the phase-1
    floor seeds `tile_seed` (log2 units, +inf forces dense completion; None: -inf)."""
    lib = load()
    rec, deg = records(scene, cam)
    n = scene.n
    tx, ty = (cam.W + ts - 1) // ts, (cam.H + ts - 1) // ts
    T = tx * ty
    counts = np.zeros(T, np.uint32)
    lst = np.zeros(max(n * T, 1), np.uint32)
    c_alpha, delta2, floor_thr = cull_constants(n)
    lm_arr, lm = _fptr(tile_lmin)
    sd_arr, sd = _fptr(tile_seed)
    total = lib.fe_ref_tiles(n, np.ascontiguousarray(rec.reshape(-1) if n else np.zeros(16, np.float32)), deg,
                             cam.H, cam.W, ts, np.float32(cam.cx), np.float32(cam.cy), np.float32(scene.beta2),
                             np.float32(scene.beta3), np.float32(c_alpha), np.float32(delta2),
                             np.float32(floor_thr), lm, sd, counts, lst)
    return counts, lst[:total].copy()


def masked_bound_check(scene, cam, ts):
    """(violating (tile, Gaussian) pairs, masked (pixel, Gaussian) evaluations checked) of the
    membership's masked-logit bound against exact float64 logits over every pixel."""
    lib = load()
    rec, _ = records(scene, cam)
    checked = ctypes.c_int64(0)
    bad = lib.fe_ref_masked_bound_check(scene.n, np.ascontiguousarray(rec.reshape(-1)), cam.H, cam.W, ts,
                                        np.float32(cam.cx), np.float32(cam.cy), np.float32(scene.beta2),
                                        np.float32(scene.beta3), ctypes.byref(checked))
    return int(bad), int(checked.value)


def ray_rects(xy, ts):
    """(T, 4) float32 tile rectangles (centre x, centre y, half-width, half-height) of a ray list."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    T = (xy.shape[0] + ts * ts - 1) // (ts * ts)
    out = np.zeros((T, 4), np.float32)
    load().fe_ref_ray_rects(xy.shape[0], ts, xy.reshape(-1), out.reshape(-1))
    return out


def tiles_rays(scene, R, t, xy, ts, tile_lmin=None, tile_seed=None):
    """Membership lists of a ray-list view (records with fx = fy = 1), as `tiles`."""
    from types import SimpleNamespace
    lib = load()
    cam = SimpleNamespace(R=np.asarray(R), t=np.asarray(t), fx=1.0, fy=1.0)
    rec, deg = records(scene, cam)
    n = scene.n
    rects = ray_rects(xy, ts)
    T = rects.shape[0]
    counts = np.zeros(T, np.uint32)
    lst = np.zeros(max(n * T, 1), np.uint32)
    c_alpha, delta2, floor_thr = cull_constants(n)
    lm_arr, lm = _fptr(tile_lmin)
    sd_arr, sd = _fptr(tile_seed)
    total = lib.fe_ref_tiles_rects(n, np.ascontiguousarray(rec.reshape(-1) if n else np.zeros(16, np.float32)), deg,
                                   T, rects.reshape(-1), np.float32(scene.beta2), np.float32(scene.beta3),
                                   np.float32(c_alpha), np.float32(delta2), np.float32(floor_thr), lm, sd, counts, lst)
    return counts, lst[:total].copy()


def order(scene, cam, ray_view=False):
    """The renderer's record order of one view, restated from the inputs (fe_ref_order):
    record j is Gaussian order[j].  For a ray-list view pass ray_view=True (cam needs R, t)."""
    lib = load()
    n = scene.n
    out = np.zeros(max(n, 1), np.int32)
    fx = 1.0 if ray_view else cam.fx
    fy = 1.0 if ray_view else cam.fy
    cx = 0.0 if ray_view else cam.cx
    cy = 0.0 if ray_view else cam.cy
    H = 1 if ray_view else cam.H
    W = 1 if ray_view else cam.W
    lib.fe_ref_order(n, np.ascontiguousarray(scene.means if n else np.zeros((1, 3)), np.float32).reshape(-1),
                     np.ascontiguousarray(cam.R, np.float32).reshape(-1), np.ascontiguousarray(cam.t, np.float32),
                     np.float32(fx), np.float32(fy), np.float32(cx), np.float32(cy), H, W, int(ray_view), out)
    return out[:n]
