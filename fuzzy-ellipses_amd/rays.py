"""Multi-pose ray batches: render rays that carry their own pose, and per-ray optical flow.

Mirrors baseline/fm_render.py:151-224 (render_func_idx_quattrans and
render_func_idx_quattrans_flow) with their argument order; the reference's functions pass
their arguments to render_func_rays in the wrong order (fm_render.py:162-164, 181-183), which
is fixed here.  Inputs:

* ``pixel_posei`` (R, 4): per ray (x, y, 1, pose index); the camera-frame ray is
  (x invF, y invF, 1);
* ``poses`` (P, 7): per pose a scalar-first quaternion (4) then a translation (3); the world
  ray is the camera ray @ quat_to_rot(q), starting at the translation.

The rays of each pose become one ray-list view of a single batched launch.  Within a pose
the rays are ordered along a Morton curve of their image position so that each tile of
consecutive rays covers a compact patch (the culling bound is per tile); outputs are
scattered back to the caller's order.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from . import BETA2, BETA3, RayList, Renderer, SceneTensors, _lib, _stream, quat_to_rot


def _morton_order(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Stable order of points along a Z curve over their bounding box (16-bit grid)."""
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0xFFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
        return v
    def quant(v):
        lo, hi = float(np.min(v)), float(np.max(v))
        span = hi - lo
        return np.zeros(v.shape, np.int64) if not span > 0 else np.floor((v - lo) / span * 65535.0).astype(np.int64)
    code = spread(quant(x)) | (spread(quant(y)) << np.uint64(1))
    return np.argsort(code, kind="stable")


class _Batch:
    """The pose grouping of one call: per pose, the caller's ray indices and the ray list."""

    def __init__(self, pixel_posei, invF: float, poses, device):
        pp = np.asarray(pixel_posei.detach().cpu() if torch.is_tensor(pixel_posei) else pixel_posei,
                        dtype=np.float64)
        if pp.ndim != 2 or pp.shape[1] != 4 or pp.shape[0] < 1:
            raise ValueError("pixel_posei must be (R, 4) with R >= 1")
        ps = np.asarray(poses.detach().cpu() if torch.is_tensor(poses) else poses, dtype=np.float64)
        if ps.ndim != 2 or ps.shape[1] != 7 or ps.shape[0] < 1:
            raise ValueError("poses must be (P, 7): quaternion (4) then translation (3)")
        if not np.all(np.isfinite(pp)) or not np.all(pp[:, 2] == 1.0):
            # t* (and with it the depth softmax) depends on the ray's length, so only the
            # homogeneous form z = 1 maps onto the renderer's normalised camera rays
            raise ValueError("pixel_posei: rays must be finite homogeneous pixel coordinates (z = 1)")
        idx = np.trunc(pp[:, 3]).astype(np.int64)
        if np.any(idx < 0) or np.any(idx >= ps.shape[0]):
            raise ValueError("pixel_posei: pose index out of range")
        self.pp, self.poses, self.invF = pp, ps, float(invF)
        self.rot = np.stack([quat_to_rot(q) for q in ps[:, :4]])
        self.R = pp.shape[0]
        self.groups = []
        xs = pp[:, 0] * invF
        ys = pp[:, 1] * invF
        for p in np.unique(idx):
            sel = np.nonzero(idx == p)[0]
            sel = sel[_morton_order(xs[sel], ys[sel])]
            xy = torch.tensor(np.stack([xs[sel], ys[sel]], 1), dtype=torch.float32, device=device).contiguous()
            self.groups.append((int(p), sel, RayList(self.rot[p], ps[p, 4:], xy)))


def _render(batch: _Batch, means, prec_full, weights_log, beta_2, beta_3, device, renderer: Optional[Renderer]):
    scene = SceneTensors(means, prec_full, weights_log, device)
    r = renderer if renderer is not None else Renderer(device=device)
    views = [(0, rl, beta_2, beta_3) for (_, _, rl) in batch.groups]
    outs = r.forward([scene], views)
    depth = torch.empty(batch.R, dtype=torch.float32, device=scene.means.device)
    alpha = torch.empty(batch.R, dtype=torch.float32, device=scene.means.device)
    for (_, sel, _), (d, a) in zip(batch.groups, outs):
        ix = torch.as_tensor(sel, device=scene.means.device)
        depth[ix] = d
        alpha[ix] = a
    return depth, alpha


def render_func_idx_quattrans(means, prec_full, weights_log, pixel_posei, invF, poses, beta_2=BETA2, beta_3=BETA3,
                              device="cuda", renderer: Optional[Renderer] = None):
    """(depth (R,), alpha (R,), None, None) for rays with per-ray poses (fm_render.py:151-166).

    The reference also returns the (N, R) softmax weights (``w``); they are not materialised here."""
    dev = torch.device(device)
    b = _Batch(pixel_posei, invF, poses, dev)
    d, a = _render(b, means, prec_full, weights_log, beta_2, beta_3, dev, renderer)
    return d, a, None, None


def ray_flow(pixel_posei: torch.Tensor, invF: float, rot: torch.Tensor, trans: torch.Tensor,
             depth: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Per-ray flow to the next and previous pose (fm_render.py:186-221) on the device."""
    n = int(pixel_posei.shape[0])
    dev = depth.device
    for t in (pixel_posei, rot, trans, depth):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.device != dev:
            raise ValueError("ray_flow: inputs must be contiguous float32 tensors on one device")
    fn = torch.empty((n, 2), dtype=torch.float32, device=dev)
    fp = torch.empty((n, 2), dtype=torch.float32, device=dev)
    lib = _lib.load()
    _lib.check(lib.fe_ray_flow(_stream(dev), n, pixel_posei.data_ptr(), float(invF), int(rot.shape[0]),
                               rot.data_ptr(), trans.data_ptr(), depth.data_ptr(), fn.data_ptr(), fp.data_ptr()),
               "fe_ray_flow")
    return fn, fp


def render_func_idx_quattrans_flow(means, prec_full, weights_log, pixel_posei, invF, poses, beta_2=BETA2,
                                   beta_3=BETA3, device="cuda", renderer: Optional[Renderer] = None):
    """(depth, alpha, None, None, flow_next (R,2), flow_prev (R,2)) as fm_render.py:170-224."""
    dev = torch.device(device)
    b = _Batch(pixel_posei, invF, poses, dev)
    d, a = _render(b, means, prec_full, weights_log, beta_2, beta_3, dev, renderer)
    pp = torch.tensor(b.pp, dtype=torch.float32, device=d.device)
    rot = torch.tensor(b.rot.reshape(-1, 9), dtype=torch.float32, device=d.device)
    trans = torch.tensor(b.poses[:, 4:], dtype=torch.float32, device=d.device).contiguous()
    fn, fp = ray_flow(pp, b.invF, rot, trans, d)
    return d, a, None, None, fn, fp
