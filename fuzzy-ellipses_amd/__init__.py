"""fuzzy_ellipses_amd -- MI355X-native fuzzy-metaball rasterizer (host-side Python API).

This mirrors the reference's caller-facing interface for the render path:

* ``render_func_quat`` / ``render_camera``: the arguments of ``baseline/fm_render.py:113``
  (means, precision factor, log-weights, camera, pose, beta2, beta3) returning depth and
  alpha images computed by the HIP kernels;
* ``rasterize``: the argument list of the reference binding
  ``fuzzy_ellipses/__init__.py:189-236`` (scales/quaternions, camera rays, intrinsics),
  through the reference-shaped custom-call target ``rasterize_gaussians_fwd`` (``xla.py``),
  with the binding's ``nan_to_num`` applied to the outputs;
* ``Renderer``: batched forward/backward over many scenes and views in one launch per
  kernel, with a reusable device workspace.

PyTorch is used only for device memory and the current HIP stream.  All arithmetic runs in
``lib/libfe_amd.so`` through its C ABI; if that library is missing, calls raise.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import FeCamera, FeError, FeOptions, FeScene, FeSceneGrad, FeStats, FeView, FeViewCotangent
from .scenes import BETA2, BETA3, Camera, Scene

__all__ = [
    "Renderer", "SceneTensors", "render_camera", "render_func_quat", "rasterize", "quat_to_rot",
    "prec_from_scale_quat", "FeError", "Camera", "Scene", "BETA2", "BETA3", "RayList",
    "render_func_idx_quattrans", "render_func_idx_quattrans_flow", "render_features",
]

MODES = {"culled": _lib.FE_MODE_CULLED, "dense": _lib.FE_MODE_DENSE}


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class SceneTensors:
    """Gaussians of one scene resident on the device (float32, contiguous), optionally with
    per-Gaussian features (N, C) (colour, descriptors, ...): the views of such a scene also
    render the features weighted by render_func_rays' depth weights (fm_render.py:96-109)."""

    def __init__(self, means, prec, weights_log, device="cuda", features=None):
        dev = torch.device(device)
        self.means = torch.as_tensor(means, dtype=torch.float32, device=dev).contiguous()
        self.prec = torch.as_tensor(prec, dtype=torch.float32, device=dev).contiguous()
        self.weights_log = torch.as_tensor(weights_log, dtype=torch.float32, device=dev).contiguous()
        n = self.means.shape[0]
        if self.means.shape != (n, 3) or self.prec.shape != (n, 3, 3) or self.weights_log.shape != (n,):
            raise ValueError("SceneTensors: expected means (N,3), prec (N,3,3), weights_log (N,)")
        self.features = None
        if features is not None:
            self.features = torch.as_tensor(features, dtype=torch.float32, device=dev).contiguous()
            if self.features.dim() != 2 or self.features.shape[0] != n or not 1 <= self.features.shape[1] <= 65536:
                raise ValueError("SceneTensors: expected features (N, C) with 1 <= C <= 65536")

    @classmethod
    def from_scene(cls, scene: Scene, device="cuda", features=None) -> "SceneTensors":
        return cls(scene.means, scene.prec, scene.weights_log, device, features=features)

    @property
    def n(self) -> int:
        return int(self.means.shape[0])

    @property
    def n_features(self) -> int:
        return 0 if self.features is None else int(self.features.shape[1])

    def fe(self) -> FeScene:
        return FeScene(self.n, _ptr(self.means), _ptr(self.prec), _ptr(self.weights_log), _ptr(self.features),
                       self.n_features)


class RayList:
    """The rays of one pose for a ray-list view: camera-plane coordinates ``xy`` ((n, 2) float32
    device tensor; ray k has camera-frame direction (xy[k, 0], xy[k, 1], 1)) and the pose
    ``R`` (rows = camera axes in world coordinates), ``t`` (ray origin)."""

    def __init__(self, R, t, xy: torch.Tensor):
        self.R = np.asarray(R, dtype=np.float64).reshape(3, 3)
        self.t = np.asarray(t, dtype=np.float64).reshape(3)
        if xy.dtype != torch.float32 or xy.dim() != 2 or xy.shape[1] != 2 or not xy.is_contiguous():
            raise ValueError("RayList: xy must be a contiguous (n, 2) float32 tensor")
        if xy.shape[0] < 1:
            raise ValueError("RayList: needs at least one ray")
        self.xy = xy

    @property
    def n(self) -> int:
        return int(self.xy.shape[0])


def fe_camera(cam) -> FeCamera:
    if isinstance(cam, RayList):
        c = FeCamera()
        R = cam.R.astype(np.float32).reshape(9)
        for i in range(9):
            c.R[i] = float(R[i])
        for i in range(3):
            c.t[i] = float(np.float32(cam.t[i]))
        c.fx = c.fy = 1.0
        c.height, c.width = 1, cam.n
        return c
    c = FeCamera()
    R = np.asarray(cam.R, dtype=np.float32).reshape(9)
    t = np.asarray(cam.t, dtype=np.float32).reshape(3)
    for i in range(9):
        c.R[i] = float(R[i])
    for i in range(3):
        c.t[i] = float(t[i])
    c.fx, c.fy, c.cx, c.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    c.height, c.width = int(cam.H), int(cam.W)
    return c


def _fe_view(si, cam, b2, b3, depth_ptr, alpha_ptr, feat_ptr=None) -> FeView:
    v = FeView(int(si), fe_camera(cam), float(b2), float(b3), depth_ptr, alpha_ptr)
    if isinstance(cam, RayList):
        v.rays, v.n_rays = cam.xy.data_ptr(), cam.n
    v.features_out = feat_ptr
    return v


def _n_pix(cam) -> int:
    return cam.n if isinstance(cam, RayList) else int(cam.H) * int(cam.W)


class Renderer:
    """Batched renderer over many scenes and views: one launch per kernel per call.

    ``tile`` in {8, 16, 32} pixels, ``batch`` (LDS depth, Gaussians per staged batch) in
    {32, 64, 128, 256}, ``mode`` "culled" (error-bounded tile culling) or "dense".  The
    device workspace is owned by the renderer and reused across calls; forward with
    ``save_for_backward=True`` keeps what ``backward`` needs in it.
    """

    def __init__(self, tile: int = 16, batch: int = 64, mode: str = "culled", device="cuda",
                 alpha_eps: float = 0.0, depth_eps: float = 0.0):
        self.lib = _lib.load()
        self.alpha_eps, self.depth_eps = float(alpha_eps), float(depth_eps)
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.tile, self.batch, self.mode = int(tile), int(batch), MODES[mode]
        self._ws: Optional[torch.Tensor] = None
        self._last = None       # forward call state kept for backward
        self._cache = None      # (key, FeScene[], FeView[]) of the last call

    def _opt(self, save: bool, reuse: bool = False) -> FeOptions:
        return FeOptions(self.tile, self.batch, self.mode, int(save),
                         _lib.FE_FLAG_REUSE_DESCRIPTORS if reuse else 0, self.alpha_eps, self.depth_eps)

    # -- planning -----------------------------------------------------------------------
    @staticmethod
    def _view_key(view):
        """Every value the C structs copy from a view: a cached struct is reused only when all
        of them are equal (object ids are not used: CPython reuses them after a free)."""
        si, cam, b2, b3 = view
        if isinstance(cam, RayList):
            ck = ("rays", cam.R.astype(np.float32).tobytes(), cam.t.astype(np.float32).tobytes(),
                  cam.xy.data_ptr(), cam.n)
        else:
            ck = (np.asarray(cam.R, dtype=np.float32).tobytes(), np.asarray(cam.t, dtype=np.float32).tobytes(),
                  int(cam.H), int(cam.W), float(np.float32(cam.fx)), float(np.float32(cam.fy)),
                  float(np.float32(cam.cx)), float(np.float32(cam.cy)))
        return (int(si), ck, float(np.float32(b2)), float(np.float32(b3)))

    def _structs(self, scenes: Sequence[SceneTensors], views, outs):
        key = (tuple((s.n, _ptr(s.means), _ptr(s.prec), _ptr(s.weights_log), _ptr(s.features), s.n_features)
                     for s in scenes),
               tuple(self._view_key(v) for v in views), tuple(tuple(_ptr(x) for x in o) for o in outs))
        if self._cache is not None and self._cache[0] == key:
            return self._cache[1], self._cache[2]
        sc = (FeScene * len(scenes))(*[s.fe() for s in scenes])
        vw = (FeView * len(views))()
        for i, (si, cam, b2, b3) in enumerate(views):
            o = outs[i]
            vw[i] = _fe_view(si, cam, b2, b3, _ptr(o[0]), _ptr(o[1]), _ptr(o[2]) if len(o) > 2 else None)
        self._cache = (key, sc, vw)
        return sc, vw

    def _check_image(self, t, cam, what, channels=1):
        """A per-view image buffer: contiguous float32 on the render device with `channels`
        values per pixel (H*W) or ray (n); the kernels index it by pixel."""
        n = _n_pix(cam) * channels
        if channels < 1:
            raise FeError(f"{what}: the view's scene has no features")
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_contiguous() \
                or t.device != self.device or t.numel() != n:
            raise FeError(f"{what}: expected a contiguous float32 tensor of {n} values on {self.device}")

    def workspace_bytes(self, scenes, views, save_for_backward=False) -> int:
        sc = (FeScene * len(scenes))(*[s.fe() for s in scenes])
        vw = (FeView * len(views))()
        for i, (si, cam, b2, b3) in enumerate(views):
            vw[i] = _fe_view(si, cam, b2, b3, None, None)
        opt = self._opt(save_for_backward)
        nbytes = ctypes.c_size_t(0)
        _lib.check(self.lib.fe_plan(sc, len(scenes), vw, len(views), ctypes.byref(opt), ctypes.byref(nbytes)),
                   "fe_plan")
        return int(nbytes.value)

    def _workspace(self, nbytes: int) -> torch.Tensor:
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = None
            self._ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def alloc_outputs(self, views, scenes=None):
        """(depth, alpha) per view, plus the feature image (.., C) for views whose scene has
        features (when ``scenes`` is given)."""
        shape = lambda cam: (cam.n,) if isinstance(cam, RayList) else (cam.H, cam.W)
        out = []
        for (si, cam, _, _) in views:
            o = (torch.empty(shape(cam), dtype=torch.float32, device=self.device),
                 torch.empty(shape(cam), dtype=torch.float32, device=self.device))
            if scenes is not None and scenes[si].n_features > 0:
                o = o + (torch.empty(shape(cam) + (scenes[si].n_features,), dtype=torch.float32, device=self.device),)
            out.append(o)
        return out

    # -- forward ------------------------------------------------------------------------
    def forward(self, scenes: Sequence[SceneTensors], views, save_for_backward: bool = False,
                outs=None, reuse_descriptors: bool = False) -> List[Tuple[torch.Tensor, torch.Tensor]]:
        """Render every view.  ``views`` = [(scene_index, Camera or RayList, beta2, beta3), ...].

        Returns [(depth, alpha), ...] float32 device tensors, (H, W) for a Camera and (n,)
        for a RayList.  With
        ``reuse_descriptors`` (same scenes, views, outputs and options as the previous call)
        the call only enqueues kernels and may be captured into a HIP graph.
        """
        if outs is None:
            outs = self.alloc_outputs(views, scenes)
        elif len(outs) != len(views):
            raise FeError("forward: outs needs one (depth, alpha) pair per view")
        else:
            for o, (si, cam, _, _) in zip(outs, views):
                self._check_image(o[0], cam, "forward: depth output")
                self._check_image(o[1], cam, "forward: alpha output")
                if len(o) > 2:
                    self._check_image(o[2], cam, "forward: feature output", scenes[si].n_features)
        sc, vw = self._structs(scenes, views, outs)
        opt = self._opt(save_for_backward)
        nbytes = ctypes.c_size_t(0)
        _lib.check(self.lib.fe_plan(sc, len(scenes), vw, len(views), ctypes.byref(opt), ctypes.byref(nbytes)),
                   "fe_plan")
        if reuse_descriptors and (self._ws is None or self._ws.numel() < nbytes.value):
            raise FeError("forward: reuse_descriptors needs a prior call with the same arguments")
        ws = self._workspace(int(nbytes.value))
        opt = self._opt(save_for_backward, reuse_descriptors)
        _lib.check(self.lib.fe_forward(_stream(self.device), sc, len(scenes), vw, len(views), ctypes.byref(opt),
                                       ws.data_ptr(), ws.numel()), "fe_forward")
        self._last = (list(scenes), sc, vw, outs, list(views)) if save_for_backward else None
        return outs

    def stats(self) -> dict:
        st = FeStats()
        _lib.check(self.lib.fe_read_stats(_stream(self.device), self._ws.data_ptr(), self._ws.numel(),
                                          ctypes.byref(st)), "fe_read_stats")
        return {"pairs_phase1": st.pairs_phase1, "pairs_phase2": st.pairs_phase2, "evals": st.evals,
                "evals_phase1": st.evals_phase1, "floor_tiles": st.floor_tiles,
                "floor_unresolved": st.floor_unresolved}

    def debug_state(self) -> dict:
        """Host copies of the last forward's workspace internals (tests only): per-Gaussian
        records and their order (record j of a view is Gaussian order[j]), tile ranges and
        lists of both phases (record ids), per-tile phase-1 thresholds and floor seeds, and the
        tile dispatch orders of the phase-1 blend and of the backward (2, T)."""
        torch.cuda.synchronize(self.device)
        ws = self._ws.cpu().numpy()
        h = ws[:8 * 14].view(np.uint64)
        magic, o_cnt, o_off1, o_off2, T, mode, o_rec, o_list, o_lmin, n_rec, o_st, n_pix, o_ord, o_tord = \
            [int(x) for x in h]
        out = {"records": ws[o_rec:o_rec + 64 * n_rec].view(np.float32).reshape(n_rec, 16), "mode": mode,
               "order": ws[o_ord:o_ord + 4 * n_rec].view(np.int32).copy()}
        if mode == _lib.FE_MODE_CULLED:
            off1 = ws[o_off1:o_off1 + 8 * (T + 1)].view(np.uint64).astype(np.int64)
            off2 = ws[o_off2:o_off2 + 8 * (T + 1)].view(np.uint64).astype(np.int64)
            total = int(off2[-1])
            out.update(off1=off1, off2=off2, list=ws[o_list:o_list + 4 * total].view(np.uint32).copy(),
                       tile_lmin=ws[o_lmin:o_lmin + 4 * T].view(np.float32).copy(),
                       tile_seed=ws[o_lmin + 4 * T:o_lmin + 8 * T].view(np.float32).copy(),
                       tile_order=ws[o_tord:o_tord + 8 * T].view(np.int32).reshape(2, T).copy())
        return out

    # -- backward -----------------------------------------------------------------------
    def backward(self, d_depth: Sequence[Optional[torch.Tensor]], d_alpha: Sequence[Optional[torch.Tensor]],
                 grads=None, reuse_descriptors: bool = False,
                 d_features: Optional[Sequence[Optional[torch.Tensor]]] = None):
        """Gradients of sum_v <d_depth_v, depth_v> + <d_alpha_v, alpha_v> (+ <d_features_v,
        features_v> for views that rendered features) with respect to every scene's (means,
        prec upper triangle, weights_log[, features]), from the last forward with
        save_for_backward=True (its feature images untouched since).  ``grads`` (optional) =
        [(d_means, d_prec, d_w[, d_features]), ...] buffers to accumulate into; by default fresh
        zeroed tensors are returned, with d_features for scenes that have features."""
        if self._last is None:
            raise FeError("backward: the last forward was not run with save_for_backward=True")
        scenes, sc, vw, outs, views = self._last
        if len(d_depth) != len(outs) or len(d_alpha) != len(outs):
            raise FeError("backward: need one depth and one alpha cotangent (or None) per view")
        if d_features is not None and len(d_features) != len(outs):
            raise FeError("backward: d_features needs one entry (or None) per view")
        cot = (FeViewCotangent * len(outs))()
        for i in range(len(outs)):
            dd, da = d_depth[i], d_alpha[i]
            df = None if d_features is None else d_features[i]
            si, cam = views[i][0], views[i][1]
            for t, what in ((dd, "depth"), (da, "alpha")):
                if t is not None:
                    self._check_image(t, cam, f"backward: {what} cotangent of view {i}")
            if df is not None:
                if len(outs[i]) < 3:
                    raise FeError(f"backward: view {i} rendered no features")
                self._check_image(df, cam, f"backward: feature cotangent of view {i}", scenes[si].n_features)
            cot[i] = FeViewCotangent(_ptr(dd), _ptr(da), _ptr(df))
        if grads is None:
            grads = [(torch.zeros_like(s.means), torch.zeros_like(s.prec), torch.zeros_like(s.weights_log))
                     + ((torch.zeros_like(s.features),) if s.features is not None else ())
                     for s in scenes]
        if len(grads) != len(scenes):
            raise FeError("backward: need one (d_means, d_prec, d_weights_log) triple per scene")
        for s, g in zip(scenes, grads):
            refs = (s.means, s.prec, s.weights_log) + ((s.features,) if len(g) == 4 else ())
            if len(g) not in (3, 4) or (len(g) == 4 and s.features is None):
                raise FeError("backward: need one (d_means, d_prec, d_weights_log[, d_features]) tuple per scene")
            for t, ref in zip(g, refs):
                if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_contiguous() \
                        or t.device != self.device or t.shape != ref.shape:
                    raise FeError(f"backward: gradient buffers must be contiguous float32 tensors on "
                                  f"{self.device} shaped like the scene's (means, prec, weights_log, features)")
        gs = (FeSceneGrad * len(scenes))(*[FeSceneGrad(*[_ptr(x) for x in g]) for g in grads])
        opt = self._opt(True, reuse_descriptors)
        ws = self._ws
        _lib.check(self.lib.fe_backward(_stream(self.device), sc, len(scenes), vw, len(outs), ctypes.byref(opt),
                                        cot, gs, ws.data_ptr(), ws.numel()), "fe_backward")
        return grads


# ------------------------------------------------------------------------------------------
# reference-shaped entry points
# ------------------------------------------------------------------------------------------
_default = {}


def _renderer(tile=16, batch=64, mode="culled", device="cuda") -> Renderer:
    key = (tile, batch, mode, str(device))
    if key not in _default:
        _default[key] = Renderer(tile, batch, mode, device)
    return _default[key]


def quat_to_rot(q) -> np.ndarray:
    """Scalar-first quaternion -> rotation (baseline/util_render.py:51-76 semantics)."""
    w, x, y, z = [float(v) for v in q]
    nq = w * w + x * x + y * y + z * z
    if not nq > 1e-12:
        return np.eye(3)
    s = 2.0 / nq
    X, Y, Z = x * s, y * s, z * s
    return np.array([[1.0 - (y * Y + z * Z), x * Y - w * Z, x * Z + w * Y],
                     [x * Y + w * Z, 1.0 - (x * X + z * Z), y * Z - w * X],
                     [x * Z - w * Y, y * Z + w * X, 1.0 - (x * X + y * Y)]])


def render_camera(means, prec_full, weights_log, camera: Camera, beta_2=BETA2, beta_3=BETA3,
                  tile=16, batch=64, mode="culled", device="cuda"):
    """Depth and alpha (H, W) for one pinhole camera; same math as render_func_rays."""
    st = means if isinstance(means, SceneTensors) else SceneTensors(means, prec_full, weights_log, device)
    r = _renderer(tile, batch, mode, device)
    (d, a), = r.forward([st], [(0, camera, beta_2, beta_3)])
    return d, a


def render_features(means, prec_full, weights_log, features, camera: Camera, beta_2=BETA2, beta_3=BETA3,
                    tile=16, batch=64, mode="culled", device="cuda"):
    """Depth (H, W), alpha (H, W) and the feature image (H, W, C) = w @ features per pixel, w the
    normalised depth weights render_func_rays returns for "color, DINO features, or any other
    per-Gaussian property" (baseline/fm_render.py:96-109), without materialising w (N, P)."""
    st = SceneTensors(means, prec_full, weights_log, device, features=features)
    r = _renderer(tile, batch, mode, device)
    (d, a, f), = r.forward([st], [(0, camera, beta_2, beta_3)])
    return d, a, f


MAX_WEIGHT_ENTRIES = 1 << 28


def render_func_quat(means, prec_full, weights_log, intrinsics, quat, t, beta_2=BETA2, beta_3=BETA3,
                     return_weights=False, **kw):
    """baseline/fm_render.py:113-126 with the camera rays given by pinhole intrinsics
    (H, W, fx, fy, cx, cy) instead of an explicit (H*W, 3) ray array.  Returns
    (depth (H*W,), alpha (H*W,), None, w) like the reference; w, the (N, P) normalised weights,
    is materialised only with ``return_weights`` (as the feature image of the identity
    features, N*P <= 2^28), else None -- use render_features to contract it with
    per-Gaussian properties instead."""
    H, W, fx, fy, cx, cy = intrinsics
    cam = Camera(R=quat_to_rot(quat), t=np.asarray(t, dtype=np.float64), H=int(H), W=int(W), fx=float(fx),
                 fy=float(fy), cx=float(cx), cy=float(cy))
    if not return_weights:
        d, a = render_camera(means, prec_full, weights_log, cam, beta_2, beta_3, **kw)
        return d.reshape(-1), a.reshape(-1), None, None
    n = int(np.asarray(weights_log).shape[0]) if not isinstance(weights_log, torch.Tensor) else int(weights_log.shape[0])
    if n * cam.H * cam.W > MAX_WEIGHT_ENTRIES or n > 65536:
        raise ValueError("render_func_quat: the (N, P) weights are too large to materialise; use render_features")
    dev = kw.get("device", "cuda")
    eye = torch.eye(max(n, 1), dtype=torch.float32, device=dev)[:n]
    d, a, f = render_features(means, prec_full, weights_log, eye, cam, beta_2, beta_3, **kw)
    return d.reshape(-1), a.reshape(-1), None, f.reshape(-1, n).t()


def prec_from_scale_quat(scales, quats) -> np.ndarray:
    """Upper-triangular U with U U^T = (R diag(s^2) R^T)^-1 for each Gaussian (float64 host math).

    The reference binding takes scales and quaternions (fuzzy_ellipses/__init__.py:189-203);
    the renderer's parameter is the precision factor of baseline/fm_render.py:14.
    """
    scales = np.asarray(scales, dtype=np.float64)
    quats = np.asarray(quats, dtype=np.float64)
    R = np.stack([quat_to_rot(q) for q in quats])
    cov = np.einsum("nij,nj,nkj->nik", R, scales ** 2, R)
    C = np.linalg.cholesky(cov)
    return np.triu(np.transpose(np.linalg.inv(C), (0, 2, 1)))


def projection_matrix(width, height, fx, fy, cx, cy, znear, zfar) -> np.ndarray:
    """The reference binding's projection matrix (fuzzy_ellipses/__init__.py:36-76, transposed
    as it is there; cx and cy do not enter).  It is passed to the custom call, which accepts
    and ignores it like the view matrix."""
    tan_x = math.tan(math.atan(width / 2 / fx))
    tan_y = math.tan(math.atan(height / 2 / fy))
    top, right = tan_y * znear, tan_x * znear
    bottom, left = -top, -right
    P = np.array([[2.0 * znear / (right - left), 0.0, (right + left) / (right - left), 0.0],
                  [0.0, 2.0 * znear / (top - bottom), (top + bottom) / (top - bottom), 0.0],
                  [0.0, 0.0, zfar / (zfar - znear), -(zfar * znear) / (zfar - znear)],
                  [0.0, 0.0, 1.0, 0.0]])
    return P.T


def rasterize(means3D, scales, rotations, weights, camera_rays, camera_rot, camera_trans, image_height,
              image_width, fx, fy, cx, cy, near=0.01, far=100.0, device="cuda"):
    """Argument list and results of the reference binding (fuzzy_ellipses/__init__.py:189-236).

    Runs the reference's custom-call target ``rasterize_gaussians_fwd`` of libfe_amd.so with
    the reference's 11 buffers and 20-byte descriptor (``xla.call_custom_target``): scales and
    quaternions become the precision factor on the device, ``camera_rays`` (H*W, 3) may be any
    camera-frame rays with third component 1 (rows with another third component get a NaN
    depth, zeroed here), beta2=21.4 and beta3=2.66 as the reference fixes them
    (rasterizer_impl.cu:196-197).  Returns the binding's nan_to_num'd (depth, alpha) of shape
    (H, W) as float32 device tensors.
    """
    from . import xla
    H, W = int(image_height), int(image_width)
    dev = torch.device(device)
    if dev.type == "cuda" and dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    f32 = lambda a: torch.as_tensor(np.asarray(a) if not isinstance(a, torch.Tensor) else a,
                                    dtype=torch.float32, device=dev).contiguous()
    means, w, sc, q = f32(means3D), f32(weights), f32(scales), f32(rotations)
    rays, R, t = f32(camera_rays), f32(camera_rot), f32(camera_trans)
    n = int(w.shape[0])
    if means.shape != (n, 3) or sc.shape != (n, 3) or q.shape != (n, 4):
        raise ValueError("rasterize: expected means3D (N,3), scales (N,3), rotations (N,4), weights (N,)")
    if rays.numel() != 3 * H * W or R.numel() != 9 or t.numel() != 3:
        raise ValueError("rasterize: expected camera_rays (H*W,3), camera_rot (3,3), camera_trans (3,)")
    view = f32(np.eye(4))
    proj = f32(projection_matrix(W, H, fx, fy, cx, cy, near, far))
    tan_fovx = math.tan(math.atan(W / 2 / fx) * 2.0)
    tan_fovy = math.tan(math.atan(H / 2 / fy) * 2.0)
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    alpha = torch.empty((H, W), dtype=torch.float32, device=dev)
    opaque = xla.build_gaussian_rasterize_descriptor(H, W, n, tan_fovx, tan_fovy)
    bufs = [means, w, sc, q, rays, R, t, view, proj, depth, alpha]
    xla.call_custom_target(_stream(dev), [b.data_ptr() for b in bufs], opaque)
    d = torch.nan_to_num(depth, nan=0.0, posinf=10.0, neginf=0.0)
    a = torch.nan_to_num(alpha, nan=0.0, posinf=1.0, neginf=0.0)
    return d, a


from .rays import render_func_idx_quattrans, render_func_idx_quattrans_flow  # noqa: E402
