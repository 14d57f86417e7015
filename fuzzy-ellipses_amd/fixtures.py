"""The reference's on-disk scene format: a directory `<W>_<H>_<N>/` of raw float32 `.bin` files.

Written by baseline/ycb_fmb_rotscale.py:239-279 and read by tests/test_jax.py:19-49 and the
native harness (tests/test_rasterize_main.cu).  Every file is a flat little-endian array with
no header; shapes follow from `width_height_gaussians.bin` (int32 [W, H, N]):

    zs.bin, alphas.bin          (H*W,)   rendered depth and alpha (ground truth)
    means.bin                   (N, 3)
    quats.bin                   (N, 4)   scalar-first rotation of each Gaussian
    scales.bin                  (N, 3)   per-axis standard deviations
    covs.bin                    (N, 3, 3) R diag(s)^2 R^T
    precs.bin                   (N, 3, 3) lower Cholesky factor L of the precision, L L^T = cov^-1
    weights.bin                 (N,)     log-weights
    camera_rays.bin             (H*W, 3) ((x - cx)/fx, (y - cy)/fy, 1), row-major y*W + x
    camera_rot.bin              (3, 3)   camera rotation R (world ray = camera ray @ R)
    camera_trans.bin            (3,)     camera centre t
    camera_view_matrix.bin      (4, 4)   transpose of the inverse camera-to-world matrix
    camera_proj_matrix.bin      (4, 4)   view matrix @ projection matrix
    tan_fovs.bin                (2,)     tan of the full horizontal and vertical fields of view
    intrinsics.bin              (6,)     fx, fy, cx, cy, near, far
    camera_rays_xfm.bin         (H*W, 3) camera_rays @ R

`read_fixture` loads a directory; `write_fixture` writes one in the same layout;
`render_fixture` renders a loaded fixture through `rasterize` (the reference binding's
argument list) and reports the error against the stored images, as tests/test_jax.py does.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict

import numpy as np

FLOAT_FILES = {
    "zs": (-1,), "alphas": (-1,), "means": (-1, 3), "quats": (-1, 4), "scales": (-1, 3), "covs": (-1, 3, 3),
    "precs": (-1, 3, 3), "weights": (-1,), "camera_rays": (-1, 3), "camera_rot": (3, 3), "camera_trans": (3,),
    "camera_view_matrix": (4, 4), "camera_proj_matrix": (4, 4), "tan_fovs": (2,), "intrinsics": (6,),
    "camera_rays_xfm": (-1, 3),
}


@dataclasses.dataclass
class Fixture:
    width: int
    height: int
    n: int
    arrays: Dict[str, np.ndarray]

    def __getitem__(self, k: str) -> np.ndarray:
        return self.arrays[k]

    @property
    def intrinsics(self):
        fx, fy, cx, cy, near, far = [float(v) for v in self.arrays["intrinsics"]]
        return fx, fy, cx, cy, near, far


def folder_name(width: int, height: int, n: int) -> str:
    """The reference's directory naming, `<W>_<H>_<N>` (ycb_fmb_rotscale.py, test_jax.py:17)."""
    return f"{int(width)}_{int(height)}_{int(n)}"


def read_fixture(directory: str) -> Fixture:
    """Load every `.bin` file present; `width_height_gaussians.bin` is required."""
    whn = np.fromfile(os.path.join(directory, "width_height_gaussians.bin"), dtype=np.int32)
    if whn.size != 3:
        raise ValueError(f"{directory}: width_height_gaussians.bin must hold 3 int32 values")
    W, H, N = (int(v) for v in whn)
    arrays = {}
    for name, shape in FLOAT_FILES.items():
        path = os.path.join(directory, name + ".bin")
        if not os.path.exists(path):
            continue
        a = np.fromfile(path, dtype=np.float32)
        arr = a.reshape(shape)
        lead = {"zs": H * W, "alphas": H * W, "camera_rays": H * W, "camera_rays_xfm": H * W}.get(name, N)
        if shape[0] == -1 and arr.shape[0] != lead:
            raise ValueError(f"{directory}: {name}.bin holds {arr.shape[0]} rows, expected {lead}")
        arrays[name] = arr
    return Fixture(W, H, N, arrays)


def projection_matrix(width, height, fx, fy, near, far) -> np.ndarray:
    """Perspective matrix of utils/utils.py:212-250 (row-vector convention, transposed)."""
    tx = width / (2.0 * fx)           # tan of half the horizontal field of view
    ty = height / (2.0 * fy)
    top, right = ty * near, tx * near
    P = np.zeros((4, 4))
    P[0, 0] = 2.0 * near / (2.0 * right)
    P[1, 1] = 2.0 * near / (2.0 * top)
    P[2, 2] = far / (far - near)
    P[2, 3] = -(far * near) / (far - near)
    P[3, 2] = 1.0
    return P.T


def write_fixture(directory: str, means, quats, scales, weights_log, R, t, width, height, fx, fy, cx, cy,
                  near: float = 0.01, far: float = 10.0, depth=None, alpha=None) -> str:
    """Write a scene and camera (and optionally its rendered depth/alpha) in the reference layout."""
    from . import quat_to_rot
    os.makedirs(directory, exist_ok=True)
    means = np.asarray(means, np.float64).reshape(-1, 3)
    quats = np.asarray(quats, np.float64).reshape(-1, 4)
    scales = np.asarray(scales, np.float64).reshape(-1, 3)
    N = means.shape[0]
    rots = np.stack([quat_to_rot(q) for q in quats]) if N else np.zeros((0, 3, 3))
    rs = rots * scales[:, None, :]
    cov = np.einsum("nij,nkj->nik", rs, rs)
    precs = np.linalg.cholesky(np.linalg.inv(cov)) if N else np.zeros((0, 3, 3))
    R = np.asarray(R, np.float64).reshape(3, 3)
    t = np.asarray(t, np.float64).reshape(3)
    ys, xs = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    rays = np.stack([(xs - cx) / fx, (ys - cy) / fy, np.ones_like(xs)], -1).reshape(-1, 3)
    cam4 = np.eye(4)
    cam4[:3, :3] = R
    cam4[:3, 3] = t
    view = np.linalg.inv(cam4).T
    proj = view @ projection_matrix(width, height, fx, fy, near, far)
    fovx = 2.0 * math.atan(width / 2.0 / fx)
    fovy = 2.0 * math.atan(height / 2.0 / fy)
    out = {
        "means": means, "quats": quats, "scales": scales, "covs": cov, "precs": precs,
        "weights": np.asarray(weights_log, np.float64).reshape(-1), "camera_rays": rays, "camera_rot": R,
        "camera_trans": t, "camera_view_matrix": view, "camera_proj_matrix": proj,
        "tan_fovs": np.array([math.tan(fovx), math.tan(fovy)]),
        "intrinsics": np.array([fx, fy, cx, cy, near, far]), "camera_rays_xfm": rays @ R,
        "zs": np.zeros(height * width) if depth is None else np.asarray(depth).reshape(-1),
        "alphas": np.zeros(height * width) if alpha is None else np.asarray(alpha).reshape(-1),
    }
    for name, arr in out.items():
        np.asarray(arr, dtype=np.float32).ravel().tofile(os.path.join(directory, name + ".bin"))
    np.asarray([width, height, N], dtype=np.int32).tofile(os.path.join(directory, "width_height_gaussians.bin"))
    return directory


def render_fixture(fx_: Fixture, device="cuda"):
    """Render a loaded fixture with `rasterize` (the reference binding's argument list).

    tests/test_jax.py:73-90 passes (width, height) into the binding's (image_height,
    image_width) slots; the call here passes them in the binding's own order.
    Returns (depth (H,W), alpha (H,W), max |depth - zs|, max |alpha - alphas|)."""
    from . import rasterize
    fx, fy, cx, cy, near, far = fx_.intrinsics
    d, a = rasterize(fx_["means"], fx_["scales"], fx_["quats"], fx_["weights"],
                     fx_["camera_rays"].reshape(fx_.height, fx_.width, 3), fx_["camera_rot"], fx_["camera_trans"],
                     fx_.height, fx_.width, fx, fy, cx, cy, near, far, device=device)
    d = d.cpu().numpy().reshape(-1)
    a = a.cpu().numpy().reshape(-1)
    ed = float(np.max(np.abs(d - fx_["zs"]))) if "zs" in fx_.arrays else math.nan
    ea = float(np.max(np.abs(a - fx_["alphas"]))) if "alphas" in fx_.arrays else math.nan
    return d.reshape(fx_.height, fx_.width), a.reshape(fx_.height, fx_.width), ed, ea
