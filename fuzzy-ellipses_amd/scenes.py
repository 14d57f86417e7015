"""Seeded synthetic scenes and cameras for the fuzzy-metaball rasterizer.

Conventions follow the reference renderer:

* a camera ray for pixel (x, y) is r_cam = ((x - cx)/fx, (y - cy)/fy, 1), pixels in
  row-major order y*W + x (baseline/ycb_fmb_rotscale.py:175-188);
* the world ray is the row vector r_cam @ R and every ray starts at the camera centre t
  (baseline/fm_render.py:113-120, utils/sfs_utils.py:187-192), so the rows of R are the
  camera axes expressed in world coordinates and depth along a ray is camera z;
* the precision factor is U = (C^-1)^T with C = chol(Sigma) lower, so U is upper
  triangular and U U^T = Sigma^-1; the renderer only reads triu(U)
  (baseline/fm_render.py:14, utils/sfs_utils.py:207-213 uses the transposed factor).

Everything is float64 numpy; callers cast to float32 for the GPU path.  Random draws use
numpy's PCG64 generator in a fixed order (means, rotations, scales, weights) so a seed
fixes a scene.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

BETA2 = 21.4
BETA3 = 2.66


@dataclasses.dataclass
class Camera:
    R: np.ndarray          # (3, 3) rows = camera x, y, z axes in world coordinates
    t: np.ndarray          # (3,) camera centre (ray origin)
    H: int
    W: int
    fx: float
    fy: float
    cx: float
    cy: float

    @property
    def n_pixels(self) -> int:
        return self.H * self.W


@dataclasses.dataclass
class Scene:
    means: np.ndarray        # (N, 3)
    prec: np.ndarray         # (N, 3, 3) upper-triangular precision factor U
    weights_log: np.ndarray  # (N,)
    cameras: List[Camera]
    beta2: float = BETA2
    beta3: float = BETA3

    @property
    def n(self) -> int:
        return int(self.means.shape[0])

    @property
    def prec_upper6(self) -> np.ndarray:
        """(N, 6) packed upper triangle [u00, u01, u02, u11, u12, u22]."""
        U = self.prec
        return np.stack([U[:, 0, 0], U[:, 0, 1], U[:, 0, 2],
                         U[:, 1, 1], U[:, 1, 2], U[:, 2, 2]], axis=1)


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform rotations on SO(3) from normalised 4-D Gaussian quaternions (w, x, y, z)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def precision_factor(rot: np.ndarray, scales: np.ndarray) -> np.ndarray:
    """Upper factor U with U U^T = (R diag(s^2) R^T)^-1, batched over the first axis."""
    cov = np.einsum("nij,nj,nkj->nik", rot, scales ** 2, rot)
    C = np.linalg.cholesky(cov)
    Cinv = np.linalg.inv(C)
    U = np.transpose(Cinv, (0, 2, 1))
    return np.triu(U)


def logsumexp(x: np.ndarray) -> float:
    m = np.max(x)
    return float(m + np.log(np.sum(np.exp(x - m))))


def random_gaussians(rng: np.random.Generator, n: int, shell: bool = False):
    """Gaussian mixture generator.

    means: uniform in [-0.5, 0.5]^3, or (shell=True) n/|n| + N(0, 0.02^2) on the unit
    sphere; rotations uniform; per-axis scales log-uniform in [0.03, 0.15]; log-weights
    N(0, sigma=0.5) then log-softmaxed so sum(exp(w)) = 1.
    """
    if shell:
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        means = d + rng.normal(scale=0.02, size=(n, 3))
    else:
        means = rng.uniform(-0.5, 0.5, size=(n, 3))
    rot = random_rotations(rng, n)
    scales = np.exp(rng.uniform(np.log(0.03), np.log(0.15), size=(n, 3)))
    w = rng.normal(scale=0.5, size=n)
    w = w - logsumexp(w)
    return means, precision_factor(rot, scales), w


def look_at(centre: np.ndarray) -> np.ndarray:
    """Rotation whose rows are the camera axes of a camera at `centre` looking at the origin."""
    z = -centre / np.linalg.norm(centre)
    up = np.array([0.0, 0.0, 1.0])
    if abs(float(np.dot(up, z))) > 0.99:
        up = np.array([0.0, 1.0, 0.0])
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=0)


def fibonacci_dirs(n: int) -> np.ndarray:
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = (np.arange(n)) * math.pi * (3.0 - math.sqrt(5.0))
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def random_dir(rng: np.random.Generator) -> np.ndarray:
    d = rng.normal(size=3)
    return d / np.linalg.norm(d)


def make_camera(centre, H, W, f) -> Camera:
    centre = np.asarray(centre, dtype=np.float64)
    return Camera(R=look_at(centre), t=centre, H=int(H), W=int(W), fx=float(f), fy=float(f),
                  cx=W / 2.0, cy=H / 2.0)


def camera_rays(cam: Camera) -> np.ndarray:
    """(H*W, 3) camera-frame rays ((x-cx)/fx, (y-cy)/fy, 1), row-major y*W + x."""
    ys, xs = np.meshgrid(np.arange(cam.H, dtype=np.float64), np.arange(cam.W, dtype=np.float64),
                         indexing="ij")
    r = np.stack([(xs - cam.cx) / cam.fx, (ys - cam.cy) / cam.fy, np.ones_like(xs)], axis=-1)
    return r.reshape(-1, 3)


# ----------------------------------------------------------------------------------------
# The five workloads.  Seeds default to the workload index.
# ----------------------------------------------------------------------------------------

def single_view_scene(seed: int = 0, n: int = 40, H: int = 64, W: int = 64, f: float = 64.0,
                      dist: float = 2.5) -> Scene:
    """N=40 Gaussians in the unit cube, one 64x64 camera (f=64) at distance 2.5."""
    rng = np.random.default_rng(seed)
    means, prec, w = random_gaussians(rng, n)
    cam = make_camera(dist * random_dir(rng), H, W, f)
    return Scene(means, prec, w, [cam])


def shell_scene(seed: int = 1, n: int = 4096, n_views: int = 16, H: int = 480, W: int = 640,
                f: float = 600.0, radius: float = 3.0) -> Scene:
    """N=4096 Gaussians on a unit sphere shell, 16 fibonacci views at radius 3, 480x640, f=600."""
    rng = np.random.default_rng(seed)
    means, prec, w = random_gaussians(rng, n, shell=True)
    cams = [make_camera(radius * d, H, W, f) for d in fibonacci_dirs(n_views)]
    return Scene(means, prec, w, cams)


RESOLUTIONS = [(64, 64), (128, 128), (240, 320), (480, 640)]


def batched_scenes(seed: int = 2, n_scenes: int = 256, n_lo: int = 16, n_hi: int = 2048,
                   resolutions=RESOLUTIONS, max_views: int = 8) -> List[Scene]:
    """256 scenes, N ~ U{16..2048}, resolution from `resolutions`, 1-8 views at radius U[2.5, 3.5].

    Focal length f = W * 600/640 (same field of view as the 480x640, f=600 camera).
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_scenes):
        n = int(rng.integers(n_lo, n_hi + 1))
        means, prec, w = random_gaussians(rng, n)
        H, W = resolutions[int(rng.integers(len(resolutions)))]
        nv = int(rng.integers(1, max_views + 1))
        cams = []
        for _v in range(nv):
            rad = float(rng.uniform(2.5, 3.5))
            cams.append(make_camera(rad * random_dir(rng), H, W, W * 600.0 / 640.0))
        out.append(Scene(means, prec, w, cams))
    return out


def silhouette_problem(seed: int = 3, n: int = 1024, n_views: int = 16, H: int = 240, W: int = 320,
                       f: float = 300.0, radius: float = 3.0, weight_shift: Optional[float] = None):
    """Shape-from-silhouette workload: an initial scene and a target scene (seed+1) sharing
    16 fibonacci views at 240x320, focal length f = W * 600/640.

    With log-softmaxed weights a mixture's alpha stays near sum(exp(w)) / n (about 0.02 here),
    so no pixel would pass the 0.5 silhouette threshold; both scenes' log-weights are
    therefore shifted by `weight_shift` (default ln n: weights N(0, 0.5) before
    normalisation), which makes the silhouettes opaque as in the reference notebook's fits.
    """
    shift = math.log(n) if weight_shift is None else float(weight_shift)
    rng = np.random.default_rng(seed)
    means, prec, w = random_gaussians(rng, n)
    rng_t = np.random.default_rng(seed + 1)
    tm, tp, tw = random_gaussians(rng_t, n)
    cams = [make_camera(radius * d, H, W, f) for d in fibonacci_dirs(n_views)]
    return Scene(means, prec, w + shift, cams), Scene(tm, tp, tw + shift, cams)


SERVE_RESOLUTIONS = [(128, 128), (240, 320), (480, 640)]


def resident_scenes(seed: int = 4, n_scenes: int = 64, n_lo: int = 64, n_hi: int = 4096) -> List[Scene]:
    """64 resident scenes for serving, N ~ U{64..4096}; cameras are attached per request."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_scenes):
        n = int(rng.integers(n_lo, n_hi + 1))
        means, prec, w = random_gaussians(rng, n)
        out.append(Scene(means, prec, w, []))
    return out
