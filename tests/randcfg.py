"""Random small render configurations for the property tests (forward and backward parity).

A configuration is drawn from a seed (numpy PCG64): scene kind (unit cube, unit-sphere shell,
or a tight cluster), Gaussian count (log-uniform), a scale factor on every Gaussian (0.3 to 4
times the generator's 0.03-0.15, log-uniform), log-weight spread and shift, the betas (beta3
log-uniform), tile size, LDS batch depth, dense or culled membership, and one to three pinhole
views, each with its own distance to the origin (0.05 to 5, log-uniform: from inside the cloud
to far outside it), size, focal lengths (5 to 300 pixels, log-uniform; aspect 0.5-2) and
principal point (up to 30% off centre).  Every array the renderer receives is rounded to
float32 first and the oracle is given the rounded values, so the comparisons measure the
renderer's arithmetic, not input rounding.
"""
from __future__ import annotations

import numpy as np


def _logu(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def draw(seed, max_n=300, max_h=72, max_w=96, max_views=3):
    """The configuration of `seed` (a dict of plain values)."""
    rng = np.random.default_rng(seed)
    cfg = {
        "seed": int(seed),
        "kind": str(rng.choice(["cube", "shell", "cluster"])),
        "n": int(_logu(rng, 1, max_n + 1)) if max_n > 1 else 1,
        "scale": _logu(rng, 0.3, 4.0),
        "w_sigma": float(rng.choice([0.0, 0.5, 2.0, 5.0])),
        "w_shift": float(rng.uniform(-10.0, 10.0)),
        "beta2": float(rng.uniform(0.0, 40.0)),
        "beta3": _logu(rng, 0.1, 10.0),
        "tile": int(rng.choice([8, 16, 32])),
        "batch": int(rng.choice([32, 64, 128, 256])),
        "mode": str(rng.choice(["dense", "culled"])),
        "views": [],
    }
    cfg["n"] = min(cfg["n"], max_n)
    for _ in range(int(rng.integers(1, max_views + 1))):
        cfg["views"].append({"dist": _logu(rng, 0.05, 5.0), "H": int(rng.integers(1, max_h + 1)),
                             "W": int(rng.integers(1, max_w + 1)), "f": _logu(rng, 5.0, 300.0),
                             "aspect": _logu(rng, 0.5, 2.0), "cx_off": float(rng.uniform(-0.3, 0.3)),
                             "cy_off": float(rng.uniform(-0.3, 0.3))})
    return cfg


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def build(cfg, sc):
    """(scene, cameras) of a drawn configuration; `sc` is the package's scenes module."""
    rng = np.random.default_rng([cfg["seed"], 1])
    n = cfg["n"]
    means, prec, _ = sc.random_gaussians(rng, n, shell=cfg["kind"] == "shell")
    if cfg["kind"] == "cluster":
        means = means * 0.1 + rng.uniform(-0.5, 0.5, size=3)
    prec = prec / cfg["scale"]
    w = rng.normal(scale=cfg["w_sigma"], size=n) if cfg["w_sigma"] > 0 else np.zeros(n)
    w = w - sc.logsumexp(w) + cfg["w_shift"]
    cams = []
    for v in cfg["views"]:
        cam = sc.make_camera(v["dist"] * sc.random_dir(rng), v["H"], v["W"], v["f"])
        cam.fy = v["f"] * v["aspect"]
        cam.cx = v["W"] / 2.0 + v["cx_off"] * v["W"]
        cam.cy = v["H"] / 2.0 + v["cy_off"] * v["H"]
        cam.R, cam.t = _f32(cam.R), _f32(cam.t)
        cam.fx, cam.fy, cam.cx, cam.cy = (float(np.float32(x)) for x in (cam.fx, cam.fy, cam.cx, cam.cy))
        cams.append(cam)
    scene = sc.Scene(_f32(means), _f32(prec), _f32(w), cams, beta2=float(np.float32(cfg["beta2"])),
                     beta3=float(np.float32(cfg["beta3"])))
    return scene, cams


def ray_lists(cfg, cams, max_rays=3000):
    """One list of camera-plane rays (m, 2) float32 per camera: sub-pixel positions over the
    image, or over three times its extent (rays far off the view), in Morton order or in
    random order; from the configuration's seed."""
    from fuzzy_ellipses_amd.rays import _morton_order
    rng = np.random.default_rng([cfg["seed"], 7])
    out = []
    for cam in cams:
        m = int(rng.integers(1, max_rays + 1))
        wide = bool(rng.integers(0, 2))
        lo_x, hi_x = (-1.5 * cam.W, 2.5 * cam.W) if wide else (-0.5, cam.W - 0.5)
        lo_y, hi_y = (-1.5 * cam.H, 2.5 * cam.H) if wide else (-0.5, cam.H - 0.5)
        x, y = rng.uniform(lo_x, hi_x, m), rng.uniform(lo_y, hi_y, m)
        xy = np.stack([(x - cam.cx) / cam.fx, (y - cam.cy) / cam.fy], 1).astype(np.float32)
        if rng.integers(0, 2):
            xy = xy[_morton_order(xy[:, 0].astype(np.float64), xy[:, 1].astype(np.float64))]
        out.append(np.ascontiguousarray(xy))
    return out
