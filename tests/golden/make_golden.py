"""Writes tests/golden/single_view_seed0.npz: inputs and FP64 oracle outputs of the
40-Gaussian, 64x64 single-view scene (seed 0).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from conftest import load_package  # noqa: E402

load_package()
import fuzzy_ellipses_amd.scenes as sc  # noqa: E402
from oracle import fm_oracle as fo  # noqa: E402

scene = sc.single_view_scene()
cam = scene.cameras[0]
d, a = fo.render_camera(scene.means, scene.prec, scene.weights_log, cam.R, cam.t, sc.camera_rays(cam),
                        scene.beta2, scene.beta3)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "single_view_seed0.npz")
np.savez(out, means=scene.means, prec=scene.prec, weights_log=scene.weights_log, R=cam.R, t=cam.t,
         intrinsics=np.array([cam.H, cam.W, cam.fx, cam.fy, cam.cx, cam.cy]),
         beta=np.array([scene.beta2, scene.beta3]), depth=d, alpha=a)
print(out)
