"""A scene written in the reference's `.bin` fixture layout, read back and rendered through
`rasterize` (the reference binding's argument list), against the FP64 oracle's images stored
in the fixture, as tests/test_jax.py does with the reference's own fixtures."""
import numpy as np
import pytest

from oracle import ref

pytestmark = pytest.mark.gpu


def test_fixture_renders_to_stored_images(fea, scenes_mod, tmp_path):
    from fuzzy_ellipses_amd import fixtures, prec_from_scale_quat
    rng = np.random.default_rng(12)
    N, W, H = 150, 64, 48
    # the reference fixture writer's scene: identity rotations, isotropic scale 1/sqrt(5e3),
    # equal weights scaled by 80 (ycb_fmb_rotscale.py:145-156); points in front of the camera
    means = rng.uniform(-0.08, 0.08, size=(N, 3)) + np.array([0.0, 0.0, 0.6])
    quats = np.tile([1.0, 0.0, 0.0, 0.0], (N, 1))
    scales = np.full((N, 3), 1.0 / np.sqrt(5e3))
    wl = np.log(np.ones(N) / N) + np.log(80.0)
    fx, fy, cx, cy = 70.0, 71.0, 31.5, 23.5
    U = prec_from_scale_quat(scales, quats)
    cam = scenes_mod.Camera(R=np.eye(3), t=np.zeros(3), H=H, W=W, fx=fx, fy=fy, cx=cx, cy=cy)
    scene = scenes_mod.Scene(means, U, wl, [cam])
    d64, a64 = ref.render_f64(scene, cam)
    d = tmp_path / fixtures.folder_name(W, H, N)
    fixtures.write_fixture(str(d), means, quats, scales, wl, np.eye(3), np.zeros(3), W, H, fx, fy, cx, cy,
                           near=0.01, far=3.0, depth=d64, alpha=a64)
    fx_ = fixtures.read_fixture(str(d))
    depth, alpha, ed, ea = fixtures.render_fixture(fx_, device="cuda:0")
    print(f"fixture {d.name}: max |depth - zs| {ed:.2e} (depth max {np.max(d64):.3f}), max |alpha - alphas| {ea:.2e}")
    assert depth.shape == (H, W)
    assert ed <= 1e-4 * np.max(np.abs(d64)) and ea <= 1e-4
