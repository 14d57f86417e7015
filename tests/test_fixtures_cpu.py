"""The reference's `<W>_<H>_<N>/*.bin` scene format (baseline/ycb_fmb_rotscale.py:239-279,
tests/test_jax.py:19-49): write/read round trip, layouts and derived arrays."""
import math
import os

import numpy as np


def _scene(n=5, seed=0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-0.2, 0.2, size=(n, 3)) + np.array([0, 0, 1.0])
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    s = np.exp(rng.uniform(np.log(0.02), np.log(0.08), size=(n, 3)))
    w = np.log(np.ones(n) / n) + np.log(80.0)
    return means, q, s, w


def test_round_trip_and_layout(fea, tmp_path):
    from fuzzy_ellipses_amd import fixtures, quat_to_rot
    means, q, s, w = _scene()
    W, H = 40, 30
    d = tmp_path / fixtures.folder_name(W, H, 5)
    assert d.name == "40_30_5"
    depth = np.linspace(0, 1, H * W)
    fixtures.write_fixture(str(d), means, q, s, w, np.eye(3), np.zeros(3), W, H, 50.0, 52.0, 19.5, 14.5,
                           near=0.01, far=5.0, depth=depth, alpha=1 - depth)
    # raw little-endian float32 files without headers, int32 dimensions
    assert os.path.getsize(d / "means.bin") == 5 * 3 * 4
    assert os.path.getsize(d / "camera_rays.bin") == H * W * 3 * 4
    assert np.fromfile(d / "width_height_gaussians.bin", dtype=np.int32).tolist() == [W, H, 5]
    fx = fixtures.read_fixture(str(d))
    assert (fx.width, fx.height, fx.n) == (W, H, 5)
    assert np.array_equal(fx["means"], means.astype(np.float32))
    assert np.array_equal(fx["zs"], depth.astype(np.float32))
    # pixel rays: ((x - cx)/fx, (y - cy)/fy, 1), row-major y*W + x
    k = 7 * W + 11
    assert np.allclose(fx["camera_rays"][k], [(11 - 19.5) / 50.0, (7 - 14.5) / 52.0, 1.0])
    # covariances and precision Cholesky factors agree with the scales and quaternions
    R0 = quat_to_rot(q[0])
    cov0 = R0 @ np.diag(s[0] ** 2) @ R0.T
    assert np.allclose(fx["covs"][0], cov0, rtol=1e-5, atol=1e-9)
    L = fx["precs"][0].astype(np.float64)
    assert np.allclose(np.triu(L, 1), 0) and np.allclose(L @ L.T, np.linalg.inv(cov0), rtol=1e-4)
    # field of view and intrinsics
    assert np.allclose(fx["tan_fovs"], [math.tan(2 * math.atan(W / 2 / 50.0)), math.tan(2 * math.atan(H / 2 / 52.0))])
    assert np.allclose(fx["intrinsics"], [50.0, 52.0, 19.5, 14.5, 0.01, 5.0])


def test_projection_matrix_maps_frustum_to_clip(fea):
    """The perspective matrix sends the near plane to depth 0 and the far plane to 1 after
    the divide, and the image corners to x, y = +-1."""
    from fuzzy_ellipses_amd import fixtures
    W, H, fx, fy, near, far = 64, 48, 60.0, 62.0, 0.1, 7.0
    P = fixtures.projection_matrix(W, H, fx, fy, near, far)
    for z, want in ((near, 0.0), (far, 1.0)):
        x = np.array([W / 2 / fx * z, H / 2 / fy * z, z, 1.0]) @ P
        assert np.allclose(x[:3] / x[3], [1.0, 1.0, want])


def test_reader_rejects_inconsistent_sizes(fea, tmp_path):
    import pytest
    from fuzzy_ellipses_amd import fixtures
    means, q, s, w = _scene()
    d = tmp_path / "f"
    fixtures.write_fixture(str(d), means, q, s, w, np.eye(3), np.zeros(3), 8, 6, 8.0, 8.0, 4.0, 3.0)
    np.zeros(4, np.float32).tofile(d / "weights.bin")
    with pytest.raises(ValueError):
        fixtures.read_fixture(str(d))
