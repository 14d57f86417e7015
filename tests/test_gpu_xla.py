"""The reference-shaped custom-call target `rasterize_gaussians_fwd`, driven exactly as XLA's
legacy GPU custom call drives it: a void** of 11 device pointers in the reference's order
(rasterize_main.cu:54-68) and the 20-byte opaque descriptor (rasterize_main.h:13-19).

Scales and quaternions become the precision factor on the device; the oracle here gets its
factor by an independent route (scipy rotation, the precision R diag(1/s^2) R^T and a
reversed Cholesky), and renders every given ray in float64 (oracle/fm_oracle.py)."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import fm_oracle

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
B2, B3 = 21.4, 2.66   # the reference custom call's fixed betas (rasterizer_impl.cu:196-197)


def _gaussians(rng, n):
    means = rng.uniform(-0.5, 0.5, size=(n, 3))
    quats = rng.normal(size=(n, 4))
    quats *= rng.uniform(0.5, 2.0, size=(n, 1))        # unnormalised: quat_to_rot divides by |q|^2
    scales = np.exp(rng.uniform(np.log(0.03), np.log(0.15), size=(n, 3)))
    wl = rng.normal(0.0, 0.5, size=n)
    wl = wl - np.log(np.sum(np.exp(wl)))
    return means, quats, scales, wl


def _precision_factor(quats, scales):
    """Upper-triangular U with U U^T = R diag(1/s^2) R^T, by a route independent of the
    library's (scipy rotation from the scalar-last quaternion; Cholesky of the reversed
    matrix, reversed back)."""
    R = Rotation.from_quat(np.roll(quats, -1, axis=1)).as_matrix()
    lam = np.einsum("nij,nj,nkj->nik", R, 1.0 / scales ** 2, R)
    J = np.eye(3)[::-1]
    L = np.linalg.cholesky(J @ lam @ J)
    U = J @ L @ J
    assert np.allclose(np.triu(U), U) and np.allclose(U @ np.transpose(U, (0, 2, 1)), lam, rtol=1e-10)
    return U


def _camera(rng, dist=2.5):
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    centre = dist * d
    z = -centre / np.linalg.norm(centre)
    up = np.array([0.0, 0.0, 1.0])
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z]), centre


def _grid_rays(H, W, fx, fy, cx, cy):
    ys, xs = np.mgrid[0:H, 0:W]
    return np.stack([(xs - cx) / fx, (ys - cy) / fy, np.ones_like(xs, dtype=np.float64)], -1).reshape(-1, 3)


def _call(fea, means, quats, scales, wl, rays, R, t, H, W, fx, fy, status=True, opaque=None):
    from fuzzy_ellipses_amd import xla
    f32 = lambda a: torch.tensor(np.asarray(a, dtype=np.float32), device=DEV).contiguous()
    n = len(wl)
    bufs = [f32(means), f32(wl), f32(scales), f32(quats), f32(rays), f32(R), f32(t), f32(np.eye(4)),
            f32(fea.projection_matrix(W, H, fx, fy, W / 2, H / 2, 0.01, 100.0)),
            torch.full((H, W), -7.0, device=DEV), torch.full((H, W), -7.0, device=DEV)]
    if opaque is None:
        opaque = xla.build_gaussian_rasterize_descriptor(H, W, n, np.tan(2 * np.arctan(W / 2 / fx)),
                                                         np.tan(2 * np.arctan(H / 2 / fy)))
    xla.call_custom_target(torch.cuda.current_stream().cuda_stream, [b.data_ptr() for b in bufs], opaque,
                           status=status)
    torch.cuda.synchronize()
    return bufs[9].cpu().numpy().ravel(), bufs[10].cpu().numpy().ravel()


def _errors(d, a, od, oa):
    ed = float(np.max(np.abs(d - od) / np.abs(od)))
    ea = float(np.max(np.abs(a - oa)))
    return ed, ea


@pytest.mark.parametrize("n, H, W", [(40, 64, 64), (300, 50, 70), (1024, 120, 160)])
def test_custom_call_matches_oracle_on_the_pinhole_grid(fea, n, H, W):
    rng = np.random.default_rng(100 + n)
    means, quats, scales, wl = _gaussians(rng, n)
    R, t = _camera(rng)
    fx = fy = float(W)
    rays = _grid_rays(H, W, fx, fy, W / 2, H / 2)
    d, a = _call(fea, means, quats, scales, wl, rays, R, t, H, W, fx, fy)
    U = _precision_factor(quats, scales)
    od, oa = fm_oracle.render_camera(means, U, wl, R, t, rays, B2, B3)
    ed, ea = _errors(d, a, od, oa)
    print(f"custom call n={n} {H}x{W}: depth rel {ed:.2e}, alpha abs {ea:.2e}")
    assert ed <= 1e-4 and ea <= 1e-4


def test_custom_call_matches_render_of_the_same_scene(fea, scenes_mod):
    """The custom call and the pinhole renderer agree on one scene (the host precision factor
    for the renderer), to the parity tolerance."""
    rng = np.random.default_rng(7)
    n, H, W = 500, 96, 128
    means, quats, scales, wl = _gaussians(rng, n)
    R, t = _camera(rng)
    fx = fy = 128.0
    d, a = _call(fea, means, quats, scales, wl, _grid_rays(H, W, fx, fy, W / 2, H / 2), R, t, H, W, fx, fy)
    cam = scenes_mod.Camera(R=R, t=t, H=H, W=W, fx=fx, fy=fy, cx=W / 2, cy=H / 2)
    rd, ra = fea.render_camera(means, fea.prec_from_scale_quat(scales, quats), wl, cam, B2, B3, device=DEV)
    rd, ra = rd.cpu().numpy().ravel(), ra.cpu().numpy().ravel()
    assert np.max(np.abs(d - rd) / np.abs(rd)) <= 1e-5 and np.max(np.abs(a - ra)) <= 1e-6


def test_custom_call_renders_arbitrary_rays(fea):
    """Rays that are not a pinhole grid (random camera-plane points, third component 1) and a
    shape that is not a multiple of the tile."""
    rng = np.random.default_rng(11)
    n, H, W = 200, 37, 53
    means, quats, scales, wl = _gaussians(rng, n)
    R, t = _camera(rng)
    xy = rng.uniform(-0.4, 0.4, size=(H * W, 2))
    rays = np.concatenate([xy, np.ones((H * W, 1))], 1)
    d, a = _call(fea, means, quats, scales, wl, rays, R, t, H, W, 50.0, 50.0)
    od, oa = fm_oracle.render_camera(means, _precision_factor(quats, scales), wl, R, t, rays, B2, B3)
    ed, ea = _errors(d, a, od, oa)
    print(f"custom call, random rays {H}x{W}: depth rel {ed:.2e}, alpha abs {ea:.2e}")
    assert ed <= 1e-4 and ea <= 1e-4


def test_custom_call_rays_not_on_the_camera_plane(fea):
    """A ray (x, y, z) with z != 1: alpha is the oracle's for that ray; depth is NaN."""
    rng = np.random.default_rng(12)
    n, H, W = 120, 32, 32
    means, quats, scales, wl = _gaussians(rng, n)
    R, t = _camera(rng)
    rays = _grid_rays(H, W, 32.0, 32.0, 16.0, 16.0)
    k = rng.choice(H * W, 40, replace=False)
    rays[k] *= rng.uniform(0.5, 2.0, size=(40, 1))
    d, a = _call(fea, means, quats, scales, wl, rays, R, t, H, W, 32.0, 32.0)
    od, oa = fm_oracle.render_camera(means, _precision_factor(quats, scales), wl, R, t, rays, B2, B3)
    flagged = rays[:, 2] != np.float32(1.0)
    assert np.all(np.isnan(d[flagged])) and np.all(np.isfinite(d[~flagged]))
    assert np.max(np.abs(a - oa)) <= 1e-4
    assert np.max(np.abs(d[~flagged] - od[~flagged]) / np.abs(od[~flagged])) <= 1e-4


def test_custom_call_zero_gaussians_and_errors(fea):
    from fuzzy_ellipses_amd import FeError, xla
    rng = np.random.default_rng(3)
    H, W = 20, 24
    R, t = _camera(rng)
    rays = _grid_rays(H, W, 24.0, 24.0, 12.0, 12.0)
    e3, e4 = np.zeros((0, 3)), np.zeros((0, 4))
    d, a = _call(fea, e3, e4, e3, np.zeros(0), rays, R, t, H, W, 24.0, 24.0)
    assert np.all(d == 0.0) and np.all(a == 0.0)        # the reference's zeroed outputs
    means, quats, scales, wl = _gaussians(rng, 10)
    with pytest.raises(FeError):
        _call(fea, means, quats, scales, wl, rays, R, t, H, W, 24.0, 24.0, opaque=b"\0" * 19)
    bad = xla.build_gaussian_rasterize_descriptor(H, W, -1, 1.0, 1.0)
    with pytest.raises(FeError):
        _call(fea, means, quats, scales, wl, rays, R, t, H, W, 24.0, 24.0, opaque=bad)
    # the void target itself: a failed call leaves NaN outputs
    d, a = _call(fea, means, quats, scales, wl, rays, R, t, H, W, 24.0, 24.0, status=False, opaque=bad)
    assert np.all(np.isnan(d)) and np.all(np.isnan(a))


def test_custom_call_repeated_calls_grow_and_reuse_scratch(fea):
    """Calls of growing and shrinking size on one stream give each its own oracle result
    (the library's scratch grows between them)."""
    for n, H, W, seed in [(30, 16, 16, 1), (400, 80, 96, 2), (60, 24, 40, 3), (400, 80, 96, 2)]:
        rng = np.random.default_rng(seed)
        means, quats, scales, wl = _gaussians(rng, n)
        R, t = _camera(rng)
        rays = _grid_rays(H, W, float(W), float(W), W / 2, H / 2)
        d, a = _call(fea, means, quats, scales, wl, rays, R, t, H, W, float(W), float(W))
        od, oa = fm_oracle.render_camera(means, _precision_factor(quats, scales), wl, R, t, rays, B2, B3)
        ed, ea = _errors(d, a, od, oa)
        assert ed <= 1e-4 and ea <= 1e-4, (n, H, W, ed, ea)


def test_rasterize_shim_matches_oracle(fea):
    """The reference binding's argument list (fuzzy_ellipses/__init__.py:189-236) through the
    custom call, with nan_to_num applied as the binding does."""
    rng = np.random.default_rng(21)
    n, H, W = 256, 60, 80
    means, quats, scales, wl = _gaussians(rng, n)
    R, t = _camera(rng)
    fx, fy, cx, cy = 80.0, 82.0, 39.5, 30.5
    rays = _grid_rays(H, W, fx, fy, cx, cy)
    d, a = fea.rasterize(means, scales, quats, wl, rays, R, t, H, W, fx, fy, cx, cy, 0.01, 100.0, device=DEV)
    assert d.shape == (H, W) and a.shape == (H, W)
    od, oa = fm_oracle.render_camera(means, _precision_factor(quats, scales), wl, R, t, rays, B2, B3)
    od, oa = fm_oracle.nan_to_num_outputs(od, oa)
    ed, ea = _errors(d.cpu().numpy().ravel(), a.cpu().numpy().ravel(), od, oa)
    assert ed <= 1e-4 and ea <= 1e-4
