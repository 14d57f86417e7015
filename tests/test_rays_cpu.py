"""Multi-pose ray batches on the CPU: the oracle's render_func_idx_quattrans and flow
(restating baseline/fm_render.py:151-224), pinned by special cases that reduce to the
already pinned single-pose render and by closed forms of the flow; and the ray-list tile
membership restatement, checked for conservativeness against brute force."""
import numpy as np
import pytest

from oracle import fm_oracle as fo
from oracle import ref

LOG2E = 1.4426950408889634


def _quat_of(R):
    """A scalar-first quaternion of rotation R (Shepperd's method)."""
    from scipy.spatial.transform import Rotation
    x, y, z, w = Rotation.from_matrix(R).as_quat()
    return np.array([w, x, y, z])


def test_single_pose_equals_camera_render(scenes_mod):
    scene = scenes_mod.single_view_scene(seed=3, n=30, H=20, W=24, f=24.0)
    cam = scene.cameras[0]
    invF = 1.0 / cam.fx
    ys, xs = np.divmod(np.arange(cam.H * cam.W), cam.W)
    pp = np.stack([xs - cam.cx, ys - cam.cy, np.ones_like(xs), np.zeros_like(xs)], 1).astype(np.float64)
    poses = np.concatenate([_quat_of(cam.R.T), cam.t])[None]
    R = fo.quat_to_rot(poses[0, :4])
    d1, a1 = fo.render_func_idx_quattrans(scene.means, scene.prec, scene.weights_log, pp, invF, poses,
                                          scene.beta2, scene.beta3)
    d2, a2 = fo.render_camera(scene.means, scene.prec, scene.weights_log, R, cam.t,
                              scenes_mod.camera_rays(cam), scene.beta2, scene.beta3)
    assert np.allclose(d1, d2, rtol=1e-12, atol=1e-12) and np.allclose(a1, a2, rtol=1e-12, atol=1e-14)


def test_flow_closed_forms():
    rng = np.random.default_rng(0)
    invF = 1 / 50.0
    pp = np.stack([rng.uniform(-20, 20, 40), rng.uniform(-15, 15, 40), np.ones(40), np.zeros(40)], 1)
    pp[20:, 3] = 1
    depth = rng.uniform(1.0, 3.0, 40)
    ident = np.array([1.0, 0, 0, 0])
    # identical poses: no motion
    poses = np.stack([np.concatenate([ident, [0.1, 0.2, 0.3]])] * 2)
    fn, fp = fo.idx_quattrans_flow(pp, invF, poses, depth)
    assert np.abs(fn).max() < 1e-9 and np.abs(fp).max() < 1e-9
    # next camera translated by dx along its x axis: flow_x = dx / (depth invF), flow_y = 0
    dx = 0.05
    poses = np.stack([np.concatenate([ident, [0, 0, 0]]), np.concatenate([ident, [dx, 0, 0]])])
    fn, fp = fo.idx_quattrans_flow(pp, invF, poses, depth)
    first = pp[:, 3] == 0
    assert np.allclose(fn[first, 0], dx / (depth[first] * invF)) and np.allclose(fn[first, 1], 0)
    assert np.allclose(fp[first], 0)                         # previous of pose 0 is pose 0
    assert np.allclose(fp[~first, 0], -dx / (depth[~first] * invF))
    assert np.allclose(fn[~first], 0)                        # next of the last pose is itself
    # a rotation about the camera centre moves pixels independently of depth
    from scipy.spatial.transform import Rotation
    q = Rotation.from_rotvec([0.0, 0.05, 0.02]).as_quat()
    poses = np.stack([np.concatenate([ident, [0, 0, 0]]), np.concatenate([[q[3], q[0], q[1], q[2]], [0, 0, 0]])])
    f1, _ = fo.idx_quattrans_flow(pp, invF, poses, depth)
    f2, _ = fo.idx_quattrans_flow(pp, invF, poses, 2.5 * depth)
    assert np.allclose(f1, f2, atol=1e-9) and np.abs(f1[first]).max() > 0.1


def test_ray_rects_cover_their_rays():
    rng = np.random.default_rng(1)
    xy = np.concatenate([rng.normal(size=(1000, 2)) * 0.3, rng.uniform(-1e3, 1e3, size=(37, 2))]).astype(np.float32)
    ts = 8
    rects = ref.ray_rects(xy, ts).astype(np.float64)
    for t, r in enumerate(rects):
        seg = xy[t * 64:(t + 1) * 64].astype(np.float64)
        assert np.all(np.abs(seg[:, 0] - r[0]) <= r[2]) and np.all(np.abs(seg[:, 1] - r[1]) <= r[3])
        ext = (seg.max(0) - seg.min(0)) / 2
        assert r[2] <= ext[0] * (1 + 1e-6) + np.abs(seg[:, 0]).max() * 2e-6


@pytest.mark.parametrize("ts", [8, 16])
def test_ray_list_membership_is_conservative(scenes_mod, ts):
    """Phase 1 lists every (ray, Gaussian) pair the alpha bound needs, for rays drawn at random
    sub-pixel positions and grouped along a Z curve as the host layer does."""
    from fuzzy_ellipses_amd.rays import _morton_order
    scene = scenes_mod.single_view_scene(seed=5, n=50, H=64, W=64, f=64.0)
    cam = scene.cameras[0]
    rng = np.random.default_rng(2)
    m = 700
    x = rng.uniform(-0.5, cam.W - 0.5, m)
    y = rng.uniform(-0.5, cam.H - 0.5, m)
    o = _morton_order(x, y)
    xy = np.stack([(x[o] - cam.cx) / cam.fx, (y[o] - cam.cy) / cam.fy], 1).astype(np.float32)
    rays = np.concatenate([xy.astype(np.float64), np.ones((m, 1))], 1) @ cam.R.astype(np.float32).astype(np.float64)
    zs, d2 = fo.pair_terms(rays, np.repeat(cam.t[None], m, 0), scene.means, scene.prec, scene.weights_log)
    c_alpha, _, _ = ref.cull_constants(scene.n)
    counts, lst = ref.tiles_rays(scene, cam.R, cam.t, xy, ts)
    T = len(counts)
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    member = np.zeros((T, scene.n), bool)
    for t in range(T):
        member[t, lst[off[t]:off[t + 1]]] = True
    tile = np.arange(m) // (ts * ts)
    need = d2 >= -c_alpha + 0.5
    assert np.all(member[tile][need]), "ray-list phase 1 misses a needed pair"
    assert member.mean() < 1.0
