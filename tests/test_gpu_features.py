"""The weights output of render_func_rays (baseline/fm_render.py:96-109) on the GPU: feature
images F = w @ f against the FP64 oracle's weights, the materialised (N, P) weights, and the
backward of F (geometry through the weights, and the features) against float64 autograd.
Tolerances: F within 1e-4 of max |f| (the depth's relative bound), weights within 2e-5
absolute, gradients within 1e-3 of each array's largest magnitude."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import fm_oracle as fo
from oracle import grad_oracle as go

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_F(scene, cam, f, pixels=None):
    rays = go.pinhole_rays(cam, pixels)
    W = fo.render_weights(rays, np.asarray(cam.t)[None], scene.means, scene.prec, scene.weights_log, scene.beta2,
                          scene.beta3)
    return W @ f, W


def _render(fea, scene, cams, f, mode="culled", tile=16, save=False):
    st = fea.SceneTensors.from_scene(scene, DEV, features=f)
    r = fea.Renderer(tile=tile, batch=64, mode=mode, device=DEV)
    outs = r.forward([st], [(0, c, scene.beta2, scene.beta3) for c in cams], save_for_backward=save)
    torch.cuda.synchronize()
    return r, st, outs


def _cube_inside(scenes_mod):
    rng = np.random.default_rng(7)
    means, prec, w = scenes_mod.random_gaussians(rng, 200)
    cam = scenes_mod.make_camera(np.array([0.12, -0.07, 0.18]), 64, 64, 40.0)
    return scenes_mod.Scene(means, prec, w, [cam])


def _shell_inside(scenes_mod):
    rng = np.random.default_rng(8)
    means, prec, w = scenes_mod.random_gaussians(rng, 1024, shell=True)
    cam = scenes_mod.make_camera(np.array([0.3, -0.2, 0.4]), 48, 48, 30.0)
    return scenes_mod.Scene(means, prec, w, [cam])


@pytest.mark.parametrize("mode,tile", [("dense", 16), ("culled", 8), ("culled", 16), ("culled", 32)])
def test_feature_image_matches_oracle(fea, scenes_mod, mode, tile):
    scene = scenes_mod.single_view_scene(seed=3)
    cam = scene.cameras[0]
    f = np.random.default_rng(1).normal(size=(scene.n, 5))      # 5 channels: a full and a partial block
    _, _, outs = _render(fea, scene, [cam], f, mode, tile)
    d, a, F = outs[0]
    assert F.shape == (cam.H, cam.W, 5)
    oF, _ = _oracle_F(scene, cam, f)
    err = float(np.max(np.abs(_np(F).reshape(-1, 5) - oF))) / np.abs(f).max()
    print(f"features {mode} tile {tile}: max |F - w f| / max|f| = {err:.2e}")
    assert err <= 1e-4


@pytest.mark.parametrize("which", ["camera_inside_cube", "floor_visible_shell", "zero_diagonal", "beta3_negative"])
def test_feature_image_special_views(fea, scenes_mod, which):
    """Masked pairs (camera inside the cloud), the visible 1e-20 floor, a degenerate record and a
    view outside the culling bounds' assumptions: F against the oracle."""
    if which == "camera_inside_cube":
        scene = _cube_inside(scenes_mod)
    elif which == "floor_visible_shell":
        scene = _shell_inside(scenes_mod)
    elif which == "zero_diagonal":
        scene = scenes_mod.single_view_scene(seed=14)
        prec = scene.prec.copy()
        prec[5, 1, 1] = 0.0
        scene = dataclasses.replace(scene, prec=prec)
    else:
        scene = dataclasses.replace(scenes_mod.single_view_scene(seed=14), beta3=-0.5)
    cam = scene.cameras[0]
    f = np.random.default_rng(2).uniform(0.0, 1.0, size=(scene.n, 3))
    for mode in ("dense", "culled"):
        _, _, outs = _render(fea, scene, [cam], f, mode)
        oF, _ = _oracle_F(scene, cam, f)
        err = float(np.max(np.abs(_np(outs[0][2]).reshape(-1, 3) - oF)))
        print(f"features {which} {mode}: max abs err {err:.2e}")
        assert err <= 1e-4


def test_feature_image_shell_view(fea, scenes_mod):
    """RGB features on the N=4096 shell at 480x640, every 4th row checked."""
    scene = scenes_mod.shell_scene(n_views=2)
    cam = scene.cameras[1]
    f = np.random.default_rng(3).uniform(0.0, 1.0, size=(scene.n, 3))
    _, _, outs = _render(fea, scene, [cam], f)
    pix = np.arange(cam.H * cam.W).reshape(cam.H, cam.W)[::4].ravel()
    oF, _ = _oracle_F(scene, cam, f, pixels=pix)
    err = float(np.max(np.abs(_np(outs[0][2]).reshape(-1, 3)[pix] - oF)))
    print(f"features shell 480x640: max abs err {err:.2e} over {pix.size} pixels")
    assert err <= 1e-4


def test_weights_materialised(fea, scenes_mod):
    """render_func_quat(..., return_weights=True) gives the reference's fourth output w (N, P)."""
    from scipy.spatial.transform import Rotation
    scene = scenes_mod.single_view_scene(seed=4)
    cam = scene.cameras[0]
    x, y, z, w = Rotation.from_matrix(cam.R).as_quat()
    d, a, _, W = fea.render_func_quat(scene.means, scene.prec, scene.weights_log,
                                      (cam.H, cam.W, cam.fx, cam.fy, cam.cx, cam.cy), [w, x, y, z], cam.t,
                                      scene.beta2, scene.beta3, return_weights=True, device=DEV)
    torch.cuda.synchronize()
    assert W.shape == (scene.n, cam.H * cam.W)
    _, oW = _oracle_F(scene, cam, np.eye(scene.n))
    err = float(np.max(np.abs(_np(W) - oW.T)))
    colsum = float(np.max(np.abs(_np(W).sum(0) - 1.0)))
    print(f"weights (N, P): max abs err {err:.2e}, columns sum to 1 within {colsum:.1e}")
    assert err <= 2e-5 and colsum <= 1e-5
    # depth is the weights' average of t* (fm_render.py:102)
    zs, _ = fo.pair_terms(go.pinhole_rays(cam), np.asarray(cam.t)[None], scene.means, scene.prec, scene.weights_log)
    np.testing.assert_allclose((_np(W).T * zs).sum(1), _np(d), rtol=1e-4)


def test_feature_image_ray_list_and_batch(fea, scenes_mod):
    """A ray-list view, and a batch mixing scenes with and without features: each view's images
    equal its own call bit for bit, features included."""
    from oracle import ref
    scene = scenes_mod.single_view_scene(seed=5, n=60, H=40, W=40, f=40.0)
    cam = scene.cameras[0]
    rng = np.random.default_rng(6)
    xy = rng.uniform(-0.45, 0.45, size=(900, 2)).astype(np.float32)
    rl = fea.RayList(cam.R, cam.t, torch.tensor(xy, device=DEV))
    f = rng.normal(size=(scene.n, 2))
    st = fea.SceneTensors.from_scene(scene, DEV, features=f)
    plain = fea.SceneTensors.from_scene(scenes_mod.single_view_scene(seed=9), DEV)
    views = [(0, rl, scene.beta2, scene.beta3), (1, scenes_mod.single_view_scene(seed=9).cameras[0], 21.4, 2.66),
             (0, cam, scene.beta2, scene.beta3)]
    r = fea.Renderer(device=DEV)
    outs = r.forward([st, plain], views)
    torch.cuda.synchronize()
    assert len(outs[0]) == 3 and len(outs[1]) == 2 and len(outs[2]) == 3
    rays = np.concatenate([xy.astype(np.float64), np.ones((len(xy), 1))], 1) @ np.asarray(cam.R)
    W = fo.render_weights(rays, np.asarray(cam.t)[None], scene.means, scene.prec, scene.weights_log, scene.beta2,
                          scene.beta3)
    assert float(np.max(np.abs(_np(outs[0][2]) - W @ f))) <= 1e-4
    for k, (si, c, b2, b3) in enumerate(views):
        r1 = fea.Renderer(device=DEV)
        o1 = r1.forward([[st, plain][si]], [(0, c, b2, b3)])[0]
        torch.cuda.synchronize()
        for x, y in zip(outs[k], o1):
            assert torch.equal(x, y)


@pytest.mark.parametrize("case", ["single", "camera_inside_cube", "floor_visible_shell", "zero_diagonal",
                                  "beta3_negative"])
@pytest.mark.parametrize("mode", ["dense", "culled"])
def test_feature_backward(fea, scenes_mod, case, mode):
    """Gradients of <gD, depth> + <gA, alpha> + <gF, F> with respect to means, precision factor,
    log-weights and features against float64 autograd of the restatement."""
    if case == "single":
        scene = scenes_mod.single_view_scene(seed=3)
    elif case == "camera_inside_cube":
        scene = _cube_inside(scenes_mod)
    elif case == "floor_visible_shell":
        scene = _shell_inside(scenes_mod)
    elif case == "zero_diagonal":
        scene = scenes_mod.single_view_scene(seed=14)
        prec = scene.prec.copy()
        prec[5, 1, 1] = 0.0
        scene = dataclasses.replace(scene, prec=prec)
    else:
        scene = dataclasses.replace(scenes_mod.single_view_scene(seed=14), beta3=-0.5)
    cam = scene.cameras[0]
    rng = np.random.default_rng(41)
    C = 6
    f = rng.uniform(-1.0, 1.0, size=(scene.n, C))
    P = cam.H * cam.W
    gd = rng.normal(size=P).astype(np.float32)
    ga = rng.normal(size=P).astype(np.float32)
    gf = rng.normal(size=(P, C)).astype(np.float32)
    r, st, outs = _render(fea, scene, [cam], f, mode, save=True)
    t = lambda x, shape: torch.tensor(x.reshape(shape), device=DEV)
    (gm, gp, gw, gF), = r.backward([t(gd, (cam.H, cam.W))], [t(ga, (cam.H, cam.W))],
                                   d_features=[t(gf, (cam.H, cam.W, C))])
    torch.cuda.synchronize()
    om, op, ow, of = go.feature_grads(scene, np.asarray(f, np.float32).astype(np.float64), [cam],
                                      [gd.astype(np.float64)], [ga.astype(np.float64)], [gf.astype(np.float64)])
    errs = []
    for got, want in ((_np(gm), om), (np.triu(_np(gp)), op), (_np(gw), ow), (_np(gF), of)):
        errs.append(float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
    print(f"feature backward {case} {mode}: means {errs[0]:.2e} prec {errs[1]:.2e} w {errs[2]:.2e} "
          f"features {errs[3]:.2e}")
    assert max(errs) <= 1e-3
