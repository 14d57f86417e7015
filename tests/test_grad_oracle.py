"""Pins the float64 autograd gradient oracle (oracle/grad_oracle.py) with central finite
differences of the FP64 forward oracle (eps = 1e-4), on random parameters."""
import numpy as np

from oracle import fm_oracle as fo
from oracle import grad_oracle as go


def _loss(scene, cam, gd, ga, means, prec, w):
    d, a = fo.render_camera(means, prec, w, cam.R, cam.t, go.pinhole_rays(cam) @ np.asarray(cam.R).T,
                            scene.beta2, scene.beta3)
    return float(np.sum(d * gd) + np.sum(a * ga))


def test_forward_matches_numpy_oracle(scenes_mod):
    import torch
    scene = scenes_mod.single_view_scene(seed=12, H=16, W=16, f=16.0)
    cam = scene.cameras[0]
    rays = go.pinhole_rays(cam)
    d, a = go.render_rays_torch(torch.tensor(rays), torch.tensor(np.repeat(cam.t[None], len(rays), 0)),
                                torch.tensor(scene.means), torch.tensor(scene.prec),
                                torch.tensor(scene.weights_log), scene.beta2, scene.beta3)
    od, oa = fo.render_func_rays(rays, cam.t[None], scene.means, scene.prec, scene.weights_log,
                                 scene.beta2, scene.beta3)
    np.testing.assert_allclose(d.numpy(), od, rtol=1e-12)
    np.testing.assert_allclose(a.numpy(), oa, rtol=1e-12, atol=1e-16)


def test_autograd_matches_central_differences(scenes_mod):
    scene = scenes_mod.single_view_scene(seed=13, H=12, W=12, f=12.0)
    cam = scene.cameras[0]
    rng = np.random.default_rng(14)
    gd = rng.normal(size=cam.H * cam.W)
    ga = rng.normal(size=cam.H * cam.W) * 10
    dm, dp, dw = go.grads(scene, [cam], [gd], [ga])
    eps = 1e-4
    checked = 0
    for kind in ("means", "prec", "w"):
        for _ in range(8):
            m, p, w = scene.means.copy(), scene.prec.copy(), scene.weights_log.copy()
            j = int(rng.integers(scene.n))
            if kind == "means":
                i = int(rng.integers(3)); arr, g = m, dm[j, i]; idx = (j, i)
            elif kind == "prec":
                r, c = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)][int(rng.integers(6))]
                arr, g = p, dp[j, r, c]; idx = (j, r, c)
            else:
                arr, g = w, dw[j]; idx = (j,)
            base = arr[idx]
            arr[idx] = base + eps
            lp = _loss(scene, cam, gd, ga, m, p, w)
            arr[idx] = base - eps
            lm = _loss(scene, cam, gd, ga, m, p, w)
            arr[idx] = base
            fd = (lp - lm) / (2 * eps)
            # central differences carry O(eps^2) truncation error on this steep softmax
            assert abs(fd - g) <= 2e-4 * max(1.0, abs(g)), (kind, idx, fd, g)
            checked += 1
    assert checked == 24


def _feat_loss(scene, cam, f, gf, means, prec, w):
    rays = go.pinhole_rays(cam)
    W = fo.render_weights(rays, np.asarray(cam.t)[None], means, prec, w, scene.beta2, scene.beta3)
    return float(np.sum((W @ f) * gf))


def test_feature_autograd_matches_central_differences(scenes_mod):
    """The feature image F = w @ f of the autograd restatement equals the numpy oracle's, and
    its gradients (geometry through the weights, and the features themselves) match central
    differences of the numpy oracle."""
    import torch
    scene = scenes_mod.single_view_scene(seed=15, n=12, H=10, W=10, f=10.0)
    cam = scene.cameras[0]
    rng = np.random.default_rng(16)
    f = rng.normal(size=(scene.n, 3))
    gf = rng.normal(size=(cam.H * cam.W, 3))
    rays = go.pinhole_rays(cam)
    _, _, F = go.render_rays_torch(torch.tensor(rays), torch.tensor(np.repeat(cam.t[None], len(rays), 0)),
                                   torch.tensor(scene.means), torch.tensor(scene.prec),
                                   torch.tensor(scene.weights_log), scene.beta2, scene.beta3,
                                   features=torch.tensor(f))
    W = fo.render_weights(rays, cam.t[None], scene.means, scene.prec, scene.weights_log, scene.beta2, scene.beta3)
    np.testing.assert_allclose(F.numpy(), W @ f, rtol=1e-12, atol=1e-14)
    zero = np.zeros(cam.H * cam.W)
    dm, dp, dw, df = go.feature_grads(scene, f, [cam], [zero], [zero], [gf])
    np.testing.assert_allclose(df, W.T @ gf, rtol=1e-12)      # dL/df = w^T gF exactly
    eps = 1e-4
    for kind, arr_of, g_of in (("means", lambda m, p, w: m, dm), ("prec", lambda m, p, w: p, dp),
                               ("w", lambda m, p, w: w, dw)):
        for _ in range(6):
            m, p, w = scene.means.copy(), scene.prec.copy(), scene.weights_log.copy()
            arr = arr_of(m, p, w)
            j = int(rng.integers(scene.n))
            if kind == "means":
                idx = (j, int(rng.integers(3)))
            elif kind == "prec":
                idx = (j,) + [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)][int(rng.integers(6))]
            else:
                idx = (j,)
            base = arr[idx]
            arr[idx] = base + eps
            lp = _feat_loss(scene, cam, f, gf, m, p, w)
            arr[idx] = base - eps
            lm = _feat_loss(scene, cam, f, gf, m, p, w)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g_of[idx]) <= 2e-4 * max(1.0, abs(g_of[idx])), (kind, idx, fd, g_of[idx])
