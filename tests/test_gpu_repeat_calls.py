"""Repeated calls through the shared renderers: each call's result is its own.

The Python layer caches the C descriptor structs between calls; a cached struct must never
be reused for a call whose scene sizes, buffers or camera differ (freed objects' ids and freed
device blocks' addresses are reused).  Also the argument checks of Renderer.forward/backward."""
import numpy as np
import pytest
import torch

from oracle import fm_oracle as fo
from oracle import ref

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _check(d, a, od, oa):
    d, a = d.detach().cpu().numpy().ravel(), a.detach().cpu().numpy().ravel()
    ed = float(np.max(np.abs(d - od) / np.abs(od)))
    ea = float(np.max(np.abs(a - oa)))
    assert ed <= 1e-4 and ea <= 1e-4, (ed, ea)


def test_render_camera_twice_with_new_scene_and_pose(fea, scenes_mod):
    for seed, n in ((0, 40), (5, 90), (6, 40)):
        scene = scenes_mod.single_view_scene(seed=seed, n=n)
        cam = scene.cameras[0]
        # a fresh SceneTensors inside the call every time, dropped right after
        d, a = fea.render_camera(scene.means, scene.prec, scene.weights_log, cam, scene.beta2, scene.beta3,
                                 device=DEV)
        torch.cuda.synchronize()
        od, oa = ref.render_f64(scene, cam)
        _check(d, a, od, oa)
        del d, a


def test_camera_changed_in_place_is_rendered(fea, scenes_mod):
    scene = scenes_mod.single_view_scene(seed=2)
    st = fea.SceneTensors.from_scene(scene, DEV)
    r = fea.Renderer(device=DEV)
    cam = scene.cameras[0]
    outs = r.alloc_outputs([(0, cam, scene.beta2, scene.beta3)])
    r.forward([st], [(0, cam, scene.beta2, scene.beta3)], outs=outs)
    other = scenes_mod.make_camera(np.array([0.3, -2.4, 0.5]), cam.H, cam.W, cam.fx)
    cam.R, cam.t = other.R, other.t                    # the same Camera object, a new pose
    r.forward([st], [(0, cam, scene.beta2, scene.beta3)], outs=outs)
    torch.cuda.synchronize()
    od, oa = ref.render_f64(scene, other)
    _check(outs[0][0], outs[0][1], od, oa)


def test_render_func_idx_quattrans_twice_on_one_renderer(fea, scenes_mod):
    from scipy.spatial.transform import Rotation
    scene = scenes_mod.shell_scene(n=256, n_views=1, H=8, W=8, f=8.0)
    r = fea.Renderer(device=DEV)
    rng = np.random.default_rng(4)
    for k in range(3):
        c = 3.0 * np.array([np.cos(1.1 * k), np.sin(1.1 * k), 0.2 * k])
        Rc = scenes_mod.look_at(c)
        x, y, z, w = Rotation.from_matrix(Rc).as_quat()
        poses = np.concatenate([[w, x, y, z], c])[None]
        R = 500 + 300 * k
        pp = np.stack([rng.uniform(-60, 60, R), rng.uniform(-60, 60, R), np.ones(R), np.zeros(R)], 1)
        d, a, _, _ = fea.render_func_idx_quattrans(scene.means, scene.prec, scene.weights_log, pp, 1.0 / 100.0,
                                                   poses, scene.beta2, scene.beta3, device=DEV, renderer=r)
        torch.cuda.synchronize()
        od, oa = fo.render_func_idx_quattrans(scene.means, scene.prec, scene.weights_log, pp, 1.0 / 100.0, poses,
                                              scene.beta2, scene.beta3)
        _check(d, a, od, oa)


def test_renderer_rejects_mismatched_buffers(fea, scenes_mod):
    from fuzzy_ellipses_amd import FeError
    scene = scenes_mod.single_view_scene(seed=1)
    cam = scene.cameras[0]
    st = fea.SceneTensors.from_scene(scene, DEV)
    r = fea.Renderer(device=DEV)
    views = [(0, cam, scene.beta2, scene.beta3)]
    P = cam.H * cam.W
    ok = torch.zeros(P, device=DEV)
    with pytest.raises(FeError):
        r.forward([st], views, outs=[(torch.empty(P - 1, device=DEV), torch.empty(P, device=DEV))])
    with pytest.raises(FeError):
        r.forward([st], views, outs=[(torch.empty(P, device=DEV, dtype=torch.float64), torch.empty(P, device=DEV))])
    r.forward([st], views, save_for_backward=True)
    with pytest.raises(FeError):
        r.backward([torch.zeros(P // 2, device=DEV)], [ok])
    with pytest.raises(FeError):
        r.backward([ok.double()], [ok])
    with pytest.raises(FeError):
        r.backward([ok], [ok.cpu()])
    with pytest.raises(FeError):
        r.backward([ok], [ok], grads=[(torch.zeros(scene.n, 3, device=DEV), torch.zeros(scene.n, 3, 3, device=DEV),
                                       torch.zeros(scene.n, device=DEV, dtype=torch.float64))])
    with pytest.raises(FeError):
        r.backward([ok], [ok], grads=[(torch.zeros(scene.n - 1, 3, device=DEV),
                                       torch.zeros(scene.n, 3, 3, device=DEV), torch.zeros(scene.n, device=DEV))])
    g = r.backward([ok], [None])
    torch.cuda.synchronize()
    assert all(torch.all(x == 0) for x in g[0])         # zero cotangents give zero gradients
