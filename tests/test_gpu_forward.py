"""Forward parity on the GPU: HIP path through the C ABI vs the FP64 oracle.

Tolerances (stated per the path's contract): alpha within 1e-4 absolute, depth within
1e-4 relative, against baseline/fm_render.py semantics evaluated in float64.
"""
import numpy as np
import pytest

from oracle import fm_oracle as fo

pytestmark = pytest.mark.gpu

ALPHA_TOL = 1e-4
DEPTH_RTOL = 1e-4


def _oracle_view(sc, scene, cam, pixels=None):
    rays = sc.camera_rays(cam)
    return fo.render_camera(scene.means, scene.prec, scene.weights_log, cam.R, cam.t, rays,
                            scene.beta2, scene.beta3, pixels=pixels)


def _errors(d, a, od, oa):
    ea = float(np.max(np.abs(a - oa)))
    ed = float(np.max(np.abs(d - od) / np.maximum(np.abs(od), 1e-30)))
    return ed, ea


@pytest.mark.parametrize("mode", ["dense", "culled"])
@pytest.mark.parametrize("tile", [8, 16, 32])
def test_single_view_full_image(fea, scenes_mod, mode, tile):
    import torch
    sc = scenes_mod
    scene = sc.single_view_scene()
    cam = scene.cameras[0]
    r = fea.Renderer(tile=tile, batch=64, mode=mode)
    st = fea.SceneTensors.from_scene(scene)
    (d, a), = r.forward([st], [(0, cam, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    od, oa = _oracle_view(sc, scene, cam)
    ed, ea = _errors(d.cpu().numpy().ravel(), a.cpu().numpy().ravel(), od, oa)
    print(f"single view {mode} tile={tile}: max depth rel err {ed:.3e}, max alpha abs err {ea:.3e}")
    assert ea <= ALPHA_TOL and ed <= DEPTH_RTOL
    if mode == "culled":
        assert r.stats()["floor_unresolved"] == 0


@pytest.mark.parametrize("view", [0, 5, 11])
def test_shell_view_sampled_pixels(fea, scenes_mod, view):
    import torch
    sc = scenes_mod
    scene = sc.shell_scene()
    cam = scene.cameras[view]
    r = fea.Renderer(tile=16, batch=64, mode="culled")
    st = fea.SceneTensors.from_scene(scene)
    (d, a), = r.forward([st], [(0, cam, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    stats = r.stats()
    rng = np.random.default_rng(100 + view)
    pix = np.concatenate([rng.choice(cam.H * cam.W, 1500, replace=False),
                          np.arange(0, cam.W), (cam.H - 1) * cam.W + np.arange(cam.W)])
    od, oa = _oracle_view(sc, scene, cam, pixels=pix)
    d = d.cpu().numpy().ravel()[pix]
    a = a.cpu().numpy().ravel()[pix]
    ed, ea = _errors(d, a, od, oa)
    print(f"shell view {view}: max depth rel err {ed:.3e}, max alpha abs err {ea:.3e}, stats {stats}")
    assert ea <= ALPHA_TOL and ed <= DEPTH_RTOL
    assert stats["floor_unresolved"] == 0 and stats["floor_tiles"] == 0


def test_shell_culled_matches_dense_full_view(fea, scenes_mod):
    import torch
    sc = scenes_mod
    scene = sc.shell_scene()
    cam = scene.cameras[3]
    st = fea.SceneTensors.from_scene(scene)
    outs = {}
    for mode in ("dense", "culled"):
        r = fea.Renderer(tile=16, batch=64, mode=mode)
        (d, a), = r.forward([st], [(0, cam, scene.beta2, scene.beta3)])
        torch.cuda.synchronize()
        outs[mode] = (d.cpu().numpy(), a.cpu().numpy())
    ed, ea = _errors(outs["culled"][0], outs["culled"][1], outs["dense"][0].astype(np.float64),
                     outs["dense"][1].astype(np.float64))
    print(f"culled vs dense full view: depth rel {ed:.3e}, alpha abs {ea:.3e}")
    assert ea <= ALPHA_TOL and ed <= DEPTH_RTOL


@pytest.mark.parametrize("mode", ["dense", "culled"])
@pytest.mark.parametrize("kind", ["nan_mean", "zero_diag", "inf_weight"])
def test_degenerate_inputs_follow_oracle(fea, scenes_mod, mode, kind):
    """Degenerate Gaussians make the view dense with the oracle's NaN/Inf semantics: a NaN
    logit anywhere leaves only the 1e-20 weight floor (depth = mean of nan_to_num(t*)), an
    infinite logit gets weight nan_to_num(exp(inf - inf)) = 0, and alpha propagates NaN."""
    import copy
    import torch
    sc = scenes_mod
    scene = copy.deepcopy(sc.single_view_scene(seed=4, n=30, H=24, W=32, f=30.0))
    if kind == "nan_mean":
        scene.means[3, 1] = np.nan
    elif kind == "zero_diag":
        scene.prec[5, 1, 1] = 0.0
    else:
        scene.weights_log[7] = np.inf
    cam = scene.cameras[0]
    r = fea.Renderer(tile=16, batch=64, mode=mode)
    (d, a), = r.forward([fea.SceneTensors.from_scene(scene)], [(0, cam, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    od, oa = _oracle_view(sc, scene, cam)
    d, a = d.cpu().numpy().ravel().astype(np.float64), a.cpu().numpy().ravel().astype(np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(oa)) and np.array_equal(np.isnan(d), np.isnan(od))
    fd, fa = np.isfinite(od), np.isfinite(oa)
    assert fd.any()
    ed = _errors(d[fd], a[fd], od[fd], np.zeros(int(fd.sum())) + a[fd])[0]
    ea = float(np.max(np.abs(a[fa] - oa[fa]))) if fa.any() else 0.0
    print(f"degenerate {kind} {mode}: NaN alpha px {int(np.isnan(oa).sum())}, depth rel {ed:.2e}, alpha abs {ea:.2e}")
    assert ea <= ALPHA_TOL and ed <= DEPTH_RTOL
    if mode == "culled":
        assert r.stats()["floor_unresolved"] == 0


def _inside_scene(sc, kind):
    """Cameras inside the Gaussian cloud: pairs behind the camera (t* <= 0, masked from the
    depth softmax but counted in alpha and in the max logit).
      cube  : 200 Gaussians in the unit cube; the max logit is a masked pair at almost every
              pixel, the 1e-20 floor stays invisible
      centre: 1024 on the unit shell, camera near the centre; the floor dominates every pixel
      offc  : the same shell, camera off centre; floor and listed weights are comparable"""
    if kind == "cube":
        rng = np.random.default_rng(7)
        means, prec, w = sc.random_gaussians(rng, 200)
        centre = (0.12, -0.07, 0.18)
    else:
        rng = np.random.default_rng(8)
        means, prec, w = sc.random_gaussians(rng, 1024, shell=True)
        centre = (0.02, -0.03, 0.05) if kind == "centre" else (0.3, -0.2, 0.4)
    cam = sc.make_camera(np.array(centre), 64, 64, 40.0)
    return sc.Scene(means, prec, w, [cam])


@pytest.mark.parametrize("mode", ["dense", "culled"])
@pytest.mark.parametrize("kind", ["cube", "centre", "offc"])
def test_camera_inside_cloud(fea, scenes_mod, mode, kind):
    import torch
    sc = scenes_mod
    scene = _inside_scene(sc, kind)
    cam = scene.cameras[0]
    r = fea.Renderer(tile=16, batch=64, mode=mode)
    st = fea.SceneTensors.from_scene(scene)
    (d, a), = r.forward([st], [(0, cam, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    od, oa = _oracle_view(sc, scene, cam)
    ed, ea = _errors(d.cpu().numpy().ravel(), a.cpu().numpy().ravel(), od, oa)
    stats = r.stats()
    print(f"camera inside ({kind}) {mode}: max depth rel err {ed:.3e}, max alpha abs err {ea:.3e}, "
          f"floor tiles {stats.get('floor_tiles')}")
    assert ea <= ALPHA_TOL and ed <= DEPTH_RTOL
    if mode == "culled":
        assert stats["floor_unresolved"] == 0
