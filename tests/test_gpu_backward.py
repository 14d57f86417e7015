"""Backward parity on the GPU: hand-written backward kernel (fe_backward) vs float64 autograd
of the restated reference math (oracle/grad_oracle.py).  Tolerance: max abs error of each
gradient array <= 1e-3 of that array's max magnitude (float32 accumulation over pixels)."""
import numpy as np
import pytest

from oracle import grad_oracle as go

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-3


def _gpu_grads(fea, scene, cams, gds, gas, mode, tile=16, batch=64):
    import torch
    r = fea.Renderer(tile=tile, batch=batch, mode=mode)
    st = fea.SceneTensors.from_scene(scene)
    r.forward([st], [(0, c, scene.beta2, scene.beta3) for c in cams], save_for_backward=True)
    g = r.backward([torch.tensor(x.reshape(c.H, c.W), dtype=torch.float32, device="cuda") for x, c in zip(gds, cams)],
                   [torch.tensor(x.reshape(c.H, c.W), dtype=torch.float32, device="cuda") for x, c in zip(gas, cams)])
    torch.cuda.synchronize()
    dm, dp, dw = [x.cpu().numpy().astype(np.float64) for x in g[0]]
    return dm, np.triu(dp), dw


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("mode,tile", [("dense", 16), ("culled", 16), ("culled", 8), ("culled", 32)])
def test_backward_single_view(fea, scenes_mod, mode, tile):
    scene = scenes_mod.single_view_scene()
    cam = scene.cameras[0]
    rng = np.random.default_rng(31)
    gd = rng.normal(size=cam.H * cam.W)
    ga = rng.normal(size=cam.H * cam.W) * 10
    ref = go.grads(scene, [cam], [gd], [ga])
    got = _gpu_grads(fea, scene, [cam], [gd], [ga], mode, tile=tile)
    errs = [_rel(g, r) for g, r in zip(got, ref)]
    print(f"backward {mode} tile={tile}: rel err means {errs[0]:.2e} prec {errs[1]:.2e} w {errs[2]:.2e}")
    assert max(errs) <= GRAD_RTOL


def test_backward_two_views_1024(fea, scenes_mod):
    """Silhouette-workload shapes (N=1024) at reduced resolution, two views summed."""
    init, _target = scenes_mod.silhouette_problem(n_views=16, H=120, W=160, f=150.0)
    cams = [init.cameras[2], init.cameras[9]]
    rng = np.random.default_rng(32)
    gds = [rng.normal(size=c.H * c.W) for c in cams]
    gas = [rng.normal(size=c.H * c.W) * 10 for c in cams]
    ref = go.grads(init, cams, gds, gas)
    got = _gpu_grads(fea, init, cams, gds, gas, "culled")
    errs = [_rel(g, r) for g, r in zip(got, ref)]
    print(f"backward N=1024 2 views: rel err means {errs[0]:.2e} prec {errs[1]:.2e} w {errs[2]:.2e}")
    assert max(errs) <= GRAD_RTOL


@pytest.mark.parametrize("mode", ["dense", "culled"])
def test_backward_camera_inside_cube(fea, scenes_mod, mode):
    """200 Gaussians in the unit cube seen from inside it: about 30% of the pairs lie behind
    the camera, so tiles take the masked blend; the 1e-20 floor stays invisible (S >= 1e-6)."""
    sc = scenes_mod
    rng = np.random.default_rng(7)
    means, prec, w = sc.random_gaussians(rng, 200)
    cam = sc.make_camera(np.array([0.12, -0.07, 0.18]), 64, 64, 40.0)
    scene = sc.Scene(means, prec, w, [cam])
    rng = np.random.default_rng(33)
    gd = rng.normal(size=cam.H * cam.W)
    ga = rng.normal(size=cam.H * cam.W) * 10
    ref = go.grads(scene, [cam], [gd], [ga])
    got = _gpu_grads(fea, scene, [cam], [gd], [ga], mode)
    errs = [_rel(g, r) for g, r in zip(got, ref)]
    print(f"backward camera inside {mode}: rel err means {errs[0]:.2e} prec {errs[1]:.2e} w {errs[2]:.2e}")
    assert max(errs) <= GRAD_RTOL


@pytest.mark.parametrize("mode", ["dense", "culled"])
@pytest.mark.parametrize("centre", [(0.02, -0.03, 0.05), (0.3, -0.2, 0.4)])
def test_backward_weight_floor(fea, scenes_mod, mode, centre):
    """1024 Gaussians on the unit shell seen from near its centre: the pairs behind the camera
    hold the maximum logit and the 1e-20 weight floor is visible at every pixel (near the
    centre it dominates; off centre it is comparable to the listed weights).  Autograd of the
    reference then has two more terms: dD/dt = 1e-20/(S + 1e-20 n) for every pair, and the
    gradient through the maximum logit of jax_stable_exp."""
    sc = scenes_mod
    rng = np.random.default_rng(8)
    means, prec, w = sc.random_gaussians(rng, 1024, shell=True)
    cam = sc.make_camera(np.array(centre), 64, 64, 40.0)
    scene = sc.Scene(means, prec, w, [cam])
    rng = np.random.default_rng(34)
    gd = rng.normal(size=cam.H * cam.W)
    ga = rng.normal(size=cam.H * cam.W) * 10
    ref = go.grads(scene, [cam], [gd], [ga])
    got = _gpu_grads(fea, scene, [cam], [gd], [ga], mode)
    errs = [_rel(g, r) for g, r in zip(got, ref)]
    print(f"backward weight floor {centre} {mode}: rel err means {errs[0]:.2e} prec {errs[1]:.2e} w {errs[2]:.2e}")
    assert max(errs) <= GRAD_RTOL


def _direct_form_case(scenes_mod, case):
    """Scenes whose views take the direct-form forward and backward (no blend records): a
    Gaussian with a zero diagonal entry of U (a degenerate record: the view is rendered densely
    with the NaN-safe blend), the camera at a Gaussian's centre (|q|^2 = 0), beta3 = 0 and
    beta3 < 0 (the blend records' running maximum needs beta3 > 0)."""
    import dataclasses
    scene = scenes_mod.single_view_scene(seed=14)
    cam = scene.cameras[0]
    if case == "zero_diagonal":
        prec = scene.prec.copy()
        prec[5, 1, 1] = 0.0
        return dataclasses.replace(scene, prec=prec), cam
    if case == "camera_at_centre":
        # a camera centre exact in float32, so the oracle and the GPU both see q = 0 exactly
        cam = dataclasses.replace(cam, t=np.asarray(cam.t, np.float32).astype(np.float64))
        means = scene.means.copy()
        means[9] = cam.t
        return dataclasses.replace(scene, means=means, cameras=[cam]), cam
    if case == "beta3_zero":
        return dataclasses.replace(scene, beta3=0.0), cam
    return dataclasses.replace(scene, beta3=-0.5), cam


@pytest.mark.parametrize("mode", ["dense", "culled"])
@pytest.mark.parametrize("case", ["zero_diagonal", "camera_at_centre", "beta3_zero", "beta3_negative"])
def test_backward_direct_form(fea, scenes_mod, case, mode):
    scene, cam = _direct_form_case(scenes_mod, case)
    rng = np.random.default_rng(35)
    gd = rng.normal(size=cam.H * cam.W)
    ga = rng.normal(size=cam.H * cam.W) * 10
    ref = go.grads(scene, [cam], [gd], [ga])
    assert all(np.all(np.isfinite(r)) for r in ref), "the reference gradient should be finite here"
    got = _gpu_grads(fea, scene, [cam], [gd], [ga], mode)
    assert all(np.all(np.isfinite(g)) for g in got)
    # the forward of the same view against the FP64 oracle
    import torch
    from oracle import ref as oref
    r = fea.Renderer(tile=16, batch=64, mode=mode)
    (d, a), = r.forward([fea.SceneTensors.from_scene(scene)], [(0, cam, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    od, oa = oref.render_f64(scene, cam)
    ed = float(np.max(np.abs(d.cpu().numpy().ravel() - od) / np.abs(od)))
    ea = float(np.max(np.abs(a.cpu().numpy().ravel() - oa)))
    print(f"forward direct form {case} {mode}: depth rel {ed:.2e}, alpha abs {ea:.2e}")
    assert ed <= 1e-4 and ea <= 1e-4
    errs = [_rel(g, r) for g, r in zip(got, ref)]
    print(f"backward direct form {case} {mode}: rel err means {errs[0]:.2e} prec {errs[1]:.2e} w {errs[2]:.2e}")
    assert max(errs) <= GRAD_RTOL
