"""Multi-pose ray batches on the GPU: ray-list views against the FP64 oracle, bit-exact
ray-list tile keys, render_func_idx_quattrans[_flow] against the oracle's restatement of
baseline/fm_render.py:151-224, and gradients through ray-list views."""
import numpy as np
import pytest

from oracle import fm_oracle as fo
from oracle import ref

pytestmark = pytest.mark.gpu

TOL_DEPTH_REL = 1e-4
TOL_ALPHA_ABS = 1e-4


def _sub_pixel_rays(cam, m, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, cam.W - 0.5, m)
    y = rng.uniform(-0.5, cam.H - 0.5, m)
    return np.stack([(x - cam.cx) / cam.fx, (y - cam.cy) / cam.fy], 1)


def _oracle_rays(scene, cam, xy):
    xy32 = np.asarray(xy, np.float32).astype(np.float64)
    rays = np.concatenate([xy32, np.ones((len(xy32), 1))], 1) @ cam.R
    return fo.render_func_rays(rays, cam.t[None], scene.means, scene.prec, scene.weights_log, scene.beta2,
                               scene.beta3)


def _errs(d, a, d64, a64):
    d, a = d.cpu().numpy().astype(np.float64), a.cpu().numpy().astype(np.float64)
    ed = np.max(np.abs(d - d64) / np.maximum(np.abs(d64), 1e-6))
    ea = np.max(np.abs(a - a64))
    return ed, ea


@pytest.mark.parametrize("ts,order", [(16, "raster"), (8, "morton"), (32, "random")])
def test_ray_list_view_matches_fp64(fea, scenes_mod, ts, order):
    import torch
    from fuzzy_ellipses_amd.rays import _morton_order
    scene = scenes_mod.single_view_scene(seed=6, n=80, H=48, W=64, f=60.0)
    cam = scene.cameras[0]
    if order == "raster":        # the whole pixel grid, as a list
        ys, xs = np.divmod(np.arange(cam.H * cam.W), cam.W)
        xy = np.stack([(xs - cam.cx) / cam.fx, (ys - cam.cy) / cam.fy], 1)
    else:
        xy = _sub_pixel_rays(cam, 1500, 7)
        if order == "morton":
            xy = xy[_morton_order(xy[:, 0], xy[:, 1])]
    rl = fea.RayList(cam.R, cam.t, torch.tensor(xy, dtype=torch.float32, device="cuda:0").contiguous())
    r = fea.Renderer(tile=ts, batch=64, mode="culled")
    (d, a), = r.forward([fea.SceneTensors.from_scene(scene)], [(0, rl, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    d64, a64 = _oracle_rays(scene, cam, xy)
    ed, ea = _errs(d, a, d64, a64)
    print(f"ray list ts={ts} {order}: {len(xy)} rays, depth rel {ed:.2e}, alpha abs {ea:.2e}, stats {r.stats()}")
    assert ed < TOL_DEPTH_REL and ea < TOL_ALPHA_ABS


def test_ray_list_keys_bit_exact(fea, scenes_mod):
    import torch
    from fuzzy_ellipses_amd.rays import _morton_order
    scene = scenes_mod.shell_scene(n=1024, n_views=2, H=120, W=160, f=150.0)
    cam = scene.cameras[1]
    xy = _sub_pixel_rays(cam, 5000, 8)
    xy = xy[_morton_order(xy[:, 0], xy[:, 1])].astype(np.float32)
    rl = fea.RayList(cam.R, cam.t, torch.tensor(xy, device="cuda:0").contiguous())
    r = fea.Renderer(tile=16, batch=64, mode="culled")
    r.forward([fea.SceneTensors.from_scene(scene)], [(0, rl, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    st = r.debug_state()
    from conftest import in_record_order
    cpu_order = ref.order(scene, cam, ray_view=True)     # the record order, from the inputs
    np.testing.assert_array_equal(st["order"], cpu_order)
    scene = in_record_order(scene, cpu_order)
    off1, off2, lst = st["off1"], st["off2"], st["list"]
    c1, l1 = ref.tiles_rays(scene, cam.R, cam.t, xy, 16)
    np.testing.assert_array_equal(np.diff(off1), c1.astype(np.int64))
    np.testing.assert_array_equal(lst[:off1[-1]], l1)
    c2, l2 = ref.tiles_rays(scene, cam.R, cam.t, xy, 16, tile_lmin=st["tile_lmin"], tile_seed=st["tile_seed"])
    np.testing.assert_array_equal(np.diff(off2), c2.astype(np.int64))
    np.testing.assert_array_equal(lst[off2[0]:off2[-1]], l2)
    print(f"ray-list keys: phase-1 pairs {int(off1[-1])}, phase-2 pairs {int(off2[-1] - off1[-1])}")


def test_nonfinite_ray_makes_the_view_dense(fea, scenes_mod):
    import torch
    scene = scenes_mod.single_view_scene(seed=2, n=40, H=32, W=32, f=32.0)
    cam = scene.cameras[0]
    xy = _sub_pixel_rays(cam, 300, 9)
    xy[17] = [np.nan, 0.1]
    rl = fea.RayList(cam.R, cam.t, torch.tensor(xy, dtype=torch.float32, device="cuda:0").contiguous())
    r = fea.Renderer(tile=16, batch=64, mode="culled")
    (d, a), = r.forward([fea.SceneTensors.from_scene(scene)], [(0, rl, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    d64, a64 = _oracle_rays(scene, cam, xy)
    ok = np.arange(len(xy)) != 17
    ed, ea = _errs(d[ok], a[ok], d64[ok], a64[ok])
    assert ed < TOL_DEPTH_REL and ea < TOL_ALPHA_ABS
    assert np.isnan(a[17].item()) and np.isnan(a64[17])
    assert d[17].item() == pytest.approx(d64[17], abs=1e-6)


def _poses(rng, P):
    from scipy.spatial.transform import Rotation
    out = []
    for p in range(P):
        c = 3.0 * np.array([np.cos(0.2 * p), np.sin(0.2 * p), 0.3])
        from fuzzy_ellipses_amd.scenes import look_at
        Rc = look_at(c)                                 # rows = camera axes
        x, y, z, w = Rotation.from_matrix(Rc).as_quat()     # quat_to_rot(q) = Rc (world = cam @ Rc)
        q = np.array([w, x, y, z])
        assert np.allclose(fo.quat_to_rot(q), Rc, atol=1e-9)
        out.append(np.concatenate([q, c]))
    return np.stack(out)


def test_render_func_idx_quattrans_flow_matches_oracle(fea, scenes_mod):
    import torch
    scene = scenes_mod.shell_scene(n=512, n_views=1, H=8, W=8, f=8.0)
    rng = np.random.default_rng(10)
    P, R = 4, 6000
    poses = _poses(rng, P)
    invF = 1.0 / 120.0
    pp = np.stack([rng.uniform(-80, 80, R), rng.uniform(-60, 60, R), np.ones(R), rng.integers(0, P, R)], 1)
    d, a, _, _, fn, fp = fea.render_func_idx_quattrans_flow(scene.means, scene.prec, scene.weights_log, pp, invF,
                                                            poses, scene.beta2, scene.beta3, device="cuda:0")
    torch.cuda.synchronize()
    d64, a64 = fo.render_func_idx_quattrans(scene.means, scene.prec, scene.weights_log, pp, invF, poses,
                                            scene.beta2, scene.beta3)
    ed, ea = _errs(d, a, d64, a64)
    # the flow kernel against the oracle's flow on the same (GPU) depth
    fn64, fp64 = fo.idx_quattrans_flow(pp, invF, poses, d.cpu().numpy().astype(np.float64))
    ef = max(np.max(np.abs(fn.cpu().numpy() - fn64)), np.max(np.abs(fp.cpu().numpy() - fp64)))
    scale = max(np.abs(fn64).max(), np.abs(fp64).max())
    print(f"idx_quattrans: {R} rays over {P} poses, depth rel {ed:.2e}, alpha abs {ea:.2e}, "
          f"flow max abs err {ef:.2e} (max |flow| {scale:.1f} px)")
    assert ed < TOL_DEPTH_REL and ea < TOL_ALPHA_ABS
    assert ef <= 1e-5 * max(scale, 1.0) + 1e-4
    # without flow: same images
    d2, a2, _, _ = fea.render_func_idx_quattrans(scene.means, scene.prec, scene.weights_log, pp, invF, poses,
                                                 scene.beta2, scene.beta3, device="cuda:0")
    assert torch.equal(d, d2) and torch.equal(a, a2)


def test_gradients_through_ray_list_views(fea, scenes_mod):
    import torch
    from oracle.grad_oracle import render_rays_torch
    scene = scenes_mod.single_view_scene(seed=11, n=24, H=24, W=24, f=24.0)
    cam = scene.cameras[0]
    xy = _sub_pixel_rays(cam, 700, 12).astype(np.float32)
    rl = fea.RayList(cam.R, cam.t, torch.tensor(xy, device="cuda:0").contiguous())
    r = fea.Renderer(tile=16, batch=64, mode="culled")
    st = fea.SceneTensors.from_scene(scene)
    r.forward([st], [(0, rl, scene.beta2, scene.beta3)], save_for_backward=True)
    g = torch.Generator().manual_seed(3)
    gD = torch.randn(len(xy), generator=g)
    gA = torch.randn(len(xy), generator=g)
    (gm, gp, gw), = r.backward([gD.cuda()], [gA.cuda()])
    torch.cuda.synchronize()
    f64 = lambda a: torch.tensor(np.asarray(a, np.float32).astype(np.float64))
    rays = torch.cat([f64(xy), torch.ones(len(xy), 1, dtype=torch.float64)], 1) @ f64(cam.R)
    means = f64(scene.means).requires_grad_(True)
    prec = f64(np.triu(scene.prec)).requires_grad_(True)
    wl = f64(scene.weights_log).requires_grad_(True)
    dd, aa = render_rays_torch(rays, f64(cam.t)[None].expand(len(xy), 3), means, prec, wl, scene.beta2, scene.beta3)
    ((dd * gD.double()).sum() + (aa * gA.double()).sum()).backward()
    for got, want, name in ((gm, means.grad, "means"), (gp, torch.triu(prec.grad), "prec"), (gw, wl.grad, "w")):
        err = (got.cpu().double() - want).abs().max().item() / want.abs().max().item()
        print(f"ray-list grad {name}: {err:.2e}")
        assert err < 1e-4, name
