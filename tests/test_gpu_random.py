"""Property tests on the GPU over random small configurations (tests/randcfg.py): scene kind,
Gaussian count and size, log-weights, betas, tile size, LDS batch depth, dense or culled
membership, and one to three views per call from inside the cloud to far outside it, wide to
narrow, with off-centre principal points, each also as a ray list (random sub-pixel rays over
the image or three times its extent, Morton or random order) in the same call, drawn from fixed
seeds so that every run checks the same configurations.

Forward, against the FP64 oracle (oracle/fe_ref.cpp for pixel grids, oracle/fm_oracle.py for
ray lists); every fifth configuration also renders each view alone and requires the result of
the mixed call bit for bit:
  * alpha within 1e-4 absolute, NaN where the oracle has NaN;
  * depth within 1e-4 relative, or, at pixels where float32 cannot resolve the reference
    formula to that budget, within 8 u kappa (u = 2^-24, kappa the oracle's first-order
    float32 sensitivity, oracle/fm_oracle.py::depth_condition, pinned on the CPU by
    tests/test_condition_cpu.py against the reference's float32 restatement);
  * no pixel whose 1e-20 weight floor the culling left unresolved.

Backward (pixel grids, and in a second test ray lists), against float64 autograd of the
restated reference (oracle/grad_oracle.py), per gradient array: max error within 1e-3 of the array's largest magnitude, or within 16 times the
error of float32 autograd of the same restatement (the reference's own precision; gradients
of scenes of a few tiny Gaussians far from most rays are only resolved that well), and for
the log-weights, whose true gradient can vanish while its per-pixel terms do not, within
1e-7 of the sum of the magnitudes of those terms.  Configurations of up to 120 Gaussians and
two 40x48 views keep the float64 autograd quick.
"""
import numpy as np
import pytest

import randcfg
from oracle import fm_oracle as fo
from oracle import grad_oracle as go
from oracle import ref

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def _depth_alpha_ok(d, a, od, oa, kap):
    """None if (d, a) meet the bounds against the oracle's (od, oa), else what failed."""
    d = d.cpu().numpy().ravel().astype(np.float64)
    a = a.cpu().numpy().ravel().astype(np.float64)
    if not (np.array_equal(np.isnan(a), np.isnan(oa)) and np.array_equal(np.isnan(d), np.isnan(od))):
        return "NaN pattern"
    fa = np.isfinite(oa)
    ea = float(np.max(np.abs(a[fa] - oa[fa]), initial=0.0))
    if ea > 1e-4:
        return f"alpha {ea:.2e}"
    fd = np.isfinite(od) & np.isfinite(kap)
    q = np.abs(d[fd] - od[fd]) / np.maximum(1e-4 * np.abs(od[fd]), 8.0 * U * kap[fd])
    if q.size and float(np.max(q)) > 1.0:
        return f"depth {float(np.max(q)):.2f} x allowed"
    return None


def _forward_case(fea, sc, cfg, lone):
    """None if the configuration renders within the bounds, else what failed."""
    import torch
    scene, cams = randcfg.build(cfg, sc)
    xys = randcfg.ray_lists(cfg, cams)
    views = [(0, c, scene.beta2, scene.beta3) for c in cams]
    views += [(0, fea.RayList(c.R, c.t, torch.tensor(xy, device="cuda").contiguous()), scene.beta2, scene.beta3)
              for c, xy in zip(cams, xys)]
    st = fea.SceneTensors.from_scene(scene)
    r = fea.Renderer(tile=cfg["tile"], batch=cfg["batch"], mode=cfg["mode"])
    outs = r.forward([st], views)
    torch.cuda.synchronize()
    b3, b2 = scene.beta3, scene.beta2
    for v, cam in enumerate(cams):
        od, oa = ref.render_f64(scene, cam)
        rays = sc.camera_rays(cam) @ np.asarray(cam.R)
        kap = fo.depth_condition(rays, cam.t, scene.means, scene.prec, scene.weights_log, b2, b3)
        why = _depth_alpha_ok(*outs[v], od, oa, kap)
        if why:
            return f"view {v}: {why}"
    for v, (cam, xy) in enumerate(zip(cams, xys)):
        rays = np.concatenate([xy.astype(np.float64), np.ones((len(xy), 1))], 1) @ np.asarray(cam.R)
        od, oa = fo.render_func_rays(rays, cam.t[None], scene.means, scene.prec, scene.weights_log, b2, b3)
        kap = fo.depth_condition(rays, cam.t, scene.means, scene.prec, scene.weights_log, b2, b3)
        why = _depth_alpha_ok(*outs[len(cams) + v], od, oa, kap)
        if why:
            return f"ray list {v}: {why}"
    if cfg["mode"] == "culled" and r.stats()["floor_unresolved"] != 0:
        return "floor unresolved"
    if lone:
        for v, view in enumerate(views):
            (d1, a1), = r.forward([st], [view])
            torch.cuda.synchronize()
            d0, a0 = outs[v]
            same = (torch.equal(torch.nan_to_num(d0, 7.0), torch.nan_to_num(d1, 7.0))
                    and torch.equal(torch.nan_to_num(a0, 7.0), torch.nan_to_num(a1, 7.0)))
            if not same:
                return f"view {v} alone differs from the mixed call"
    return None


def test_forward_random_configs(fea, scenes_mod):
    bad = []
    for seed in range(150):
        cfg = randcfg.draw(seed)
        why = _forward_case(fea, scenes_mod, cfg, lone=seed % 5 == 0)
        if why is not None:
            bad.append((seed, why, cfg))
    print(f"forward: {150 - len(bad)}/150 random configurations (pixel grids and ray lists) within bounds")
    assert not bad, bad[:5]


def test_random_multi_scene_calls_equal_lone_views(fea, scenes_mod):
    """Groups of four random configurations in one call (their scenes, pixel-grid views and ray
    lists; the group's first configuration sets tile, batch depth and mode): every view equals
    its own single-view call bit for bit."""
    import torch
    bad = []
    for g in range(25):
        cfgs = [randcfg.draw(1000 + 4 * g + k) for k in range(4)]
        base = cfgs[0]
        scenes, views = [], []
        for i, cfg in enumerate(cfgs):
            scene, cams = randcfg.build(cfg, scenes_mod)
            scenes.append(fea.SceneTensors.from_scene(scene))
            views += [(i, c, scene.beta2, scene.beta3) for c in cams]
            views += [(i, fea.RayList(c.R, c.t, torch.tensor(xy, device="cuda").contiguous()), scene.beta2, scene.beta3)
                      for c, xy in zip(cams, randcfg.ray_lists(cfg, cams, max_rays=800))]
        r = fea.Renderer(tile=base["tile"], batch=base["batch"], mode=base["mode"])
        outs = r.forward(scenes, views)
        torch.cuda.synchronize()
        for v, (i, cam, b2, b3) in enumerate(views):
            (d1, a1), = r.forward([scenes[i]], [(0, cam, b2, b3)])
            torch.cuda.synchronize()
            d0, a0 = outs[v]
            if not (torch.equal(torch.nan_to_num(d0, 7.0), torch.nan_to_num(d1, 7.0))
                    and torch.equal(torch.nan_to_num(a0, 7.0), torch.nan_to_num(a1, 7.0))):
                bad.append((g, v))
    print(f"multi-scene calls: {25 - len({g for g, _ in bad})}/25 groups equal their lone views bit for bit")
    assert not bad, bad[:5]


def _grads_ok(got, ref64, ref32_fn, mass):
    """None if each gradient array meets the bounds against float64 autograd, else what failed."""
    ref32 = None
    for k, name in enumerate(("means", "prec", "weights_log")):
        a, b = got[k], ref64[k]
        fin = np.isfinite(b)
        if not np.array_equal(np.isfinite(a), fin):
            return f"{name}: non-finite pattern"
        if not fin.any():
            continue
        err = float(np.max(np.abs(a[fin] - b[fin])))
        allowed = 1e-3 * float(np.max(np.abs(b[fin])))
        if k == 2:
            allowed = max(allowed, 1e-7 * mass)
        if err <= allowed:
            continue
        if ref32 is None:
            ref32 = ref32_fn()
        c = ref32[k]
        fin2 = fin & np.isfinite(c)
        err32 = float(np.max(np.abs(c[fin2] - b[fin2])))
        if err > 16.0 * err32:
            return f"{name}: error {err:.2e}, allowed {allowed:.2e}, float32 autograd {err32:.2e}"
    return None


def _run_backward(fea, scene, views, gds, gas, cfg):
    import torch
    r = fea.Renderer(tile=cfg["tile"], batch=cfg["batch"], mode=cfg["mode"])
    r.forward([fea.SceneTensors.from_scene(scene)], views, save_for_backward=True)
    shape = [(c.H, c.W) if not isinstance(c, fea.RayList) else (c.n,) for _, c, _, _ in views]
    g = r.backward([torch.tensor(x.reshape(s), dtype=torch.float32, device="cuda") for x, s in zip(gds, shape)],
                   [torch.tensor(x.reshape(s), dtype=torch.float32, device="cuda") for x, s in zip(gas, shape)])
    torch.cuda.synchronize()
    got = [x.cpu().numpy().astype(np.float64) for x in g[0]]
    got[1] = np.triu(got[1])
    return got


def _cotangents(cfg, sizes):
    rng = np.random.default_rng([cfg["seed"], 2])
    gds = [rng.normal(size=m).astype(np.float32).astype(np.float64) for m in sizes]
    gas = [(10.0 * rng.normal(size=m)).astype(np.float32).astype(np.float64) for m in sizes]
    return gds, gas


def _mass(scene, depths, gds, gas):
    """The log-weight gradient's per-pixel terms: alpha's exp(-E) e^{d2} (at most |gA|) and
    depth's b3 w (z - D), at the scale of |D|."""
    return float(sum(np.sum(np.abs(ga)) + scene.beta3 * np.sum(np.abs(gd) * np.abs(np.nan_to_num(d)))
                     for d, gd, ga in zip(depths, gds, gas)))


def _backward_case(fea, sc, cfg):
    import torch
    scene, cams = randcfg.build(cfg, sc)
    gds, gas = _cotangents(cfg, [c.H * c.W for c in cams])
    got = _run_backward(fea, scene, [(0, c, scene.beta2, scene.beta3) for c in cams], gds, gas, cfg)
    ref64 = go.grads(scene, cams, gds, gas)
    mass = _mass(scene, [ref.render_f64(scene, c)[0] for c in cams], gds, gas)
    return _grads_ok(got, ref64, lambda: go.grads(scene, cams, gds, gas, dtype=torch.float32), mass)


def _backward_rays_case(fea, sc, cfg):
    import torch
    scene, cams = randcfg.build(cfg, sc)
    xys = randcfg.ray_lists(cfg, cams, max_rays=1500)
    ray_views = [(np.concatenate([xy.astype(np.float64), np.ones((len(xy), 1))], 1) @ np.asarray(c.R), c.t)
                 for c, xy in zip(cams, xys)]
    gds, gas = _cotangents(cfg, [len(xy) for xy in xys])
    views = [(0, fea.RayList(c.R, c.t, torch.tensor(xy, device="cuda").contiguous()), scene.beta2, scene.beta3)
             for c, xy in zip(cams, xys)]
    got = _run_backward(fea, scene, views, gds, gas, cfg)
    ref64 = go.ray_grads(scene, ray_views, gds, gas)
    depths = [fo.render_func_rays(rays, o[None], scene.means, scene.prec, scene.weights_log, scene.beta2,
                                  scene.beta3)[0] for rays, o in ray_views]
    return _grads_ok(got, ref64, lambda: go.ray_grads(scene, ray_views, gds, gas, dtype=torch.float32),
                     _mass(scene, depths, gds, gas))


def test_backward_random_configs(fea, scenes_mod):
    bad = []
    for seed in range(80):
        cfg = randcfg.draw(seed, max_n=120, max_h=40, max_w=48, max_views=2)
        why = _backward_case(fea, scenes_mod, cfg)
        if why is not None:
            bad.append((seed, why, cfg))
    print(f"backward: {80 - len(bad)}/80 random configurations within bounds")
    assert not bad, bad[:5]


def test_backward_random_ray_lists(fea, scenes_mod):
    bad = []
    for seed in range(80):
        cfg = randcfg.draw(seed, max_n=120, max_h=40, max_w=48, max_views=2)
        why = _backward_rays_case(fea, scenes_mod, cfg)
        if why is not None:
            bad.append((seed, why, cfg))
    print(f"backward through ray lists: {80 - len(bad)}/80 random configurations within bounds")
    assert not bad, bad[:5]
