"""Property tests on the GPU over random small configurations (tests/randcfg.py): scene kind,
Gaussian count and size, log-weights, betas, tile size, LDS batch depth, dense or culled
membership, and one to three views per call from inside the cloud to far outside it, wide to
narrow, with off-centre principal points, drawn from fixed seeds so that every run checks the
same configurations.

Forward, against the FP64 oracle (oracle/fe_ref.cpp):
  * alpha within 1e-4 absolute, NaN where the oracle has NaN;
  * depth within 1e-4 relative, or, at pixels where float32 cannot resolve the reference
    formula to that budget, within 8 u kappa (u = 2^-24, kappa the oracle's first-order
    float32 sensitivity, oracle/fm_oracle.py::depth_condition, pinned on the CPU by
    tests/test_condition_cpu.py against the reference's float32 restatement);
  * no pixel whose 1e-20 weight floor the culling left unresolved.

Backward, against float64 autograd of the restated reference (oracle/grad_oracle.py), per
gradient array: max error within 1e-3 of the array's largest magnitude, or within 16 times the
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


def _forward_case(fea, sc, cfg):
    """None if the configuration renders within the bounds, else what failed."""
    import torch
    scene, cams = randcfg.build(cfg, sc)
    r = fea.Renderer(tile=cfg["tile"], batch=cfg["batch"], mode=cfg["mode"])
    outs = r.forward([fea.SceneTensors.from_scene(scene)], [(0, c, scene.beta2, scene.beta3) for c in cams])
    torch.cuda.synchronize()
    for v, ((d, a), cam) in enumerate(zip(outs, cams)):
        od, oa = ref.render_f64(scene, cam)
        d = d.cpu().numpy().ravel().astype(np.float64)
        a = a.cpu().numpy().ravel().astype(np.float64)
        if not (np.array_equal(np.isnan(a), np.isnan(oa)) and np.array_equal(np.isnan(d), np.isnan(od))):
            return f"view {v}: NaN pattern"
        fa = np.isfinite(oa)
        ea = float(np.max(np.abs(a[fa] - oa[fa]), initial=0.0))
        if ea > 1e-4:
            return f"view {v}: alpha {ea:.2e}"
        rays = sc.camera_rays(cam) @ np.asarray(cam.R)
        kap = fo.depth_condition(rays, cam.t, scene.means, scene.prec, scene.weights_log, scene.beta2,
                                 scene.beta3)
        fd = np.isfinite(od) & np.isfinite(kap)
        allowed = np.maximum(1e-4 * np.abs(od[fd]), 8.0 * U * kap[fd])
        q = np.abs(d[fd] - od[fd]) / allowed
        if q.size and float(np.max(q)) > 1.0:
            return f"view {v}: depth {float(np.max(q)):.2f} x allowed"
    if cfg["mode"] == "culled" and r.stats()["floor_unresolved"] != 0:
        return "floor unresolved"
    return None


def test_forward_random_configs(fea, scenes_mod):
    bad = []
    for seed in range(150):
        cfg = randcfg.draw(seed)
        why = _forward_case(fea, scenes_mod, cfg)
        if why is not None:
            bad.append((seed, why, cfg))
    print(f"forward: {150 - len(bad)}/150 random configurations within bounds")
    assert not bad, bad[:5]


def _backward_case(fea, sc, cfg):
    import torch
    scene, cams = randcfg.build(cfg, sc)
    rng = np.random.default_rng([cfg["seed"], 2])
    gds = [rng.normal(size=c.H * c.W).astype(np.float32).astype(np.float64) for c in cams]
    gas = [(10.0 * rng.normal(size=c.H * c.W)).astype(np.float32).astype(np.float64) for c in cams]
    r = fea.Renderer(tile=cfg["tile"], batch=cfg["batch"], mode=cfg["mode"])
    r.forward([fea.SceneTensors.from_scene(scene)], [(0, c, scene.beta2, scene.beta3) for c in cams],
              save_for_backward=True)
    g = r.backward([torch.tensor(x.reshape(c.H, c.W), dtype=torch.float32, device="cuda") for x, c in zip(gds, cams)],
                   [torch.tensor(x.reshape(c.H, c.W), dtype=torch.float32, device="cuda") for x, c in zip(gas, cams)])
    torch.cuda.synchronize()
    got = [x.cpu().numpy().astype(np.float64) for x in g[0]]
    got[1] = np.triu(got[1])
    ref64 = go.grads(scene, cams, gds, gas)
    # the log-weight gradient's per-pixel terms: alpha's exp(-E) e^{d2} (at most |gA|) and
    # depth's b3 w (z - D)
    mass = 0.0
    for cam, gd, ga in zip(cams, gds, gas):
        od, _ = ref.render_f64(scene, cam)
        mass += float(np.sum(np.abs(ga)) + scene.beta3 * np.sum(np.abs(gd) * np.abs(np.nan_to_num(od))))
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
            ref32 = go.grads(scene, cams, gds, gas, dtype=torch.float32)
        c = ref32[k]
        fin2 = fin & np.isfinite(c)
        err32 = float(np.max(np.abs(c[fin2] - b[fin2])))
        if err > 16.0 * err32:
            return f"{name}: error {err:.2e}, allowed {allowed:.2e}, float32 autograd {err32:.2e}"
    return None


def test_backward_random_configs(fea, scenes_mod):
    bad = []
    for seed in range(80):
        cfg = randcfg.draw(seed, max_n=120, max_h=40, max_w=48, max_views=2)
        why = _backward_case(fea, scenes_mod, cfg)
        if why is not None:
            bad.append((seed, why, cfg))
    print(f"backward: {80 - len(bad)}/80 random configurations within bounds")
    assert not bad, bad[:5]
