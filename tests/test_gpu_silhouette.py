"""Shape-from-silhouette caller on the GPU: the loss gradient from the hand-written backward
vs central finite differences (eps = 1e-4) of the float64 loss, and Adam lowers the loss."""
import numpy as np
import pytest

from oracle import ref

pytestmark = pytest.mark.gpu


def _loss64(scene, cams, masks, tdepths):
    tot = 0.0
    for cam, m, td in zip(cams, masks, tdepths):
        d, a = ref.render_f64(scene, cam)
        a = np.clip(a, 3e-8, 1 - 3e-8)
        bce = -(m * np.log(a) + (1 - m) * np.log1p(-a)).mean()
        l1 = (m * np.abs(d - td)).sum() / max(m.sum(), 1.0)
        tot += (bce + 0.1 * l1) / len(cams)
    return tot


def test_loss_gradient_matches_finite_differences(fea, scenes_mod):
    import copy
    import torch
    from fuzzy_ellipses_amd.silhouette import SilhouetteFit
    init, target = scenes_mod.silhouette_problem(n=128, n_views=3, H=48, W=64, f=60.0)
    fit = SilhouetteFit(init, target, tile=16, batch=64)
    masks = [m.cpu().numpy().astype(np.float64).ravel() for m in fit.masks]
    tds = [t.cpu().numpy().astype(np.float64).ravel() for t in fit.tdepths]
    assert all(m.sum() > 20 for m in masks), "silhouettes should not be empty"
    loss, (gm, gp, gw) = fit.loss_and_grads()
    torch.cuda.synchronize()
    gm, gp, gw = gm.cpu().numpy(), gp.cpu().numpy(), gw.cpu().numpy()
    base = copy.deepcopy(init)
    l0 = _loss64(base, base.cameras, masks, tds)
    assert abs(float(loss) - l0) <= 1e-4 * abs(l0)
    rng = np.random.default_rng(21)
    eps = 1e-4
    errs = []
    for k in range(16):
        sc = copy.deepcopy(init)
        j = int(rng.integers(sc.n))
        kind = k % 3
        if kind == 0:
            i = int(rng.integers(3)); arr, g, idx = sc.means, gm[j, i], (j, i)
        elif kind == 1:
            r_, c_ = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)][int(rng.integers(6))]
            arr, g, idx = sc.prec, gp[j, r_, c_], (j, r_, c_)
        else:
            arr, g, idx = sc.weights_log, gw[j], (j,)
        v = arr[idx]
        arr[idx] = v + eps
        lp = _loss64(sc, sc.cameras, masks, tds)
        arr[idx] = v - eps
        lm = _loss64(sc, sc.cameras, masks, tds)
        fd = (lp - lm) / (2 * eps)
        errs.append(abs(fd - g) / max(abs(fd), 1e-3 * max(np.abs(gm).max(), np.abs(gw).max())))
    print("silhouette FD rel errors: max %.2e median %.2e" % (max(errs), float(np.median(errs))))
    assert max(errs) <= 2e-2 and float(np.median(errs)) <= 2e-3


def test_adam_lowers_loss(fea, scenes_mod):
    from fuzzy_ellipses_amd.silhouette import SilhouetteFit
    init, target = scenes_mod.silhouette_problem(n=256, n_views=4, H=60, W=80, f=75.0)
    fit = SilhouetteFit(init, target, lr=1e-2)
    out = fit.run(steps=40)
    print("silhouette: loss %.4f -> %.4f in %d steps (%.2f ms/step)" %
          (out["loss_first"], out["loss_last"], out["steps"], out["ms_per_step"]))
    assert out["loss_last"] < 0.9 * out["loss_first"]
