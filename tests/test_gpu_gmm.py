"""GMM point-cloud log-likelihood on the GPU (fe_gmm_forward / fe_gmm_backward) against the
float64 restatement of baseline/fm_render.py:231-259 and its autograd gradients."""
import numpy as np
import pytest

from oracle import gmm_oracle as go

pytestmark = pytest.mark.gpu


def _problem(scenes_mod, seed, n, m):
    rng = np.random.default_rng(seed)
    means, U, wl = scenes_mod.random_gaussians(rng, n)
    wl = wl + rng.normal(size=n)
    pts = rng.uniform(-0.7, 0.7, size=(m, 3))
    return means, U, wl, pts


@pytest.mark.parametrize("n,m", [(1, 1), (37, 1000), (300, 5000)])
def test_forward_matches_float64(fea, scenes_mod, n, m):
    import torch
    from fuzzy_ellipses_amd.gmm import GMM
    means, U, wl, pts = _problem(scenes_mod, n + m, n, m)
    g = GMM(means, U, wl, device="cuda:0")
    p32 = torch.tensor(pts, dtype=torch.float32, device="cuda:0")
    nll, lse, res = g.forward(p32, return_res=True)
    nll2, lse2, _ = g.forward(p32)
    torch.cuda.synchronize()
    # the oracle on the float32-rounded inputs the kernel sees
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    want_nll, want_res = go.log_likelihood(f(means), f(U), f(wl), f(pts))
    want_lse = go.point_lse(f(means), f(U), f(wl), f(pts))
    e_lse = np.max(np.abs(lse.cpu().numpy() - want_lse) / np.maximum(1.0, np.abs(want_lse)))
    e_res = np.max(np.abs(res.cpu().numpy() - want_res) / np.maximum(1.0, np.abs(want_res)))
    print(f"gmm n={n} m={m}: lse rel err {e_lse:.2e}, res rel err {e_res:.2e}, nll {float(nll):.6f} vs {want_nll:.6f}")
    assert e_lse < 5e-6 and e_res < 5e-6
    assert abs(float(nll) - want_nll) <= 5e-6 * max(1.0, abs(want_nll))
    assert torch.equal(lse, lse2) and float(nll) == float(nll2), "forward must be deterministic"


def test_backward_matches_autograd(fea, scenes_mod):
    import torch
    from fuzzy_ellipses_amd.gmm import GMM
    n, m = 70, 6000                        # several point chunks, Gaussians across waves
    means, U, wl, pts = _problem(scenes_mod, 5, n, m)
    g = GMM(means, U, wl, device="cuda:0")
    p32 = torch.tensor(pts, dtype=torch.float32, device="cuda:0")
    g.forward(p32)
    gm, gp, gw = g.backward(1.0)
    gm2, gp2, gw2 = g.backward(2.0, (gm.clone(), gp.clone(), gw.clone()))
    torch.cuda.synchronize()
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    wm, wp, ww = go.gradients(f(means), f(U), f(wl), f(pts))
    for got, want, name in ((gm, wm, "means"), (gp, wp, "prec"), (gw, ww, "weights_log")):
        got = got.cpu().numpy().astype(np.float64)
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"gmm grad {name}: max err / max |grad| = {err:.2e}")
        assert err < 2e-5, name
    assert np.allclose(gm2.cpu().numpy(), 3 * gm.cpu().numpy(), rtol=1e-6, atol=1e-9), "accumulates d_nll-scaled"
    assert np.all(np.tril(gp.cpu().numpy(), -1) == 0), "only the upper triangle is written"


def test_fit_lowers_nll(fea, scenes_mod):
    import torch
    from fuzzy_ellipses_amd.gmm import GMM, GMMFit
    rng = np.random.default_rng(9)
    t_means, t_U, t_wl = scenes_mod.random_gaussians(rng, 8)
    cov = np.linalg.inv(np.einsum("nij,nkj->nik", t_U, t_U))
    comp = rng.choice(8, size=20000, p=np.exp(t_wl) / np.exp(t_wl).sum())
    pts = t_means[comp] + np.einsum("mij,mj->mi", np.linalg.cholesky(cov[comp]), rng.normal(size=(20000, 3)))
    means, U, wl = scenes_mod.random_gaussians(np.random.default_rng(10), 16)
    fit = GMMFit(GMM(means, U * 0.5, wl, device="cuda:0"), pts, lr=2e-2)
    out = fit.run(150)
    print("gmm fit: nll %.4f -> %.4f (%.3f ms/step)" % (out["nll_first"], out["nll_last"], out["ms_per_step"]))
    assert out["nll_last"] < out["nll_first"] - 0.5
