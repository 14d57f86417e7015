"""The GMM log-likelihood oracle (restating baseline/fm_render.py:231-259), pinned against
scikit-learn's GaussianMixture, a closed form and the reference's own conventions."""
import math

import numpy as np
import pytest

from oracle import gmm_oracle as go


def _random_mixture(scenes_mod, seed, n):
    rng = np.random.default_rng(seed)
    means, U, w = scenes_mod.random_gaussians(rng, n)
    return means, U, w + rng.normal(size=n), rng


def test_matches_sklearn_score_samples(scenes_mod):
    from sklearn.mixture import GaussianMixture
    means, U, wl, rng = _random_mixture(scenes_mod, 0, 7)
    pts = rng.uniform(-0.7, 0.7, size=(200, 3))
    gm = GaussianMixture(n_components=7, covariance_type="full")
    w = np.exp(wl) / np.exp(wl).sum()
    gm.weights_, gm.means_, gm.precisions_cholesky_ = w, means, U
    gm.covariances_ = np.linalg.inv(np.einsum("nij,nkj->nik", U, U))
    want = gm.score_samples(pts)
    got = go.point_lse(means, U, wl, pts)
    assert np.max(np.abs(got - want)) < 1e-10
    nll, res = go.log_likelihood(means, U, wl, pts)
    assert res.shape == (7, 200)
    assert abs(nll + want.mean()) < 1e-10


def test_single_isotropic_closed_form():
    s = 0.3
    mu = np.array([[0.1, -0.2, 0.05]])
    U = np.eye(3)[None] / s
    pts = np.random.default_rng(1).normal(size=(50, 3))
    d2 = ((pts - mu) ** 2).sum(1)
    want = -0.5 * d2 / s ** 2 - 3 * math.log(s) - 1.5 * math.log(2 * math.pi)
    assert np.allclose(go.point_lse(mu, U, np.zeros(1), pts), want, atol=1e-12)


def test_lower_triangle_ignored_and_weight_shift_invariant(scenes_mod):
    means, U, wl, rng = _random_mixture(scenes_mod, 2, 5)
    pts = rng.uniform(-0.5, 0.5, size=(40, 3))
    full = U + np.tril(rng.normal(size=(5, 3, 3)), -1)
    a, _ = go.log_likelihood(means, U, wl, pts)
    b, _ = go.log_likelihood(means, full, wl + 3.7, pts)
    assert abs(a - b) < 1e-12


def test_gradient_restatement_matches_central_differences(scenes_mod):
    means, U, wl, rng = _random_mixture(scenes_mod, 3, 4)
    pts = rng.uniform(-0.5, 0.5, size=(30, 3))
    gm, gp, gw = go.gradients(means, U, wl, pts)
    eps = 1e-6
    f = lambda m, p, w: go.log_likelihood(m, p, w, pts)[0]
    for arr, g, idx in [(means, gm, (1, 2)), (U, gp, (2, 0, 1)), (U, gp, (3, 2, 2)), (wl, gw, (0,))]:
        a = arr.copy()
        a[idx] += eps
        args_p = [a if arr is x else x for x in (means, U, wl)]
        a2 = arr.copy()
        a2[idx] -= eps
        args_m = [a2 if arr is x else x for x in (means, U, wl)]
        fd = (f(*args_p) - f(*args_m)) / (2 * eps)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd))
