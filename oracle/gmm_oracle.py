"""FP64 restatement of the reference GMM log-likelihood -- TEST INFRASTRUCTURE ONLY.

Only tests may import this module, as the checker of the HIP kernels behind
fe_gmm_forward / fe_gmm_backward.  It restates baseline/fm_render.py:231-259
(``log_likelihood(params, points)``):

* prec = triu(prec_full); the factor used is its transpose (fm_render.py:233,238);
* weights = exp(weights_log) normalised to sum 1 (fm_render.py:234-235);
* per Gaussian, div = prod |diag(prec)| (fm_render.py:239);
* res[i, j] = -0.5 (|prec_i^T (mu_i - x_j)|^2 + 3 ln 2pi) + ln w_i + ln div_i (fm_render.py:241-249);
* the loss is -mean_j logsumexp_i res[i, j] (fm_render.py:257-259).

Pinning (tests/test_gmm_cpu.py): per-point log-likelihoods equal scikit-learn's
GaussianMixture.score_samples with the same means, weights and precision Cholesky factors
(an independent implementation of the textbook density); a single isotropic Gaussian has
the closed form -0.5 |x - mu|^2 / s^2 - 3 ln s - 1.5 ln 2pi; the lower triangle is ignored;
a common shift of the log-weights leaves the loss unchanged.  The gradient restatement is
torch float64 autograd of the same formula (checked against central differences).
"""
from __future__ import annotations

import math

import numpy as np


def log_likelihood(means, prec_full, weights_log, points):
    """(nll, res) in float64: nll = -mean_j logsumexp_i res[i, j], res (N, M)."""
    means = np.asarray(means, np.float64)
    U = np.triu(np.asarray(prec_full, np.float64))
    wl = np.asarray(weights_log, np.float64)
    pts = np.asarray(points, np.float64)
    w = np.exp(wl - wl.max())
    w = w / w.sum()
    logdiv = np.log(np.abs(np.diagonal(U, axis1=1, axis2=2))).sum(1)        # (N,)
    p = means[:, None, :] - pts[None, :, :]                                  # (N,M,3)
    z = np.einsum("nab,nma->nmb", U, p)                                      # U^T p
    res = -0.5 * ((z * z).sum(-1) + 3.0 * math.log(2.0 * math.pi)) + (np.log(w) + logdiv)[:, None]
    mx = res.max(0)
    lse = mx + np.log(np.exp(res - mx[None, :]).sum(0))
    return -lse.mean(), res


def point_lse(means, prec_full, weights_log, points):
    _, res = log_likelihood(means, prec_full, weights_log, points)
    mx = res.max(0)
    return mx + np.log(np.exp(res - mx[None, :]).sum(0))


def nll_torch(means, prec_full, weights_log, points):
    """The same loss as a torch float64 expression (for autograd gradients)."""
    import torch
    U = torch.triu(prec_full)
    logw = weights_log - torch.logsumexp(weights_log, 0)
    logdiv = torch.log(torch.abs(torch.diagonal(U, dim1=1, dim2=2))).sum(1)
    p = means[:, None, :] - points[None, :, :]
    z = torch.einsum("nab,nma->nmb", U, p)
    res = -0.5 * ((z * z).sum(-1) + 3.0 * math.log(2.0 * math.pi)) + (logw + logdiv)[:, None]
    return -torch.logsumexp(res, 0).mean()


def gradients(means, prec_full, weights_log, points):
    """d nll / d(means, prec (upper triangle; lower entries zero), weights_log), float64."""
    import torch
    m = torch.tensor(np.asarray(means, np.float64), requires_grad=True)
    P = torch.tensor(np.triu(np.asarray(prec_full, np.float64)), requires_grad=True)
    w = torch.tensor(np.asarray(weights_log, np.float64), requires_grad=True)
    x = torch.tensor(np.asarray(points, np.float64))
    nll_torch(m, P, w, x).backward()
    return m.grad.numpy(), np.triu(P.grad.numpy()), w.grad.numpy()
