"""torch.float64 autograd restatement of render_func_rays -- TEST INFRASTRUCTURE ONLY.

Stands in for "JAX autodiff of the baseline" (jax is not available): the same per-pair
math as baseline/fm_render.py:11-109 (and oracle/fm_oracle.py) written with torch ops so
reverse-mode autodiff yields reference gradients with respect to the means, the upper
triangle of the precision factor and the log-weights.  Only tests and the smoke check use it.
"""
from __future__ import annotations

import numpy as np
import torch


def render_rays_torch(rays, starts, means, prec_full, weights_log, beta_2, beta_3, features=None):
    """(depth, alpha) for rays (P,3) starting at starts (P,3); float64 torch tensors.  With
    ``features`` (N, C) also the feature image w @ features (P, C), w the normalised weights
    render_func_rays returns (fm_render.py:96-109)."""
    prc = torch.triu(prec_full).transpose(1, 2)                     # (N,3,3) lower
    p = means[None, :, :] - starts[:, None, :]                      # (P,N,3)
    projp = torch.einsum("nij,pnj->pni", prc, p)
    pr = torch.einsum("nij,pj->pni", prc, rays)
    vsv = (pr * pr).sum(-1)
    psv = (projp * pr).sum(-1)
    res = psv / vsv
    v = rays[:, None, :] * res[..., None] - p
    pv = torch.einsum("nij,pnj->pni", prc, v)
    d0 = (pv * pv).sum(-1)
    d2 = -0.5 * d0 + weights_log[None, :]
    alpha = 1 - torch.exp(-torch.exp(d2).sum(1))
    sig1 = (res > 0).to(res.dtype)
    logit = -res * beta_2 + beta_3 * d2
    se = torch.exp(logit - logit.max(1, keepdim=True).values)
    w = sig1 * torch.nan_to_num(se) + 1e-20
    w = w / w.sum(1, keepdim=True)
    depth = (w * torch.nan_to_num(res)).sum(1)
    if features is not None:
        return depth, alpha, w @ features
    return depth, alpha


def pinhole_rays(cam, pixels=None):
    H, W = cam.H, cam.W
    idx = np.arange(H * W) if pixels is None else np.asarray(pixels)
    x = (idx % W).astype(np.float64)
    y = (idx // W).astype(np.float64)
    rc = np.stack([(x - cam.cx) / cam.fx, (y - cam.cy) / cam.fy, np.ones_like(x)], 1)
    return rc @ np.asarray(cam.R, dtype=np.float64)


def feature_grads(scene, features, cams, d_depth, d_alpha, d_features, chunk=2048):
    """Gradients of sum_v <d_depth_v, depth_v> + <d_alpha_v, alpha_v> + <d_features_v, F_v>
    (F = w @ features) with respect to (means, prec upper triangle, weights_log, features),
    float64 autograd."""
    means = torch.tensor(scene.means, dtype=torch.float64, requires_grad=True)
    prec = torch.tensor(scene.prec, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(scene.weights_log, dtype=torch.float64, requires_grad=True)
    f = torch.tensor(np.asarray(features, dtype=np.float64), requires_grad=True)
    for v, cam in enumerate(cams):
        rays_all = pinhole_rays(cam)
        gd = np.asarray(d_depth[v], dtype=np.float64).reshape(-1)
        ga = np.asarray(d_alpha[v], dtype=np.float64).reshape(-1)
        gf = np.asarray(d_features[v], dtype=np.float64).reshape(rays_all.shape[0], -1)
        for s in range(0, rays_all.shape[0], chunk):
            e = min(rays_all.shape[0], s + chunk)
            rays = torch.tensor(rays_all[s:e])
            starts = torch.tensor(np.repeat(np.asarray(cam.t, dtype=np.float64)[None], e - s, 0))
            d, a, F = render_rays_torch(rays, starts, means, prec, w, scene.beta2, scene.beta3, features=f)
            loss = (d * torch.tensor(gd[s:e])).sum() + (a * torch.tensor(ga[s:e])).sum() + \
                (F * torch.tensor(gf[s:e])).sum()
            loss.backward()
    return means.grad.numpy(), np.triu(prec.grad.numpy()), w.grad.numpy(), f.grad.numpy()


def grads(scene, cams, d_depth, d_alpha, chunk=2048, pixels=None, dtype=torch.float64):
    """Gradients of sum_v <d_depth_v, depth_v> + <d_alpha_v, alpha_v> (float64).

    With ``pixels`` (one index array per view) only those pixels are rendered, and
    d_depth[v] / d_alpha[v] hold their cotangents (the others being zero): the render is
    separable per pixel, so this is the gradient of the full sum with zeros elsewhere.
    ``dtype=torch.float32`` evaluates the same autograd in float32, the precision the
    reference's JAX path differentiates in by default: its distance from the float64 result
    measures how well float32 can resolve these gradients at all (inputs, rays and the
    cotangents are rounded to float32 first).
    Returns (d_means (N,3), d_prec (N,3,3) -- upper triangle, d_weights_log (N,)) in float64."""
    def leaf(x):
        return torch.tensor(np.asarray(x, dtype=np.float64), dtype=dtype, requires_grad=True)
    means, prec, w = leaf(scene.means), leaf(scene.prec), leaf(scene.weights_log)
    for v, cam in enumerate(cams):
        rays_all = pinhole_rays(cam, None if pixels is None else pixels[v])
        gd = np.asarray(d_depth[v], dtype=np.float64).reshape(-1)
        ga = np.asarray(d_alpha[v], dtype=np.float64).reshape(-1)
        for s in range(0, rays_all.shape[0], chunk):
            e = min(rays_all.shape[0], s + chunk)
            rays = torch.tensor(rays_all[s:e], dtype=dtype)
            starts = torch.tensor(np.repeat(np.asarray(cam.t, dtype=np.float64)[None], e - s, 0), dtype=dtype)
            d, a = render_rays_torch(rays, starts, means, prec, w, scene.beta2, scene.beta3)
            loss = (d * torch.tensor(gd[s:e], dtype=dtype)).sum() + (a * torch.tensor(ga[s:e], dtype=dtype)).sum()
            loss.backward()
    return (means.grad.double().numpy(), np.triu(prec.grad.double().numpy()), w.grad.double().numpy())


def ray_grads(scene, ray_views, d_depth, d_alpha, dtype=torch.float64):
    """grads() for views given as explicit rays: ray_views = [(rays (m,3) world directions,
    origin (3,)), ...] with one cotangent pair per view (the ray-list views of the renderer;
    their rays are (x, y, 1) @ R of the pose)."""
    def leaf(x):
        return torch.tensor(np.asarray(x, dtype=np.float64), dtype=dtype, requires_grad=True)
    means, prec, w = leaf(scene.means), leaf(scene.prec), leaf(scene.weights_log)
    for (rays, origin), gd, ga in zip(ray_views, d_depth, d_alpha):
        r = torch.tensor(np.asarray(rays, dtype=np.float64), dtype=dtype)
        st = torch.tensor(np.repeat(np.asarray(origin, dtype=np.float64)[None], r.shape[0], 0), dtype=dtype)
        d, a = render_rays_torch(r, st, means, prec, w, scene.beta2, scene.beta3)
        loss = (d * torch.tensor(np.asarray(gd, dtype=np.float64), dtype=dtype)).sum() + \
            (a * torch.tensor(np.asarray(ga, dtype=np.float64), dtype=dtype)).sum()
        loss.backward()
    return (means.grad.double().numpy(), np.triu(prec.grad.double().numpy()), w.grad.double().numpy())
