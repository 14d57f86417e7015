"""FP64 CPU restatement of the reference fuzzy-metaball renderer -- TEST INFRASTRUCTURE ONLY.

This module is the correctness oracle.  Only tests, the smoke check and the CPU-baseline
leg of the benchmark may import it, and only as the checker.  Nothing in the product
(the HIP path behind the C ABI) calls it.

It restates, in numpy float64, the dense per-(ray, Gaussian) math of
``baseline/fm_render.py:11-109`` (``render_func_rays``) and the pose wrapper
``baseline/fm_render.py:113-126`` (``render_func_quat``) together with
``baseline/util_render.py:5-9`` (``jax_stable_exp``) and ``:51-76`` (``quat_to_rot``).
Work is chunked over pixels; every reduction runs over the Gaussian axis, so chunking
does not change the result.

Pinning (tests/test_oracle.py): closed forms for isotropic Gaussians, a brute-force
line search for d0 in long double, single- and two-Gaussian closed forms for alpha and
depth, the all-masked mean, invariants (translation, rotation, permutation, log-weight
shift, ray scaling) and the quaternion map against a textbook rotation.
"""
from __future__ import annotations

import numpy as np


def quat_to_rot(q) -> np.ndarray:
    """Scalar-first quaternion to rotation, normalised by |q|^2; identity if |q|^2 <= 1e-12.

    Follows baseline/util_render.py:51-76.
    """
    w, x, y, z = [float(v) for v in q]
    nq = w * w + x * x + y * y + z * z
    if not nq > 1e-12:
        return np.eye(3)
    s = 2.0 / nq
    X, Y, Z = x * s, y * s, z * s
    wX, wY, wZ = w * X, w * Y, w * Z
    xX, xY, xZ = x * X, x * Y, x * Z
    yY, yZ, zZ = y * Y, y * Z, z * Z
    return np.array([
        [1.0 - (yY + zZ), xY - wZ, xZ + wY],
        [xY + wZ, 1.0 - (xX + zZ), yZ - wX],
        [xZ - wY, yZ + wX, 1.0 - (xX + yY)],
    ])


def pair_terms(rays, starts, means, prec_full, weights_log):
    """Per (ray, Gaussian) t* ("zs") and d2 ("stds"), each (P, N).

    baseline/fm_render.py:22-79: prc = triu(prec)^T, p = mean - start, projp = prc p,
    vsv = |prc r|^2, psv = projp . prc r, res = psv / vsv, v = r res - p,
    d0 = |prc v|^2, d2 = -0.5 d0 + w.
    """
    prc = np.transpose(np.triu(np.asarray(prec_full, dtype=np.float64)), (0, 2, 1))  # (N,3,3) lower
    r = np.asarray(rays, dtype=np.float64)
    t = np.asarray(starts, dtype=np.float64)
    mu = np.asarray(means, dtype=np.float64)
    w = np.asarray(weights_log, dtype=np.float64)
    p = mu[None, :, :] - t[:, None, :]                       # (P,N,3)
    projp = np.einsum("nij,pnj->pni", prc, p)
    pr = np.einsum("nij,pj->pni", prc, r)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        vsv = np.sum(pr * pr, axis=-1)
        psv = np.sum(projp * pr, axis=-1)
        res = psv / vsv
        v = r[:, None, :] * res[..., None] - p
        pv = np.einsum("nij,pnj->pni", prc, v)
        d0 = np.sum(pv * pv, axis=-1)
        d2 = -0.5 * d0 + w[None, :]
    return res, d2


def combine(zs, stds, beta_2, beta_3):
    """Per-ray reduction over the Gaussian axis (axis 1): baseline/fm_render.py:88-109.

    alpha = 1 - exp(-sum exp(d2)); weights = (t*>0) nan_to_num(stable_exp(-b2 t* + b3 d2))
    + 1e-20, normalised; depth = sum weights * nan_to_num(t*).
    """
    with np.errstate(over="ignore", invalid="ignore"):
        alpha = 1.0 - np.exp(-np.sum(np.exp(stds), axis=1))
        sig1 = zs > 0
        logit = -zs * beta_2 + beta_3 * stds
        se = np.exp(logit - np.max(logit, axis=1, keepdims=True))   # util_render.py:5-9
        wgt = sig1 * np.nan_to_num(se) + 1e-20
        wgt = wgt / np.sum(wgt, axis=1, keepdims=True)
        depth = np.sum(wgt * np.nan_to_num(zs), axis=1)
    return depth, alpha


def weights(zs, stds, beta_2, beta_3):
    """The normalised per-(ray, Gaussian) weights w that render_func_rays returns as its fourth
    output (baseline/fm_render.py:93-99), (P, N): (t*>0) nan_to_num(stable_exp(l)) + 1e-20 with
    l = -b2 t* + b3 d2, normalised over the Gaussians."""
    with np.errstate(over="ignore", invalid="ignore"):
        sig1 = zs > 0
        logit = -zs * beta_2 + beta_3 * stds
        se = np.exp(logit - np.max(logit, axis=1, keepdims=True))   # util_render.py:5-9
        wgt = sig1 * np.nan_to_num(se) + 1e-20
        return wgt / np.sum(wgt, axis=1, keepdims=True)


def render_weights(rays, starts, means, prec_full, weights_log, beta_2, beta_3):
    """(P, N) weights w of P rays (fm_render.py:96-109); features f (N, C) give w @ f."""
    rays = np.asarray(rays, dtype=np.float64)
    starts = np.broadcast_to(np.asarray(starts, dtype=np.float64), rays.shape)
    zs, d2 = pair_terms(rays, starts, means, prec_full, weights_log)
    return weights(zs, d2, beta_2, beta_3)


def render_func_rays(rays, starts, means, prec_full, weights_log, beta_2, beta_3, chunk=None):
    """Dense FP64 render of P rays -> (depth (P,), alpha (P,)).

    ``rays`` and ``starts`` are (P, 3); the reference packs them as (P, 2, 3)
    (baseline/fm_render.py:11, :116-120).
    """
    rays = np.asarray(rays, dtype=np.float64)
    starts = np.broadcast_to(np.asarray(starts, dtype=np.float64), rays.shape)
    P = rays.shape[0]
    N = np.asarray(means).shape[0]
    if chunk is None:
        chunk = max(1, int(2_000_000 // max(N, 1)))
    depth = np.empty(P)
    alpha = np.empty(P)
    for s in range(0, P, chunk):
        e = min(P, s + chunk)
        zs, d2 = pair_terms(rays[s:e], starts[s:e], means, prec_full, weights_log)
        depth[s:e], alpha[s:e] = combine(zs, d2, beta_2, beta_3)
    return depth, alpha


def render_func_quat(means, prec_full, weights_log, camera_rays, quat, t, beta_2, beta_3):
    """baseline/fm_render.py:113-126: world rays = camera_rays @ R(quat), all starting at t."""
    R = quat_to_rot(quat)
    rays = np.asarray(camera_rays, dtype=np.float64) @ R
    return render_func_rays(rays, np.asarray(t, dtype=np.float64)[None, :], means, prec_full,
                            weights_log, beta_2, beta_3)


def render_camera(means, prec_full, weights_log, R, t, camera_rays, beta_2, beta_3, pixels=None):
    """Render with an explicit rotation R (rows = camera axes): rays = camera_rays @ R.

    ``pixels`` optionally selects a subset of pixel indices (the render is separable
    per pixel, so a subset is exact).
    """
    cr = np.asarray(camera_rays, dtype=np.float64)
    if pixels is not None:
        cr = cr[pixels]
    rays = cr @ np.asarray(R, dtype=np.float64)
    return render_func_rays(rays, np.asarray(t, dtype=np.float64)[None, :], means, prec_full,
                            weights_log, beta_2, beta_3)


def depth_condition(rays, starts, means, prec_full, weights_log, beta_2, beta_3, t_error=True):
    """First-order sensitivity of each ray's depth to float32 rounding of the per-pair math,
    kappa (P,): a float32 evaluation of render_func_rays, by any arrangement of its
    operations, moves depth by about 2^-24 * (a small constant) * kappa, however exact the
    evaluation is otherwise.  Test infrastructure: it sets how far float32 can be expected to
    follow the float64 result where the 1e-4 relative depth budget is out of float32's reach
    (logits of thousands: pixels far, in whitened units, from every Gaussian; t* near 0 on
    rays almost perpendicular to a Gaussian's direction).

    With s_j the unmasked softmax weights and f the 1e-20 floor (fm_render.py:93-99), depth
    D = sum (s_j + f) z_j / sum (s_j + f), z = nan_to_num(t*).  A logit error dl_j moves D by
    s_j (z_j - D) dl_j, an error in the maximum logit (util_render.py:5-9) by
    -sum_j s_j (z_j - D) dl_max, and an error dz_j by (s_j + f) dz_j.  The error scales are
      dz_j  = |q_j| / |A_j r|           (t* = (q.Ar)/|Ar|^2, q = A(mu - o): the cancellation
                                         in q.Ar when the ray is nearly perpendicular to q)
      dl_j  = b3 (|w_j| + d0_j + |q_j| sqrt(d0_j)) + b2 dz_j
                                        (d2 = w - d0/2, d0 = |A(t* r - p)|^2: rounding of
                                         t* r - p at the scale of |q|)
    and kappa = sum s_j |z_j - D| dl_j + sum (s_j + f) dz_j + |sum s_j (z_j - D)| dl_max,
    all in float64 from baseline/fm_render.py:22-99's quantities.  ``t_error=False`` leaves
    the t* terms out (dz = 0); the tests use it to show that a kappa missing them fails.
    """
    rays = np.asarray(rays, dtype=np.float64)
    starts = np.broadcast_to(np.asarray(starts, dtype=np.float64), rays.shape)
    mu = np.asarray(means, dtype=np.float64)
    w = np.asarray(weights_log, dtype=np.float64)
    zs, d2 = pair_terms(rays, starts, mu, prec_full, w)
    prc = np.transpose(np.triu(np.asarray(prec_full, dtype=np.float64)), (0, 2, 1))
    q = np.einsum("nij,pnj->pni", prc, mu[None, :, :] - starts[:, None, :])
    qn = np.sqrt(np.sum(q * q, axis=-1))                                      # (P,N)
    pr = np.einsum("nij,pj->pni", prc, rays)
    with np.errstate(all="ignore"):
        dz = qn / np.sqrt(np.sum(pr * pr, axis=-1)) if t_error else np.zeros_like(qn)
        d0 = np.maximum(2.0 * (w[None, :] - d2), 0.0)
        logit = -zs * beta_2 + beta_3 * d2
        lmax = np.max(np.where(np.isnan(logit), -np.inf, logit), axis=1, keepdims=True)
        se = (zs > 0) * np.nan_to_num(np.exp(logit - lmax))
        den = np.sum(se, axis=1, keepdims=True) + 1e-20 * zs.shape[1]
        s = se / den
        f = 1e-20 / den
        z = np.nan_to_num(zs)
        D = np.sum((s + f) * z, axis=1, keepdims=True)
        dl = beta_3 * (np.abs(w)[None, :] + d0 + qn * np.sqrt(d0)) + beta_2 * dz
        am = np.argmax(np.where(np.isnan(logit), -np.inf, logit), axis=1)
        dl_max = dl[np.arange(dl.shape[0]), am]
        return (np.sum(s * np.abs(z - D) * dl, axis=1) + np.sum((s + f) * dz, axis=1)
                + np.abs(np.sum(s * (z - D), axis=1)) * dl_max)


def idx_quattrans_rays(pixel_posei, invF, poses):
    """World rays and starts of baseline/fm_render.py:151-166: per ray, the camera-frame ray
    (x invF, y invF, z) @ R(pose), starting at the pose's translation."""
    pp = np.asarray(pixel_posei, dtype=np.float64)
    ps = np.asarray(poses, dtype=np.float64)
    rot = np.stack([quat_to_rot(q) for q in ps[:, :4]])
    idx = pp[:, 3].astype(np.int64)
    cam = pp[:, :3] * np.array([invF, invF, 1.0])
    rays = np.einsum("rj,rjk->rk", cam, rot[idx])
    return rays, ps[idx, 4:], rot


def render_func_idx_quattrans(means, prec_full, weights_log, pixel_posei, invF, poses, beta_2, beta_3):
    """(depth, alpha) of baseline/fm_render.py:151-166 (arguments passed to render_func_rays
    in its own order)."""
    rays, starts, _ = idx_quattrans_rays(pixel_posei, invF, poses)
    return render_func_rays(rays, starts, means, prec_full, weights_log, beta_2, beta_3)


def idx_quattrans_flow(pixel_posei, invF, poses, depth):
    """(flow_next, flow_prev), each (R, 2), of baseline/fm_render.py:186-221."""
    pp = np.asarray(pixel_posei, dtype=np.float64)
    ps = np.asarray(poses, dtype=np.float64)
    rot = np.stack([quat_to_rot(q) for q in ps[:, :4]])
    P = ps.shape[0]
    idx = pp[:, 3].astype(np.int64)
    ip = np.minimum(idx + 1, P - 1)
    im = np.maximum(idx - 1, 0)
    cam = pp[:, :3] * np.array([invF, invF, 1.0]) * np.asarray(depth, dtype=np.float64)[:, None]
    world = np.einsum("rj,rjk->rk", cam, rot[idx]) + ps[idx, 4:]

    def proj(j):
        c = np.einsum("rj,rkj->rk", world - ps[j, 4:], rot[j])
        return -(c[:, :2] / (c[:, 2:3] * invF) - pp[:, :2])
    return proj(ip), proj(im)


def nan_to_num_outputs(depth, alpha):
    """Output clean-up of the reference binding (fuzzy_ellipses/__init__.py:233-234)."""
    return (np.nan_to_num(depth, posinf=10.0, neginf=0.0),
            np.nan_to_num(alpha, posinf=1.0, neginf=0.0))
