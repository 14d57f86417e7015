"""Pins the FP64 oracle (oracle/fm_oracle.py, oracle/fe_ref.cpp) without importing the reference.

The reference holds no golden vectors for this path, so each formula is checked against
what mathematics fixes: closed forms, a brute-force line search in long double, special
cases, invariants, and a textbook rotation.  Each check is chosen so that a plausible
mistake (transposed precision factor, dropped mask, wrong sign of beta2, missing weight
floor, max over unmasked only) fails it.
"""
import numpy as np
import pytest

from oracle import fm_oracle as fo
from oracle import ref

B2, B3 = 21.4, 2.66


def _rng(s=0):
    return np.random.default_rng(s)


def _iso(n, rng, s=0.1):
    means = rng.uniform(-0.5, 0.5, size=(n, 3))
    prec = np.repeat((np.eye(3) / s)[None], n, 0)
    w = rng.normal(size=n)
    return means, prec, w


def test_quat_to_rot_textbook(scenes_mod):
    from scipy.spatial.transform import Rotation
    rng = _rng(1)
    for _ in range(20):
        q = rng.normal(size=4)
        R = fo.quat_to_rot(q)
        Rt = Rotation.from_quat(q / np.linalg.norm(q), scalar_first=True).as_matrix()
        np.testing.assert_allclose(R, Rt, atol=1e-12)
        np.testing.assert_allclose(fo.quat_to_rot(3.7 * q), R, atol=1e-12)   # normalised by |q|^2
    np.testing.assert_array_equal(fo.quat_to_rot([0, 0, 0, 1e-7]), np.eye(3))


def test_isotropic_closed_form():
    rng = _rng(2)
    means, prec, w = _iso(7, rng, s=0.2)
    rays = rng.normal(size=(11, 3))
    o = rng.normal(size=3)
    zs, d2 = fo.pair_terms(rays, np.repeat(o[None], 11, 0), means, prec, w)
    p = means[None] - o[None, None]
    r = rays[:, None, :]
    t_ref = np.sum(p * r, -1) / np.sum(r * r, -1)
    d0_ref = np.sum(np.cross(p, r) ** 2, -1) / (np.sum(r * r, -1) * 0.2 ** 2)
    np.testing.assert_allclose(zs, t_ref, rtol=1e-12)
    np.testing.assert_allclose(d2, w[None] - 0.5 * d0_ref, rtol=1e-10, atol=1e-10)


def test_line_search_long_double(scenes_mod):
    """d0 is the minimum over t of the Mahalanobis distance^2 under Sigma = R diag(s^2) R^T;
    brute-force grid + golden-section search in long double, Sigma built from R and s
    directly (independent of the precision-factor convention)."""
    sc = scenes_mod
    rng = _rng(3)
    means = rng.uniform(-0.5, 0.5, size=(5, 3))
    rot = sc.random_rotations(rng, 5)
    s = np.exp(rng.uniform(np.log(0.03), np.log(0.15), size=(5, 3)))
    w = rng.normal(size=5)
    prec = sc.precision_factor(rot, s)
    sig_inv = np.linalg.inv(np.einsum("nij,nj,nkj->nik", rot, s ** 2, rot))
    o = np.array([0.3, -2.5, 0.4])
    rays = rng.normal(size=(6, 3)) * 0.2 + (-o)
    zs, d2 = fo.pair_terms(rays, np.repeat(o[None], 6, 0), means, prec, w)
    for i in range(6):
        for j in range(5):
            L = sig_inv[j].astype(np.longdouble)
            r = rays[i].astype(np.longdouble)
            p = (means[j] - o).astype(np.longdouble)
            ts = np.linspace(-1, 6, 20001).astype(np.longdouble)
            v = ts[:, None] * r[None] - p[None]
            f = np.einsum("ti,ij,tj->t", v, L, v)
            k = int(np.argmin(f))
            lo, hi = ts[max(k - 1, 0)], ts[min(k + 1, len(ts) - 1)]
            for _ in range(80):   # golden-section refinement
                m1 = lo + (hi - lo) * 0.381966
                m2 = lo + (hi - lo) * 0.618034
                f1 = (m1 * r - p) @ L @ (m1 * r - p)
                f2 = (m2 * r - p) @ L @ (m2 * r - p)
                if f1 < f2:
                    hi = m2
                else:
                    lo = m1
            tmin = (lo + hi) / 2
            dmin = (tmin * r - p) @ L @ (tmin * r - p)
            assert abs(float(tmin) - zs[i, j]) < 1e-8 * max(1.0, abs(zs[i, j]))
            assert abs(float(w[j] - 0.5 * dmin) - d2[i, j]) < 1e-8 * max(1.0, float(dmin))


def test_generator_precision_matches_covariance(scenes_mod):
    sc = scenes_mod
    rng = _rng(4)
    R = sc.random_rotations(rng, 6)
    s = np.exp(rng.uniform(np.log(0.03), np.log(0.15), size=(6, 3)))
    U = sc.precision_factor(R, s)
    assert np.allclose(U, np.triu(U))
    cov = np.einsum("nij,nj,nkj->nik", R, s ** 2, R)
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", U, U), np.linalg.inv(cov), rtol=1e-9, atol=1e-6)
    # an anisotropic factor is not symmetric, so a transposed use would change densities
    assert not np.allclose(U, np.transpose(U, (0, 2, 1)))


def test_single_gaussian_alpha_and_depth():
    rng = _rng(5)
    means, prec, w = _iso(1, rng, s=0.15)
    o = np.array([0.0, 0.0, -3.0])
    rays = rng.normal(size=(9, 3)) * 0.1 + np.array([0, 0, 1.0])
    zs, d2 = fo.pair_terms(rays, np.repeat(o[None], 9, 0), means, prec, w)
    d, a = fo.render_func_rays(rays, o[None], means, prec, w, B2, B3)
    np.testing.assert_allclose(a, 1 - np.exp(-np.exp(d2[:, 0])), rtol=1e-13)
    np.testing.assert_allclose(d, zs[:, 0], rtol=1e-13)   # t* > 0: softmax of one = t*


def test_behind_camera_and_all_masked():
    """A pixel whose only Gaussians have t* <= 0 gets the mean of t* (weight floor)."""
    means = np.array([[0.0, 0.0, -1.0], [0.1, 0.0, -2.0]])
    prec = np.repeat((np.eye(3) / 0.2)[None], 2, 0)
    w = np.array([-0.3, 0.2])
    rays = np.array([[0.0, 0.0, 1.0], [0.05, 0.0, 1.0]])
    zs, d2 = fo.pair_terms(rays, np.zeros((2, 3)), means, prec, w)
    assert np.all(zs < 0)
    d, a = fo.render_func_rays(rays, np.zeros((1, 3)), means, prec, w, B2, B3)
    np.testing.assert_allclose(d, zs.mean(1), rtol=1e-12)
    np.testing.assert_allclose(a, 1 - np.exp(-np.exp(d2).sum(1)), rtol=1e-12)


def test_two_gaussian_softmax_and_max_over_all():
    """Closed form D = (sum s z + 1e-20 sum z') / (sum s + 1e-20 n) with the max over ALL
    logits (a masked Gaussian with the largest logit shrinks s but cancels except for the
    floor)."""
    means = np.array([[0.0, 0.0, 1.0], [0.02, 0.0, 1.3], [0.0, 0.0, -0.5]])
    prec = np.repeat((np.eye(3) / 0.3)[None], 3, 0)
    w = np.array([0.1, -0.2, 3.0])
    rays = np.array([[0.0, 0.0, 1.0], [0.01, 0.0, 1.0]])
    zs, d2 = fo.pair_terms(rays, np.zeros((2, 3)), means, prec, w)
    l = -B2 * zs + B3 * d2
    L = l.max(1, keepdims=True)
    s = (zs > 0) * np.exp(l - L)
    dref = ((s * zs).sum(1) + 1e-20 * zs.sum(1)) / (s.sum(1) + 1e-20 * 3)
    d, _ = fo.render_func_rays(rays, np.zeros((1, 3)), means, prec, w, B2, B3)
    np.testing.assert_allclose(d, dref, rtol=1e-12)
    # the masked third Gaussian has the largest logit: its floor shift is visible
    assert np.all(np.argmax(l, 1) == 2)


def test_invariants(scenes_mod):
    sc = scenes_mod
    scene = sc.single_view_scene(seed=7, H=16, W=16, f=16.0)
    cam = scene.cameras[0]
    rays = sc.camera_rays(cam)
    d0, a0 = fo.render_camera(scene.means, scene.prec, scene.weights_log, cam.R, cam.t, rays, B2, B3)
    # translation of means and camera together
    sh = np.array([0.3, -1.2, 2.0])
    d1, a1 = fo.render_camera(scene.means + sh, scene.prec, scene.weights_log, cam.R, cam.t + sh, rays, B2, B3)
    np.testing.assert_allclose(d1, d0, rtol=1e-9)
    np.testing.assert_allclose(a1, a0, atol=1e-12)
    # rotation Q: mu -> Q mu, Sigma -> Q Sigma Q^T, camera R -> R Q^T, t -> Q t
    Q = sc.random_rotations(_rng(8), 1)[0]
    Ui = np.linalg.inv(np.transpose(scene.prec, (0, 2, 1)))    # C with C C^T = Sigma
    cov = np.einsum("nij,nkj->nik", Ui, Ui)
    cov_q = np.einsum("ij,njk,lk->nil", Q, cov, Q)
    Uq = np.triu(np.transpose(np.linalg.inv(np.linalg.cholesky(cov_q)), (0, 2, 1)))
    d2_, a2_ = fo.render_camera(scene.means @ Q.T, Uq, scene.weights_log, cam.R @ Q.T, Q @ cam.t, rays, B2, B3)
    np.testing.assert_allclose(d2_, d0, rtol=1e-9)
    np.testing.assert_allclose(a2_, a0, atol=1e-12)
    # permutation of Gaussians
    perm = _rng(9).permutation(scene.n)
    d3, a3 = fo.render_camera(scene.means[perm], scene.prec[perm], scene.weights_log[perm], cam.R, cam.t,
                              rays, B2, B3)
    np.testing.assert_allclose(d3, d0, rtol=1e-12)
    np.testing.assert_allclose(a3, a0, rtol=1e-12)
    # log-weight shift: depth invariant, E scales by e^c
    d4, a4 = fo.render_camera(scene.means, scene.prec, scene.weights_log + 1.5, cam.R, cam.t, rays, B2, B3)
    np.testing.assert_allclose(d4, d0, rtol=1e-10)
    np.testing.assert_allclose(-np.log1p(-a4), -np.log1p(-a0) * np.exp(1.5), rtol=1e-6, atol=1e-14)
    # ray scaling r -> k r: alpha invariant, t* -> t*/k
    k = 2.5
    zs0, _ = fo.pair_terms(rays @ cam.R, np.repeat(cam.t[None], len(rays), 0), scene.means, scene.prec,
                           scene.weights_log)
    zs1, _ = fo.pair_terms(k * rays @ cam.R, np.repeat(cam.t[None], len(rays), 0), scene.means, scene.prec,
                           scene.weights_log)
    np.testing.assert_allclose(zs1, zs0 / k, rtol=1e-10)
    _, a5 = fo.render_camera(scene.means, scene.prec, scene.weights_log, cam.R, cam.t, k * rays, B2, B3)
    np.testing.assert_allclose(a5, a0, atol=1e-12)


def test_lower_triangle_ignored(scenes_mod):
    sc = scenes_mod
    scene = sc.single_view_scene(seed=10, H=8, W=8, f=8.0)
    cam = scene.cameras[0]
    rays = sc.camera_rays(cam)
    junk = scene.prec + np.tril(_rng(11).normal(size=(scene.n, 3, 3)), -1)
    d0, a0 = fo.render_camera(scene.means, scene.prec, scene.weights_log, cam.R, cam.t, rays, B2, B3)
    d1, a1 = fo.render_camera(scene.means, junk, scene.weights_log, cam.R, cam.t, rays, B2, B3)
    np.testing.assert_array_equal(d0, d1)
    np.testing.assert_array_equal(a0, a1)


def test_c_oracle_matches_numpy(scenes_mod):
    sc = scenes_mod
    scene = sc.single_view_scene()
    cam = scene.cameras[0]
    d1, a1 = fo.render_camera(scene.means, scene.prec, scene.weights_log, cam.R, cam.t, sc.camera_rays(cam),
                              scene.beta2, scene.beta3)
    d2, a2 = ref.render_f64(scene, cam)
    np.testing.assert_allclose(d2, d1, rtol=1e-12)
    np.testing.assert_allclose(a2, a1, rtol=1e-12, atol=1e-15)
    # the float32 CPU port stays inside the tolerance of the path
    d3, a3 = ref.render_f32(scene, cam)
    assert np.max(np.abs(d3 - d1) / np.abs(d1)) < 1e-4
    assert np.max(np.abs(a3 - a1)) < 1e-4


def test_golden_fixture(scenes_mod):
    """Regression against the committed FP64 fixture of the 40-Gaussian 64x64 scene."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "single_view_seed0.npz")
    g = np.load(path)
    scene = scenes_mod.single_view_scene()
    np.testing.assert_array_equal(scene.means, g["means"])
    np.testing.assert_array_equal(scene.prec, g["prec"])
    d, a = ref.render_f64(scene, scene.cameras[0])
    np.testing.assert_allclose(d, g["depth"], rtol=1e-12)
    np.testing.assert_allclose(a, g["alpha"], rtol=1e-12, atol=1e-15)


def test_weights_output_closed_forms():
    """render_func_rays' fourth output w (fm_render.py:96-109): columns sum to one, depth is
    w @ nan_to_num(t*), two unmasked Gaussians give the closed-form softmax of their logits,
    and an all-masked ray gives the uniform 1e-20 floor (1/N each)."""
    rng = np.random.default_rng(5)
    n = 6
    means = rng.uniform(-0.3, 0.3, size=(n, 3)) + np.array([0.0, 0.0, 3.0])
    U = np.tile(np.eye(3) * 4.0, (n, 1, 1))
    wl = rng.normal(size=n)
    rays = np.concatenate([rng.uniform(-0.1, 0.1, size=(7, 2)), np.ones((7, 1))], 1)
    starts = np.zeros((7, 3))
    w = fo.render_weights(rays, starts, means, U, wl, 21.4, 2.66)
    np.testing.assert_allclose(w.sum(1), 1.0, rtol=1e-14)
    zs, d2 = fo.pair_terms(rays, starts, means, U, wl)
    d, _ = fo.render_func_rays(rays, starts, means, U, wl, 21.4, 2.66)
    np.testing.assert_allclose((w * np.nan_to_num(zs)).sum(1), d, rtol=1e-12)
    # two Gaussians in front: softmax of l = b3 d2 - b2 t (the floor is 1e-20 of it)
    two = fo.render_weights(rays[:1], starts[:1], means[:2], U[:2], wl[:2], 21.4, 2.66)
    l = 2.66 * d2[0, :2] - 21.4 * zs[0, :2]
    e = np.exp(l - l.max())
    np.testing.assert_allclose(two[0], e / e.sum(), rtol=1e-12)
    # every Gaussian behind the ray start: only the floor, uniform
    back = fo.render_weights(-rays[:1], starts[:1], means, U, wl, 21.4, 2.66)
    np.testing.assert_allclose(back[0], np.full(n, 1.0 / n), rtol=1e-12)
