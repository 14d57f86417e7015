"""Tile membership (CPU restatement oracle/fe_ref.cpp): conservativeness against brute force.

Phase 1 must list every (pixel, Gaussian) pair whose density reaches exp(-c_alpha); the
union of phase 1 and 2 must list every pair whose depth logit is within the depth margin of
the pixel's log-sum-exp.  Both are checked against float64 values of every pair.
"""
import dataclasses

import numpy as np
import pytest

from oracle import fm_oracle as fo
from oracle import ref

LOG2E = 1.4426950408889634


def _pairs(scene, cam):
    rays = scene_rays = None
    import fuzzy_ellipses_amd.scenes as sc
    rays = sc.camera_rays(cam) @ cam.R
    zs, d2 = fo.pair_terms(rays, np.repeat(cam.t[None], len(rays), 0), scene.means, scene.prec, scene.weights_log)
    return zs, d2


def _tile_of(cam, ts):
    ys, xs = np.divmod(np.arange(cam.H * cam.W), cam.W)
    tx = (cam.W + ts - 1) // ts
    return (ys // ts) * tx + xs // ts


def _member(counts, lst, T, n):
    m = np.zeros((T, n), bool)
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
    for t in range(T):
        ids = lst[off[t]:off[t + 1]]
        assert np.all(np.diff(ids.astype(np.int64)) > 0), "tile list not strictly ascending"
        m[t, ids] = True
    return m, off


@pytest.mark.parametrize("seed,ts", [(0, 8), (0, 16), (21, 32), (22, 16)])
def test_phase1_and_phase2_are_conservative(scenes_mod, seed, ts):
    sc = scenes_mod
    scene = sc.single_view_scene(seed=seed, n=60, H=64, W=80, f=64.0)
    cam = scene.cameras[0]
    zs, d2 = _pairs(scene, cam)
    c_alpha, delta2, _ = ref.cull_constants(scene.n)
    T = ((cam.W + ts - 1) // ts) * ((cam.H + ts - 1) // ts)
    counts, lst = ref.tiles(scene, cam, ts)
    assert counts.sum() == len(lst)
    m1, _ = _member(counts, lst, T, scene.n)
    tile = _tile_of(cam, ts)
    need_alpha = d2 >= -c_alpha + 0.5        # margin for float32 vs float64 bounds
    assert np.all(m1[tile][need_alpha]), "phase 1 misses a pair the alpha bound needs"
    # per-tile threshold from the phase-1 lists (float64), then phase 2
    l2 = LOG2E * (scene.beta3 * d2 - scene.beta2 * zs)
    listed = m1[tile]
    lm = np.where(listed & (zs > 0), l2, -np.inf)
    mx = lm.max(1)
    with np.errstate(invalid="ignore", divide="ignore"):
        lse = mx + np.log2(np.sum(np.exp2(lm - mx[:, None]), 1))
    lse[~np.isfinite(mx)] = -np.inf
    tile_lmin = np.full(T, np.inf)
    np.minimum.at(tile_lmin, tile, lse)
    counts2, lst2 = ref.tiles(scene, cam, ts, tile_lmin=tile_lmin.astype(np.float32))
    m2, _ = _member(counts2, lst2, T, scene.n)
    assert not np.any(m1 & m2), "phase 2 repeats a phase-1 pair"
    union = (m1 | m2)[tile]
    # the guarantee: every dropped unmasked pair's logit is below the log2-sum-exp of the
    # final listed set minus the margin (so dropped weight <= n * 2^-delta2 of the listed)
    lu = np.where(union & (zs > 0), l2, -np.inf)
    mu_ = lu.max(1)
    with np.errstate(invalid="ignore", divide="ignore"):
        lse_u = mu_ + np.log2(np.sum(np.exp2(lu - mu_[:, None]), 1))
    lse_u[~np.isfinite(mu_)] = -np.inf
    need_depth = (zs > 0) & (l2 >= lse_u[:, None] - delta2 + 0.5)
    assert np.all(union[need_depth]), "phase 1+2 miss a pair the depth bound needs"
    # pixels whose phase-1 coverage was empty still end up with a finite listed mass
    assert np.all(np.isfinite(lse_u[~np.isfinite(lse)])), "uncovered pixels stay uncovered"
    # and culling actually removes work on these scenes
    assert union.mean() < 1.0


def test_degenerate_gaussian_forces_dense(scenes_mod):
    sc = scenes_mod
    scene = sc.single_view_scene(seed=3, n=20, H=32, W=32, f=32.0)
    scene.prec[5, 1, 1] = 0.0
    cam = scene.cameras[0]
    rec, deg = ref.records(scene, cam)
    assert deg == 1
    counts, lst = ref.tiles(scene, cam, 16)
    assert np.all(counts == scene.n)


def _floor_scene(sc, kind):
    """Scenes where masked pairs hold large logits: cameras inside a cube of Gaussians and
    inside a shell of them (every ray meets Gaussians behind the camera), and a wide view."""
    if kind == "cube_inside":
        rng = np.random.default_rng(7)
        m, p, w = sc.random_gaussians(rng, 200)
        cam = sc.make_camera(np.array([0.12, -0.07, 0.18]), 64, 64, 40.0)
    elif kind == "shell_centre":
        rng = np.random.default_rng(8)
        m, p, w = sc.random_gaussians(rng, 512, shell=True)
        cam = sc.make_camera(np.array([0.02, -0.03, 0.05]), 64, 64, 40.0)
    elif kind == "shell_offc":
        rng = np.random.default_rng(8)
        m, p, w = sc.random_gaussians(rng, 512, shell=True)
        cam = sc.make_camera(np.array([0.3, -0.2, 0.4]), 64, 64, 40.0)
    else:
        rng = np.random.default_rng(9)
        m, p, w = sc.random_gaussians(rng, 300)
        cam = sc.make_camera(3.0 * sc.random_dir(rng), 128, 128, 120.0)
    return sc.Scene(m, p, w, [cam]), cam


@pytest.mark.parametrize("kind", ["cube_inside", "shell_centre", "shell_offc", "wide_fov"])
@pytest.mark.parametrize("ts", [8, 16])
def test_dense_completion_covers_the_weight_floor(scenes_mod, kind, ts):
    """Where phase 2 leaves a tile incomplete, the 1e-20 weight floor (summed over all n
    Gaussians in the reference's depth) is negligible at every pixel: the max logit over ALL
    pairs, masked ones included, stays below floor_thr above the log2-sum-exp of the listed
    unmasked pairs.  Phase-1 inputs (thresholds, floor seeds) are computed in float64 from the
    phase-1 lists, as the phase-1 blend computes them."""
    sc = scenes_mod
    scene, cam = _floor_scene(sc, kind)
    zs, d2 = _pairs(scene, cam)
    _, _, floor_thr = ref.cull_constants(scene.n)
    T = ((cam.W + ts - 1) // ts) * ((cam.H + ts - 1) // ts)
    tile = _tile_of(cam, ts)
    counts, lst = ref.tiles(scene, cam, ts)
    m1, _ = _member(counts, lst, T, scene.n)
    l2 = LOG2E * (scene.beta3 * d2 - scene.beta2 * zs)

    def lse_of(listed):
        mx = np.where(listed, l2, -np.inf).max(1)            # max over listed, masked included
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            s = np.sum(np.where(listed & (zs > 0), np.exp2(l2 - mx[:, None]), 0.0), 1)
            return mx, s, mx + np.log2(s)

    mx, s1, lse1 = lse_of(m1[tile])
    lse1 = np.where(s1 > 0, lse1, -np.inf)
    tile_lmin = np.full(T, np.inf)
    np.minimum.at(tile_lmin, tile, lse1)
    with np.errstate(divide="ignore"):
        gap = np.where(s1 > 0, -np.log2(np.where(s1 > 0, s1, 1.0)), -np.inf)
    tgap = np.full(T, -np.inf)
    np.maximum.at(tgap, tile, gap)
    tm0 = np.full(T, -np.inf)
    np.maximum.at(tm0, tile, np.where(s1 > 0, -np.inf, mx))
    seed = np.where(tgap >= floor_thr, np.inf, tm0)
    counts2, lst2 = ref.tiles(scene, cam, ts, tile_lmin=tile_lmin.astype(np.float32),
                              tile_seed=seed.astype(np.float32))
    m2, _ = _member(counts2, lst2, T, scene.n)
    union = m1 | m2
    complete = union.all(1)
    mtrue = l2.max(1)
    _, _, lse_u = lse_of(union[tile])
    excess = mtrue - lse_u - floor_thr                   # >= 0: the floor is visible there
    bad = (~complete[tile]) & ~(excess < 0.25)
    print(f"{kind} ts={ts}: {int(complete.sum())}/{T} tiles completed densely, "
          f"{int(bad.sum())} pixels with a visible floor on an incomplete tile")
    assert not np.any(bad)


@pytest.mark.parametrize("kind", ["cube_inside", "shell_centre", "wide_fov"])
@pytest.mark.parametrize("ts", [8, 16, 32])
def test_masked_logit_bound_holds(scenes_mod, kind, ts):
    """The membership's bound on the depth logit of a pair where it is masked (t* <= 0), which
    decides when a tile is completed densely for the 1e-20 floor, is at least the exact float64
    logit at every pixel of the tile where the pair is masked."""
    sc = scenes_mod
    scene, cam = _floor_scene(sc, kind)
    bad, checked = ref.masked_bound_check(scene, cam, ts)
    print(f"{kind} ts={ts}: {checked} masked (pixel, Gaussian) logits checked, {bad} violating pairs")
    assert checked > 0 and bad == 0


def _morton_f64(scene, cam, ray_view=False):
    """Z-order code of every projected centre in float64 (an independent statement of the
    record-order key: 10 bits per axis over [-W/2, 3W/2] x [-H/2, 3H/2] pixels, x bits even),
    with the distance of each scaled fraction to the nearest integer (where float32 rounding
    may land on the other side)."""
    p = (np.asarray(scene.means, np.float64) - np.asarray(cam.t, np.float64)) @ np.asarray(cam.R, np.float64).T
    z = p[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        if ray_view:
            u, v = (p[:, 0] / z + 2.0) / 4.0, (p[:, 1] / z + 2.0) / 4.0
        else:
            u = (cam.fx * p[:, 0] / z + cam.cx + cam.W / 2) / (2 * cam.W)
            v = (cam.fy * p[:, 1] / z + cam.cy + cam.H / 2) / (2 * cam.H)
    su, sv = np.clip(u, 0, 1) * 1023.0, np.clip(v, 0, 1) * 1023.0
    iu, iv = np.floor(su).astype(np.int64), np.floor(sv).astype(np.int64)
    code = np.zeros(scene.n, np.int64)
    for b in range(10):
        code |= ((iu >> b) & 1) << (2 * b) | ((iv >> b) & 1) << (2 * b + 1)
    code[~(z > 0)] = 0xFFFFF
    edge = np.minimum(np.abs(su - np.rint(su)), np.abs(sv - np.rint(sv)))
    return code, np.where(z > 0, edge, 1.0)


@pytest.mark.parametrize("which", ["shell", "single", "rays", "inside"])
def test_record_order_restatement(scenes_mod, which):
    """fe_ref_order (the renderer's record order) against the float64 statement: each chunk of
    1024 Gaussians is a permutation of its own indices, sorted by code then index, wherever the
    float64 code is not within rounding of a cell edge."""
    if which == "inside":     # camera inside a cube: centres behind it get the last code
        scene = scenes_mod.single_view_scene(seed=4, n=2500)
        cam = dataclasses.replace(scene.cameras[0], t=np.array([0.05, -0.02, 0.1]))
    elif which == "single":
        scene = scenes_mod.single_view_scene(seed=1, n=3000)
        cam = scene.cameras[0]
    else:
        scene = scenes_mod.shell_scene()
        cam = scene.cameras[3]
    order = ref.order(scene, cam, ray_view=(which == "rays"))
    code, edge = _morton_f64(scene, cam, ray_view=(which == "rays"))
    n = scene.n
    sure = edge > 1e-4
    for c0 in range(0, n, 1024):
        o = order[c0:c0 + 1024]
        np.testing.assert_array_equal(np.sort(o), np.arange(c0, min(n, c0 + 1024)))
        key = code[o] * 4096 + o
        ok = sure[o]
        ks = key[ok]
        assert np.all(np.diff(ks) > 0), "record order is not ascending in (code, index)"
    if which == "inside":
        assert np.any(code == 0xFFFFF) and np.any(code < 0xFFFFF)   # both kinds present
