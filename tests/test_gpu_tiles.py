"""Bit-exact tile keys on the GPU: the record order, per-Gaussian records, per-tile ranges and
sorted Gaussian lists of both culling phases equal the CPU restatement (oracle/fe_ref.cpp),
computed from the inputs alone (phase 2 given the GPU's per-tile float thresholds)."""
import numpy as np
import pytest

from conftest import in_record_order
from oracle import ref

pytestmark = pytest.mark.gpu


def _check(fea, scene, cam, ts):
    import torch
    r = fea.Renderer(tile=ts, batch=64, mode="culled")
    r.forward([fea.SceneTensors.from_scene(scene)], [(0, cam, scene.beta2, scene.beta3)])
    torch.cuda.synchronize()
    st = r.debug_state()
    # the record order (Z order of projected centres in chunks of 1024), restated on the CPU
    cpu_order = ref.order(scene, cam)
    np.testing.assert_array_equal(st["order"], cpu_order)
    # the GPU lists record ids; the restatement runs on the Gaussians in the CPU's order
    scene = in_record_order(scene, cpu_order)
    rec_cpu, _ = ref.records(scene, cam)
    np.testing.assert_array_equal(st["records"].view(np.uint32), rec_cpu.view(np.uint32))
    off1, off2, lst = st["off1"], st["off2"], st["list"]
    T = len(off1) - 1
    c1, l1 = ref.tiles(scene, cam, ts)
    np.testing.assert_array_equal(np.diff(off1), c1.astype(np.int64))
    np.testing.assert_array_equal(lst[:off1[-1]], l1)
    # phase 2 given the GPU's per-tile thresholds (they come from the phase-1 blend)
    c2, l2 = ref.tiles(scene, cam, ts, tile_lmin=st["tile_lmin"], tile_seed=st["tile_seed"])
    np.testing.assert_array_equal(np.diff(off2), c2.astype(np.int64))
    assert off2[0] == off1[-1]
    np.testing.assert_array_equal(lst[off2[0]:off2[-1]], l2)
    # ranges partition [0, K) and every list is strictly ascending
    for t in range(T):
        seg = lst[off1[t]:off1[t + 1]].astype(np.int64)
        assert np.all(np.diff(seg) > 0)
    _check_order(st)
    return int(off1[-1]), int(off2[-1] - off1[-1])


def _check_order(st):
    """The blends' tile dispatch orders: permutations of the tiles, longest list first up to
    the bucket resolution (a key may exceed the one before it by at most 1/8 of it, plus 1)."""
    off1, off2, order = st["off1"], st["off2"], st["tile_order"]
    T = len(off1) - 1
    n1 = np.diff(off1)
    for row, key in ((0, n1), (1, n1 + np.diff(off2))):
        o = order[row].astype(np.int64)
        np.testing.assert_array_equal(np.sort(o), np.arange(T))
        k = key[o].astype(np.float64)
        assert np.all(k[1:] <= k[:-1] * 1.125 + 1.0), f"order {row} is not longest-first"


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_keys_single_view(fea, scenes_mod, ts):
    scene = scenes_mod.single_view_scene()
    k1, k2 = _check(fea, scene, scene.cameras[0], ts)
    print(f"single view ts={ts}: phase-1 pairs {k1}, phase-2 pairs {k2}")


@pytest.mark.parametrize("view", [0, 7])
def test_keys_shell_view(fea, scenes_mod, view):
    scene = scenes_mod.shell_scene()
    k1, k2 = _check(fea, scene, scene.cameras[view], 16)
    print(f"shell view {view}: phase-1 pairs {k1}, phase-2 pairs {k2}")


@pytest.mark.parametrize("ts,H,W,f", [(16, 128, 128, 120.0), (16, 240, 320, 300.0), (8, 128, 128, 120.0)])
def test_keys_chunked_membership(fea, scenes_mod, ts, H, W, f):
    """One small view of N=4096 (9, 35 and 16 membership groups): each group's Gaussians are
    shared out over several blocks, whose partial counts, flags and bounds a finishing kernel
    merges, and the phase-2 lower bounds of tiles without a phase-1 threshold come from a
    chunked pre-pass.  The lists still equal the restatement."""
    rng = np.random.default_rng(31)
    means, prec, w = scenes_mod.random_gaussians(rng, 4096)
    cam = scenes_mod.make_camera(3.0 * scenes_mod.random_dir(rng), H, W, f)
    scene = scenes_mod.Scene(means, prec, w, [cam])
    k1, k2 = _check(fea, scene, cam, ts)
    print(f"chunked membership ts={ts} {H}x{W}: phase-1 pairs {k1}, phase-2 pairs {k2}")


def test_dispatch_order_all_shell_views(fea, scenes_mod):
    """With 16 views in one call the orders run over every view's tiles."""
    import torch
    scene = scenes_mod.shell_scene()
    r = fea.Renderer(tile=16, batch=64, mode="culled")
    r.forward([fea.SceneTensors.from_scene(scene)], [(0, c, scene.beta2, scene.beta3) for c in scene.cameras])
    torch.cuda.synchronize()
    st = r.debug_state()
    _check_order(st)
    k = np.diff(st["off1"])[st["tile_order"][0]]
    print(f"16 views: {len(k)} tiles, phase-1 lengths in dispatch order {k[:3]} ... {k[-3:]}")


def test_keys_and_order_on_81920_tiles(fea, scenes_mod):
    """81920 tiles (a 2048x2560 view at tile 8; the scan's blocks span several buckets and
    views' worth of tiles): lists, ranges and both dispatch orders as for small batches."""
    scene = scenes_mod.single_view_scene(seed=5, n=64, H=2048, W=2560, f=2000.0)
    cam = scene.cameras[0]
    assert (cam.H // 8) * (cam.W // 8) > 65536
    k1, k2 = _check(fea, scene, cam, 8)
    print(f"2048x2560 tile 8: phase-1 pairs {k1}, phase-2 pairs {k2}")


@pytest.mark.parametrize("betas", [(21.4, -0.5), (-3.0, 2.66)])
def test_keys_betas_outside_the_bounds(fea, scenes_mod, betas):
    """A view with beta3 < 0 or beta2 < 0 lists every pair (the bounds assume both >= 0), on
    the GPU and in the restatement."""
    import dataclasses
    scene = scenes_mod.single_view_scene(seed=6)
    scene = dataclasses.replace(scene, beta2=betas[0], beta3=betas[1])
    k1, k2 = _check(fea, scene, scene.cameras[0], 16)
    assert k1 == scene.n * 16 and k2 == 0
