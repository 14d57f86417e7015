"""Split tile lists: the blend cuts each tile list of a view of fewer than 4096 16-pixel tiles
(every view on 8-pixel tiles) into fixed contiguous ranges (two on 16-pixel tiles, four on
8-pixel tiles), blends them from separate states and merges the states.  On 16-pixel tiles the
launcher runs the ranges side by side for grids up to 8 times what the GPU holds at once (28
waves per CU; 57344 waves on 256 CUs) and one after another beyond that (8-pixel tiles always
side by side); a batch with no split view runs the plain kernel.  The arithmetic is the same
either way, so a view rendered inside a batch past that size, or beside large views, equals the
view rendered alone, bit for bit."""
import numpy as np
import pytest

from oracle import ref

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("mode,n,tile", [("culled", 4096, 16), ("dense", 1024, 16), ("culled", 4096, 8)])
def test_large_batch_equals_lone_view(fea, scenes_mod, mode, n, tile):
    import torch
    # 256 views of 128x128 around a cloud that fills them (lists of thousands of records): at
    # tile 16, 16384 tiles of 4 waves when split side by side (65536 waves: one after another);
    # one view alone: 64 tiles (side by side).  At tile 8: 65536 and 256 tiles.
    rng = np.random.default_rng(21)
    means, prec, w = scenes_mod.random_gaussians(rng, n)
    cams = [scenes_mod.make_camera(3.0 * scenes_mod.random_dir(rng), 128, 128, 120.0) for _ in range(256)]
    scene = scenes_mod.Scene(means, prec, w, cams)
    st = fea.SceneTensors.from_scene(scene)
    views = [(0, c, scene.beta2, scene.beta3) for c in cams]
    r = fea.Renderer(tile=tile, batch=64, mode=mode)
    outs = r.forward([st], views)
    torch.cuda.synchronize()
    for k in (0, 97, 255):
        r1 = fea.Renderer(tile=tile, batch=64, mode=mode)
        (d1, a1), = r1.forward([st], [views[k]])
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_np(outs[k][0]), _np(d1))
        np.testing.assert_array_equal(_np(outs[k][1]), _np(a1))
    cam = cams[97]
    pix = rng.choice(cam.H * cam.W, 200, replace=False)
    od, oa = ref.render_f64(scene, cam, pixels=pix)
    d = _np(outs[97][0]).ravel()[pix]
    a = _np(outs[97][1]).ravel()[pix]
    print(f"split placement {mode} n={n} tile {tile}: batch == lone views; alpha err {np.max(np.abs(a - oa)):.1e}, "
          f"depth rel err {np.max(np.abs(d - od) / np.abs(od)):.1e}")
    assert np.max(np.abs(a - oa)) <= 1e-4
    assert np.max(np.abs(d - od) / np.abs(od)) <= 1e-4


def test_mixed_split_and_unsplit_views(fea, scenes_mod):
    """A batch of split views (128x128 and 480x640) and unsplit large ones (1024x1024, 4096
    tiles): every view equals its lone render bit for bit (the lone 1024x1024 view runs the plain
    kernel, the batch the split one)."""
    import torch
    rng = np.random.default_rng(23)
    means, prec, w = scenes_mod.random_gaussians(rng, 2048)
    cams = [scenes_mod.make_camera(3.0 * scenes_mod.random_dir(rng), 128, 128, 120.0) for _ in range(3)]
    cams += [scenes_mod.make_camera(3.0 * scenes_mod.random_dir(rng), 480, 640, 600.0)]
    cams += [scenes_mod.make_camera(3.0 * scenes_mod.random_dir(rng), 1024, 1024, 960.0) for _ in range(2)]
    scene = scenes_mod.Scene(means, prec, w, cams)
    st = fea.SceneTensors.from_scene(scene)
    views = [(0, c, scene.beta2, scene.beta3) for c in cams]
    for mode in ("culled", "dense"):
        r = fea.Renderer(tile=16, batch=64, mode=mode)
        outs = r.forward([st], views)
        torch.cuda.synchronize()
        for k in range(len(views)):
            r1 = fea.Renderer(tile=16, batch=64, mode=mode)
            (d1, a1), = r1.forward([st], [views[k]])
            torch.cuda.synchronize()
            np.testing.assert_array_equal(_np(outs[k][0]), _np(d1))
            np.testing.assert_array_equal(_np(outs[k][1]), _np(a1))

