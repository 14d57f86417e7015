"""Split tile lists: the blend cuts each 16-pixel tile's list into two fixed ranges, blends them
from separate states and merges the states.  The launcher runs the two ranges side by side for
grids up to 8 times what the GPU holds at once (28 waves per CU; 57344 waves on 256 CUs) and one
after another beyond that.  The arithmetic is the same either way, so a view rendered inside a
batch past that size equals the view rendered alone (a grid far below it), bit for bit."""
import numpy as np
import pytest

from oracle import ref

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("mode,n", [("culled", 4096), ("dense", 512)])
def test_large_batch_equals_lone_view(fea, scenes_mod, mode, n):
    import torch
    # 16 views of 480x640 at tile 16: 19200 tiles, 4 waves each when split side by side
    # (76800 waves); one view alone: 1200 tiles (4800 waves)
    scene = scenes_mod.shell_scene(n=n)
    st = fea.SceneTensors.from_scene(scene)
    views = [(0, c, scene.beta2, scene.beta3) for c in scene.cameras]
    r = fea.Renderer(tile=16, batch=64, mode=mode)
    outs = r.forward([st], views)
    torch.cuda.synchronize()
    rng = np.random.default_rng(11)
    for k in (0, 7, 15):
        r1 = fea.Renderer(tile=16, batch=64, mode=mode)
        (d1, a1), = r1.forward([st], [views[k]])
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_np(outs[k][0]), _np(d1))
        np.testing.assert_array_equal(_np(outs[k][1]), _np(a1))
    cam = scene.cameras[7]
    pix = rng.choice(cam.H * cam.W, 300, replace=False)
    od, oa = ref.render_f64(scene, cam, pixels=pix)
    d = _np(outs[7][0]).ravel()[pix]
    a = _np(outs[7][1]).ravel()[pix]
    print(f"split placement {mode}: batch == lone views; alpha err {np.max(np.abs(a - oa)):.1e}, "
          f"depth rel err {np.max(np.abs(d - od) / np.abs(od)):.1e}")
    assert np.max(np.abs(a - oa)) <= 1e-4
    assert np.max(np.abs(d - od) / np.abs(od)) <= 1e-4
