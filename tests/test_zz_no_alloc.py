"""fe_render_fwd_noalloc (the entry point that allocates nothing: its workspace is the
library's static per-GPU scratch array) against fe_render_fwd with a caller workspace, through
the C ABI: bit-identical images on two calls in a row, the second with a smaller input, both
within the parity tolerance of the FP64 oracle; views needing more than the scratch are
refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import ref

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _render(fea, lib, scene, cam, noalloc):
    from fuzzy_ellipses_amd import _lib
    st = fea.SceneTensors.from_scene(scene, DEV)
    depth = torch.full((cam.H, cam.W), -1.0, device=DEV)
    alpha = torch.full((cam.H, cam.W), -1.0, device=DEV)
    c = fea.fe_camera(cam)
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        status = lib.fe_render_fwd_noalloc(s, scene.n, st.means.data_ptr(), st.prec.data_ptr(),
                                           st.weights_log.data_ptr(), ctypes.byref(c), scene.beta2, scene.beta3,
                                           depth.data_ptr(), alpha.data_ptr())
    else:
        nbytes = ctypes.c_size_t(0)
        _lib.check(lib.fe_render_workspace_bytes(scene.n, cam.H, cam.W, 16, ctypes.byref(nbytes)), "workspace")
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=DEV)
        status = lib.fe_render_fwd(s, scene.n, st.means.data_ptr(), st.prec.data_ptr(), st.weights_log.data_ptr(),
                                   ctypes.byref(c), scene.beta2, scene.beta3, depth.data_ptr(), alpha.data_ptr(),
                                   ws.data_ptr(), nbytes.value)
    _lib.check(status, "fe_render_fwd_noalloc" if noalloc else "fe_render_fwd")
    torch.cuda.synchronize()
    return depth.cpu().numpy().ravel(), alpha.cpu().numpy().ravel()


def test_noalloc_equals_workspace_entry_point(fea, scenes_mod):
    from fuzzy_ellipses_amd import _lib
    lib = _lib.load()
    shell = scenes_mod.shell_scene(n_views=2)            # N=4096 at 480x640: 27.6 MiB of scratch
    small = scenes_mod.single_view_scene(seed=3)         # N=40 at 64x64, after it
    rng = np.random.default_rng(0)
    for scene, cam in ((shell, shell.cameras[1]), (small, small.cameras[0])):
        d0, a0 = _render(fea, lib, scene, cam, noalloc=False)
        d1, a1 = _render(fea, lib, scene, cam, noalloc=True)
        np.testing.assert_array_equal(d1.view(np.uint32), d0.view(np.uint32))
        np.testing.assert_array_equal(a1.view(np.uint32), a0.view(np.uint32))
        P = cam.H * cam.W
        pix = np.sort(rng.choice(P, min(P, 2000), replace=False))
        od, oa = ref.render_f64(scene, cam, pixels=pix)
        ed = float(np.max(np.abs(d1[pix] - od) / np.abs(od)))
        ea = float(np.max(np.abs(a1[pix] - oa)))
        print(f"noalloc N={scene.n} {cam.H}x{cam.W}: bitwise equal to fe_render_fwd; "
              f"depth rel {ed:.2e}, alpha abs {ea:.2e}")
        assert ed <= 1e-4 and ea <= 1e-4


def test_noalloc_refuses_views_beyond_the_scratch(fea, scenes_mod):
    from fuzzy_ellipses_amd import _lib
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.fe_render_workspace_bytes(8192, 480, 640, 16, ctypes.byref(nbytes)), "workspace")
    assert nbytes.value > _lib.FE_NOALLOC_SCRATCH_BYTES
    cam = fea.fe_camera(scenes_mod.make_camera(np.array([0.0, 0.0, 3.0]), 480, 640, 600.0))
    dummy = torch.empty(16, device=DEV)
    status = lib.fe_render_fwd_noalloc(torch.cuda.current_stream().cuda_stream, 8192, dummy.data_ptr(),
                                       dummy.data_ptr(), dummy.data_ptr(), ctypes.byref(cam), 21.4, 2.66,
                                       dummy.data_ptr(), dummy.data_ptr())
    assert status == 2     # FE_ERR_WORKSPACE, before any device work
    assert b"FE_NOALLOC_SCRATCH_BYTES" in lib.fe_last_error_string()
