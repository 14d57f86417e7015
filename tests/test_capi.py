"""C ABI checks that need no GPU: the library loads, exports every declared entry point,
and the host-only planner validates its arguments and sizes the workspace."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(fe_\w+|rasterize_gaussians_fwd)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol(fea):
    from fuzzy_ellipses_amd import _lib
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 9
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= set(declared)
    assert lib.fe_version() >= 101


def _one_view(fea, H=480, W=640, n=4096, tile=16, batch=64, mode=0, save=0):
    from fuzzy_ellipses_amd import _lib
    sc = (_lib.FeScene * 1)(_lib.FeScene(n, None, None, None))
    cam = _lib.FeCamera()
    cam.fx = cam.fy = 600.0
    cam.cx, cam.cy = W / 2, H / 2
    cam.height, cam.width = H, W
    vw = (_lib.FeView * 1)(_lib.FeView(0, cam, 21.4, 2.66, None, None))
    opt = _lib.FeOptions(tile, batch, mode, save, 0)
    nbytes = ctypes.c_size_t(0)
    st = _lib.load().fe_plan(sc, 1, vw, 1, ctypes.byref(opt), ctypes.byref(nbytes))
    return st, nbytes.value


def test_plan_sizes_workspace_for_worst_case(fea):
    st, nbytes = _one_view(fea)
    assert st == 0
    tiles = 30 * 40
    # lists sized for every (tile, Gaussian) pair; records 64 B per Gaussian
    assert nbytes >= tiles * 4096 * 4 + 4096 * 64
    st_d, nb_dense = _one_view(fea, mode=1)
    assert st_d == 0 and nb_dense < nbytes
    st_s, nb_save = _one_view(fea, save=1)
    assert st_s == 0 and nb_save > nbytes


@pytest.mark.parametrize("kw", [dict(tile=12), dict(batch=48), dict(mode=7), dict(H=0), dict(n=-1)])
def test_plan_rejects_bad_arguments(fea, kw):
    from fuzzy_ellipses_amd import _lib
    st, _ = _one_view(fea, **kw)
    assert st == 1
    assert _lib.load().fe_last_error_string().decode().startswith("fe_plan")


def test_plan_zero_gaussians_ok(fea):
    st, nbytes = _one_view(fea, n=0, H=8, W=8)
    assert st == 0 and nbytes > 0


def test_forward_rejects_small_workspace(fea):
    """fe_forward validates the workspace before touching the device."""
    from fuzzy_ellipses_amd import _lib
    lib = _lib.load()
    sc = (_lib.FeScene * 1)(_lib.FeScene(4, None, None, None))
    cam = _lib.FeCamera()
    cam.fx = cam.fy = 8.0
    cam.height = cam.width = 8
    vw = (_lib.FeView * 1)(_lib.FeView(0, cam, 21.4, 2.66, None, None))
    opt = _lib.FeOptions(8, 32, 0, 0, 0)
    st = lib.fe_forward(None, sc, 1, vw, 1, ctypes.byref(opt), None, 0)
    assert st == 2  # FE_ERR_WORKSPACE


def test_python_api_fails_loudly_without_library(fea, monkeypatch, tmp_path):
    from fuzzy_ellipses_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.FeError):
        _lib.load(str(tmp_path / "missing.so"))


def test_error_budgets_validated_and_monotone(fea):
    from fuzzy_ellipses_amd import _lib
    lib = _lib.load()
    sc = (_lib.FeScene * 1)(_lib.FeScene(100, None, None, None))
    cam = _lib.FeCamera()
    cam.fx = cam.fy = 64.0
    cam.height = cam.width = 64
    vw = (_lib.FeView * 1)(_lib.FeView(0, cam, 21.4, 2.66, None, None))
    nbytes = ctypes.c_size_t(0)
    bad = _lib.FeOptions(16, 64, 0, 0, 0, 1.5, 0.0)
    assert lib.fe_plan(sc, 1, vw, 1, ctypes.byref(bad), ctypes.byref(nbytes)) == 1
    ok = _lib.FeOptions(16, 64, 0, 0, 0, 1e-3, 1e-4)
    assert lib.fe_plan(sc, 1, vw, 1, ctypes.byref(ok), ctypes.byref(nbytes)) == 0


def test_looser_alpha_budget_lists_fewer_pairs(fea, scenes_mod):
    """The CPU restatement of the membership with the planner's constants: a looser budget
    never lists more pairs (the keep region shrinks monotonically)."""
    import numpy as np
    from oracle import ref
    scene = scenes_mod.single_view_scene(seed=4, n=80, H=64, W=64, f=64.0)
    cam = scene.cameras[0]
    rec, deg = ref.records(scene, cam)
    lib = ref.load()
    counts = []
    for eps in (1e-7, 2e-5, 1e-3):
        c_alpha, _, _ = ref.cull_constants(scene.n, alpha_eps=eps)
        T = 16
        c = np.zeros(T, np.uint32)
        lst = np.zeros(scene.n * T, np.uint32)
        tot = lib.fe_ref_tiles(scene.n, np.ascontiguousarray(rec.reshape(-1)), deg, 64, 64, 16, np.float32(cam.cx),
                               np.float32(cam.cy), np.float32(scene.beta2), np.float32(scene.beta3),
                               np.float32(c_alpha), np.float32(1.0), np.float32(40.0), None, None, c, lst)
        counts.append(tot)
    assert counts[0] >= counts[1] >= counts[2]


def test_silhouette_and_adam_entry_points_validate(fea):
    """The caller-side entry points validate before touching the device."""
    from fuzzy_ellipses_amd import _lib
    lib = _lib.load()
    assert ctypes.sizeof(_lib.FeSilView) == 56
    nb = ctypes.c_size_t(0)
    assert lib.fe_silhouette_workspace_bytes(16, ctypes.byref(nb)) == 0
    assert nb.value >= 16 * 56 + 2 * 16 * 64 * 4
    assert lib.fe_silhouette_workspace_bytes(0, ctypes.byref(nb)) == 1
    views = (_lib.FeSilView * 1)(_lib.FeSilView(None, None, None, None, None, None, 10))
    assert lib.fe_silhouette_loss(None, views, 1, 0.1, 3e-8, 8, 0, None, 0) == 1      # null buffers
    good = (_lib.FeSilView * 1)(_lib.FeSilView(8, 8, 8, 8, 8, 8, 10))
    assert lib.fe_silhouette_loss(None, good, 1, 0.1, 0.7, 8, 0, None, 0) == 1        # clip out of range
    assert lib.fe_silhouette_loss(None, good, 1, 0.1, 3e-8, 8, 0, None, 0) == 2       # no workspace
    assert lib.fe_adam_step(None, 10, 8, 8, 8, 8, 1e-2, 0.9, 0.999, 1e-8, 0) == 1     # step counts from 1
    assert lib.fe_adam_step(None, 10, 8, 8, 8, 8, 1e-2, 1.0, 0.999, 1e-8, 1) == 1     # beta1 = 1
    assert lib.fe_adam_step(None, 0, None, None, None, None, 1e-2, 0.9, 0.999, 1e-8, 1) == 0  # empty: no-op


def test_profile_selection_is_host_only_and_empty_when_nothing_ran(fea):
    """fe_profile_enable accepts 0, 1 and FE_PROFILE_SELECT | bits; with no launch since the
    reset every kernel reads zero time and zero launches."""
    from fuzzy_ellipses_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "fuzzy_ellipses_amd.h")).read()
    assert f"#define FE_PROFILE_SELECT 0x{_lib.FE_PROFILE_SELECT:08x}" in hdr
    _lib.profile_read(reset=True)
    _lib.profile_enable(True, kernels=["bwd_blend", "blend1"])
    _lib.profile_enable(True)
    _lib.profile_enable(False)
    kt = _lib.profile_read(reset=True)
    assert set(kt) == set(_lib.KERNELS) and all(v == (0.0, 0) for v in kt.values())
    with pytest.raises(ValueError):
        _lib.profile_enable(True, kernels=["no_such_kernel"])


def test_plan_rejects_2_31_tiles(fea):
    """One block per tile and int32 tile indices in the dispatch order: the planner refuses a
    batch of 2^31 tiles (here 128 ray-list views of 2^30 rays at tile 8, scenes of n = 0)."""
    from fuzzy_ellipses_amd import _lib
    sc = (_lib.FeScene * 1)(_lib.FeScene(0, None, None, None))
    cam = _lib.FeCamera()
    cam.fx = cam.fy = 1.0

    def plan(nv):
        vw = (_lib.FeView * nv)(*[_lib.FeView(0, cam, 1.0, 1.0, None, None, None, 8, 1 << 30) for _ in range(nv)])
        opt = _lib.FeOptions(8, 64, 0, 0, 0)
        nbytes = ctypes.c_size_t(0)
        return _lib.load().fe_plan(sc, 1, vw, nv, ctypes.byref(opt), ctypes.byref(nbytes))

    assert plan(127) == 0
    assert plan(128) != 0
    assert b"2^31 tiles" in _lib.load().fe_last_error_string()


def test_custom_call_registration_and_descriptor(fea):
    """The reference binding's `_C` surface (ext.cpp:22-47): a capsule named
    "xla._CUSTOM_CALL_TARGET" around `rasterize_gaussians_fwd`, and the 20-byte descriptor
    {int32 H, W, P; float tan_fovx, tan_fovy} (rasterize_main.h:13-19)."""
    import struct
    from fuzzy_ellipses_amd import xla
    regs = xla.registrations()
    assert set(regs) == {"rasterize_gaussians_fwd"}
    cap = regs["rasterize_gaussians_fwd"]
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    assert get_name(cap) == b"xla._CUSTOM_CALL_TARGET"
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    assert get_ptr(cap, b"xla._CUSTOM_CALL_TARGET") == xla.target_address() != 0
    d = xla.build_gaussian_rasterize_descriptor(480, 640, 4096, 1.5, 0.75)
    assert len(d) == 20
    assert struct.unpack("<iiiff", d) == (480, 640, 4096, 1.5, 0.75)


def test_custom_call_rejects_bad_descriptor_without_gpu(fea):
    """Argument checks run before any device work: a descriptor of the wrong size or a
    negative count is refused with a status (the void target then touches no buffer)."""
    from fuzzy_ellipses_amd import _lib, xla
    lib = _lib.load()
    bufs = (ctypes.c_void_p * 11)()
    op = ctypes.create_string_buffer(b"\0" * 19, 19)
    assert lib.fe_rasterize_gaussians_fwd(None, bufs, op, 19) == 1
    bad = xla.build_gaussian_rasterize_descriptor(4, 4, -1, 1.0, 1.0)
    assert lib.fe_rasterize_gaussians_fwd(None, bufs, bad, 20) == 1
    assert b"rasterize_gaussians_fwd" in lib.fe_last_error_string()
