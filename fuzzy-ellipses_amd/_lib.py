"""ctypes binding of libfe_amd.so (the C ABI declared in include/fuzzy_ellipses_amd.h).

The library is loaded from this package's lib/ directory.  There is no fallback: if the
library is missing or does not export the full ABI, every render call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfe_amd.so")

FE_OK = 0
FE_MODE_CULLED = 0
FE_MODE_DENSE = 1

EXPORTS = [
    "fe_version", "fe_status_string", "fe_last_error_string", "fe_plan", "fe_forward",
    "fe_backward", "fe_read_stats", "fe_render_workspace_bytes", "fe_render_fwd", "fe_profile_enable",
    "fe_profile_read", "fe_silhouette_workspace_bytes", "fe_silhouette_loss", "fe_adam_step",
    "fe_gmm_workspace_bytes", "fe_gmm_forward", "fe_gmm_backward", "fe_ray_flow",
    "fe_rasterize_gaussians_fwd", "rasterize_gaussians_fwd", "fe_render_fwd_noalloc",
]
KERNELS = ["preprocess", "mask1", "scan1", "write1", "blend1", "mask2", "scan2", "write2", "blend2",
           "blend_dense", "bwd_blend", "bwd_finalize"]


class FeCamera(ctypes.Structure):
    _fields_ = [("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("height", ctypes.c_int32), ("width", ctypes.c_int32)]


class FeScene(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("means", ctypes.c_void_p), ("prec", ctypes.c_void_p),
                ("weights_log", ctypes.c_void_p), ("features", ctypes.c_void_p), ("n_features", ctypes.c_int32)]


class FeView(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int32), ("cam", FeCamera), ("beta2", ctypes.c_float),
                ("beta3", ctypes.c_float), ("depth", ctypes.c_void_p), ("alpha", ctypes.c_void_p),
                ("pose_device", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("n_rays", ctypes.c_int64),
                ("features_out", ctypes.c_void_p)]


class FeSceneGrad(ctypes.Structure):
    _fields_ = [("d_means", ctypes.c_void_p), ("d_prec", ctypes.c_void_p),
                ("d_weights_log", ctypes.c_void_p), ("d_features", ctypes.c_void_p)]


class FeViewCotangent(ctypes.Structure):
    _fields_ = [("d_depth", ctypes.c_void_p), ("d_alpha", ctypes.c_void_p), ("d_features", ctypes.c_void_p)]


class FeOptions(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int32), ("batch", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("save_for_backward", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("alpha_eps", ctypes.c_float), ("depth_eps", ctypes.c_float)]


FE_FLAG_REUSE_DESCRIPTORS = 1
FE_NOALLOC_SCRATCH_BYTES = 32 << 20


class FeSilView(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_void_p), ("depth", ctypes.c_void_p), ("mask", ctypes.c_void_p),
                ("target_depth", ctypes.c_void_p), ("d_alpha", ctypes.c_void_p), ("d_depth", ctypes.c_void_p),
                ("n_pixels", ctypes.c_int64)]


class FeStats(ctypes.Structure):
    _fields_ = [("pairs_phase1", ctypes.c_int64), ("pairs_phase2", ctypes.c_int64),
                ("evals", ctypes.c_int64), ("evals_phase1", ctypes.c_int64), ("floor_tiles", ctypes.c_int64),
                ("floor_unresolved", ctypes.c_int64)]


class FeKernelTimes(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 12), ("launches", ctypes.c_int64 * 12)]


class FeError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library once; raise FeError if it is missing or incomplete."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FeError(f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [s for s in EXPORTS if not hasattr(lib, s)]
    if missing:
        raise FeError(f"{path} does not export {missing}")
    lib.fe_version.restype = ctypes.c_int32
    lib.fe_status_string.restype = ctypes.c_char_p
    lib.fe_status_string.argtypes = [ctypes.c_int]
    lib.fe_last_error_string.restype = ctypes.c_char_p
    lib.fe_plan.restype = ctypes.c_int
    lib.fe_plan.argtypes = [ctypes.POINTER(FeScene), ctypes.c_int32, ctypes.POINTER(FeView), ctypes.c_int32,
                            ctypes.POINTER(FeOptions), ctypes.POINTER(ctypes.c_size_t)]
    lib.fe_forward.restype = ctypes.c_int
    lib.fe_forward.argtypes = [ctypes.c_void_p, ctypes.POINTER(FeScene), ctypes.c_int32, ctypes.POINTER(FeView),
                               ctypes.c_int32, ctypes.POINTER(FeOptions), ctypes.c_void_p, ctypes.c_size_t]
    lib.fe_backward.restype = ctypes.c_int
    lib.fe_backward.argtypes = [ctypes.c_void_p, ctypes.POINTER(FeScene), ctypes.c_int32, ctypes.POINTER(FeView),
                                ctypes.c_int32, ctypes.POINTER(FeOptions), ctypes.POINTER(FeViewCotangent),
                                ctypes.POINTER(FeSceneGrad), ctypes.c_void_p, ctypes.c_size_t]
    lib.fe_read_stats.restype = ctypes.c_int
    lib.fe_read_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(FeStats)]
    lib.fe_render_workspace_bytes.restype = ctypes.c_int
    lib.fe_render_workspace_bytes.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_size_t)]
    lib.fe_render_fwd.restype = ctypes.c_int
    lib.fe_render_fwd.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.POINTER(FeCamera), ctypes.c_float, ctypes.c_float,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.fe_profile_enable.restype = ctypes.c_int
    lib.fe_profile_enable.argtypes = [ctypes.c_int32]
    lib.fe_profile_read.restype = ctypes.c_int
    lib.fe_profile_read.argtypes = [ctypes.POINTER(FeKernelTimes), ctypes.c_int32]
    lib.fe_silhouette_workspace_bytes.restype = ctypes.c_int
    lib.fe_silhouette_workspace_bytes.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t)]
    lib.fe_silhouette_loss.restype = ctypes.c_int
    lib.fe_silhouette_loss.argtypes = [ctypes.c_void_p, ctypes.POINTER(FeSilView), ctypes.c_int32, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_size_t]
    lib.fe_adam_step.restype = ctypes.c_int
    lib.fe_adam_step.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_int32]
    lib.fe_gmm_workspace_bytes.restype = ctypes.c_int
    lib.fe_gmm_workspace_bytes.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    lib.fe_gmm_forward.restype = ctypes.c_int
    lib.fe_gmm_forward.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_size_t]
    lib.fe_gmm_backward.restype = ctypes.c_int
    lib.fe_gmm_backward.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.fe_ray_flow.restype = ctypes.c_int
    lib.fe_ray_flow.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_float, ctypes.c_int32,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.fe_render_fwd_noalloc.restype = ctypes.c_int
    lib.fe_render_fwd_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.POINTER(FeCamera), ctypes.c_float, ctypes.c_float,
                                          ctypes.c_void_p, ctypes.c_void_p]
    xla_args = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]
    lib.fe_rasterize_gaussians_fwd.restype = ctypes.c_int
    lib.fe_rasterize_gaussians_fwd.argtypes = xla_args
    lib.rasterize_gaussians_fwd.restype = None
    lib.rasterize_gaussians_fwd.argtypes = xla_args
    _lib = lib
    return lib


FE_PROFILE_SELECT = 0x40000000


def profile_enable(on: bool, kernels=None) -> None:
    """Time every kernel launch (or only those named in `kernels`, from KERNELS) from now on."""
    if on and kernels is not None:
        bits = 0
        for k in kernels:
            bits |= 1 << KERNELS.index(k)
        check(load().fe_profile_enable(FE_PROFILE_SELECT | bits), "fe_profile_enable")
    else:
        check(load().fe_profile_enable(1 if on else 0), "fe_profile_enable")


def profile_read(reset: bool = True) -> dict:
    """{kernel: (total_ms, launches)} of the event-timed launches since the last reset."""
    kt = FeKernelTimes()
    check(load().fe_profile_read(ctypes.byref(kt), 1 if reset else 0), "fe_profile_read")
    return {k: (kt.ms[i], kt.launches[i]) for i, k in enumerate(KERNELS)}


def check(status: int, what: str = "") -> None:
    if status != FE_OK:
        lib = load()
        raise FeError(f"{what}: {lib.fe_status_string(status).decode()} "
                      f"({lib.fe_last_error_string().decode()})")
