"""The reference binding's native module, served by libfe_amd.so.

The reference's ``fuzzy_ellipses/__init__.py`` imports a compiled ``_C`` with two functions
(``ext.cpp:22-47``):

* ``registrations()`` -> ``{"rasterize_gaussians_fwd": PyCapsule}``, the capsule named
  ``"xla._CUSTOM_CALL_TARGET"`` that ``xla_client.register_custom_call_target`` takes
  (``fuzzy_ellipses/__init__.py:22-23``);
* ``build_gaussian_rasterize_descriptor(H, W, P, tan_fovx, tan_fovy)`` -> the 20-byte
  opaque descriptor (``rasterize_main.h:13-19``, ``bindings.h:22-37``).

This module provides both over ``libfe_amd.so``'s ``rasterize_gaussians_fwd`` (same
signature, buffers and descriptor), so the reference's lowering binds it unchanged with
``from fuzzy_ellipses_amd import xla as _C``.  ``call_custom_target`` drives the target
exactly as XLA's legacy custom call does (a ``void**`` of 11 device pointers and the
opaque bytes) for callers without JAX.
"""
from __future__ import annotations

import ctypes
import struct
from typing import Dict, Sequence

from . import _lib

TARGET = "rasterize_gaussians_fwd"
CAPSULE_NAME = b"xla._CUSTOM_CALL_TARGET"
DESCRIPTOR_FORMAT = "<iiiff"     # int32 H, W, P; float tan_fovx, tan_fovy (20 bytes)
N_BUFFERS = 11


def build_gaussian_rasterize_descriptor(image_height: int, image_width: int, P: int, tan_fovx: float,
                                        tan_fovy: float) -> bytes:
    """The opaque bytes of RasterizeDescriptor (ext.cpp:34-42): a bit copy of the struct."""
    return struct.pack(DESCRIPTOR_FORMAT, int(image_height), int(image_width), int(P), float(tan_fovx),
                       float(tan_fovy))


def target_address() -> int:
    """Address of the C function ``rasterize_gaussians_fwd`` in libfe_amd.so."""
    lib = _lib.load()
    return ctypes.cast(lib.rasterize_gaussians_fwd, ctypes.c_void_p).value


def registrations() -> Dict[str, object]:
    """{target name: PyCapsule named "xla._CUSTOM_CALL_TARGET"} (ext.cpp:22-32)."""
    new = ctypes.pythonapi.PyCapsule_New
    new.restype = ctypes.py_object
    new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    return {TARGET: new(target_address(), CAPSULE_NAME, None)}


def call_custom_target(stream: int, buffers: Sequence[int], opaque: bytes, status: bool = True) -> None:
    """Call the target as XLA's legacy GPU custom call does: ``buffers`` are the 9 input and 2
    output device pointers in the reference's order (rasterize_main.cu:54-68).  With
    ``status`` the status-returning twin is called and a failure raises FeError; without it
    the void target itself is called (a failure then leaves NaN outputs)."""
    if len(buffers) != N_BUFFERS:
        raise _lib.FeError(f"{TARGET}: expected {N_BUFFERS} buffers, got {len(buffers)}")
    lib = _lib.load()
    arr = (ctypes.c_void_p * N_BUFFERS)(*[ctypes.c_void_p(int(b)) if b else None for b in buffers])
    op = ctypes.create_string_buffer(opaque, len(opaque))
    if status:
        _lib.check(lib.fe_rasterize_gaussians_fwd(ctypes.c_void_p(stream), arr, op, len(opaque)), TARGET)
    else:
        lib.rasterize_gaussians_fwd(ctypes.c_void_p(stream), arr, op, len(opaque))
