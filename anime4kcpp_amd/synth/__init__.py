"""Seeded synthetic flat-shaded images (polygons + Bezier strokes + noise).

Wraps ``libac_synth.so`` (built from synth.c by anime4kcpp_amd.build).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libac_synth.so")
_lib = None


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run anime4kcpp_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.ac_synth_flat_shaded.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.ac_synth_flat_shaded_u8.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.ac_synth_randint.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.ac_synth_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


def flat_shaded(seed, w, h, channels=1, dtype=np.float32):
    """H x W (x C) image in [0,1] (float32), its u8 quantisation round(255 v) or its u16
    quantisation round(65535 v)."""
    shape = (h, w) if channels == 1 else (h, w, channels)
    dtype = np.dtype(dtype)
    if dtype == np.uint8:
        out = np.empty(shape, np.uint8)
        rc = _L().ac_synth_flat_shaded_u8(seed, w, h, channels, out.ctypes.data)
    elif dtype in (np.float32, np.uint16):
        out = np.empty(shape, np.float32)
        rc = _L().ac_synth_flat_shaded(seed, w, h, channels, out.ctypes.data)
    else:
        raise ValueError(f"flat_shaded: unsupported dtype {dtype}")
    if rc != 0:
        raise ValueError(f"ac_synth_flat_shaded({seed}, {w}, {h}, {channels}) -> {rc}")
    if dtype == np.uint16:
        out = np.rint(out.astype(np.float64) * 65535.0).astype(np.uint16)
    return out


def randint(seed, n, lo, hi):
    out = np.empty(n, np.int32)
    if _L().ac_synth_randint(seed, n, lo, hi, out.ctypes.data) != 0:
        raise ValueError("randint")
    return out


def uniform(seed, n):
    out = np.empty(n, np.float64)
    if _L().ac_synth_uniform(seed, n, out.ctypes.data) != 0:
        raise ValueError("uniform")
    return out
