"""ctypes access to the oracle (oracle/liboracle.so) and, where it was built, to the
reference's own CPU kernels (oracle/_ref/libacref.so).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libacref.so")

U8, U16, F16, F32 = 0x001, 0x002, 0x202, 0x204
NP_TYPE = {U8: np.uint8, U16: np.uint16, F16: np.float16, F32: np.float32}
TYPE_OF = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.float16): F16, np.dtype(np.float32): F32}
ORDER_FMA, ORDER_GENERIC = 0, 1
# ResizeModes (core/include/AC/Core/Image.hpp:272-291)
RESIZE_MODES = ["POINT", "CATMULL_ROM", "MITCHELL_NETRAVALI", "BICUBIC_0_60", "BICUBIC_0_75", "BICUBIC_0_100",
                "BICUBIC_20_50", "SOFTCUBIC50", "SOFTCUBIC75", "SOFTCUBIC100", "LANCZOS2", "LANCZOS3", "LANCZOS4",
                "SPLINE16", "SPLINE36", "SPLINE64", "BILINEAR"]
RESIZE_POINT, RESIZE_CATMULL_ROM, RESIZE_LANCZOS4, RESIZE_SPLINE36, RESIZE_BILINEAR = 0, 1, 12, 14, 16
# (b, c) of the cubic modes
RESIZE_BC = {1: (0, 1 / 2), 2: (1 / 3, 1 / 3), 3: (0, 3 / 5), 4: (0, 3 / 4), 5: (0, 1), 6: (1 / 5, 1 / 2),
             7: (1 / 2, 1 / 2), 8: (3 / 4, 1 / 4), 9: (1, 0)}
RESIZE_LANCZOS = {10: 2, 11: 3, 12: 4}
RESIZE_SPLINE = {13: 2, 14: 3, 15: 4}

_orc = None
_ref = None
P = ctypes.c_void_p
L = ctypes.c_long


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: make -C oracle")
        o = ctypes.CDLL(ORACLE_SO)
        o.orc_acnet_pass.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int]
        o.orc_acnet_pass_f64.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int]
        o.orc_to_float.argtypes = [P, ctypes.c_int]
        o.orc_to_float.restype = ctypes.c_float
        o.orc_from_float.argtypes = [ctypes.c_float, P, ctypes.c_int]
        o.orc_rgb2yuv.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, L]
        o.orc_yuv2rgb.argtypes = [P, L, P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L]
        o.orc_resize2x.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, ctypes.c_int]
        o.orc_resize.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]
        o.orc_acnet8_pass_f64.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, P, P, ctypes.c_int, ctypes.c_int]
        o.orc_acnet8_pass.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, P, P, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]
        o.orc_artcnn_pass.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, P, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        o.orc_fsrcnnx_pass.argtypes = [P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, P, P, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int]
        o.orc_filter_weight.argtypes = [ctypes.c_int, ctypes.c_float]
        o.orc_filter_weight.restype = ctypes.c_float
        o.orc_filter_support.argtypes = [ctypes.c_int]
        o.orc_filter_support.restype = ctypes.c_float
        _orc = o
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        r = ctypes.CDLL(REF_SO)
        r.acref_pass.argtypes = [ctypes.c_int, P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, ctypes.c_int, P, P]
        r.acref_has_arch.argtypes = [ctypes.c_int]
        r.acref_artcnn.argtypes = [ctypes.c_int, P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, P,
                                   ctypes.c_int, ctypes.c_int]
        r.acref_fsrcnnx.argtypes = [ctypes.c_int, P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, P, P,
                                    ctypes.c_int, ctypes.c_int]
        r.acref_pass8.argtypes = [ctypes.c_int, P, L, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, L, P, P, P, ctypes.c_int]
        _ref = r
    return _ref


def oracle_pass(src, kernels, biases, out_dtype=None, order=ORDER_FMA, threads=0):
    """src: 2-D numpy array (u8/u16/f16/f32) -> 2x output in out_dtype (default src dtype)."""
    src = np.ascontiguousarray(src)
    h, w = src.shape
    out_dtype = np.dtype(out_dtype or src.dtype)
    dst = np.zeros((2 * h, 2 * w), out_dtype)
    rc = orc().orc_acnet_pass(src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype],
                              dst.ctypes.data, dst.strides[0], TYPE_OF[out_dtype],
                              kernels.ctypes.data, biases.ctypes.data, order, threads)
    assert rc == 0, rc
    return dst


def oracle_pass_f64(src_float, kernels, biases, threads=0):
    """src_float: toFloat'ed input as float32 -> saturated float64 2x output."""
    s = np.ascontiguousarray(src_float, dtype=np.float32)
    h, w = s.shape
    dst = np.zeros((2 * h, 2 * w), np.float64)
    rc = orc().orc_acnet_pass_f64(s.ctypes.data, w, h, dst.ctypes.data, kernels.ctypes.data, biases.ctypes.data, threads)
    assert rc == 0, rc
    return dst


def ref_pass(arch, src, kernels, biases, out_dtype=None):
    src = np.ascontiguousarray(src)
    h, w = src.shape
    out_dtype = np.dtype(out_dtype or src.dtype)
    dst = np.zeros((2 * h, 2 * w), out_dtype)
    rc = ref().acref_pass(arch, src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype],
                          dst.ctypes.data, dst.strides[0], TYPE_OF[out_dtype], kernels.ctypes.data, biases.ctypes.data)
    assert rc == 0, rc
    return dst


def to_float(src):
    """toFloat of a whole array (Util.hpp:49-61), via the oracle's scalar routine."""
    src = np.ascontiguousarray(src)
    if src.dtype == np.float32:
        return src.copy()
    if src.dtype == np.uint8:
        return (src.astype(np.float32) / np.float32(255.0)).astype(np.float32)
    if src.dtype == np.uint16:
        return (src.astype(np.float32) / np.float32(65535.0)).astype(np.float32)
    return src.astype(np.float32)


def resize(src, dw, dh, mode):
    """The oracle's resize of an H x W (x C) array to dw x dh."""
    src = np.ascontiguousarray(src)
    h, w = src.shape[:2]
    c = 1 if src.ndim == 2 else src.shape[2]
    dst = np.zeros((dh, dw) + (() if c == 1 else (c,)), src.dtype)
    assert orc().orc_resize(src.ctypes.data, src.strides[0], w, h, c, TYPE_OF[src.dtype], dst.ctypes.data,
                            dst.strides[0], dw, dh, mode) == 0
    return dst


def oracle_pass8(src, kernels, biases, alphas, blocks, order=ORDER_FMA, threads=0):
    """ACNet<8> 2x pass of a 2-D array (u8/u16/f32); output of the same type."""
    src = np.ascontiguousarray(src)
    h, w = src.shape
    dst = np.zeros((2 * h, 2 * w), src.dtype)
    k, b, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases, alphas))
    assert orc().orc_acnet8_pass(src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype], dst.ctypes.data,
                                 dst.strides[0], k.ctypes.data, b.ctypes.data, a.ctypes.data, blocks, order, threads) == 0
    return dst


def oracle_pass8_f64(src_float, kernels, biases, alphas, blocks, threads=0):
    """ACNet<8> in float64: src_float is the toFloat'ed input (float32) -> saturated 2x output."""
    s = np.ascontiguousarray(src_float, dtype=np.float32)
    h, w = s.shape
    dst = np.zeros((2 * h, 2 * w), np.float64)
    k, b, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases, alphas))
    assert orc().orc_acnet8_pass_f64(s.ctypes.data, w, h, dst.ctypes.data, k.ctypes.data, b.ctypes.data, a.ctypes.data,
                                     blocks, threads) == 0
    return dst


def ref_pass8(arch, src, kernels, biases, alphas, blocks):
    """The same pass through the reference's own per-ISA kernels (oracle/_ref)."""
    src = np.ascontiguousarray(src)
    h, w = src.shape
    dst = np.zeros((2 * h, 2 * w), src.dtype)
    k, b, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases, alphas))
    rc = ref().acref_pass8(arch, src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype], dst.ctypes.data,
                           dst.strides[0], k.ctypes.data, b.ctypes.data, a.ctypes.data, blocks)
    assert rc == 0, rc
    return dst


def oracle_artcnn(src, kernels, biases, f, blocks=4, order=ORDER_FMA, threads=0):
    """ArtCNN<f> 2x pass of a 2-D array (u8/u16/f32); output of the same type."""
    src = np.ascontiguousarray(src)
    h, w = src.shape
    dst = np.zeros((2 * h, 2 * w), src.dtype)
    k, b = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases))
    assert orc().orc_artcnn_pass(src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype], dst.ctypes.data,
                                 dst.strides[0], k.ctypes.data, b.ctypes.data, f, blocks, order, threads) == 0
    return dst


def ref_artcnn(arch, src, kernels, biases, f, blocks=4):
    src = np.ascontiguousarray(src)
    h, w = src.shape
    dst = np.zeros((2 * h, 2 * w), src.dtype)
    k, b = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases))
    rc = ref().acref_artcnn(arch, src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype], dst.ctypes.data,
                            dst.strides[0], k.ctypes.data, b.ctypes.data, f, blocks)
    assert rc == 0, rc
    return dst


def oracle_fsrcnnx(src, kernels, biases, alphas, f, blocks=4, order=ORDER_FMA, threads=0):
    """FSRCNNX<f> 2x pass of a 2-D array (u8/u16/f32); output of the same type."""
    src = np.ascontiguousarray(src)
    h, w = src.shape
    dst = np.zeros((2 * h, 2 * w), src.dtype)
    k, b, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases, alphas))
    assert orc().orc_fsrcnnx_pass(src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype], dst.ctypes.data,
                                  dst.strides[0], k.ctypes.data, b.ctypes.data, a.ctypes.data, f, blocks, order,
                                  threads) == 0
    return dst


def ref_fsrcnnx(arch, src, kernels, biases, alphas, f, blocks=4):
    src = np.ascontiguousarray(src)
    h, w = src.shape
    dst = np.zeros((2 * h, 2 * w), src.dtype)
    k, b, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (kernels, biases, alphas))
    rc = ref().acref_fsrcnnx(arch, src.ctypes.data, src.strides[0], w, h, TYPE_OF[src.dtype], dst.ctypes.data,
                             dst.strides[0], k.ctypes.data, b.ctypes.data, a.ctypes.data, f, blocks)
    assert rc == 0, rc
    return dst
