"""Scalar arithmetic the device kernels rely on, checked exhaustively on the CPU.

toFloat (core/internal/AC/Core/Internal/Util.hpp:49-61) divides an unsigned sample by its
maximum with IEEE float division.  The kernels compute the same correctly rounded quotient
as q0 = v * RN(1/d), then q = fma(fma(-q0, d, v), RN(1/d), q0).  This test evaluates that
sequence with the C library's fmaf for every u8 and u16 value and compares it with
numpy's float32 division (IEEE, correctly rounded)."""
import ctypes
import ctypes.util

import numpy as np


def _fmaf():
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    f = libm.fmaf
    f.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    f.restype = ctypes.c_float
    return f


def _check(maxv):
    fmaf = _fmaf()
    d = np.float32(maxv)
    r = np.float32(1.0) / d
    v = np.arange(maxv + 1, dtype=np.float32)
    exact = v / d
    q0 = v * r
    got = np.empty_like(exact)
    for i in range(maxv + 1):
        e = fmaf(float(-q0[i]), float(d), float(v[i]))
        got[i] = fmaf(e, float(r), float(q0[i]))
    assert np.array_equal(got.view(np.uint32), exact.view(np.uint32))
    # the uncorrected product is not enough: it differs for some inputs
    assert not np.array_equal(q0, exact)


def test_u8_quotient_exact():
    _check(255)


def test_u16_quotient_exact():
    _check(65535)
