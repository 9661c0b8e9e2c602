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


def test_u8_from_float_fma():
    """fromFloat<u8> (Util.hpp:63-78) is (unsigned)(s * 255 + 0.5f) with the multiply and the
    add rounded separately; the kernels compute it as one FMA.  Each form is within one ulp
    of the exact s * 255 + 0.5 (at most 2^-16 below 256), so the truncations can only part
    where the exact value lies within 2^-15 of an integer n, i.e. s within 2^-15 / 255 of
    (n - 0.5) / 255.  Every float in a window of +-4096 ulps around each of those 255 points
    (several times wider than needed) is compared here, with the FMA evaluated exactly in
    float64 (s * 255 + 0.5 needs at most 41 significant bits for s >= 2^-9) and rounded once
    to float32."""
    centres = ((np.arange(1, 256, dtype=np.float64) - 0.5) / 255.0).astype(np.float32).view(np.uint32)
    offs = np.arange(-4096, 4097, dtype=np.int64)
    bits = (centres.astype(np.int64)[:, None] + offs[None, :]).ravel()
    s = bits.astype(np.uint32).view(np.float32)
    s = s[(s >= 0) & (s <= 1)]
    assert s.min() >= 2.0 ** -9
    separate = ((s * np.float32(255)) + np.float32(0.5)).astype(np.uint32)
    fused = (s.astype(np.float64) * 255.0 + 0.5).astype(np.float32).astype(np.uint32)
    assert np.array_equal(separate, fused)
    # the windows do hold inputs whose two float results differ before the truncation
    sep_f = (s * np.float32(255)) + np.float32(0.5)
    fused_f = (s.astype(np.float64) * 255.0 + 0.5).astype(np.float32)
    assert (sep_f != fused_f).any()
    # and both land on each boundary: codes n - 1 and n occur in every window
    assert len(np.unique(separate)) == 256
