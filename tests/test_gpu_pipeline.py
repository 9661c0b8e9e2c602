"""The whole of Processor::process(src, dst, factor) on the device (processAll): colour split,
chained 2x passes, resize of the non-power-of-two remainder (Catmull-Rom), UV resize by the
full factor, merge -- against the same steps composed from the oracle
(core/src/processor/Processor.cpp:199-276).  The colour and resize steps use the same
arithmetic on both sides; differences come only from the 2x passes (1 LSB on a few pixels),
which the merge can spread to one more LSB."""
import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import processor as P
from anime4kcpp_amd import synth
from test_gpu_fp32 import check_int

pytestmark = pytest.mark.gpu


def ceil_log2(f):
    import math
    return 1 if f <= 2.0 else math.ceil(math.log2(f))


def oracle_pipeline(src, factor, k, b, pass_fn=None):
    """Processor.cpp:199-276 with dst empty, composed from oracle steps; pass_fn (default: the
    ACNetLegacy pass with k, b) is one 2x pass of a 2-D array."""
    if pass_fn is None:
        pass_fn = lambda a: OL.oracle_pass(a, k, b)   # noqa: E731
    h, w = src.shape[:2]
    c = 1 if src.ndim == 2 else src.shape[2]
    t = OL.TYPE_OF[src.dtype]
    power = ceil_log2(factor)
    fxy = factor / (1 << power)
    if c > 1:
        y = np.zeros((h, w), src.dtype)
        uv = np.zeros((h, w, c - 1), src.dtype)
        assert OL.orc().orc_rgb2yuv(src.ctypes.data, src.strides[0], w, h, c, t, y.ctypes.data, y.strides[0],
                                    uv.ctypes.data, uv.strides[0]) == 0
    else:
        y = src
    out = y
    for _ in range(power):
        out = pass_fn(out)
    ph, pw = out.shape
    fw, fh = int(pw * fxy), int(ph * fxy)
    if fxy != 1.0:
        r = np.zeros((fh, fw), src.dtype)
        assert OL.orc().orc_resize(out.ctypes.data, out.strides[0], pw, ph, 1, t, r.ctypes.data, r.strides[0], fw, fh, 1) == 0
        out = r
    if c == 1:
        return out
    uvr = np.zeros((fh, fw, c - 1), src.dtype)
    assert OL.orc().orc_resize(uv.ctypes.data, uv.strides[0], w, h, c - 1, t, uvr.ctypes.data, uvr.strides[0], fw, fh, 1) == 0
    rgb = np.zeros((fh, fw, c), src.dtype)
    assert OL.orc().orc_yuv2rgb(out.ctypes.data, out.strides[0], uvr.ctypes.data, uvr.strides[0], fw, fh, c, t,
                                rgb.ctypes.data, rgb.strides[0]) == 0
    return rgb


@pytest.fixture(scope="module")
def proc(cuda):
    p = P.Processor("hip", 0, "acnet-legacy-hdn1")
    assert p.ok(), p.error()
    return p


@pytest.mark.parametrize("c", [3, 4])
@pytest.mark.parametrize("factor", [2.0, 4.0])
def test_rgb_u8(proc, weights, c, factor):
    src = (synth.flat_shaded(31 + c, 97, 61, channels=c) * 255 + 0.5).astype(np.uint8)
    if c == 4:
        src[..., 3] = np.clip(src[..., 3], 40, 255)        # away from alpha ~ 0, where RGB is undefined
    rc, out = proc.process(src, factor)
    assert rc == 0, proc.error()
    want = oracle_pipeline(src, factor, *weights["hdn1"])
    assert out.shape == want.shape
    check_int(out, want, frac=0.999, maxlsb=2)


@pytest.mark.parametrize("factor", [3.0, 1.5, 2.5])
def test_gray_non_power_of_two(proc, weights, factor):
    src = synth.flat_shaded(41, 83, 57, dtype=np.uint8)
    rc, out = proc.process(src, factor)
    assert rc == 0, proc.error()
    want = oracle_pipeline(src, factor, *weights["hdn1"])
    assert out.shape == want.shape == (int((57 << ceil_log2(factor)) * factor / (1 << ceil_log2(factor))),
                                       int((83 << ceil_log2(factor)) * factor / (1 << ceil_log2(factor))))
    check_int(out, want, frac=0.999, maxlsb=2)


def test_rgb_f32(proc, weights):
    src = synth.flat_shaded(42, 64, 48, channels=3)
    rc, out = proc.process(src, 2.0)
    assert rc == 0, proc.error()
    want = oracle_pipeline(src, 2.0, *weights["hdn1"])
    assert np.abs(out - want).max() < 2e-5


def test_rgb_preallocated(proc, weights):
    src = (synth.flat_shaded(43, 40, 30, channels=3) * 255 + 0.5).astype(np.uint8)
    big = np.zeros((60, 100, 3), np.uint8)
    rc, _ = proc.process(src, 2.0, dst=big[:, 10:90])
    assert rc == 0, proc.error()
    check_int(big[:, 10:90], oracle_pipeline(src, 2.0, *weights["hdn1"]), frac=0.999, maxlsb=2)
    assert big[:, :10].max() == 0 and big[:, 90:].max() == 0


def test_reference_processor_test_shapes(proc):
    """tests/core/src/ProcessorTest.cpp:58-104 on the HIP processor: 32x32x3 u8 filled with
    128, factor 2 -> 64x64x3 of the same type; preallocated dst; factor 4 on 16x16x3."""
    src = np.full((32, 32, 3), 128, np.uint8)
    rc, out = proc.process(src, 2.0)
    assert rc == 0 and out.shape == (64, 64, 3) and out.dtype == np.uint8
    pre = np.zeros((64, 64, 3), np.uint8)
    rc, _ = proc.process(src, 2.0, dst=pre)
    assert rc == 0 and np.array_equal(pre, out)
    rc, out4 = proc.process(np.zeros((16, 16, 3), np.uint8), 4.0)
    assert rc == 0 and out4.shape == (64, 64, 3)


def test_wrong_preallocated_size_is_an_error(proc):
    src = np.zeros((10, 10, 3), np.uint8)
    rc, _ = proc.process(src, 2.0, dst=np.zeros((20, 21, 3), np.uint8))
    assert rc == -256 and "wrong size" in proc.error()
