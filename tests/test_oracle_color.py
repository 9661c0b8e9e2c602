"""Pins for the oracle's colour split/merge and resize (CPU).

Colour (core/src/ImageProcess.cpp:38-61, :114-139, :191-215, :276-308): a numpy float32
restatement with the same rounding order, exhaustive-ish over u8 triples, and the
reference's own round-trip bound (tests/core/src/ImageProcessTest.cpp:52-79: gray
within +-1).  Resize: stb_image_resize2 is absent, so parity is unpinned; the oracle is
pinned to closed-form 2x Catmull-Rom / bilinear taps, partition of unity and exact
reproduction of linear ramps in the interior."""
import ctypes

import numpy as np
import pytest

import oracle_lib as OL

f32 = np.float32


def np_from_float_u8(v):
    s = np.where(v < 0, f32(0), np.where(v < 1, v, f32(1))).astype(np.float32)
    return (s * f32(255) + f32(0.5)).astype(np.float32).astype(np.uint8)


def np_rgb2yuv(rgb, alpha=None):
    r, g, b = (rgb[..., i].astype(np.float32) / f32(255) for i in range(3))
    if alpha is not None:
        a = alpha.astype(np.float32) / f32(255)
        r, g, b = (r * a).astype(np.float32), (g * a).astype(np.float32), (b * a).astype(np.float32)
    y = ((f32(0.299) * r + f32(0.587) * g).astype(np.float32) + f32(0.114) * b).astype(np.float32)
    u = (f32(0.564) * (b - y).astype(np.float32) + f32(0.5)).astype(np.float32)
    v = (f32(0.713) * (r - y).astype(np.float32) + f32(0.5)).astype(np.float32)
    return np_from_float_u8(y), np_from_float_u8(u), np_from_float_u8(v)


def orc_split(rgb):
    h, w, c = rgb.shape
    rgb = np.ascontiguousarray(rgb)
    y = np.zeros((h, w), np.uint8)
    uv = np.zeros((h, w, c - 1), np.uint8)
    assert OL.orc().orc_rgb2yuv(rgb.ctypes.data, rgb.strides[0], w, h, c, OL.U8, y.ctypes.data, y.strides[0],
                                uv.ctypes.data, uv.strides[0]) == 0
    return y, uv


def orc_merge(y, uv):
    h, w = y.shape
    c = uv.shape[2] + 1
    out = np.zeros((h, w, c), np.uint8)
    assert OL.orc().orc_yuv2rgb(y.ctypes.data, y.strides[0], uv.ctypes.data, uv.strides[0], w, h, c, OL.U8,
                                out.ctypes.data, out.strides[0]) == 0
    return out


def test_split_closed_form_u8():
    vals = np.arange(0, 256, 15, dtype=np.uint8)
    r, g, b = np.meshgrid(vals, vals, vals, indexing="ij")
    rgb = np.stack([r, g, b], -1).reshape(-1, 1, 3)
    rgb = np.ascontiguousarray(rgb.reshape(len(vals) ** 2, len(vals), 3))
    y, uv = orc_split(rgb)
    yy, uu, vv = np_rgb2yuv(rgb)
    assert np.array_equal(y, yy) and np.array_equal(uv[..., 0], uu) and np.array_equal(uv[..., 1], vv)


def test_split_rgba_premultiplies():
    rng = np.random.default_rng(3)
    rgba = rng.integers(0, 256, (17, 23, 4), dtype=np.uint8)
    y, uva = orc_split(rgba)
    yy, uu, vv = np_rgb2yuv(rgba[..., :3], rgba[..., 3])
    assert np.array_equal(y, yy) and np.array_equal(uva[..., 0], uu) and np.array_equal(uva[..., 1], vv)
    assert np.array_equal(uva[..., 2], rgba[..., 3])


@pytest.mark.parametrize("c", [3, 4])
def test_gray_round_trip_within_one(c):
    for level in range(0, 256, 5):
        img = np.full((4, 4, c), level, np.uint8)
        if c == 4:
            img[..., 3] = 255
        y, uv = orc_split(img)
        back = orc_merge(y, uv)
        assert np.abs(back[..., :3].astype(int) - level).max() <= 1


def resize(src, dw, dh, mode):
    src = np.ascontiguousarray(src)
    h, w = src.shape[:2]
    c = 1 if src.ndim == 2 else src.shape[2]
    dst = np.zeros((dh, dw) + (() if c == 1 else (c,)), src.dtype)
    assert OL.orc().orc_resize(src.ctypes.data, src.strides[0], w, h, c, OL.TYPE_OF[src.dtype], dst.ctypes.data,
                               dst.strides[0], dw, dh, mode) == 0
    return dst


def closed_2x(x, mode):
    """2x: even outputs sample i - 1/4, odd i + 1/4; taps (1/4, 3/4) bilinear or
    (-3/128, 29/128, 111/128, -9/128) Catmull-Rom, edge clamped, rows then columns."""
    h, w = x.shape
    if mode == OL.RESIZE_BILINEAR:
        even, odd = [(-1, 0.25), (0, 0.75)], [(0, 0.75), (1, 0.25)]
    else:
        even = [(-2, -0.0234375), (-1, 0.2265625), (0, 0.8671875), (1, -0.0703125)]
        odd = [(-1, -0.0703125), (0, 0.8671875), (1, 0.2265625), (2, -0.0234375)]
    tmp = np.zeros((h, 2 * w), np.float32)
    for X in range(2 * w):
        for off, wt in (odd if X & 1 else even):
            tmp[:, X] = (tmp[:, X] + f32(wt) * x[:, min(max(X // 2 + off, 0), w - 1)]).astype(np.float32)
    out = np.zeros((2 * h, 2 * w), np.float32)
    for Y in range(2 * h):
        for off, wt in (odd if Y & 1 else even):
            out[Y] = (out[Y] + f32(wt) * tmp[min(max(Y // 2 + off, 0), h - 1)]).astype(np.float32)
    return np.clip(out, 0, 1)


@pytest.mark.parametrize("mode", [OL.RESIZE_BILINEAR, OL.RESIZE_CATMULL_ROM])
def test_resize_2x_closed_form(mode):
    x = np.random.default_rng(5).random((9, 13), dtype=np.float32)
    assert np.array_equal(resize(x, 26, 18, mode), closed_2x(x, mode))


@pytest.mark.parametrize("mode", [OL.RESIZE_BILINEAR, OL.RESIZE_CATMULL_ROM])
@pytest.mark.parametrize("dw,dh", [(40, 30), (26, 18), (10, 9), (7, 5), (52, 36)])
def test_resize_partition_of_unity_and_symmetry(mode, dw, dh):
    const = np.full((9, 13), 0.4, np.float32)
    assert np.abs(resize(const, dw, dh, mode) - 0.4).max() < 1e-6
    # mirroring the input mirrors the output (pixel-centre alignment, symmetric filters)
    x = np.random.default_rng(7).random((9, 13), dtype=np.float32)
    a = resize(x, dw, dh, mode)
    b = resize(np.ascontiguousarray(x[:, ::-1]), dw, dh, mode)[:, ::-1]
    assert np.abs(a - b).max() < 4e-6   # taps summed in mirrored order


@pytest.mark.parametrize("mode", [OL.RESIZE_BILINEAR, OL.RESIZE_CATMULL_ROM])
@pytest.mark.parametrize("dw,dh", [(40, 30), (26, 18), (52, 36), (17, 11)])
def test_resize_up_reproduces_ramps(mode, dw, dh):
    """Upscaling interpolates: a horizontal ramp inside (0, 1) comes back exactly at every
    output pixel whose taps stay inside the image."""
    ramp = np.tile(np.linspace(0.1, 0.9, 13, dtype=np.float32), (9, 1))
    out = resize(ramp, dw, dh, mode)
    sx = (np.arange(dw) + 0.5) * 13 / dw - 0.5
    want = 0.1 + 0.8 * sx / 12
    inner = (sx > 1.0) & (sx < 11.0)
    assert np.abs(out[:, inner] - want[inner]).max() < 2e-6


# ---- every ResizeModes filter (ImageResize.cpp:53-134, 193-263) ----

def fw(mode, x):
    return float(OL.orc().orc_filter_weight(mode, float(x)))


XS = np.linspace(-4.5, 4.5, 721)


@pytest.mark.parametrize("mode", sorted(OL.RESIZE_BC))
def test_cubic_filters_match_mitchell_netravali(mode):
    """The (b, c) family as Mitchell and Netravali publish it (Reconstruction Filters in
    Computer Graphics, 1988, eq. 8), in float64, against the oracle's Horner form."""
    B, C = OL.RESIZE_BC[mode]

    def k(x):
        x = abs(x)
        if x < 1:
            return ((12 - 9 * B - 6 * C) * x ** 3 + (-18 + 12 * B + 6 * C) * x ** 2 + (6 - 2 * B)) / 6
        if x < 2:
            return ((-B - 6 * C) * x ** 3 + (6 * B + 30 * C) * x ** 2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) / 6
        return 0.0

    assert max(abs(fw(mode, x) - k(x)) for x in XS) < 2e-6
    assert abs(fw(mode, 0) - (1 - B / 3)) < 1e-6 and abs(fw(mode, 1) - B / 6) < 1e-6 and fw(mode, 2) == 0


@pytest.mark.parametrize("mode", sorted(OL.RESIZE_LANCZOS))
def test_lanczos_matches_numpy_sinc(mode):
    a = OL.RESIZE_LANCZOS[mode]
    want = np.where(np.abs(XS) < a, np.sinc(XS) * np.sinc(XS / a), 0.0)
    got = np.array([fw(mode, x) for x in XS])
    assert np.abs(got - want).max() < 2e-6
    assert OL.orc().orc_filter_support(mode) == a


@pytest.mark.parametrize("mode", sorted(OL.RESIZE_SPLINE))
def test_spline_kernels_interpolate_and_reproduce_lines(mode):
    """Spline16/36/64 are interpolating piecewise cubics: k(0) = 1, k(n) = 0, continuous at
    the knots, and at every phase the integer shifts sum to 1 and reproduce a line."""
    s = OL.RESIZE_SPLINE[mode]
    assert OL.orc().orc_filter_support(mode) == s
    assert fw(mode, 0) == 1.0
    for n in range(1, s + 1):
        assert abs(fw(mode, n)) < 1e-6 and abs(fw(mode, -n)) < 1e-6
        assert abs(fw(mode, n - 1e-4) - fw(mode, n + 1e-4)) < 1e-3
    for t in np.linspace(0, 1, 17):
        taps = [(n, fw(mode, t - n)) for n in range(-s, s + 2)]
        assert abs(sum(w for _, w in taps) - 1) < 2e-6
        assert abs(sum(w * n for n, w in taps) - t) < 2e-6


def test_bilinear_and_point_kernels():
    for x in XS:
        assert abs(fw(OL.RESIZE_BILINEAR, x) - max(0.0, 1 - abs(x))) < 1e-7
        assert fw(OL.RESIZE_POINT, x) == 1.0
    assert OL.orc().orc_filter_support(OL.RESIZE_POINT) == 0.5
    assert OL.orc().orc_filter_support(OL.RESIZE_BILINEAR) == 1.0


def test_point_sampling_known_answer():
    """tests/core/src/ImageResizeTest.cpp:46-78: 2x2 -> 4x4 RESIZE_POINT repeats pixels."""
    src = np.array([[10, 20], [30, 40]], np.uint8)
    want = np.array([[10, 10, 20, 20], [10, 10, 20, 20], [30, 30, 40, 40], [30, 30, 40, 40]], np.uint8)
    assert np.array_equal(OL.resize(src, 4, 4, OL.RESIZE_POINT), want)


@pytest.mark.parametrize("mode", range(17))
@pytest.mark.parametrize("dw,dh", [(39, 27), (7, 5), (4, 3)])
def test_every_mode_keeps_constants_and_mirrors(mode, dw, dh):
    const = np.full((9, 13), 0.4, np.float32)
    assert np.abs(OL.resize(const, dw, dh, mode) - 0.4).max() < 2e-6
    x = np.random.default_rng(11).random((9, 13), dtype=np.float32)
    a = OL.resize(x, dw, dh, mode)
    b = OL.resize(np.ascontiguousarray(x[:, ::-1]), dw, dh, mode)[:, ::-1]
    if mode == OL.RESIZE_POINT:
        return   # ties (centre exactly between two pixels) go to the upper one: not mirror-symmetric
    assert np.abs(a - b).max() < 5e-6


@pytest.mark.parametrize("mode", [m for m in range(17) if m not in (2, 6, 7, 8, 9)])
def test_interpolating_modes_pass_samples_through_at_3x(mode):
    """At 3x every third output pixel sits on a source centre; filters with k(0) = 1 and
    k(n) = 0 (b = 0 cubics, Lanczos, splines, tent, point) reproduce the source there."""
    x = np.random.default_rng(13).random((7, 9), dtype=np.float32)
    out = OL.resize(x, 27, 21, mode)
    assert np.abs(out[1::3, 1::3] - x).max() < 1e-6


# ---- an independent pin: the published BT.601 definition -----------------------------------
# The reference's constants (ImageProcess.cpp:38-61, :191-215) are the BT.601 ones rounded to
# three decimals: luma weights Kr = 0.299, Kg = 0.587, Kb = 0.114 (ITU-R BT.601), chroma
# scales 0.5 / (1 - Kb) and 0.5 / (1 - Kr), and the inverse matrix of that transform.  Deriving
# every coefficient from the three published luma weights (not from the oracle's own table)
# and comparing on f32 images, where no rounding to u8 can hide a misread digit, pins each of
# the nine coefficients to within the reference's three-decimal rounding.
KR, KG, KB = 0.299, 0.587, 0.114


def bt601_forward(r, g, b):
    y = KR * r + KG * g + KB * b
    return y, 0.5 / (1 - KB) * (b - y) + 0.5, 0.5 / (1 - KR) * (r - y) + 0.5


def bt601_inverse(y, u, v):
    u, v = u - 0.5, v - 0.5
    r = y + 2 * (1 - KR) * v
    b = y + 2 * (1 - KB) * u
    g = y - 2 * (1 - KB) * KB / KG * u - 2 * (1 - KR) * KR / KG * v
    return r, g, b


def orc_split_f32(rgb):
    h, w, c = rgb.shape
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    y = np.zeros((h, w), np.float32)
    uv = np.zeros((h, w, c - 1), np.float32)
    assert OL.orc().orc_rgb2yuv(rgb.ctypes.data, rgb.strides[0], w, h, c, OL.F32, y.ctypes.data, y.strides[0],
                                uv.ctypes.data, uv.strides[0]) == 0
    return y, uv


def orc_merge_f32(y, uv):
    h, w = y.shape
    c = uv.shape[2] + 1
    y = np.ascontiguousarray(y, dtype=np.float32)
    uv = np.ascontiguousarray(uv, dtype=np.float32)
    out = np.zeros((h, w, c), np.float32)
    assert OL.orc().orc_yuv2rgb(y.ctypes.data, y.strides[0], uv.ctypes.data, uv.strides[0], w, h, c, OL.F32,
                                out.ctypes.data, out.strides[0]) == 0
    return out


def unit_colours():
    """The eight corners of the RGB cube plus a seeded spread inside it (f32)."""
    corners = np.array([[r, g, b] for r in (0, 1) for g in (0, 1) for b in (0, 1)], np.float64)
    rng = np.random.default_rng(601)
    inner = rng.uniform(0.05, 0.95, (64, 3))
    return np.concatenate([corners, inner]).reshape(8, 9, 3).astype(np.float32)


def test_split_matches_bt601_definition():
    rgb = unit_colours()
    y, uv = orc_split_f32(rgb)
    ey, eu, ev = bt601_forward(*(rgb[..., i].astype(np.float64) for i in range(3)))
    sat = lambda a: np.clip(a, 0.0, 1.0)   # noqa: E731  (fromFloat saturates f32 too)
    # three-decimal constants: |coefficient error| <= 5e-4 per unit of input
    assert np.abs(y - sat(ey)).max() < 1.5e-3
    assert np.abs(uv[..., 0] - sat(eu)).max() < 1.5e-3
    assert np.abs(uv[..., 1] - sat(ev)).max() < 1.5e-3
    # primaries single out each luma weight (any misread digit moves Y by >= 1e-3)
    assert abs(float(y[0, 4]) - KR) < 1e-6 and abs(float(y[0, 2]) - KG) < 1e-6 and abs(float(y[0, 1]) - KB) < 1e-6


def test_merge_matches_bt601_inverse():
    rng = np.random.default_rng(602)
    yuv = rng.uniform(0.2, 0.8, (6, 11, 3)).astype(np.float32)
    out = orc_merge_f32(yuv[..., 0], yuv[..., 1:])
    er, eg, eb = bt601_inverse(*(yuv[..., i].astype(np.float64) for i in range(3)))
    for got, want in ((out[..., 0], er), (out[..., 1], eg), (out[..., 2], eb)):
        want = np.clip(want, 0.0, 1.0)
        assert np.abs(got - want).max() < 2e-3
    # a chroma-only input isolates each inverse coefficient: v alone moves R by 1.403 v and
    # G by -0.714 v, u alone moves B by 1.773 u and G by -0.344 u
    y0 = np.full((1, 2), 0.5, np.float32)
    uv = np.array([[[0.5, 0.6], [0.6, 0.5]]], np.float32)
    o = orc_merge_f32(y0, uv)
    assert abs(float(o[0, 0, 0]) - 0.5 - 0.1 * 2 * (1 - KR)) < 2e-4       # R from v
    assert abs(float(o[0, 0, 1]) - 0.5 + 0.1 * 2 * (1 - KR) * KR / KG) < 2e-4
    assert abs(float(o[0, 1, 2]) - 0.5 - 0.1 * 2 * (1 - KB)) < 2e-4       # B from u
    assert abs(float(o[0, 1, 1]) - 0.5 + 0.1 * 2 * (1 - KB) * KB / KG) < 2e-4


def test_planted_coefficient_misread_fails():
    """A coefficient off in its third decimal (0.578 for 0.587) must fail the BT.601 pin
    even though the oracle's own restatement would carry the same misread."""
    rgb = unit_colours()
    r, g, b = (rgb[..., i].astype(np.float64) for i in range(3))
    bad_y = 0.299 * r + 0.578 * g + 0.114 * b
    ey, _, _ = bt601_forward(r, g, b)
    assert np.abs(np.clip(bad_y, 0, 1) - np.clip(ey, 0, 1)).max() >= 1.5e-3
