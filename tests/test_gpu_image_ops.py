"""The image helpers of the C API on the device against the oracle: ac_resize for every
ResizeModes filter, element type, channel count and scale (up, down, odd ratios, the
anisotropic case), and ac_rgb2yuv / ac_rgba2yuva / ac_yuv2rgb / ac_yuva2rgba
(binding/c/src/Binding.cpp:125-159; core/src/ImageResize.cpp, core/src/ImageProcess.cpp).

Bars: colour conversion is bit-exact (same float steps on both sides).  Resize is bit-exact
except for the Lanczos filters, whose weights go through sinf (the device's and the host's
libm may differ by an ulp): there u8/u16 outputs may differ by 1 LSB and f16/f32 by 1e-6.
Parity with stb_image_resize2 itself is unpinned (the library is not in this tree); the
oracle's filters are pinned in test_oracle_color.py."""
import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import _native
from anime4kcpp_amd import processor as P

pytestmark = pytest.mark.gpu

DTYPES = [np.uint8, np.uint16, np.float16, np.float32]


def rand_image(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.uint8:
        return rng.integers(0, 256, shape, dtype=np.uint8)
    if dtype == np.uint16:
        return rng.integers(0, 65536, shape, dtype=np.uint16)
    # float content slightly outside [0, 1]: fromFloat clamps on the way out
    return (rng.random(shape, dtype=np.float32) * 1.2 - 0.1).astype(dtype)


def compare_resize(got, want, lanczos):
    if not lanczos:
        assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} of {got.size} differ"
        return
    if np.issubdtype(got.dtype, np.integer):
        d = np.abs(got.astype(np.int64) - want.astype(np.int64))
        assert d.max() <= 1
    else:
        assert np.abs(got.astype(np.float64) - want.astype(np.float64)).max() <= 1e-6 + 1e-3 * (got.dtype == np.float16)


# from a 23 x 17 source: up, down, odd, mixed, and exactly 2x and 4x (the staged upscale kernel)
SIZES = [(40, 26), (10, 7), (33, 21), (6, 19), (101, 13), (46, 34), (92, 68)]


@pytest.mark.parametrize("mode", range(17), ids=OL.RESIZE_MODES)
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("c", [1, 3])
def test_resize_matches_oracle(cuda, mode, dtype, c):
    shape = (17, 23) if c == 1 else (17, 23, c)
    src = rand_image(shape, dtype, seed=mode * 31 + c)
    lanczos = mode in OL.RESIZE_LANCZOS
    for dw, dh in SIZES:
        got = P.resize(src, (dw, dh), mode=mode)
        want = OL.resize(src, dw, dh, mode)
        assert got.shape == want.shape
        compare_resize(got, want, lanczos)


@pytest.mark.parametrize("c", [2, 4])
def test_resize_two_and_four_channels(cuda, c):
    src = rand_image((31, 45, c), np.uint8, seed=c)
    for mode in (OL.RESIZE_CATMULL_ROM, OL.RESIZE_BILINEAR, OL.RESIZE_SPLINE36):
        for dw, dh in ((90, 62), (20, 14), (180, 124), (91, 62)):
            assert np.array_equal(P.resize(src, (dw, dh), mode=mode), OL.resize(src, dw, dh, mode))


def test_resize_host_images_beyond_one_staging_chunk(cuda):
    """Host images of several 4 MiB staging chunks each way: a pageable source whose rows are
    narrower than its stride (a column slice), its page-locked copy (torch's pinned memory)
    and a pageable destination inside a wider array; all give the oracle's pixels."""
    import torch
    wide = rand_image((1100, 1700, 3), np.uint8, seed=5)
    src = wide[:, :1500]                                     # 4.95 MB of rows 5.1 KB apart
    want = OL.resize(np.ascontiguousarray(src), 3000, 2200, OL.RESIZE_BILINEAR)
    assert np.array_equal(P.resize(src, (3000, 2200), mode=OL.RESIZE_BILINEAR), want)
    pinned = torch.from_numpy(np.ascontiguousarray(src)).pin_memory().numpy()
    assert np.array_equal(P.resize(pinned, (3000, 2200), mode=OL.RESIZE_BILINEAR), want)
    big = np.zeros((2200, 3100, 3), np.uint8)
    si, di = P.Image(src, mapped=True), P.Image(big[:, :3000], mapped=True)
    _native.check_ac(_native.lib().ac_resize(si._p, di._p, 2.0, 2.0, OL.RESIZE_BILINEAR), "ac_resize")
    assert np.array_equal(big[:, :3000], want) and not big[:, 3000:].any()


@pytest.mark.parametrize("k", [2, 4])
def test_staged_upscale_multi_tile(cuda, k):
    """Exact K-times upscales over many 64 x 32 output tiles and ragged edges, every type."""
    for dtype in DTYPES:
        src = rand_image((37, 150, 2), dtype, seed=k)
        for mode in (OL.RESIZE_CATMULL_ROM, OL.RESIZE_BILINEAR, OL.RESIZE_POINT, 12):
            got = P.resize(src, (150 * k, 37 * k), mode=mode)
            compare_resize(got, OL.resize(src, 150 * k, 37 * k, mode), mode == 12)


@pytest.mark.parametrize("k", [2, 4])
def test_u8_upscale_four_pixel_kernel(cuda, k):
    """The u8 exact-upscale kernel (four outputs per thread; 256 x 16 output tiles of 4-row
    bands for small grids, 256 x 32 of 8-row bands from 512 workgroups) for the support-1 and
    support-2 filters: chroma resize + merge into RGB and RGBA, and the plain resize of one-
    and two-channel planes, over several tiles with ragged right and bottom edges, on
    dword-aligned pitches (output width 300 / 600) and unaligned ones (134 / 268: Y, RGB and
    the last thread's four columns take the byte path), bit-exact."""
    import torch
    from anime4kcpp_amd import hip
    for w, h in ((150, 37), (67, 21)):
        W, H = k * w, k * h
        for mode in (OL.RESIZE_CATMULL_ROM, OL.RESIZE_BILINEAR, 2):   # 2: Mitchell-Netravali
            for c in (3, 4):
                ys = rand_image((H, W), np.uint8, seed=W + c)
                small = rand_image((h, w, c - 1), np.uint8, seed=w + c + mode)
                got = hip.resize_yuv2rgb(torch.from_numpy(ys).cuda(), torch.from_numpy(small).cuda(), mode).cpu().numpy()
                up = OL.resize(small, W, H, mode)
                assert same(got, oracle_merge(np.concatenate([ys[..., None], up], axis=2))), (w, k, mode, c)
            for c in (1, 2):
                plane = rand_image((h, w) if c == 1 else (h, w, c), np.uint8, seed=3 * w + c + mode)
                got = hip.resize(torch.from_numpy(plane).cuda(), (W, H), mode).cpu().numpy()
                assert np.array_equal(got, OL.resize(plane, W, H, mode)), (w, k, mode, c)
    # a grid of 512 workgroups or more takes the 8-row bands (1024 x 4096 output)
    w, h = 1024 // k, 4096 // k
    ys = rand_image((4096, 1024), np.uint8, seed=7 + k)
    small = rand_image((h, w, 2), np.uint8, seed=9 + k)
    got = hip.resize_yuv2rgb(torch.from_numpy(ys).cuda(), torch.from_numpy(small).cuda(), OL.RESIZE_CATMULL_ROM).cpu().numpy()
    up = OL.resize(small, 1024, 4096, OL.RESIZE_CATMULL_ROM)
    assert same(got, oracle_merge(np.concatenate([ys[..., None], up], axis=2)))
    got = hip.resize(torch.from_numpy(small).cuda(), (1024, 4096), OL.RESIZE_BILINEAR).cpu().numpy()
    assert np.array_equal(got, OL.resize(small, 1024, 4096, OL.RESIZE_BILINEAR))


def test_resize_large_downscale_and_limits(cuda):
    """Down by 1/7 (29 taps per axis for Catmull-Rom) matches the oracle; a downscale past
    the tap limit is refused with an error, not truncated."""
    src = rand_image((140, 210), np.uint8, seed=3)
    got = P.resize(src, (30, 20), mode=OL.RESIZE_CATMULL_ROM)
    assert np.array_equal(got, OL.resize(src, 30, 20, OL.RESIZE_CATMULL_ROM))
    with pytest.raises(RuntimeError):
        P.resize(np.zeros((8, 4000), np.uint8), (40, 8), mode=OL.RESIZE_LANCZOS4)


def test_resize_reference_cases(cuda):
    """tests/core/src/ImageResizeTest.cpp through the C API: factor 1 hands back src, the
    sizes, point sampling, an out-of-range mode means bilinear, fx/fy <= 0 take dst's size."""
    L = _native.lib()
    src = rand_image((4, 4, 3), np.uint8, seed=1)
    s = P.Image(src, mapped=True)
    d = P.Image()
    assert L.ac_resize(s._p, d._p, 1.0, 1.0, 1) == 0 and d.c.ptr == s.c.ptr
    d2 = P.Image()
    assert L.ac_resize(s._p, d2._p, 2.0, 2.0, 1) == 0
    assert (d2.c.width, d2.c.height, d2.c.channels, d2.c.element_type) == (8, 8, 3, 0x001)
    d3 = P.Image()
    assert L.ac_resize(s._p, d3._p, 0.5, 0.5, 1) == 0 and (d3.c.width, d3.c.height) == (2, 2)
    d4 = P.Image()
    assert L.ac_resize(s._p, d4._p, 2.0, 1.0, 1) == 0 and (d4.c.width, d4.c.height) == (8, 4)
    p = np.array([[10, 20], [30, 40]], np.uint8)
    q = P.resize(p, (4, 4), mode=OL.RESIZE_POINT)
    assert np.array_equal(q, np.repeat(np.repeat(p, 2, 0), 2, 1))
    assert np.array_equal(P.resize(src, (9, 7), mode=99), OL.resize(src, 9, 7, OL.RESIZE_BILINEAR))
    pre = np.zeros((8, 8, 3), np.uint8)
    pi = P.Image(pre, mapped=True)
    assert L.ac_resize(s._p, pi._p, 0.0, 0.0, 1) == 0
    assert np.array_equal(pre, OL.resize(src, 8, 8, OL.RESIZE_CATMULL_ROM))
    assert L.ac_resize(None, pi._p, 2.0, 2.0, 1) == -22


def run_colour(fn, src):
    """fn(src mapped, empty dst) through the C API; the result copied out."""
    s = P.Image(src, mapped=True)
    d = P.Image()
    assert getattr(_native.lib(), fn)(s._p, d._p) == 0
    return d.numpy()


def oracle_split(rgb):
    h, w, c = rgb.shape
    t = OL.TYPE_OF[rgb.dtype]
    y = np.zeros((h, w), rgb.dtype)
    uv = np.zeros((h, w, c - 1), rgb.dtype)
    assert OL.orc().orc_rgb2yuv(rgb.ctypes.data, rgb.strides[0], w, h, c, t, y.ctypes.data, y.strides[0],
                                uv.ctypes.data, uv.strides[0]) == 0
    return np.concatenate([y[..., None], uv], axis=2)


def oracle_merge(yuv):
    h, w, c = yuv.shape
    t = OL.TYPE_OF[yuv.dtype]
    y = np.ascontiguousarray(yuv[..., 0])
    uv = np.ascontiguousarray(yuv[..., 1:])
    out = np.zeros((h, w, c), yuv.dtype)
    assert OL.orc().orc_yuv2rgb(y.ctypes.data, y.strides[0], uv.ctypes.data, uv.strides[0], w, h, c, t,
                                out.ctypes.data, out.strides[0]) == 0
    return out


def same(a, b):
    if np.issubdtype(a.dtype, np.floating):
        return np.array_equal(a.view(np.uint16 if a.dtype == np.float16 else np.uint32),
                              b.view(np.uint16 if b.dtype == np.float16 else np.uint32))
    return np.array_equal(a, b)


@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("alpha", [False, True])
def test_colour_conversions_match_oracle(cuda, dtype, alpha):
    c = 4 if alpha else 3
    rgb = rand_image((29, 53, c), dtype, seed=c)
    fwd, inv = ("ac_rgba2yuva", "ac_yuva2rgba") if alpha else ("ac_rgb2yuv", "ac_yuv2rgb")
    yuv = run_colour(fwd, rgb)
    assert yuv.shape == rgb.shape and yuv.dtype == rgb.dtype
    assert same(yuv, oracle_split(rgb))
    back = run_colour(inv, yuv)
    assert same(back, oracle_merge(yuv))


def test_colour_in_place_and_bad_shapes(cuda):
    L = _native.lib()
    rgb = rand_image((9, 11, 3), np.uint8, seed=5)
    want = oracle_split(rgb)
    buf = rgb.copy()
    im = P.Image(buf, mapped=True)
    assert L.ac_rgb2yuv(im._p, im._p) == 0           # rgb and yuv the same image
    assert np.array_equal(buf, want)
    wrong = P.Image(np.zeros((9, 10, 3), np.uint8), mapped=True)
    assert L.ac_rgb2yuv(P.Image(rgb, mapped=True)._p, wrong._p) == -256


def test_device_stage_entry_points(cuda):
    """achip_rgb2yuv / achip_yuv2rgb / achip_resize / achip_resize_yuv2rgb on device tensors
    equal the oracle's colour conversion, resize, and resize followed by the merge."""
    import torch
    from anime4kcpp_amd import hip
    for dtype, tdt in ((np.uint8, torch.uint8), (np.float32, torch.float32)):
        for c in (3, 4):
            rgb = rand_image((70, 130, c), dtype, seed=20 + c)
            y, uv = hip.rgb2yuv(torch.from_numpy(rgb).cuda())
            want = oracle_split(rgb)
            assert same(y.cpu().numpy(), want[..., 0]) and same(uv.cpu().numpy(), want[..., 1:])
            back = hip.yuv2rgb(y, uv).cpu().numpy()
            assert same(back, oracle_merge(want))
            for mode in (OL.RESIZE_CATMULL_ROM, OL.RESIZE_BILINEAR):
                small = np.ascontiguousarray(want[:35, :65, 1:])          # chroma at half size
                ys = np.ascontiguousarray(want[:70, :130, 0])
                got = hip.resize_yuv2rgb(torch.from_numpy(ys).cuda(), torch.from_numpy(small).cuda(), mode).cpu().numpy()
                up = OL.resize(small, 130, 70, mode)
                assert same(got, oracle_merge(np.concatenate([ys[..., None], up], axis=2)))
                r = hip.resize(torch.from_numpy(small).cuda(), (97, 51), mode).cpu().numpy()
                assert same(r, OL.resize(small, 97, 51, mode))
    with pytest.raises(_native.AchipError):
        hip.resize(torch.zeros((8, 4000), dtype=torch.uint8, device="cuda"), (40, 8), 12)


def test_device_colour_four_pixel_path(cuda):
    """The u8 split/merge with every base and pitch a multiple of 4 bytes runs four pixels per
    thread with dword loads and stores; rows whose width is not a multiple of four finish
    pixel by pixel.  Both must equal the oracle bit for bit."""
    import torch
    from anime4kcpp_amd import hip
    for c in (3, 4):
        for w in (1, 2, 4, 5, 131, 132, 133, 135):
            h = 9
            wp = (w + 3) // 4 * 4 + 4                      # padded pitch: a multiple of 4 pixels
            full = rand_image((h, wp, c), np.uint8, seed=40 + w + c)
            rgb_t = torch.from_numpy(full).cuda()[:, :w]
            y_t = torch.zeros((h, wp), dtype=torch.uint8, device="cuda")[:, :w]
            uv_t = torch.zeros((h, wp, c - 1), dtype=torch.uint8, device="cuda")[:, :w]
            hip.rgb2yuv(rgb_t, y_t, uv_t)
            want = oracle_split(np.ascontiguousarray(full[:, :w]))
            assert same(y_t.cpu().numpy(), want[..., 0]) and same(uv_t.cpu().numpy(), want[..., 1:]), (c, w)
            out_t = torch.zeros((h, wp, c), dtype=torch.uint8, device="cuda")[:, :w]
            hip.yuv2rgb(y_t, uv_t, out_t)
            assert same(out_t.cpu().numpy(), oracle_merge(want)), (c, w)
