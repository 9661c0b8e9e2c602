"""Pins for the oracle (CPU restatement of the reference ACNetLegacy 2x pass).

Each check is independent of the oracle's own code path:
  - the reference's own per-ISA kernels compiled from its sources (oracle/_ref), bit-exact;
  - torch conv2d / conv_transpose2d in float64 on replicate-padded maps (library routine);
  - the constant-image closed form (each layer reduces to an 8-vector recurrence);
  - brute force on tiny images, written with per-layer np.pad(mode="edge");
  - translation equivariance away from the borders;
  - fromFloat/toFloat closed forms, exhaustively over all u8 and u16 codes;
  - a planted weight-layout mistake in the oracle must fail the torch check.
"""
import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import synth
from anime4kcpp_amd.hip import VARIANTS

rng = np.random.default_rng(12345)


# ---- torch float64 reference of the same network (library routines) ------------------

def torch_net_f64(x, k, b):
    """x: HxW float input (already toFloat'ed). Replicate padding before every conv."""
    import torch
    import torch.nn.functional as F
    t = torch.from_numpy(np.asarray(x, np.float64))[None, None]
    kk = torch.from_numpy(k.astype(np.float64))
    bb = torch.from_numpy(b.astype(np.float64))
    # L0: W[n][0][ky][kx] = k[n*9 + 3ky + kx]
    w0 = kk[:72].reshape(8, 1, 3, 3)
    t = F.relu(F.conv2d(F.pad(t, (1, 1, 1, 1), mode="replicate"), w0, bb[:8]))
    for l in range(1, 9):
        # W[n][c][ky][kx] = k[72 + 576(l-1) + n*72 + (3ky+kx)*8 + c]
        wl = kk[72 + 576 * (l - 1): 72 + 576 * l].reshape(8, 3, 3, 8).permute(0, 3, 1, 2)
        t = F.relu(F.conv2d(F.pad(t, (1, 1, 1, 1), mode="replicate"), wl, bb[8 * l: 8 * l + 8]))
    # deconv: Wt[c][0][dy][dx] = k2[(2dy+dx)*8 + c]
    wt = kk[4680:4712].reshape(2, 2, 8).permute(2, 0, 1)[:, None]
    out = F.conv_transpose2d(t, wt, stride=2)
    return out.clamp(0.0, 1.0)[0, 0].numpy()


def brute_f64(x, k, b):
    """Direct per-pixel loops with per-layer edge padding, float64."""
    h, w = x.shape
    feat = np.pad(x.astype(np.float64), 1, mode="edge")
    cur = np.zeros((h, w, 8))
    for i in range(h):
        for j in range(w):
            for n in range(8):
                s = b[n]
                for ky in range(3):
                    for kx in range(3):
                        s += feat[i + ky, j + kx] * k[n * 9 + ky * 3 + kx]
                cur[i, j, n] = max(s, 0.0)
    for l in range(1, 9):
        p = np.pad(cur, ((1, 1), (1, 1), (0, 0)), mode="edge")
        base = 72 + 576 * (l - 1)
        nxt = np.zeros_like(cur)
        for i in range(h):
            for j in range(w):
                for n in range(8):
                    s = float(b[8 * l + n])
                    for ky in range(3):
                        for kx in range(3):
                            for c in range(8):
                                s += p[i + ky, j + kx, c] * k[base + n * 72 + (ky * 3 + kx) * 8 + c]
                    nxt[i, j, n] = max(s, 0.0)
        cur = nxt
    out = np.zeros((2 * h, 2 * w))
    for i in range(h):
        for j in range(w):
            for dy in range(2):
                for dx in range(2):
                    v = sum(cur[i, j, c] * k[4680 + (dy * 2 + dx) * 8 + c] for c in range(8))
                    out[2 * i + dy, 2 * j + dx] = min(max(v, 0.0), 1.0)
    return out


# ---- element conversion closed forms ----------------------------------------------------

def test_to_float_exhaustive():
    import ctypes
    o = OL.orc()
    for v in range(256):
        x = ctypes.c_uint8(v)
        assert o.orc_to_float(ctypes.byref(x), OL.U8) == np.float32(v) / np.float32(255)
    vals = np.arange(65536, dtype=np.uint16)
    got = np.array([o.orc_to_float(ctypes.byref(ctypes.c_uint16(int(v))), OL.U16) for v in vals[::97]], np.float32)
    assert np.array_equal(got, vals[::97].astype(np.float32) / np.float32(65535))


def test_from_float_roundtrip_and_closed_form():
    import ctypes
    o = OL.orc()
    # fromFloat(toFloat(v)) == v for every u8 and u16 code
    for v in range(256):
        f = np.float32(v) / np.float32(255)
        out = ctypes.c_uint8()
        o.orc_from_float(ctypes.c_float(f), ctypes.byref(out), OL.U8)
        assert out.value == v
    for v in range(0, 65536, 13):
        f = np.float32(v) / np.float32(65535)
        out = ctypes.c_uint16()
        o.orc_from_float(ctypes.c_float(f), ctypes.byref(out), OL.U16)
        assert out.value == v
    # truncation after + 0.5 (not round-half-even), saturation, NaN -> 1 (Util.hpp:63-78)
    for f, want in [(-1.0, 0), (0.0, 0), (1.0, 255), (2.0, 255), (float("nan"), 255),
                    (0.5 / 255, 1),
                    # the product is rounded before the + 0.5: just below 0.5/255 still gives 1
                    (np.nextafter(np.float32(0.5 / 255), np.float32(0)), 1),
                    (np.float32(0.49 / 255), 0),
                    (2.5 / 255, 3), (254.5 / 255, 255)]:
        out = ctypes.c_uint8()
        o.orc_from_float(ctypes.c_float(f), ctypes.byref(out), OL.U8)
        sat = np.float32(0) if f < 0 else (np.float32(f) if f < 1 else np.float32(1))
        closed = int(np.float32(np.float32(sat) * np.float32(255)) + np.float32(0.5))
        assert out.value == want == closed, (f, out.value, want, closed)


def test_half_conversion_matches_ieee():
    import ctypes
    o = OL.orc()
    o.orc_float_to_half.argtypes = [ctypes.c_float]
    o.orc_float_to_half.restype = ctypes.c_uint16
    o.orc_half_to_float.argtypes = [ctypes.c_uint16]
    o.orc_half_to_float.restype = ctypes.c_float
    xs = np.concatenate([rng.random(2000, dtype=np.float32), rng.normal(0, 100, 2000).astype(np.float32),
                         np.array([0, -0.0, 1, 65504, 65520, 1e-8, 6e-8, 3e-5], np.float32)])
    for x in xs:
        assert o.orc_float_to_half(float(x)) == np.float32(x).astype(np.float16).view(np.uint16), x
    for hbits in range(0, 65536, 7):
        h = np.uint16(hbits).view(np.float16)
        if np.isnan(h):
            continue
        assert o.orc_half_to_float(hbits) == np.float32(h)


# ---- against the reference's own kernels ---------------------------------------------

@pytest.mark.skipif(not OL.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (17, 13), (40, 33)])
def test_oracle_bit_exact_vs_reference_kernels(weights, variant, dt, w, h):
    k, b = weights[variant]
    if dt == np.float32:
        src = rng.random((h, w), dtype=np.float32)
    else:
        src = rng.integers(0, np.iinfo(dt).max + 1, (h, w), dtype=dt)
    for arch, order in [(1, OL.ORDER_GENERIC), (4, OL.ORDER_FMA), (5, OL.ORDER_FMA)]:
        if not OL.ref().acref_has_arch(arch):
            continue
        want = OL.ref_pass(arch, src, k, b)
        got = OL.oracle_pass(src, k, b, order=order)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (arch, variant, dt)


# ---- against torch (library routine) and closed forms ---------------------------------

@pytest.mark.parametrize("variant", VARIANTS)
def test_oracle_vs_torch_f64(weights, variant):
    k, b = weights[variant]
    x = synth.flat_shaded(7, 37, 29)
    want = torch_net_f64(x, k, b)
    f64 = OL.oracle_pass_f64(x, k, b)
    assert np.abs(f64 - want).max() < 1e-12
    for order in (OL.ORDER_FMA, OL.ORDER_GENERIC):
        f32 = OL.oracle_pass(x, k, b, order=order)
        assert np.abs(f32 - want).max() < 3e-6


def test_constant_image_closed_form(weights):
    """A constant image stays constant under replicate padding: each layer is
    y = ReLU(Kbar y + b) with Kbar[n][c] = sum over taps; the output is a 2x2-periodic
    pattern o[dy][dx] = sat(k2[dy][dx] . y8)."""
    for variant in VARIANTS:
        k, b = (a.astype(np.float64) for a in weights[variant])
        for v in (0.0, 0.25, 0.5, 1.0):
            y = np.maximum(k[:72].reshape(8, 9).sum(1) * v + b[:8], 0)
            for l in range(1, 9):
                K = k[72 + 576 * (l - 1): 72 + 576 * l].reshape(8, 9, 8).sum(1)
                y = np.maximum(K @ y + b[8 * l: 8 * l + 8], 0)
            pat = np.clip(k[4680:].reshape(4, 8) @ y, 0, 1).reshape(2, 2)
            src = np.full((9, 11), v, np.float32)
            got64 = OL.oracle_pass_f64(src, *weights[variant])
            assert np.abs(got64 - np.tile(pat, (9, 11))).max() < 1e-12
            got32 = OL.oracle_pass(src, *weights[variant])
            assert np.abs(got32 - np.tile(pat, (9, 11))).max() < 2e-6


@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (3, 3), (4, 2)])
def test_brute_force_tiny(weights, w, h):
    k, b = weights["hdn0"]
    x = rng.random((h, w), dtype=np.float32)
    want = brute_f64(x, k.astype(np.float64), b.astype(np.float64))
    assert np.abs(OL.oracle_pass_f64(x, k, b) - want).max() < 1e-12
    assert np.abs(OL.oracle_pass(x, k, b) - want).max() < 3e-6


def test_translation_equivariance(weights):
    k, b = weights["hdn3"]
    big = synth.flat_shaded(9, 80, 70)
    a = OL.oracle_pass(big[0:60, 0:64], k, b)
    s = OL.oracle_pass(big[2:62, 3:67], k, b)
    # interior: input pixel (i, j) of the shifted crop is (i+2, j+3) of the first; the
    # receptive field is 19x19, so pixels >= 10 from every border are unaffected
    a_in = a[2 * 12: 2 * 48, 2 * 13: 2 * 50]
    s_in = s[2 * 10: 2 * 46, 2 * 10: 2 * 47]
    assert np.array_equal(a_in, s_in)


def test_upscale_quality(weights):
    """The trained network approximately preserves content: 2x2 mean of the output vs
    the input.  A wrong weight layout destroys this."""
    x = synth.flat_shaded(11, 96, 80)
    for variant in VARIANTS:
        out = OL.oracle_pass(x, *weights[variant]).astype(np.float64)
        down = out.reshape(80, 2, 96, 2).mean((1, 3))
        mse = ((down - x) ** 2).mean()
        psnr = 10 * np.log10(1 / mse)
        assert psnr > 30, (variant, psnr)


def test_planted_layout_mistake_is_caught(weights):
    """Reverse the channel order of one layer's kernel before it reaches the oracle:
    the torch check and the quality check must both fail."""
    k, b = weights["hdn0"]
    bad = k.copy()
    blk = bad[72 + 576 * 3: 72 + 576 * 4].reshape(8, 9, 8)
    bad[72 + 576 * 3: 72 + 576 * 4] = blk[:, :, ::-1].reshape(-1)   # channel order reversed within each tap
    x = synth.flat_shaded(13, 40, 30)
    assert np.abs(OL.oracle_pass(x, bad, b) - torch_net_f64(x, k, b)).max() > 1e-3
    out = OL.oracle_pass(x, bad, b).astype(np.float64)
    good = OL.oracle_pass(x, k, b).astype(np.float64)
    assert np.abs(out - good).max() > 1e-2
