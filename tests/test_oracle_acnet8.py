"""ACNet<8> (core/include/AC/Core/Model/ACNet.hpp:63-118, CPUProcessor.cpp:1417-1436) on the
CPU: the committed weight blobs against their manifest, and the oracle's restatement pinned
bit for bit to the reference's own per-ISA kernels (oracle/_ref) and, through a different
route, to torch's float64 conv2d / prelu / pixel_shuffle on replicate-padded maps."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import hip, synth


def test_blobs_match_manifest():
    man = json.load(open(os.path.join(hip.WEIGHTS_DIR, "MANIFEST.json")))["acnet8"]
    assert sorted(man) == sorted(hip.VARIANTS8)
    for v in hip.VARIANTS8:
        k, b, a, blocks = hip.load_weights8(v)
        assert blocks == man[v]["blocks"]
        assert k.size == 8 * 9 + 64 * 9 * blocks + 8 * 4 * 9 and b.size == 8 + 8 * blocks + 4 and a.size == 8 * (blocks + 1)
        for arr, kind in ((k, "kernels"), (b, "biases"), (a, "alphas")):
            assert hashlib.sha256(arr.astype("<f4").tobytes()).hexdigest() == man[v][kind]["sha256"]
            assert np.isfinite(arr).all()


@pytest.mark.skipif(not OL.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("variant", hip.VARIANTS8)
def test_oracle_equals_reference_kernels(variant):
    k, b, a, blocks = hip.load_weights8(variant)
    for (w, h), dt in (((1, 1), np.uint8), ((3, 2), np.uint16), ((23, 17), np.uint8), ((40, 33), np.float32)):
        x = synth.flat_shaded(w * 31 + h, w, h, dtype=dt)
        for order, arch in ((OL.ORDER_FMA, 4), (OL.ORDER_GENERIC, 1)):
            assert np.array_equal(OL.oracle_pass8(x, k, b, a, blocks, order=order), OL.ref_pass8(arch, x, k, b, a, blocks))


def torch_pass8(x, k, b, a, blocks):
    """float64 restatement with library routines: replicate padding + conv2d per layer,
    prelu, pixel_shuffle(2), nearest 2x of the source added, clamp to [0, 1]."""
    import torch
    import torch.nn.functional as F
    t = torch.from_numpy(x.astype(np.float64))[None, None]
    k = torch.from_numpy(k.astype(np.float64))
    b = torch.from_numpy(b.astype(np.float64))
    a = torch.from_numpy(a.astype(np.float64))

    def conv(inp, w, bias, cin, cout):
        # reference layout k[n * 9 * cin + (3 ky + kx) * cin + c] -> W[n][c][ky][kx]
        W = w.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
        return F.conv2d(F.pad(inp, (1, 1, 1, 1), mode="replicate"), W, bias)

    f = F.prelu(conv(t, k[:72], b[:8], 1, 8), a[:8])
    for l in range(1, blocks + 1):
        f = F.prelu(conv(f, k[72 + 576 * (l - 1):72 + 576 * l], b[8 * l:8 * l + 8], 8, 8), a[8 * l:8 * l + 8])
    kf = k[72 + 576 * blocks:]
    out = F.pixel_shuffle(conv(f, kf, b[8 + 8 * blocks:8 + 8 * blocks + 4], 8, 4), 2)
    out = out + F.interpolate(t, scale_factor=2, mode="nearest")
    return out.clamp(0, 1)[0, 0].numpy()


@pytest.mark.parametrize("variant", ["b4", "b8_hdn", "b18_box_hdn"])
def test_oracle_against_torch_float64(variant):
    k, b, a, blocks = hip.load_weights8(variant)
    x = synth.flat_shaded(7, 29, 21)
    got = OL.oracle_pass8(x, k, b, a, blocks)
    want = torch_pass8(x, k, b, a, blocks)
    assert np.abs(got.astype(np.float64) - want).max() < 5e-6


@pytest.mark.parametrize("variant", ["b4", "b8_hdn", "b18_box_hdn"])
def test_oracle_f64_against_torch_float64(variant):
    """The double-precision restatement (the reference the 16-bit forms are measured against)
    equals torch's float64 routines to rounding, and differs from a weight-order mistake."""
    k, b, a, blocks = hip.load_weights8(variant)
    x = synth.flat_shaded(9, 23, 17)
    want = torch_pass8(x, k, b, a, blocks)
    assert np.abs(OL.oracle_pass8_f64(x, k, b, a, blocks) - want).max() < 1e-12
    bad = k.copy()
    bad[72:72 + 576] = bad[72:72 + 576].reshape(8, 9, 8).transpose(0, 2, 1).reshape(-1)   # channel/tap swapped
    assert np.abs(OL.oracle_pass8_f64(x, bad, b, a, blocks) - want).max() > 1e-4


def test_planted_shuffle_mistake_fails():
    """A pixel shuffle with the sub-pixel axes swapped must not pass the torch check."""
    k, b, a, blocks = hip.load_weights8("b4_hdn")
    x = synth.flat_shaded(8, 16, 12)
    got = OL.oracle_pass8(x, k, b, a, blocks)
    bad = got.reshape(12, 2, 16, 2).transpose(0, 3, 2, 1).reshape(24, 32)
    assert np.abs(bad.astype(np.float64) - torch_pass8(x, k, b, a, blocks)).max() > 1e-3
