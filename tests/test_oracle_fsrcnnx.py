"""FSRCNNX<8>/<16> (core/include/AC/Core/Model/FSRCNNX.hpp:11-79, CPUProcessor.cpp:1629-1717)
on the CPU: the committed weight blobs against their manifest, and the oracle's restatement
pinned bit for bit to the reference's own per-ISA kernels (oracle/_ref) and, through a
different route, to torch's float64 conv2d / prelu / pixel_shuffle on replicate-padded maps."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import hip, synth


def test_blobs_match_manifest():
    man = json.load(open(os.path.join(hip.WEIGHTS_DIR, "MANIFEST.json")))["fsrcnnx"]
    assert man.pop("source").endswith("Param/FSRCNNX.p")
    assert sorted(man) == sorted(hip.VARIANTS_FSR)
    for v in hip.VARIANTS_FSR:
        k, b, a, f, blocks = hip.load_weights_fsrcnnx(v)
        assert f == man[v]["channels"] and blocks == man[v]["blocks"] == 4
        assert k.size == f * 25 + f * f * 9 * blocks + f * f + f * 4 * 9
        assert b.size == f + f * (blocks + 1) + 4 and a.size == f * (blocks + 1)
        for arr, kind in ((k, "kernels"), (b, "biases"), (a, "alphas")):
            assert hashlib.sha256(arr.astype("<f4").tobytes()).hexdigest() == man[v][kind]["sha256"]
            assert np.isfinite(arr).all()


@pytest.mark.skipif(not OL.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("variant", hip.VARIANTS_FSR)
def test_oracle_equals_reference_kernels(variant):
    k, b, a, f, blocks = hip.load_weights_fsrcnnx(variant)
    for (w, h), dt in (((1, 1), np.uint8), ((2, 5), np.uint8), ((3, 2), np.uint16), ((23, 17), np.uint8),
                       ((30, 21), np.float32)):
        x = synth.flat_shaded(w * 41 + h, w, h, dtype=dt)
        for order, arch in ((OL.ORDER_FMA, 4), (OL.ORDER_GENERIC, 1)):
            assert np.array_equal(OL.oracle_fsrcnnx(x, k, b, a, f, blocks, order=order),
                                  OL.ref_fsrcnnx(arch, x, k, b, a, f, blocks))


def torch_fsrcnnx(x, k, b, a, f, blocks, skip=True):
    """float64 restatement with library routines: replicate padding + conv2d per layer (5x5
    head, 3x3 body with prelu, the 1x1 added to the head's features then prelu), a 3x3
    F -> 4 tail, pixel_shuffle(2), clamp to [0, 1]."""
    import torch
    import torch.nn.functional as F
    t = torch.from_numpy(x.astype(np.float64))[None, None]
    k, b, a = (torch.from_numpy(v.astype(np.float64)) for v in (k, b, a))

    def conv(inp, w, bias, cin, cout, ks):
        # reference layout k[n * ks*ks * cin + (ks ky + kx) * cin + c] -> W[n][c][ky][kx]
        W = w.reshape(cout, ks, ks, cin).permute(0, 3, 1, 2)
        p = ks // 2
        return F.conv2d(F.pad(inp, (p, p, p, p), mode="replicate") if p else inp, W, bias)

    ko = 25 * f
    feat = conv(t, k[:ko], b[:f], 1, f, 5)
    y = feat
    for l in range(1, blocks + 1):
        y = F.prelu(conv(y, k[ko:ko + 9 * f * f], b[f * l:f * l + f], f, f, 3), a[f * (l - 1):f * l])
        ko += 9 * f * f
    y = conv(y, k[ko:ko + f * f], b[f * (blocks + 1):f * (blocks + 2)], f, f, 1)
    ko += f * f
    y = F.prelu(y + feat if skip else y, a[f * blocks:f * (blocks + 1)])
    out = F.pixel_shuffle(conv(y, k[ko:], b[f * (blocks + 2):], f, 4, 3), 2)
    return out.clamp(0, 1)[0, 0].numpy()


@pytest.mark.parametrize("variant", hip.VARIANTS_FSR)
def test_oracle_against_torch_float64(variant):
    k, b, a, f, blocks = hip.load_weights_fsrcnnx(variant)
    x = synth.flat_shaded(13, 27, 19)
    got = OL.oracle_fsrcnnx(x, k, b, a, f, blocks)
    want = torch_fsrcnnx(x, k, b, a, f, blocks)
    assert np.abs(got.astype(np.float64) - want).max() < 5e-6


def test_planted_skip_mistake_fails():
    """Without the head-feature skip connection the torch check must fail."""
    k, b, a, f, blocks = hip.load_weights_fsrcnnx("f8")
    x = synth.flat_shaded(14, 16, 12)
    got = OL.oracle_fsrcnnx(x, k, b, a, f, blocks)
    assert np.abs(got.astype(np.float64) - torch_fsrcnnx(x, k, b, a, f, blocks, skip=False)).max() > 1e-3
