"""ArtCNN<16>/<32> (core/include/AC/Core/Model/ArtCNN.hpp:11-56, CPUProcessor.cpp:1519-1602)
on the CPU: the committed weight blobs against their manifest, and the oracle's restatement
pinned bit for bit to the reference's own per-ISA kernels (oracle/_ref) and, through a
different route, to torch's float64 conv2d / relu / pixel_shuffle on replicate-padded maps."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import hip, synth


def test_blobs_match_manifest():
    man = json.load(open(os.path.join(hip.WEIGHTS_DIR, "MANIFEST.json")))["artcnn"]
    assert man.pop("source").endswith("Param/ArtCNN.p")
    assert sorted(man) == sorted(hip.VARIANTS_ART)
    for v in hip.VARIANTS_ART:
        k, b, f, blocks = hip.load_weights_artcnn(v)
        assert f == man[v]["channels"] and blocks == man[v]["blocks"] == 4
        assert k.size == f * 9 + f * f * 9 * (blocks + 1) + f * 4 * 9 and b.size == f + f * (blocks + 1) + 4
        for arr, kind in ((k, "kernels"), (b, "biases")):
            assert hashlib.sha256(arr.astype("<f4").tobytes()).hexdigest() == man[v][kind]["sha256"]
            assert np.isfinite(arr).all()


@pytest.mark.skipif(not OL.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("variant", hip.VARIANTS_ART)
def test_oracle_equals_reference_kernels(variant):
    k, b, f, blocks = hip.load_weights_artcnn(variant)
    for (w, h), dt in (((1, 1), np.uint8), ((3, 2), np.uint16), ((23, 17), np.uint8), ((30, 21), np.float32)):
        x = synth.flat_shaded(w * 37 + h, w, h, dtype=dt)
        for order, arch in ((OL.ORDER_FMA, 4), (OL.ORDER_GENERIC, 1)):
            assert np.array_equal(OL.oracle_artcnn(x, k, b, f, blocks, order=order), OL.ref_artcnn(arch, x, k, b, f, blocks))


def torch_artcnn(x, k, b, f, blocks, skip_residual=False):
    """float64 restatement with library routines: replicate padding + conv2d per layer, relu
    on the body, the head's features added after the last F->F layer, pixel_shuffle(2),
    clamp to [0, 1]."""
    import torch
    import torch.nn.functional as F
    t = torch.from_numpy(x.astype(np.float64))[None, None]
    k = torch.from_numpy(k.astype(np.float64))
    b = torch.from_numpy(b.astype(np.float64))

    def conv(inp, w, bias, cin, cout):
        # reference layout k[n * 9 * cin + (3 ky + kx) * cin + c] -> W[n][c][ky][kx]
        W = w.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
        return F.conv2d(F.pad(inp, (1, 1, 1, 1), mode="replicate"), W, bias)

    ko = 9 * f
    feat = conv(t, k[:ko], b[:f], 1, f)
    y = feat
    for l in range(1, blocks + 2):
        y = conv(y, k[ko:ko + 9 * f * f], b[f * l:f * l + f], f, f)
        ko += 9 * f * f
        y = y + feat if l == blocks + 1 and not skip_residual else (y if l == blocks + 1 else F.relu(y))
    out = F.pixel_shuffle(conv(y, k[ko:], b[f * (blocks + 2):], f, 4), 2)
    return out.clamp(0, 1)[0, 0].numpy()


@pytest.mark.parametrize("variant", ["f16", "f16_ds", "f32_dn"])
def test_oracle_against_torch_float64(variant):
    k, b, f, blocks = hip.load_weights_artcnn(variant)
    x = synth.flat_shaded(9, 27, 19)
    got = OL.oracle_artcnn(x, k, b, f, blocks)
    want = torch_artcnn(x, k, b, f, blocks)
    assert np.abs(got.astype(np.float64) - want).max() < 5e-6


def test_planted_residual_mistake_fails():
    """Without the head-feature skip connection the torch check must fail."""
    k, b, f, blocks = hip.load_weights_artcnn("f16")
    x = synth.flat_shaded(10, 16, 12)
    got = OL.oracle_artcnn(x, k, b, f, blocks)
    assert np.abs(got.astype(np.float64) - torch_artcnn(x, k, b, f, blocks, skip_residual=True)).max() > 1e-3
