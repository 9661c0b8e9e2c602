"""The layer-by-layer families (ACNet<8>, ArtCNN, FSRCNNX) through the whole of
Processor::process (colour split, chained 2x passes, Catmull-Rom remainder, chroma resize,
merge; core/src/processor/Processor.cpp:199-276) on RGB/RGBA u8 images, against the same steps
composed from the oracle; and each family on f16 images (fromFloat<half> of the FP32 result:
at least 99% of samples identical, the rest within one half-precision step)."""
import numpy as np
import pytest
import torch

import oracle_lib as OL
from anime4kcpp_amd import hip, synth
from anime4kcpp_amd import processor as P
from test_gpu_fp32 import check_int
from test_gpu_pipeline import oracle_pipeline

pytestmark = pytest.mark.gpu

FAMILIES = {
    "acnet-b4-hdn": ("acnet8", "b4_hdn"),
    "artcnn-c4f16-ds": ("artcnn", "f16_ds"),
    "fsrcnnx-f16": ("fsrcnnx", "f16"),
    "fsrcnnx": ("fsrcnnx", "f8"),
}


def pass_fn(family, variant):
    if family == "acnet8":
        k, b, a, n = hip.load_weights8(variant)
        return lambda x: OL.oracle_pass8(x, k, b, a, n)
    if family == "artcnn":
        k, b, f, n = hip.load_weights_artcnn(variant)
        return lambda x: OL.oracle_artcnn(x, k, b, f, n)
    k, b, a, f, n = hip.load_weights_fsrcnnx(variant)
    return lambda x: OL.oracle_fsrcnnx(x, k, b, a, f, n)


@pytest.mark.parametrize("model", list(FAMILIES))
@pytest.mark.parametrize("c,factor", [(3, 2.0), (4, 2.0), (3, 3.0), (3, 4.0)])
def test_rgb_pipeline(cuda, model, c, factor):
    proc = P.Processor("hip", 0, model)
    assert proc.ok(), proc.error()
    src = (synth.flat_shaded(50 + c, 57, 39, channels=c) * 255 + 0.5).astype(np.uint8)
    if c == 4:
        src[..., 3] = np.clip(src[..., 3], 40, 255)        # away from alpha ~ 0, where RGB is undefined
    rc, out = proc.process(src, factor)
    assert rc == 0, proc.error()
    want = oracle_pipeline(src, factor, None, None, pass_fn(*FAMILIES[model]))
    assert out.shape == want.shape
    check_int(out, want, frac=0.999, maxlsb=2)


@pytest.mark.parametrize("family,variant,cls", [("acnet8", "b8", hip.Net8), ("artcnn", "f32", hip.NetArt),
                                                ("fsrcnnx", "f8_dp", hip.NetFsr)])
def test_f16_images(cuda, family, variant, cls):
    x = synth.flat_shaded(60, 53, 37).astype(np.float16)
    net = cls(variant, 0)
    got = net.upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    net.close()
    want = pass_fn(family, variant)(x)
    assert got.dtype == want.dtype == np.float16
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    assert (d == 0).mean() >= 0.99 and d.max() <= 2.0 ** -11 * 2, (float((d == 0).mean()), float(d.max()))


@pytest.mark.parametrize("model", ["acnet-b4-hdn", "artcnn-c4f16-ds", "fsrcnnx"])
def test_threads_share_one_processor(cuda, model):
    """One layer-by-layer processor, eight threads at once: each thread's stream, buffers and
    feature-map workspace are its own (sizes differ per thread, so workspaces grow apart)."""
    import threading
    proc = P.Processor("hip", 0, model)
    assert proc.ok(), proc.error()
    run = pass_fn(*FAMILIES[model])
    srcs = [synth.flat_shaded(200 + i, 40 + 7 * i, 30 + 3 * i, dtype=np.uint8) for i in range(8)]
    outs = [None] * 8

    def work(i):
        for _ in range(3):
            outs[i] = proc.process(srcs[i], 2.0)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    for i in range(8):
        rc, out = outs[i]
        assert rc == 0, proc.error()
        check_int(out, run(srcs[i]))
