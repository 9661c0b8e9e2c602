"""ArtCNN<16>/<32> on the device (achip_pass_net; the layer-by-layer FP32 MFMA path) against
the oracle (the reference's FMA summation order) for all six weight sets: u8/u16 within the
u8 tolerance (<= 1 LSB on >= 99.9% of samples, max 2), f32 within 1e-5; the processor
factory for the reference's model strings; and the entry points' argument checks."""
import ctypes

import numpy as np
import pytest
import torch

import oracle_lib as OL
from anime4kcpp_amd import _native, hip, synth
from anime4kcpp_amd import processor as P
from test_gpu_fp32 import check_int, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nets_art(cuda):
    out = {v: hip.NetArt(v, 0) for v in hip.VARIANTS_ART}
    yield out
    for n in out.values():
        n.close()


@pytest.mark.parametrize("variant", hip.VARIANTS_ART)
@pytest.mark.parametrize("w,h", [(1, 1), (5, 3), (32, 16), (64, 16), (97, 61), (130, 35)])
def test_u8_matches_oracle(nets_art, variant, w, h):
    k, b, f, blocks = hip.load_weights_artcnn(variant)
    x = synth.flat_shaded(w * 5 + h, w, h, dtype=np.uint8)
    got = nets_art[variant].upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    check_int(got, OL.oracle_artcnn(x, k, b, f, blocks))


@pytest.mark.parametrize("variant", ["f16", "f32_ds"])
def test_u16_and_f32(nets_art, variant):
    k, b, f, blocks = hip.load_weights_artcnn(variant)
    x16 = synth.flat_shaded(3, 71, 43, dtype=np.uint16)
    got = to_host(nets_art[variant].upscale(to_dev(x16, "cuda")))
    check_int(got, OL.oracle_artcnn(x16, k, b, f, blocks))
    xf = synth.flat_shaded(4, 83, 51)
    gf = nets_art[variant].upscale(torch.from_numpy(xf).cuda()).cpu().numpy()
    assert np.abs(gf - OL.oracle_artcnn(xf, k, b, f, blocks)).max() <= 1e-5


@pytest.mark.parametrize("variant", ["f16", "f32"])
def test_1080p_against_oracle(nets_art, variant):
    k, b, f, blocks = hip.load_weights_artcnn(variant)
    x = synth.flat_shaded(12, 1920, 1080, dtype=np.uint8)
    got = nets_art[variant].upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    check_int(got, OL.oracle_artcnn(x, k, b, f, blocks))


@pytest.mark.parametrize("model,variant", [("artcnn", "f16"), ("ArtCNN-C4F16-DS", "f16_ds"), ("artcnn-c4f32-dn", "f32_dn"),
                                           ("artcnn_f32", "f32")])
def test_processor_model_strings(cuda, model, variant):
    """Processor::create parses the reference's ArtCNN names (Processor.cpp:59-77) and
    process(src, factor) chains passes on the device."""
    k, b, f, blocks = hip.load_weights_artcnn(variant)
    proc = P.Processor("hip", 0, model)
    assert proc.ok(), proc.error()
    x = synth.flat_shaded(22, 40, 28, dtype=np.uint8)
    rc, out = proc.process(x, 2.0)
    assert rc == 0
    check_int(out, OL.oracle_artcnn(x, k, b, f, blocks))
    rc, out4 = proc.process(x, 4.0)
    assert rc == 0 and out4.shape == (112, 160)
    # factor 4 is two passes with a u8 image between them: the chain equals pass 2 run on
    # the device's own pass-1 output, which is held to the u8 bar of one pass
    rc2, again = proc.process(out, 2.0)
    assert rc2 == 0 and np.array_equal(out4, again)
    check_int(out4, OL.oracle_artcnn(out, k, b, f, blocks))
    # against the oracle's own chain each side quantises its own pass-1 output, so pass 1's
    # rare 1-LSB differences enter pass 2 on one side only
    check_int(out4, OL.oracle_artcnn(OL.oracle_artcnn(x, k, b, f, blocks), k, b, f, blocks), frac=0.995, maxlsb=3)
    assert not P.Processor("hip-fp16", 0, model).ok()     # FP32 only


def test_entry_point_checks(nets_art):
    L = _native.lib()
    n = nets_art["f16"]
    src = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    dst = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    need = L.achip_net_workspace_bytes(n._h, 16, 16)
    assert need == 3 * 1024 * 16          # three 16-channel f32 maps, rows of 1024 bytes
    ws = torch.zeros(need, dtype=torch.uint8, device="cuda")
    args = (ctypes.c_void_p(src.data_ptr()), 16, 0x001, 16, 16, ctypes.c_void_p(dst.data_ptr()), 32, 0x001)
    assert L.achip_pass_net(n._h, *args, ctypes.c_void_p(ws.data_ptr()), need - 1, None) == -22   # workspace too small
    assert L.achip_pass_net(n._h, *args, ctypes.c_void_p(src.data_ptr()), need, None) == -22      # overlaps src
    assert L.achip_pass(n._h, 0, *args, None) == -22                                               # not an ACNetLegacy net
    assert L.achip_pass8(n._h, *args, ctypes.c_void_p(ws.data_ptr()), need, None) == -22          # not an ACNet<8> net
    assert L.achip_net_blocks(n._h) == 4
    assert L.achip_pass_net(n._h, *args, ctypes.c_void_p(ws.data_ptr()), need, None) == 0
    torch.cuda.synchronize()
