"""ACNet<8> on the device (achip_pass8; the layer-by-layer FP32 MFMA path) against the oracle
(the reference's FMA summation order) for all twelve weight sets: u8/u16 within the u8
tolerance (<= 1 LSB on >= 99.9% of samples, max 2), f32 within 1e-5; the processor factory
for the reference's model strings; and the entry points' argument checks."""
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
def nets8(cuda):
    out = {v: hip.Net8(v, 0) for v in hip.VARIANTS8}
    yield out
    for n in out.values():
        n.close()


@pytest.mark.parametrize("variant", hip.VARIANTS8)
@pytest.mark.parametrize("w,h", [(1, 1), (5, 3), (64, 16), (97, 61), (130, 35)])
def test_u8_matches_oracle(nets8, variant, w, h):
    k, b, a, blocks = hip.load_weights8(variant)
    x = synth.flat_shaded(w * 7 + h, w, h, dtype=np.uint8)
    got = nets8[variant].upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    check_int(got, OL.oracle_pass8(x, k, b, a, blocks))


@pytest.mark.parametrize("variant", ["b4_hdn", "b8", "b18_box"])
def test_u16_and_f32(nets8, variant):
    k, b, a, blocks = hip.load_weights8(variant)
    x16 = synth.flat_shaded(3, 77, 45, dtype=np.uint16)
    got = to_host(nets8[variant].upscale(to_dev(x16, "cuda")))
    check_int(got, OL.oracle_pass8(x16, k, b, a, blocks))
    xf = synth.flat_shaded(4, 83, 51)
    gf = nets8[variant].upscale(torch.from_numpy(xf).cuda()).cpu().numpy()
    assert np.abs(gf - OL.oracle_pass8(xf, k, b, a, blocks)).max() <= 1e-5


def test_1080p_against_oracle(nets8):
    k, b, a, blocks = hip.load_weights8("b8_hdn")
    x = synth.flat_shaded(11, 1920, 1080, dtype=np.uint8)
    got = nets8["b8_hdn"].upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    check_int(got, OL.oracle_pass8(x, k, b, a, blocks))


@pytest.mark.parametrize("model,variant", [("acnet-b4-hdn", "b4_hdn"), ("acnet", "b8"), ("ACNet-B18-Box-HDN", "b18_box_hdn"),
                                           ("acnet-box", "b8_box"),
                                           ("acnet-f8b8-hdn", "b8_hdn")])   # the reference CLI's and filters' default
def test_processor_model_strings(cuda, model, variant):
    """Processor::create parses the reference's ACNet<8> names (Processor.cpp:130-186) and
    process(src, factor) chains passes on the device."""
    k, b, a, blocks = hip.load_weights8(variant)
    proc = P.Processor("hip", 0, model)
    assert proc.ok(), proc.error()
    x = synth.flat_shaded(21, 40, 28, dtype=np.uint8)
    rc, out = proc.process(x, 2.0)
    assert rc == 0
    check_int(out, OL.oracle_pass8(x, k, b, a, blocks))
    rc, out4 = proc.process(x, 4.0)
    assert rc == 0 and out4.shape == (112, 160)
    # factor 4 is two passes with a u8 image between them: the chain equals pass 2 run on
    # the device's own pass-1 output, which is held to the u8 bar of one pass
    rc2, again = proc.process(out, 2.0)
    assert rc2 == 0 and np.array_equal(out4, again)
    check_int(out4, OL.oracle_pass8(out, k, b, a, blocks))
    # against the oracle's own chain each side quantises its own pass-1 output, so pass 1's
    # rare 1-LSB differences enter pass 2 on one side only
    check_int(out4, OL.oracle_pass8(OL.oracle_pass8(x, k, b, a, blocks), k, b, a, blocks), frac=0.995, maxlsb=3)
    assert not P.Processor("hip-fp16", 0, model).ok()     # FP32 only


def test_entry_point_checks(nets8):
    L = _native.lib()
    n8 = nets8["b4"]
    src = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    dst = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    ws = torch.zeros(L.achip_pass8_workspace_bytes(16, 16), dtype=torch.uint8, device="cuda")
    args = (ctypes.c_void_p(src.data_ptr()), 16, 0x001, 16, 16, ctypes.c_void_p(dst.data_ptr()), 32, 0x001)
    assert L.achip_pass8(n8._h, *args, ctypes.c_void_p(ws.data_ptr()), ws.numel() - 1, None) == -22   # workspace too small
    assert L.achip_pass8(n8._h, *args, ctypes.c_void_p(src.data_ptr()), ws.numel(), None) == -22      # overlaps src
    assert L.achip_pass(n8._h, 0, *args, None) == -22                                                   # not an ACNetLegacy net
    assert L.achip_net_blocks(n8._h) == 4
    assert L.achip_pass8(n8._h, *args, ctypes.c_void_p(ws.data_ptr()), ws.numel(), None) == 0
    torch.cuda.synchronize()
