"""FSRCNNX<8>/<16> on the device (achip_pass_net; the layer-by-layer FP32 MFMA path with the
conv1x1 fused into the last block) against the oracle (the reference's FMA summation order)
for all four weight sets: u8/u16 within the u8 tolerance (<= 1 LSB on >= 99.9% of samples,
max 2), f32 within 1e-5; the processor factory for the reference's model strings; and the
entry points' argument checks."""
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
def nets_fsr(cuda):
    out = {v: hip.NetFsr(v, 0) for v in hip.VARIANTS_FSR}
    yield out
    for n in out.values():
        n.close()


@pytest.mark.parametrize("variant", hip.VARIANTS_FSR)
@pytest.mark.parametrize("w,h", [(1, 1), (2, 5), (5, 3), (32, 16), (64, 16), (97, 61), (130, 35)])
def test_u8_matches_oracle(nets_fsr, variant, w, h):
    k, b, a, f, blocks = hip.load_weights_fsrcnnx(variant)
    x = synth.flat_shaded(w * 3 + h, w, h, dtype=np.uint8)
    got = nets_fsr[variant].upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    check_int(got, OL.oracle_fsrcnnx(x, k, b, a, f, blocks))


@pytest.mark.parametrize("variant", ["f8", "f16_dp"])
def test_u16_and_f32(nets_fsr, variant):
    k, b, a, f, blocks = hip.load_weights_fsrcnnx(variant)
    x16 = synth.flat_shaded(5, 71, 43, dtype=np.uint16)
    got = to_host(nets_fsr[variant].upscale(to_dev(x16, "cuda")))
    check_int(got, OL.oracle_fsrcnnx(x16, k, b, a, f, blocks))
    xf = synth.flat_shaded(6, 83, 51)
    gf = nets_fsr[variant].upscale(torch.from_numpy(xf).cuda()).cpu().numpy()
    assert np.abs(gf - OL.oracle_fsrcnnx(xf, k, b, a, f, blocks)).max() <= 1e-5


@pytest.mark.parametrize("variant", ["f8", "f16"])
def test_1080p_against_oracle(nets_fsr, variant):
    k, b, a, f, blocks = hip.load_weights_fsrcnnx(variant)
    x = synth.flat_shaded(15, 1920, 1080, dtype=np.uint8)
    got = nets_fsr[variant].upscale(torch.from_numpy(x).cuda()).cpu().numpy()
    check_int(got, OL.oracle_fsrcnnx(x, k, b, a, f, blocks))


@pytest.mark.parametrize("model,variant", [("fsrcnnx", "f8"), ("FSRCNNX-Distort-Plus", "f8_dp"), ("fsrcnnx-f16", "f16"),
                                           ("fsrcnnx_f16_dp", "f16_dp")])
def test_processor_model_strings(cuda, model, variant):
    """Processor::create parses the reference's FSRCNNX names (Processor.cpp:40-58) and
    process(src, factor) chains passes on the device."""
    k, b, a, f, blocks = hip.load_weights_fsrcnnx(variant)
    proc = P.Processor("hip", 0, model)
    assert proc.ok(), proc.error()
    x = synth.flat_shaded(23, 40, 28, dtype=np.uint8)
    rc, out = proc.process(x, 2.0)
    assert rc == 0
    check_int(out, OL.oracle_fsrcnnx(x, k, b, a, f, blocks))
    rc, out4 = proc.process(x, 4.0)
    assert rc == 0 and out4.shape == (112, 160)
    # factor 4 is two passes with a u8 image between them: the chain equals pass 2 run on
    # the device's own pass-1 output, which is held to the u8 bar of one pass
    rc2, again = proc.process(out, 2.0)
    assert rc2 == 0 and np.array_equal(out4, again)
    check_int(out4, OL.oracle_fsrcnnx(out, k, b, a, f, blocks))
    # against the oracle's own chain each side quantises its own pass-1 output, so pass 1's
    # rare 1-LSB differences enter pass 2 on one side only
    check_int(out4, OL.oracle_fsrcnnx(OL.oracle_fsrcnnx(x, k, b, a, f, blocks), k, b, a, f, blocks), frac=0.995, maxlsb=3)
    assert not P.Processor("hip-fp16", 0, model).ok()     # FP32 only


def test_entry_point_checks(nets_fsr):
    L = _native.lib()
    n = nets_fsr["f8"]
    src = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    dst = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    need = L.achip_net_workspace_bytes(n._h, 16, 16)
    assert need == 3 * 512 * 16           # three 8-channel f32 maps, rows of 512 bytes
    ws = torch.zeros(need, dtype=torch.uint8, device="cuda")
    args = (ctypes.c_void_p(src.data_ptr()), 16, 0x001, 16, 16, ctypes.c_void_p(dst.data_ptr()), 32, 0x001)
    assert L.achip_pass_net(n._h, *args, ctypes.c_void_p(ws.data_ptr()), need - 1, None) == -22   # workspace too small
    assert L.achip_pass_net(n._h, *args, ctypes.c_void_p(dst.data_ptr()), need, None) == -22      # overlaps dst
    assert L.achip_pass(n._h, 0, *args, None) == -22                                               # not an ACNetLegacy net
    assert L.achip_net_blocks(n._h) == 4
    assert L.achip_pass_net(n._h, *args, ctypes.c_void_p(ws.data_ptr()), need, None) == 0
    torch.cuda.synchronize()
    k, b, a, f, blocks = hip.load_weights_fsrcnnx("f8")
    h = ctypes.c_void_p()
    assert L.achip_net_fsrcnnx_create(0, 12, 4, k.ctypes.data, b.ctypes.data, a.ctypes.data, ctypes.byref(h)) == -22
    assert L.achip_net_fsrcnnx_create(0, 8, 1, k.ctypes.data, b.ctypes.data, a.ctypes.data, ctypes.byref(h)) == -22
