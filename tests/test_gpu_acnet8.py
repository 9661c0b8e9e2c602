"""ACNet<8> on the device (achip_pass8; the layer-by-layer FP32 MFMA path) against the oracle
(the reference's FMA summation order) for all twelve weight sets: u8/u16 within the u8
tolerance (<= 1 LSB on >= 99.9% of samples, max 2), f32 within 1e-5; the processor factory
for the reference's model strings; and the entry points' argument checks.

The 16-bit forms (achip_pass8_precision: conv3x3 layers on FP16 / BF16 MFMA, FP32
accumulation, 16-bit maps) against the FP64 restatement of the network, f32 output in [0,1]:
  FP16: max abs <= 4e-3, mean abs <= 4e-4
  BF16: max abs <= 3e-2, mean abs <= 4e-3
and u8 output >= 99 % (FP16) / >= 90 % (BF16) of pixels within 1 LSB of the FP32 oracle."""
import ctypes

import numpy as np
import pytest
import torch

import oracle_lib as OL
from anime4kcpp_amd import _native, hip, synth
from anime4kcpp_amd import processor as P
from test_gpu_fp32 import check_int, to_dev, to_host

pytestmark = pytest.mark.gpu

BOUNDS16 = {"fp16": (4e-3, 4e-4), "bf16": (3e-2, 4e-3)}
U8_AGREE = {"fp16": 0.99, "bf16": 0.90}


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
    # the 16-bit forms run the same passes on 16-bit MFMA
    for prec in ("fp16", "bf16"):
        p16 = P.Processor(f"hip-{prec}", 0, model)
        assert p16.ok(), p16.error()
        rc, o16 = p16.process(x, 2.0)
        assert rc == 0
        d = np.abs(o16.astype(int) - OL.oracle_pass8(x, k, b, a, blocks).astype(int))
        assert (d <= 1).mean() >= U8_AGREE[prec]


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


def err64(got, want):
    e = np.abs(got.astype(np.float64) - want)
    return e.max(), e.mean()


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("variant", ["b4", "b8_hdn", "b18_box_hdn"])
@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (17, 13), (64, 64), (130, 35)])
def test_16bit_error_vs_fp64(nets8, prec, variant, w, h):
    k, b, a, blocks = hip.load_weights8(variant)
    x = synth.flat_shaded(400 + w + h, w, h)
    got = nets8[variant].upscale(torch.from_numpy(x).cuda(), precision=prec).cpu().numpy()
    mx, mean = err64(got, OL.oracle_pass8_f64(x, k, b, a, blocks))
    print(f"{prec} {variant} {w}x{h}: max {mx:.3g} mean {mean:.3g}")
    assert mx <= BOUNDS16[prec][0] and mean <= BOUNDS16[prec][1]


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_16bit_1080p(nets8, prec):
    """The reference CLI's default model (B8 HDN) at 1080p: f32 error against FP64 and u8
    agreement with the FP32 oracle; a repeated pass gives the same bits."""
    k, b, a, blocks = hip.load_weights8("b8_hdn")
    net = nets8["b8_hdn"]
    xf = synth.flat_shaded(12, 1920, 1080)
    got = net.upscale(torch.from_numpy(xf).cuda(), precision=prec).cpu().numpy()
    mx, mean = err64(got, OL.oracle_pass8_f64(xf, k, b, a, blocks))
    assert mx <= BOUNDS16[prec][0] and mean <= BOUNDS16[prec][1]
    x8 = torch.from_numpy(synth.flat_shaded(13, 1920, 1080, dtype=np.uint8)).cuda()
    o1 = net.upscale(x8, precision=prec)
    o2 = net.upscale(x8, precision=prec)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)
    d = np.abs(o1.cpu().numpy().astype(int) - OL.oracle_pass8(x8.cpu().numpy(), k, b, a, blocks).astype(int))
    assert (d <= 1).mean() >= U8_AGREE[prec]


def truncated(blocks):
    """B4's weights with its last 4 - blocks blocks dropped: an ACNet<8> of `blocks` blocks."""
    k, b, a, _ = hip.load_weights8("b4")
    kk = np.concatenate([k[:72 + 576 * blocks], k[72 + 576 * 4:]])
    bb = np.concatenate([b[:8 + 8 * blocks], b[8 + 8 * 4:]])
    aa = a[:8 * (blocks + 1)].copy()
    return kk, bb, aa, blocks


@pytest.mark.parametrize("blocks", [1, 3])
def test_odd_block_counts(cuda, blocks):
    """An odd block count runs its last block with the final layer in one launch (16-bit) or
    alone (FP32); every form against the oracle / FP64."""
    wts = truncated(blocks)
    net = hip.Net8(weights=wts)
    assert net.blocks == blocks
    x8 = synth.flat_shaded(14 + blocks, 97, 61, dtype=np.uint8)
    check_int(net.upscale(torch.from_numpy(x8).cuda()).cpu().numpy(), OL.oracle_pass8(x8, *wts))
    xf = synth.flat_shaded(16 + blocks, 83, 45)
    ref = OL.oracle_pass8_f64(xf, *wts)
    assert np.abs(net.upscale(torch.from_numpy(xf).cuda()).cpu().numpy() - ref).max() <= 1e-5
    for prec in ("fp16", "bf16"):
        mx, mean = err64(net.upscale(torch.from_numpy(xf).cuda(), precision=prec).cpu().numpy(), ref)
        assert mx <= BOUNDS16[prec][0] and mean <= BOUNDS16[prec][1]
    net.close()


def test_16bit_u16_f16_and_bad_precision(nets8):
    k, b, a, blocks = hip.load_weights8("b8")
    net = nets8["b8"]
    x16 = synth.flat_shaded(15, 66, 40, dtype=np.uint16)
    got = to_host(net.upscale(to_dev(x16, "cuda"), precision="fp16"))
    d = np.abs(got.astype(np.int64) - OL.oracle_pass8(x16, k, b, a, blocks).astype(np.int64))
    assert d.max() <= 256            # 16-bit maps: within a few hundred u16 codes (~4e-3 of full scale)
    L = _native.lib()
    src = torch.zeros((8, 8), dtype=torch.uint8, device="cuda")
    dst = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    ws = torch.zeros(L.achip_pass8_workspace_bytes(8, 8), dtype=torch.uint8, device="cuda")
    assert L.achip_pass8_precision(net._h, 7, ctypes.c_void_p(src.data_ptr()), 8, 0x001, 8, 8,
                                   ctypes.c_void_p(dst.data_ptr()), 16, 0x001, ctypes.c_void_p(ws.data_ptr()),
                                   ws.numel(), None) == -22
    torch.cuda.synchronize()
