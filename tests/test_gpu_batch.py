"""Persistent batched launches (achip_pass_batch): many images of uneven sizes in one
launch with a tile work-queue.  Each image's result must be identical to the same image
processed alone (same tiles, same arithmetic), and within tolerance of the oracle."""
import ctypes

import numpy as np
import pytest
import torch

from anime4kcpp_amd import _native, hip, synth
from oracle_lib import oracle_pass
from test_gpu_fp32 import check_f32, check_int

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(cuda):
    n = hip.Net("hdn2", 0)
    yield n
    n.close()


def sizes(seed, n, lo, hi):
    return list(zip(synth.randint(seed, n, lo, hi), synth.randint(seed + 1, n, lo, hi)))


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_batch_equals_alone(net, weights, cuda, prec, dtype):
    dims = [(1, 1), (3, 2), (65, 33)] + [(int(w), int(h)) for w, h in sizes(11, 9, 16, 300)]
    srcs = [synth.flat_shaded(50 + i, w, h, dtype=dtype) for i, (w, h) in enumerate(dims)]
    td = torch.uint8 if dtype == np.uint8 else torch.float32
    dsrc = [torch.from_numpy(s).to(cuda) for s in srcs]
    batch_out = net.upscale_batch(dsrc, precision=prec)
    torch.cuda.synchronize()
    for i, s in enumerate(dsrc):
        alone = net.upscale(s, precision=prec, config=0)
        torch.cuda.synchronize()
        assert torch.equal(alone, batch_out[i]), f"image {i} {dims[i]} differs"
    if prec == "fp32":
        k, b = weights["hdn2"]
        for i, s in enumerate(srcs):
            want = oracle_pass(s, k, b)
            got = batch_out[i].cpu().numpy()
            if dtype == np.uint8:
                check_int(got, want)
            else:
                check_f32(got, want)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
@pytest.mark.parametrize("config", [0, 1])
def test_batch_one_tile(net, cuda, prec, config):
    """A batch of one image of one tile: every workgroup of the persistent grid but one draws
    an index past the end at once, and the one that runs the tile draws its next index while
    running it; the output equals the image alone."""
    for w, h in [(1, 1), (7, 5)]:
        s = torch.from_numpy(synth.flat_shaded(90 + w, w, h, dtype=np.uint8)).to(cuda)
        out = net.upscale_batch([s], precision=prec, config=config)
        alone = net.upscale(s, precision=prec, config=config)
        torch.cuda.synchronize()
        assert torch.equal(out[0], alone), f"{prec} config {config}: {w}x{h} differs"


@pytest.mark.parametrize("config", [0, 1])
def test_batch_64_uneven_u8(net, weights, cuda, config):
    """64 u8 images with sides uniform in [128, 2048] (seeded), one launch; spot-checked
    against the oracle on a few images, all compared with their alone results."""
    dims = [(int(w), int(h)) for w, h in sizes(2024, 64, 128, 2048)]
    dsrc = [torch.from_numpy(synth.flat_shaded(700 + i, w, h, dtype=np.uint8)).to(cuda) for i, (w, h) in enumerate(dims)]
    out = net.upscale_batch(dsrc, precision="fp16", config=config)
    torch.cuda.synchronize()
    for i in range(0, 64, 9):
        alone = net.upscale(dsrc[i], precision="fp16", config=config)
        torch.cuda.synchronize()
        assert torch.equal(alone, out[i])
    out32 = net.upscale_batch(dsrc[:6], precision="fp32", config=config)
    torch.cuda.synchronize()
    k, b = weights["hdn2"]
    for i in (0, 5):
        check_int(out32[i].cpu().numpy(), oracle_pass(dsrc[i].cpu().numpy(), k, b))


def test_batch_rejects_bad_workspace(net, cuda):
    L = _native.lib()
    src = torch.zeros((8, 8), dtype=torch.uint8, device=cuda)
    dst = torch.zeros((16, 16), dtype=torch.uint8, device=cuda)
    im = _native.ACHipImage(src.data_ptr(), 8, dst.data_ptr(), 16, 8, 8)
    arr = (_native.ACHipImage * 1)(im)
    need = L.achip_batch_workspace_bytes(1)
    small = torch.zeros(need - 64, dtype=torch.uint8, device=cuda)
    assert L.achip_pass_batch(net.handle, 0, 0, 1, 1, arr, 1, ctypes.c_void_p(small.data_ptr()), need - 64, None) == -22
    host = torch.zeros(need, dtype=torch.uint8)
    assert L.achip_pass_batch(net.handle, 0, 0, 1, 1, arr, 1, ctypes.c_void_p(host.data_ptr()), need, None) == -22
    assert L.achip_pass_batch(net.handle, 0, 0, 1, 1, arr, 0, None, 0, None) == -22
    torch.cuda.synchronize()


def test_batch_rejects_overlaps(net, cuda):
    """A batch whose written spans (each dst, the workspace) overlap any other image's src or
    dst, or a src overlapping the workspace, is refused before anything is launched."""
    L = _native.lib()
    pool = torch.zeros(1 << 20, dtype=torch.uint8, device=cuda)
    base = pool.data_ptr()
    need = L.achip_batch_workspace_bytes(2)
    ws = torch.zeros(need, dtype=torch.uint8, device=cuda)

    def run(images, wsp=ws.data_ptr()):
        arr = (_native.ACHipImage * len(images))(*[_native.ACHipImage(s, 8, d, 16, 8, 8) for s, d in images])
        return L.achip_pass_batch(net.handle, 0, 0, 1, 1, arr, len(images), ctypes.c_void_p(wsp), need, None)

    a_src, a_dst, b_src, b_dst = base, base + 4096, base + 8192, base + 12288
    assert run([(a_src, a_dst), (b_src, b_dst)]) == 0
    assert run([(a_src, a_dst), (a_src, b_dst)]) == 0                 # shared sources are fine
    assert run([(a_src, a_dst), (b_src, a_dst + 64)]) == -22          # dst overlaps dst
    assert b"image 0 dst overlaps image 1 dst" in L.achip_last_error()
    assert run([(a_src, a_dst), (a_dst + 32, b_dst)]) == -22          # src inside another dst
    assert run([(a_src, a_dst), (b_src, a_src)]) == -22               # dst over another src
    wsbig = torch.zeros(need + 4096, dtype=torch.uint8, device=cuda)
    w0 = (wsbig.data_ptr() + 255) // 256 * 256
    assert run([(w0 + 256, a_dst), (b_src, b_dst)], wsp=w0) == -22   # src inside the workspace
    assert b"workspace" in L.achip_last_error()
    torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_autotune_choice_is_used(net, cuda, prec):
    """achip_autotune times every configuration, and achip_pass then runs the chosen one:
    its output equals that configuration's output bit for bit.  (The size class, 2^11 x 2^5,
    is one no other test uses: the choice is remembered for the whole process.)"""
    best, ms = net.autotune(1500, 24, precision=prec, src_dtype=torch.uint8)
    assert 0 <= best < len(ms) and all(t > 0 for t in ms)
    assert ms[best] == min(ms)
    src = torch.from_numpy(synth.flat_shaded(91, 1500, 24, dtype=np.uint8)).to(cuda)
    auto = net.upscale(src, precision=prec)
    expl = net.upscale(src, precision=prec, config=best)
    torch.cuda.synchronize()
    assert torch.equal(auto, expl)


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_batch_layers_equals_fused(net, weights, cuda, dtype):
    """achip_pass_batch_layers (the layer-split form, four launches over the 62 x 16 tiles of
    every image): each image's output is the fused kernel's alone, bit for bit, for tiny,
    tile-edge and uneven sizes, and within the oracle's tolerance."""
    dims = [(1, 1), (5, 3), (62, 16), (63, 17), (130, 71)] + [(int(w), int(h)) for w, h in sizes(21, 9, 16, 400)]
    srcs = [synth.flat_shaded(80 + i, w, h, dtype=dtype) for i, (w, h) in enumerate(dims)]
    dsrc = [torch.from_numpy(s).to(cuda) for s in srcs]
    out = net.upscale_batch_layers(dsrc)
    torch.cuda.synchronize()
    k, b = weights["hdn2"]
    for i, s in enumerate(dsrc):
        alone = net.upscale(s, config=0)
        torch.cuda.synchronize()
        assert torch.equal(alone, out[i]), f"image {i} {dims[i]} differs"
        want = oracle_pass(srcs[i], k, b)
        (check_int if dtype == np.uint8 else check_f32)(out[i].cpu().numpy(), want)


def test_batch_layers_rejects_bad_arguments(net, cuda):
    """A short workspace, a dst overlapping another image's src, and a non-ACNetLegacy net are
    refused before any launch (ACHIP_EINVAL)."""
    L = _native.lib()
    srcs = [torch.zeros((40, 50), dtype=torch.uint8, device=cuda) for _ in range(2)]
    dsts = [torch.zeros((80, 100), dtype=torch.uint8, device=cuda) for _ in range(2)]
    arr = (_native.ACHipImage * 2)()
    for i in range(2):
        arr[i] = _native.ACHipImage(srcs[i].data_ptr(), 50, dsts[i].data_ptr(), 100, 50, 40)
    need = L.achip_pass_batch_layers_bytes(arr, 2)
    ws = torch.empty(need, dtype=torch.uint8, device=cuda)
    assert L.achip_pass_batch_layers(net.handle, 1, 1, arr, 2, ctypes.c_void_p(ws.data_ptr()), need - 256, None) == -22
    big = torch.zeros((80, 100), dtype=torch.uint8, device=cuda)
    arr[1] = _native.ACHipImage(big.data_ptr(), 100, big.data_ptr() + 0, 100, 50, 40)   # its src is its own dst
    assert L.achip_pass_batch_layers(net.handle, 1, 1, arr, 2, ctypes.c_void_p(ws.data_ptr()), need, None) == -22
    arr[1] = _native.ACHipImage(srcs[1].data_ptr(), 50, srcs[0].data_ptr(), 100, 25, 20)   # dst over image 0's src
    assert L.achip_pass_batch_layers(net.handle, 1, 1, arr, 2, ctypes.c_void_p(ws.data_ptr()), need, None) == -22
    n8 = hip.Net8("b4", 0)
    arr[1] = _native.ACHipImage(srcs[1].data_ptr(), 50, dsts[1].data_ptr(), 100, 50, 40)
    assert L.achip_pass_batch_layers(n8.handle, 1, 1, arr, 2, ctypes.c_void_p(ws.data_ptr()), need, None) == -22
    n8.close()
    assert L.achip_pass_batch_layers(net.handle, 1, 1, arr, 2, ctypes.c_void_p(ws.data_ptr()), need, None) == 0
    torch.cuda.synchronize()
