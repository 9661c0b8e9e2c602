"""achip_pass_batch_noalloc (include/achip.h): the batched pass whose tile queue and
descriptors live in the library's fixed device array instead of a caller workspace.  Through
the C ABI its output must equal achip_pass_batch's bit for bit, on two calls in a row (the
second with fewer, smaller images, so the array still holds the first call's descriptors and
counter), in FP32 and FP16; sizes past the array are refused."""
import numpy as np
import pytest
import torch

from anime4kcpp_amd import _native, hip, synth

pytestmark = pytest.mark.gpu

U8 = 0x001


def lib():
    return _native.lib()


def image_array(srcs, dsts):
    arr = (_native.ACHipImage * len(srcs))()
    for i, (s, d) in enumerate(zip(srcs, dsts)):
        arr[i] = _native.ACHipImage(s.data_ptr(), s.stride(0), d.data_ptr(), d.stride(0), s.shape[1], s.shape[0])
    return arr


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_noalloc_equals_batch_twice(cuda, prec):
    L = lib()
    net = hip.Net("hdn3", 0)
    p = hip.PRECISIONS[prec]
    stream = torch.cuda.current_stream().cuda_stream
    for call, (n, lo, hi, seed) in enumerate([(40, 16, 700, 31), (6, 1, 200, 47)]):
        ws = synth.randint(seed, n, lo, hi)
        hs = synth.randint(seed + 1, n, lo, hi)
        print(f"{prec} call {call}: {n} images, sides in [{lo}, {hi}]", flush=True)
        srcs = [torch.from_numpy(synth.flat_shaded(seed + i, int(w), int(h), dtype=np.uint8)).to(cuda)
                for i, (w, h) in enumerate(zip(ws, hs))]
        want = net.upscale_batch(srcs, precision=prec, config=0)
        got = [torch.full((2 * s.shape[0], 2 * s.shape[1]), 7, dtype=torch.uint8, device=cuda) for s in srcs]
        rc = L.achip_pass_batch_noalloc(net.handle, p, 0, U8, U8, image_array(srcs, got), n, stream)
        assert rc == 0, _native.last_error()
        torch.cuda.synchronize()
        for i in range(n):
            assert torch.equal(got[i], want[i]), f"call {call}: image {i} ({srcs[i].shape}) differs"
    net.close()


def test_noalloc_limits(cuda):
    L = lib()
    net = hip.Net("hdn0", 0)
    assert L.achip_noalloc_scratch_bytes() == L.achip_batch_workspace_bytes(256)
    src = torch.zeros((8, 8), dtype=torch.uint8, device=cuda)
    dsts = [torch.zeros((16, 16), dtype=torch.uint8, device=cuda) for _ in range(257)]
    arr = image_array([src] * 257, dsts)
    assert L.achip_pass_batch_noalloc(net.handle, 0, 0, U8, U8, arr, 257, None) == -22
    assert "ACHIP_NOALLOC_MAX_IMAGES" in _native.last_error()
    assert L.achip_pass_batch_noalloc(None, 0, 0, U8, U8, arr, 1, None) == -22
    assert L.achip_pass_batch_noalloc(net.handle, 0, 0, U8, U8, arr, 0, None) == -22
    # 256 images is the limit itself: accepted, and equal to the workspace form
    rc = L.achip_pass_batch_noalloc(net.handle, 0, 0, U8, U8, image_array([src] * 256, dsts[:256]), 256, None)
    assert rc == 0, _native.last_error()
    torch.cuda.synchronize()
    want = net.upscale(src)
    torch.cuda.synchronize()
    for d in dsts[:256]:
        assert torch.equal(d, want)
    net.close()
