"""Parity of the fused FP32 HIP kernel with the oracle (the reference CPU processor's
FMA-order 2x pass, pinned bit-exactly to the reference's own kernels on the CPU).

Tolerances (stated by the target): f32 output max abs error <= 1e-5; u8 output
within 1 LSB on >= 99.9 % of pixels and never more than 2 LSB.  Every case calls
the C ABI (include/achip.h) through anime4kcpp_amd.hip.
"""
import numpy as np
import pytest
import torch

from anime4kcpp_amd import hip, synth
from oracle_lib import oracle_pass, oracle_pass_f64, to_float

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5
U8_FRAC_1LSB = 0.999
U8_MAX_LSB = 2

_TORCH = {np.uint8: torch.uint8, np.uint16: torch.uint16, np.float16: torch.float16, np.float32: torch.float32}
_VIEW = {np.uint16: np.int16}


def to_dev(a, dev):
    a = np.ascontiguousarray(a)
    if a.dtype.type in _VIEW:
        return torch.from_numpy(a.view(_VIEW[a.dtype.type])).to(dev).view(_TORCH[a.dtype.type])
    return torch.from_numpy(a).to(dev)


def to_host(t):
    if t.dtype == torch.uint16:
        return t.view(torch.int16).cpu().numpy().view(np.uint16)
    return t.cpu().numpy()


def gpu_pass(net, src, out_dtype, config=None, precision="fp32"):
    dev = torch.device("cuda", net.device)
    s = to_dev(src, dev)
    d = net.upscale(s, out_dtype=_TORCH[np.dtype(out_dtype).type], config=config, precision=precision)
    torch.cuda.synchronize()
    return to_host(d)


def check_f32(got, want, tol=F32_TOL):
    err = np.abs(got.astype(np.float64) - want.astype(np.float64)).max()
    assert err <= tol, f"max abs error {err:.3g} > {tol}"
    return err


def check_int(got, want, frac=U8_FRAC_1LSB, maxlsb=U8_MAX_LSB):
    d = np.abs(got.astype(np.int64) - want.astype(np.int64))
    ok = (d <= 1).mean()
    assert ok >= frac and d.max() <= maxlsb, f"{ok * 100:.4f}% within 1 LSB, max {d.max()} LSB"
    return ok, int(d.max())


@pytest.fixture(scope="module")
def nets(cuda, weights):
    out = {v: hip.Net(v, cuda.index or 0) for v in hip.VARIANTS}
    yield out
    for n in out.values():
        n.close()


def test_device_present(cuda):
    assert hip.device_count() >= 1


@pytest.mark.parametrize("variant", hip.VARIANTS)
@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (3, 2), (17, 13), (64, 64), (65, 33), (100, 7)])
def test_small_f32(nets, weights, variant, w, h):
    src = synth.flat_shaded(1000 + w * 7 + h, w, h)
    k, b = weights[variant]
    check_f32(gpu_pass(nets[variant], src, np.float32), oracle_pass(src, k, b))


@pytest.mark.parametrize("variant", hip.VARIANTS)
@pytest.mark.parametrize("w,h", [(1, 1), (5, 4), (31, 47), (64, 64), (129, 70)])
def test_small_u8(nets, weights, variant, w, h):
    src = synth.flat_shaded(2000 + w + h, w, h, dtype=np.uint8)
    k, b = weights[variant]
    check_int(gpu_pass(nets[variant], src, np.uint8), oracle_pass(src, k, b))


def test_config1_256_f32_hdn0(nets, weights):
    """256x256 f32 luma, HDN0 (the reference's CPU-runnable case), quantised to u8 and
    compared bit-tolerantly, and f32 against 1e-5."""
    src = synth.flat_shaded(1, 256, 256)
    k, b = weights["hdn0"]
    want = oracle_pass(src, k, b)
    got = gpu_pass(nets["hdn0"], src, np.float32)
    check_f32(got, want)
    q = lambda a: (np.clip(a, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    check_int(q(got), q(want))
    check_int(gpu_pass(nets["hdn0"], src, np.uint8), oracle_pass(src, k, b, out_dtype=np.uint8))


@pytest.mark.parametrize("config", range(len(hip.configs("fp32"))) if True else [])
@pytest.mark.parametrize("w,h", [(1, 1), (33, 31), (97, 66), (200, 130)])
def test_every_config(nets, weights, config, w, h):
    src = synth.flat_shaded(3000 + config, w, h)
    k, b = weights["hdn1"]
    check_f32(gpu_pass(nets["hdn1"], src, np.float32, config=config), oracle_pass(src, k, b))


@pytest.mark.parametrize("sdt,ddt", [(np.uint8, np.float32), (np.float32, np.uint8), (np.uint16, np.uint16),
                                     (np.float16, np.float16), (np.uint8, np.uint16), (np.float32, np.float16)])
def test_element_types(nets, weights, sdt, ddt):
    f = synth.flat_shaded(77, 70, 45)
    if sdt == np.uint8:
        src = (f * 255 + 0.5).astype(np.uint8)
    elif sdt == np.uint16:
        src = (f * 65535 + 0.5).astype(np.uint16)
    else:
        src = f.astype(sdt)
    k, b = weights["gan"]
    if sdt == np.float16:   # the oracle takes f16 through its own conversion
        want = oracle_pass(src, k, b, out_dtype=ddt)
    else:
        want = oracle_pass(src, k, b, out_dtype=ddt)
    got = gpu_pass(nets["gan"], src, ddt)
    if ddt == np.float32:
        check_f32(got, want)
    elif ddt == np.float16:
        check_f32(got.astype(np.float32), want.astype(np.float32), tol=1e-3)   # one f16 ulp at 1.0
    else:
        check_int(got, want)


def test_strided_images(nets, weights, cuda):
    """Row strides larger than the row (a crop of a bigger buffer, as ac_image_map/view give)."""
    big = synth.flat_shaded(5, 300, 90, dtype=np.uint8)
    src_view = big[10:80, 25:225]
    k, b = weights["hdn2"]
    want = oracle_pass(np.ascontiguousarray(src_view), k, b)
    sdev = to_dev(big, cuda)[10:80, 25:225]
    out_big = torch.zeros((200, 500), dtype=torch.uint8, device=cuda)
    dview = out_big[30:170, 50:450]
    nets["hdn2"].upscale(sdev, dview)
    torch.cuda.synchronize()
    check_int(dview.cpu().numpy(), want)
    # nothing written outside the destination view
    ob = out_big.cpu().numpy()
    assert ob[:30].max() == 0 and ob[170:].max() == 0 and ob[:, :50].max() == 0 and ob[:, 450:].max() == 0


@pytest.mark.parametrize("variant", hip.VARIANTS)
def test_config2_1080p_u8(nets, weights, variant):
    """1920x1080 u8 -> 3840x2160 for every weight set."""
    src = synth.flat_shaded(2, 1920, 1080, dtype=np.uint8)
    k, b = weights[variant]
    ok, mx = check_int(gpu_pass(nets[variant], src, np.uint8), oracle_pass(src, k, b))
    print(f"{variant}: {ok * 100:.4f}% within 1 LSB, max {mx}")


def test_fp64_error_1080p_f32(nets, weights):
    """Error of the FP32 kernel against the FP64 restatement (reported, bounded)."""
    src = synth.flat_shaded(3, 480, 270)
    k, b = weights["hdn0"]
    got = gpu_pass(nets["hdn0"], src, np.float32)
    want = oracle_pass_f64(to_float(src), k, b)
    err = np.abs(got - want)
    print(f"fp32 vs fp64: max {err.max():.3g} mean {err.mean():.3g}")
    assert err.max() < 1e-5


def test_invalid_arguments_launch_nothing(nets, cuda):
    net = nets["hdn0"]
    src = torch.zeros((16, 16), dtype=torch.uint8, device=cuda)
    small = torch.zeros((31, 32), dtype=torch.uint8, device=cuda)   # one row short
    with pytest.raises(Exception):
        net.upscale(src, small.view(-1)[: 31 * 32].view(31, 32))
    host = torch.zeros((16, 16), dtype=torch.uint8)
    dst = torch.zeros((32, 32), dtype=torch.uint8, device=cuda)
    import ctypes
    from anime4kcpp_amd import _native
    L = _native.lib()
    rc = L.achip_pass(net.handle, 0, ctypes.c_void_p(host.data_ptr()), 16, 1, 16, 16,
                      ctypes.c_void_p(dst.data_ptr()), 32, 1, None)
    assert rc == -22 and "not" in _native.last_error()
    # a destination one row too short for 2H rows, in an allocation of exactly that size
    hiprt = ctypes.CDLL("libamdhip64.so")
    buf = ctypes.c_void_p()
    assert hiprt.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(32 * 31)) == 0
    try:
        rc = L.achip_pass(net.handle, 0, ctypes.c_void_p(src.data_ptr()), 16, 1, 16, 16, buf, 32, 1, None)
        assert rc == -22, _native.last_error()
        assert "allocation" in _native.last_error()
    finally:
        hiprt.hipFree(buf)
    torch.cuda.synchronize()


@pytest.mark.parametrize("variant", ["gan", "hdn0", "hdn3"])
@pytest.mark.parametrize("w,h", [(1, 1), (5, 3), (63, 17), (130, 71), (62, 16), (64, 18), (125, 33)])
def test_layer_split_form(nets, weights, variant, w, h):
    """achip_pass_layers (layers 0-2 in one launch, then layers 3..8 two per launch through
    f32 maps in a workspace): the fused MFMA kernel's bits in both of its configurations (the
    same instructions on the same operands), and so f32 within 1e-5 and u8 within the u8 bar
    of the oracle."""
    import ctypes
    from anime4kcpp_amd import _native
    L = _native.lib()
    net = nets[variant]
    k, b = weights[variant]
    dev = torch.device("cuda", net.device)
    ws = torch.empty(L.achip_pass8_workspace_bytes(w, h) + 256, dtype=torch.uint8, device=dev)
    wp = (ws.data_ptr() + 255) // 256 * 256
    for dt in (np.float32, np.uint8):
        src = synth.flat_shaded(60 + w, w, h, dtype=dt)
        s = to_dev(src, dev)
        d = torch.empty((2 * h, 2 * w), dtype=s.dtype, device=dev)
        es = s.element_size()
        rc = L.achip_pass_layers(net.handle, ctypes.c_void_p(s.data_ptr()), w * es, hip.elem_type(s.dtype), w, h,
                                 ctypes.c_void_p(d.data_ptr()), 2 * w * es, hip.elem_type(d.dtype), ctypes.c_void_p(wp),
                                 L.achip_pass8_workspace_bytes(w, h), None)
        assert rc == 0, _native.last_error()
        torch.cuda.synchronize()
        got = to_host(d)
        for config in (0, 1):
            assert np.array_equal(got, gpu_pass(net, src, dt, config=config)), (dt, config)
        want = oracle_pass(src, k, b)
        if dt == np.float32:
            check_f32(got, want)
        else:
            check_int(got, want)


def test_layer_split_1080p_u8(nets, weights):
    import ctypes
    from anime4kcpp_amd import _native
    L = _native.lib()
    net = nets["hdn1"]
    k, b = weights["hdn1"]
    src = synth.flat_shaded(61, 1920, 1080, dtype=np.uint8)
    s = torch.from_numpy(src).cuda()
    d = torch.empty((2160, 3840), dtype=torch.uint8, device="cuda")
    ws = torch.empty(L.achip_pass8_workspace_bytes(1920, 1080), dtype=torch.uint8, device="cuda")
    assert L.achip_pass_layers(net.handle, ctypes.c_void_p(s.data_ptr()), 1920, 0x001, 1920, 1080, ctypes.c_void_p(d.data_ptr()),
                               3840, 0x001, ctypes.c_void_p(ws.data_ptr()), ws.numel(), None) == 0
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    assert np.array_equal(got, gpu_pass(net, src, np.uint8, config=0))
    check_int(got, oracle_pass(src, k, b))
    # the workspace may not overlap src or dst, and an ACNet<8> net is refused
    assert L.achip_pass_layers(net.handle, ctypes.c_void_p(s.data_ptr()), 1920, 0x001, 1920, 1080, ctypes.c_void_p(d.data_ptr()),
                               3840, 0x001, ctypes.c_void_p(d.data_ptr()), ws.numel(), None) == -22
    n8 = hip.Net8("b4", 0)
    assert L.achip_pass_layers(n8.handle, ctypes.c_void_p(s.data_ptr()), 1920, 0x001, 1920, 1080, ctypes.c_void_p(d.data_ptr()),
                               3840, 0x001, ctypes.c_void_p(ws.data_ptr()), ws.numel(), None) == -22
    n8.close()


@pytest.mark.parametrize("w,h", [(1, 1), (130, 71), (256, 256), (960, 540), (1280, 720)])
def test_pass_ws_picks_a_form_with_the_fused_bits(nets, weights, w, h):
    """achip_pass_ws runs the layer-split form where the size rule prefers it (a workspace of
    achip_pass_ws_bytes) and the fused kernel elsewhere, or when the workspace is short: the
    output is the fused kernel's, bit for bit, and within the u8 bar of the oracle."""
    net = nets["hdn2"]
    src = synth.flat_shaded(300 + w, w, h, dtype=np.uint8)
    s = torch.from_numpy(src).cuda()
    need = net.workspace_bytes(w, h)
    tiles = ((w + 63) // 64) * ((h + 31) // 32)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert (need > 0) == (not (tiles <= cus and w * h > 256 * 256)), (need, tiles, cus)
    fused = net.upscale(s, config=0)
    for ws_bytes in (need, need // 2, 0):
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
        got = net.upscale(s, workspace=ws if ws_bytes else None)
        torch.cuda.synchronize()
        assert torch.equal(got, fused), (w, h, ws_bytes)
    assert net.workspace_bytes(w, h, precision="fp16") == 0
    check_int(fused.cpu().numpy(), oracle_pass(src, *weights["hdn2"]))
