"""The batched, precision and serving workloads at their own sizes.

- 64 u8 images with sides uniform in [128, 2048] (the benchmark's seeds) in one persistent
  FP32 launch per configuration: every image bit-identical to the same image run alone, and
  a sample within the u8 bar of the oracle (1 LSB on >= 99.9 %, max 2).
- 3840x2160 f32 luma (HDN0) to 15360x8640 by two chained 2x passes on the device, in FP32,
  FP16 and BF16.  Each pass is judged on its own input (the device's previous pass), on
  interior crops: FP32 against the oracle's FMA order within 1e-5; FP16 and BF16 against
  the FP64 oracle within the 16-bit bounds (FP16 max 4e-3 / mean 4e-4, BF16 3e-2 / 4e-3).
- Serving: 8 concurrent 1920x1080 YUV420 streams on the frame-graph path mixed with RGB
  one-offs up to 4096x4096 at factors 2 and 4: every output bit-identical to the same
  request served alone; frame luma within the u8 bar of the oracle on crops, frame chroma
  equal to the oracle's bilinear 2x; one-off RGB equal to Processor::process and, on
  crops, within the bar of the oracle's composition of Processor.cpp:199-276.

A crop's oracle runs on the crop grown by a margin the network cannot see across (nine
3x3 layers: 9 input pixels per pass), and only the crop's interior is compared.
"""
import numpy as np
import pytest
import torch

import oracle_lib as OL
from anime4kcpp_amd import hip, synth
from anime4kcpp_amd import processor as P
from test_gpu_fp32 import check_int
from test_gpu_mfma import BOUNDS, err_report
from test_gpu_pipeline import oracle_pipeline

pytestmark = pytest.mark.gpu

MARGIN = 10


def crops(w, h, n, size, seed, m=MARGIN):
    """n interior crop origins (x, y) of size x size, at least m from every edge."""
    assert w - size - m > m and h - size - m > m
    xs = synth.randint(seed, n, m, w - size - m)
    ys = synth.randint(seed + 1, n, m, h - size - m)
    return [(int(x), int(y)) for x, y in zip(xs, ys)]


def grown(a, x, y, size, m=MARGIN):
    return np.ascontiguousarray(a[y - m:y + size + m, x - m:x + size + m])


def inner(out2x, size, m=MARGIN):
    return out2x[2 * m:2 * m + 2 * size, 2 * m:2 * m + 2 * size]


def test_batch_64_uneven_fp32_all_images(cuda, weights):
    net = hip.Net("hdn0", 0)
    ws = synth.randint(2024, 64, 128, 2048)
    hs = synth.randint(2025, 64, 128, 2048)
    dims = [(int(w), int(h)) for w, h in zip(ws, hs)]
    srcs = [synth.flat_shaded(700 + i, w, h, dtype=np.uint8) for i, (w, h) in enumerate(dims)]
    dsrc = [torch.from_numpy(s).to(cuda) for s in srcs]
    k, b = weights["hdn0"]
    for config in range(len(hip.configs("fp32"))):
        out = net.upscale_batch(dsrc, precision="fp32", config=config)
        torch.cuda.synchronize()
        for i, s in enumerate(dsrc):
            alone = net.upscale(s, precision="fp32", config=config)
            torch.cuda.synchronize()
            assert torch.equal(alone, out[i]), f"config {config}: image {i} {dims[i]} differs from alone"
        if config == 0:
            for i in (0, 13, 27, 41, 63):
                check_int(out[i].cpu().numpy(), OL.oracle_pass(srcs[i], k, b))
        del out
    net.close()


@pytest.fixture(scope="module")
def chain_4k(cuda):
    """The 4K f32 input and, per precision, the device results of the two chained passes."""
    net = hip.Net("hdn0", 0)
    src = synth.flat_shaded(44, 3840, 2160)
    d0 = torch.from_numpy(src).to(cuda)
    out = {}
    for prec in ("fp32", "fp16", "bf16"):
        p1 = net.upscale(d0, precision=prec)
        p2 = net.upscale(p1, precision=prec)
        torch.cuda.synchronize()
        assert tuple(p2.shape) == (8640, 15360)
        out[prec] = (p1, p2)
    yield src, out
    net.close()


def pass_crops(src_np, p1, p2, n, size, seed):
    """(pass, input crop grown by MARGIN, device output of that crop's interior) for n crops of
    each pass; pass 2's input is the device's own pass-1 output."""
    res = []
    h, w = src_np.shape
    for x, y in crops(w, h, n, size, seed):
        res.append((1, grown(src_np, x, y, size), p1[2 * y:2 * y + 2 * size, 2 * x:2 * x + 2 * size].cpu().numpy()))
    h2, w2 = p1.shape
    for x, y in crops(w2, h2, n, size, seed + 7):
        g = p1[y - MARGIN:y + size + MARGIN, x - MARGIN:x + size + MARGIN].cpu().numpy()
        res.append((2, np.ascontiguousarray(g), p2[2 * y:2 * y + 2 * size, 2 * x:2 * x + 2 * size].cpu().numpy()))
    return res


def test_4k_chain_fp32(chain_4k, weights):
    src, out = chain_4k
    k, b = weights["hdn0"]
    p1, p2 = out["fp32"]
    for pas, g, got in pass_crops(src, p1, p2, 3, 200, 51):
        want = inner(OL.oracle_pass(g, k, b), 200)
        err = np.abs(got.astype(np.float64) - want).max()
        print(f"fp32 pass {pas}: max abs {err:.3g}")
        assert err <= 1e-5, (pas, err)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_4k_chain_16bit(chain_4k, weights, prec):
    src, out = chain_4k
    k, b = weights["hdn0"]
    p1, p2 = out[prec]
    for pas, g, got in pass_crops(src, p1, p2, 3, 200, 61):
        want = inner(OL.oracle_pass_f64(g, k, b), 200)
        mx, mean, psnr = err_report(got, want)
        print(f"{prec} pass {pas}: max {mx:.3g} mean {mean:.3g} PSNR {psnr:.1f} dB")
        assert mx <= BOUNDS[prec][0] and mean <= BOUNDS[prec][1], (pas, mx, mean)


def test_4k_chain_matches_processor_fp16(chain_4k):
    """The drop-in processor ("hip-fp16", factor 4 on a gray f32 image) chains the same two
    passes: its output equals the device chain bit for bit."""
    src, out = chain_4k
    proc = P.Processor("hip-fp16", 0, "acnet-legacy-hdn0")
    assert proc.ok(), proc.error()
    rc, got = proc.process(src, 4.0)
    assert rc == 0 and got.shape == (8640, 15360)
    assert np.array_equal(got, out["fp16"][1].cpu().numpy())


# ---- serving: 8 concurrent 1080p YUV420 streams + large RGB one-offs -----------------------

def yuv420(seed, w=1920, h=1080):
    return [synth.flat_shaded(seed, w, h, dtype=np.uint8), synth.flat_shaded(seed + 1, w // 2, h // 2, dtype=np.uint8),
            synth.flat_shaded(seed + 2, w // 2, h // 2, dtype=np.uint8)]


def rgb(seed, w, h):
    return np.ascontiguousarray(np.stack([synth.flat_shaded(seed + c, w, h, dtype=np.uint8) for c in range(3)], axis=2))


def out_planes(kind, planes, factor):
    h, w = planes[0].shape[:2]
    if kind == hip.REQ_RGB:
        return [np.empty((factor * h, factor * w, 3), np.uint8)]
    return [np.empty((factor * h, factor * w), np.uint8), np.empty((factor * h // 2, factor * w // 2), np.uint8),
            np.empty((factor * h // 2, factor * w // 2), np.uint8)]


def serve(server, reqs, together):
    outs = [out_planes(k, p, f) for k, p, f, _ in reqs]
    if together:
        tickets = [server.submit(k, p, f, s, outs[i]) for i, (k, p, f, s) in enumerate(reqs)]
        for t in tickets:
            server.wait(t, 120000)
    else:
        for i, (k, p, f, s) in enumerate(reqs):
            server.wait(server.submit(k, p, f, s, outs[i]), 120000)
    return outs


@pytest.mark.parametrize("prec", ["fp16", "fp32"])
def test_serving_8_streams_and_large_oneoffs(cuda, weights, prec):
    variant = "hdn1"
    net = hip.Net(variant, 0)
    frames = [(hip.REQ_YUV420, yuv420(4000 + 10 * s + i), 2, s) for i in range(2) for s in range(8)]
    oneoffs = [(hip.REQ_RGB, [rgb(4500, 4096, 4096)], 4, -1), (hip.REQ_RGB, [rgb(4510, 3001, 2003)], 2, -1),
               (hip.REQ_RGB, [rgb(4520, 128, 128)], 4, -1), (hip.REQ_RGB, [rgb(4530, 1234, 777)], 4, -1)]
    reqs, frame_at, oneoff_at = [], [], []
    for i, f in enumerate(frames):       # one-offs interleaved with the streams' frames
        frame_at.append(len(reqs))
        reqs.append(f)
        if i % 4 == 3:
            oneoff_at.append(len(reqs))
            reqs.append(oneoffs[i // 4])
    with hip.Server(net, prec) as srv:
        alone = serve(srv, reqs, together=False)
    with hip.Server(net, prec) as srv:
        mixed = serve(srv, reqs, together=True)
        st = srv.stats()
    assert st["failed"] == 0 and st["completed_frames"] == 16 and st["completed_oneoffs"] == 4
    assert st["graph_replays"] == 16
    for i, (a, m) in enumerate(zip(alone, mixed)):
        for pa, pm in zip(a, m):
            assert np.array_equal(pa, pm), f"request {i} differs between alone and mixed"
    del alone
    k, b = weights[variant]
    # frames: luma per crop against the oracle pass (FP32 at the u8 bar; FP16 >= 99 % within
    # 1 LSB), chroma equal to the oracle's bilinear 2x
    for idx in (0, 7, 15):
        kind, planes, _, _ = frames[idx]
        got = mixed[frame_at[idx]]
        y = planes[0]
        for x0, y0 in crops(1920, 1080, 2, 160, 70 + idx):
            want = inner(OL.oracle_pass(grown(y, x0, y0, 160), k, b), 160)
            g = got[0][2 * y0:2 * y0 + 320, 2 * x0:2 * x0 + 320]
            if prec == "fp32":
                check_int(g, want)
            else:
                assert (np.abs(g.astype(int) - want.astype(int)) <= 1).mean() >= BOUNDS["fp16"][2]
        for src, out in ((planes[1], got[1]), (planes[2], got[2])):
            want = OL.resize(src, 2 * src.shape[1], 2 * src.shape[0], OL.RESIZE_BILINEAR)
            assert np.array_equal(out, want)
    # one-offs: equal to the drop-in processor; in FP32 within the u8 bar of the oracle's
    # composition on crops (factor 2: one pass; factor 4: two passes, each output quantised)
    proc = P.Processor("hip" if prec == "fp32" else "hip-fp16", 0, f"acnet-legacy-{variant}")
    assert proc.ok(), proc.error()
    for (kind, planes, factor, _), at in zip(oneoffs, oneoff_at):
        got = mixed[at][0]
        rc, want = proc.process(planes[0], factor)
        assert rc == 0 and np.array_equal(got, want)
        del want
        if prec != "fp32":
            continue
        h, w = planes[0].shape[:2]
        m = 16      # two passes (9 + 5 input pixels) and the chroma filter's support
        size = min(96, w - 2 * m - 1, h - 2 * m - 1)
        for x0, y0 in crops(w, h, 2, size, 90 + w, m):
            g = np.ascontiguousarray(planes[0][y0 - m:y0 + size + m, x0 - m:x0 + size + m])
            o = oracle_pipeline(g, factor, k, b)[factor * m:factor * (m + size), factor * m:factor * (m + size)]
            gg = got[factor * y0:factor * (y0 + size), factor * x0:factor * (x0 + size)]
            d = np.abs(gg.astype(int) - o.astype(int))
            print(f"one-off {w}x{h} x{factor} crop ({x0},{y0}): {(d <= 1).mean() * 100:.4f}% within 1, max {d.max()}")
            check_int(gg, o)
    net.close()
