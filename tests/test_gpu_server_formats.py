"""Serving loop with the reference video module's frame formats (ACHIP_REQ_YUV and
ACHipFrameFormat; video/src/Pipeline.cpp:495-566, cli/src/Main.cpp:183-205): NV12 / P010
(interleaved chroma), planar 4:2:0 / 4:2:2 / 4:4:4, 10-bit LSB-aligned words shifted up before
the passes and back after, gray, bilinear or Catmull-Rom chroma -- on the hipGraph frame path
and the batched one-off path.  Luma is checked against the oracle pass of the shifted input
(u8/u16 tolerance), chroma is bit-exact against the oracle's resize, and every request
gives the same bits alone and in a mix of u8 and u16 work."""
import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import hip, synth
from test_gpu_fp32 import check_int

pytestmark = pytest.mark.gpu

VARIANT = "hdn2"
FORMATS = ["nv12", "p010", "yuv420p10le", "yuv444p", "nv16", "yuv422p", "gray10le", "p210"]


@pytest.fixture(scope="module")
def net(cuda):
    n = hip.Net(VARIANT, 0)
    yield n
    n.close()


def plane(seed, w, h, ch, fmt):
    """Synthetic content in the format's sample range: 8-bit, 16-bit words, or 10/12-bit in the
    low bits of a word."""
    chans = [synth.flat_shaded(seed + c, w, h) for c in range(ch)]
    v = np.stack(chans, axis=2) if ch > 1 else chans[0]
    if fmt.element_type == 0x001:
        return np.ascontiguousarray(np.round(v * 255).astype(np.uint8))
    top = (1 << (16 - fmt.shift)) - 1
    return np.ascontiguousarray(np.round(v * top).astype(np.uint16))


def frame(seed, w, h, fmt):
    planes = [plane(seed, w, h, 1, fmt)]
    if fmt.planes > 1:
        cw, chh = w >> fmt.chroma_shift_x, h >> fmt.chroma_shift_y
        cc = 2 if fmt.planes == 2 else 1
        planes += [plane(seed + 10 * (i + 1), cw, chh, cc, fmt) for i in range(fmt.planes - 1)]
    return planes


def outputs(planes, factor):
    return [np.zeros((factor * p.shape[0], factor * p.shape[1]) + p.shape[2:], p.dtype) for p in planes]


def expected_luma(y, fmt, weights, factor):
    k, b = weights[VARIANT]
    dt = y.dtype
    x = (y.astype(np.uint32) << fmt.shift).astype(dt) if fmt.shift else y
    for _ in range(factor.bit_length() - 1):
        x = OL.oracle_pass(x, k, b)
    return (x >> fmt.shift).astype(dt) if fmt.shift else x


def check_frame(got, planes, fmt, weights, factor):
    check_int(got[0], expected_luma(planes[0], fmt, weights, factor))
    mode = OL.RESIZE_CATMULL_ROM if fmt.chroma_filter == hip.CHROMA_CATMULL_ROM else OL.RESIZE_BILINEAR
    for src, out in zip(planes[1:], got[1:]):
        want = OL.resize(src, out.shape[1], out.shape[0], mode)
        assert np.array_equal(out, want), f"{np.count_nonzero(out != want)} of {out.size} chroma samples differ"


@pytest.mark.parametrize("name", FORMATS)
@pytest.mark.parametrize("cf", [hip.CHROMA_BILINEAR, hip.CHROMA_CATMULL_ROM], ids=["bilinear", "catmullrom"])
def test_frames_and_oneoffs_match_oracle(net, weights, name, cf):
    fmt = hip.frame_format(name, chroma_filter=cf)
    with hip.Server(net, "fp32", step_tiles=64) as srv:
        f0 = frame(300, 160, 96, fmt)
        got = outputs(f0, 2)
        srv.wait(srv.submit(hip.REQ_YUV, f0, 2, 3, got, fmt=fmt), 60000)          # frame (graph)
        check_frame(got, f0, fmt, weights, 2)
        f1 = frame(400, 96, 64, fmt)
        got4 = outputs(f1, 4)
        srv.wait(srv.submit(hip.REQ_YUV, f1, 4, -1, got4, fmt=fmt), 60000)        # one-off, 2 passes
        check_frame(got4, f1, fmt, weights, 4)
        assert srv.stats()["failed"] == 0


def test_mix_of_formats_equals_alone(net):
    reqs = []
    for i, name in enumerate(FORMATS):
        fmt = hip.frame_format(name, chroma_filter=i % 2)
        reqs.append((frame(10 * i, 128 + 32 * (i % 3), 72 + 16 * (i % 2), fmt), 2 if i % 3 else 4, -1, fmt))
        reqs.append((frame(10 * i + 5, 96, 48, fmt), 2, i, fmt))                  # a stream per format
    reqs.append(([np.ascontiguousarray(np.stack([synth.flat_shaded(90 + c, 90, 70, dtype=np.uint8) for c in range(3)], 2))],
                 2, -1, None))                                                       # RGB among them

    def run(srv, together):
        outs = [outputs(p, f) for p, f, _, _ in reqs]
        kinds = [hip.REQ_RGB if fm is None else hip.REQ_YUV for _, _, _, fm in reqs]
        if together:
            ts = [srv.submit(kinds[i], p, f, s, outs[i], fmt=fm) for i, (p, f, s, fm) in enumerate(reqs)]
            for t in ts:
                srv.wait(t, 60000)
        else:
            for i, (p, f, s, fm) in enumerate(reqs):
                srv.wait(srv.submit(kinds[i], p, f, s, outs[i], fmt=fm), 60000)
        return outs

    with hip.Server(net, "fp16") as srv:
        alone = run(srv, False)
    with hip.Server(net, "fp16", step_tiles=16) as srv:
        mixed = run(srv, True)
        st = srv.stats()
    assert st["failed"] == 0 and st["graph_replays"] == len(FORMATS)
    for i, (a, m) in enumerate(zip(alone, mixed)):
        for pa, pm in zip(a, m):
            assert np.array_equal(pa, pm), f"request {i} differs between alone and mixed"


def test_format_validation(net):
    y = np.zeros((48, 64), np.uint8)
    with hip.Server(net, "fp16", hbm_budget=256 << 20) as srv:
        bad = [hip.FrameFormat(4), hip.FrameFormat(1, element_type=0x204), hip.FrameFormat(1, shift=2),
               hip.FrameFormat(1, element_type=0x002, shift=16), hip.FrameFormat(3, chroma_shift_x=2),
               hip.FrameFormat(3, chroma_filter=7)]
        for fmt in bad:
            planes = [y.astype(fmt.dtype)] * max(1, min(fmt.planes, 3))
            with pytest.raises(hip._native.AchipError):
                srv.submit(hip.REQ_YUV, planes, 2, fmt=fmt)
        nv12 = hip.frame_format("nv12")
        odd = frame(1, 63, 48, hip.frame_format("gray8"))
        with pytest.raises(hip._native.AchipError):                     # 4:2:0 needs even sizes
            srv.submit(hip.REQ_YUV, [odd[0], np.zeros((24, 31, 2), np.uint8)], 2, fmt=nv12)
        f = frame(2, 64, 48, nv12)
        srv.wait(srv.submit(hip.REQ_YUV, f, 2, 7, outputs(f, 2), fmt=nv12), 60000)
        p010 = hip.frame_format("p010")
        g = frame(3, 64, 48, p010)
        with pytest.raises(hip._native.AchipError):                     # a stream keeps its format
            srv.submit(hip.REQ_YUV, g, 2, 7, fmt=p010)
        short = [f[0], f[1][:, :20]]                                    # chroma rows too short
        with pytest.raises(hip._native.AchipError):
            srv.submit(hip.REQ_YUV, [np.ascontiguousarray(p) for p in short], 2, fmt=nv12)
