"""Serving loop (achip_server_*): continuous batching of mixed requests, HBM-budget
admission and hipGraph-replayed video frames.  Each request's output in a mixed load must
equal, bit for bit, the same request processed alone -- and the drop-in processor's output
(Processor::process, core/src/processor/Processor.cpp:199-276) for the one-offs -- and stay
within the oracle tolerance."""
import numpy as np
import pytest
import torch

import oracle_lib as OL
from anime4kcpp_amd import hip, synth
from anime4kcpp_amd import processor as P
from test_gpu_fp32 import check_int
from test_gpu_pipeline import oracle_pipeline

pytestmark = pytest.mark.gpu

VARIANT = "hdn1"


@pytest.fixture(scope="module")
def net(cuda):
    n = hip.Net(VARIANT, 0)
    yield n
    n.close()


def rgb(seed, w, h):
    return np.ascontiguousarray(np.stack([synth.flat_shaded(seed + c, w, h, dtype=np.uint8) for c in range(3)], axis=2))


def yuv(seed, w, h):
    return [synth.flat_shaded(seed, w, h, dtype=np.uint8), synth.flat_shaded(seed + 1, w // 2, h // 2, dtype=np.uint8),
            synth.flat_shaded(seed + 2, w // 2, h // 2, dtype=np.uint8)]


def make_requests():
    """(kind, planes, factor, stream): one-offs of every kind and factor, interleaved with the
    frames of two video streams."""
    reqs = [
        (hip.REQ_RGB, [rgb(10, 200, 150)], 2, -1),
        (hip.REQ_LUMA, [synth.flat_shaded(20, 333, 211, dtype=np.uint8)], 4, -1),
        (hip.REQ_YUV420, yuv(30, 256, 144), 4, -1),
        (hip.REQ_RGB, [rgb(40, 97, 61)], 4, -1),
        (hip.REQ_LUMA, [synth.flat_shaded(50, 1, 1, dtype=np.uint8)], 2, -1),
        (hip.REQ_RGB, [rgb(60, 700, 390)], 2, -1),
    ]
    frames = [(hip.REQ_YUV420, yuv(100 + 10 * i + s, 320 if s == 0 else 192, 180 if s == 0 else 108), 2, s)
              for i in range(4) for s in range(2)]
    mix = []
    for i in range(max(len(reqs), len(frames))):   # interleave
        if i < len(frames):
            mix.append(frames[i])
        if i < len(reqs):
            mix.append(reqs[i])
    return mix


def out_planes(kind, planes, factor):
    h, w = planes[0].shape[:2]
    if kind == hip.REQ_RGB:
        return [np.zeros((factor * h, factor * w, 3), np.uint8)]
    if kind == hip.REQ_LUMA:
        return [np.zeros((factor * h, factor * w), np.uint8)]
    return [np.zeros((factor * h, factor * w), np.uint8), np.zeros((factor * h // 2, factor * w // 2), np.uint8),
            np.zeros((factor * h // 2, factor * w // 2), np.uint8)]


def run(server, reqs, together):
    outs = [out_planes(k, p, f) for k, p, f, _ in reqs]
    if together:
        tickets = [server.submit(k, p, f, s, outs[i]) for i, (k, p, f, s) in enumerate(reqs)]
        for t in tickets:
            server.wait(t, 60000)
    else:
        for i, (k, p, f, s) in enumerate(reqs):
            server.wait(server.submit(k, p, f, s, outs[i]), 60000)
    return outs


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_mix_equals_alone(net, prec):
    reqs = make_requests()
    with hip.Server(net, prec) as alone_srv:
        alone = run(alone_srv, reqs, together=False)
    # a small step budget spreads every pass over several steps, mixing requests in each
    with hip.Server(net, prec, step_tiles=16) as mix_srv:
        mixed = run(mix_srv, reqs, together=True)
        st = mix_srv.stats()
    assert st["failed"] == 0
    assert st["completed_frames"] == 8 and st["completed_oneoffs"] == 6
    assert st["graph_replays"] == 8
    assert st["steps"] > 6
    for i, (a, m) in enumerate(zip(alone, mixed)):
        for pa, pm in zip(a, m):
            assert np.array_equal(pa, pm), f"request {i} differs between alone and mixed"


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_oneoffs_equal_processor(net, weights, prec):
    """RGB one-offs through the server equal Processor::process on the same input (same kernels)."""
    proc = P.Processor("hip" if prec == "fp32" else "hip-fp16", 0, f"acnet-legacy-{VARIANT}")
    assert proc.ok(), proc.error()
    with hip.Server(net, prec, step_tiles=64) as srv:
        for seed, (w, h), factor in [(7, (200, 150), 2), (8, (97, 61), 4)]:
            src = rgb(seed, w, h)
            out = out_planes(hip.REQ_RGB, [src], factor)
            srv.wait(srv.submit(hip.REQ_RGB, [src], factor, -1, out), 60000)
            rc, want = proc.process(src, factor)
            assert rc == 0
            assert np.array_equal(out[0], want)
            if prec == "fp32":
                k, b = weights[VARIANT]
                check_int(out[0], oracle_pipeline(src, factor, k, b), frac=0.999, maxlsb=2)


def test_frames_match_oracle(net, weights):
    """Frame path (hipGraph replays): Y within the u8 tolerance of the oracle pass, U/V equal
    to the oracle's bilinear 2x resize (same operations, each rounded separately on both
    sides); repeated replays give identical results."""
    k, b = weights[VARIANT]
    y, u, v = yuv(500, 320, 180)
    with hip.Server(net, "fp32") as srv:
        outs = []
        for _ in range(3):
            o = out_planes(hip.REQ_YUV420, [y, u, v], 2)
            srv.wait(srv.submit(hip.REQ_YUV420, [y, u, v], 2, 5, o), 60000)
            outs.append(o)
        assert srv.stats()["graph_replays"] == 3
    for o in outs[1:]:
        for a, c in zip(outs[0], o):
            assert np.array_equal(a, c)
    check_int(outs[0][0], OL.oracle_pass(y, k, b))
    for src, got in ((u, outs[0][1]), (v, outs[0][2])):
        h, w = src.shape
        want = np.zeros((2 * h, 2 * w), np.uint8)
        assert OL.orc().orc_resize(src.ctypes.data, src.strides[0], w, h, 1, OL.U8, want.ctypes.data, want.strides[0],
                                   2 * w, 2 * h, OL.RESIZE_BILINEAR) == 0
        d = np.abs(got.astype(int) - want.astype(int))
        assert np.array_equal(got, want), f"{np.count_nonzero(d)} of {d.size} chroma samples differ, max {d.max()}"


def test_budget_admission_and_device_outputs(net):
    """A budget that holds one request at a time still completes every request in order of
    admission; outputs kept in the server are read back from device memory."""
    srcs = [synth.flat_shaded(900 + i, 640, 360, dtype=np.uint8) for i in range(4)]
    alone = [net.upscale(torch.from_numpy(s).cuda()).cpu().numpy() for s in srcs]
    budget = 4 << 20       # one request: 2 MiB-rounded blocks for 640x360 and 1280x720
    with hip.Server(net, "fp32", hbm_budget=budget, step_tiles=32) as srv:
        tickets = [srv.submit(hip.REQ_LUMA, [s], 2) for s in srcs]
        for t, want in zip(tickets, alone):
            srv.wait(t, 60000)
            got = srv.fetch(t, 0, want.shape)
            assert np.array_equal(got, want)
            srv.release(t)     # frees the budget for the next request
        st = srv.stats()
    assert st["completed_oneoffs"] == 4 and st["failed"] == 0


def test_invalid_requests(net):
    y = synth.flat_shaded(1, 64, 48, dtype=np.uint8)
    with hip.Server(net, "fp16", hbm_budget=64 << 20) as srv:
        with pytest.raises(hip._native.AchipError):
            srv.submit(hip.REQ_LUMA, [y], 3)                        # factor not a power of two
        with pytest.raises(hip._native.AchipError):
            srv.submit(hip.REQ_RGB, [rgb(5, 64, 48)], 2, stream=0)  # frames are luma / YUV
        big = synth.flat_shaded(2, 4096, 4096, dtype=np.uint8)
        with pytest.raises(hip._native.AchipError):
            srv.submit(hip.REQ_LUMA, [big], 4)                      # alone exceeds the budget
        f = yuv(3, 64, 48)
        srv.wait(srv.submit(hip.REQ_YUV420, f, 2, 1, out_planes(hip.REQ_YUV420, f, 2)), 60000)
        g = yuv(4, 96, 48)
        with pytest.raises(hip._native.AchipError):
            srv.submit(hip.REQ_YUV420, g, 2, 1)                     # a stream keeps its geometry


@pytest.mark.parametrize("model,cls,variant", [("acnet-b4-hdn", hip.Net8, "b4_hdn"), ("artcnn-c4f16-ds", hip.NetArt, "f16_ds"),
                                               ("fsrcnnx", hip.NetFsr, "f8")])
def test_layered_nets_serve(cuda, model, cls, variant):
    """ACNet<8>, ArtCNN and FSRCNNX nets in the serving loop (FP32, and FP16 for ACNet<8>): the
    mixed load of one-offs and frames gives every request the same result as alone, and RGB
    one-offs equal Processor::process with the same model."""
    net = cls(variant, 0)
    reqs = make_requests()
    with hip.Server(net, "fp32") as alone_srv:
        alone = run(alone_srv, reqs, together=False)
    with hip.Server(net, "fp32", step_tiles=16) as mix_srv:
        mixed = run(mix_srv, reqs, together=True)
        st = mix_srv.stats()
    assert st["failed"] == 0 and st["completed_frames"] == 8 and st["completed_oneoffs"] == 6
    assert st["graph_replays"] == 8 and st["steps"] >= 3      # whole passes: up to 2 per one-off
    for i, (a, m) in enumerate(zip(alone, mixed)):
        for pa, pm in zip(a, m):
            assert np.array_equal(pa, pm), f"request {i} differs between alone and mixed"
    proc = P.Processor("hip", 0, model)
    assert proc.ok(), proc.error()
    for i, (k, p, f, s) in enumerate(reqs):
        if k == hip.REQ_RGB:
            rc, want = proc.process(p[0], f)
            assert rc == 0 and np.array_equal(mixed[i][0], want), f"request {i}"
    if cls is hip.Net8:
        # ACNet<8> also serves in FP16: mixed equals alone, one-offs equal the "hip-fp16" processor
        with hip.Server(net, "fp16") as alone_srv:
            alone16 = run(alone_srv, reqs, together=False)
        with hip.Server(net, "fp16", step_tiles=16) as mix_srv:
            mixed16 = run(mix_srv, reqs, together=True)
            assert mix_srv.stats()["failed"] == 0
        for i, (a, m) in enumerate(zip(alone16, mixed16)):
            for pa, pm in zip(a, m):
                assert np.array_equal(pa, pm), f"fp16 request {i} differs between alone and mixed"
        p16 = P.Processor("hip-fp16", 0, model)
        assert p16.ok(), p16.error()
        for i, (k, p, f, s) in enumerate(reqs):
            if k == hip.REQ_RGB:
                rc, want = p16.process(p[0], f)
                assert rc == 0 and np.array_equal(mixed16[i][0], want), f"fp16 request {i}"
    else:
        with pytest.raises(Exception):
            hip.Server(net, "fp16")                   # ArtCNN and FSRCNNX run in FP32 only
    net.close()


def test_page_locked_planes_equal_pageable(net):
    """Page-locked input and output planes (torch's pinned memory) give the same outputs as
    pageable ones, for every request of the mix: the frames' planes then move by the library's
    row copy kernel through their device mapping (dense planes as one row of 16-byte units and
    a byte tail; strided views in 16- or 4-byte units per row), the one-offs' by the copy
    engines, pageable ones by the runtime's staged copies.  One more RGB input is a column
    slice of a wider pinned buffer (rows 750 bytes apart, 600 used)."""
    def pinned(a):
        return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().numpy()

    # plus two frames of a stream whose planes are not whole 16-byte units (a byte tail to copy)
    reqs = make_requests() + [(hip.REQ_YUV420, yuv(300 + i, 202, 114), 2, 2) for i in range(2)]
    # and frames whose page-locked planes are column slices of wider buffers (rows 256 / 128
    # bytes apart): 16-byte units for 192 / 96-pixel rows, 4-byte units for 200 / 100
    strided = []
    for stream, w in ((3, 192), (4, 200)):
        for i in range(2):
            planes = yuv(400 + 10 * stream + i, w, 64)
            views = []
            for pl in planes:
                buf = torch.zeros((pl.shape[0], 256 if pl.shape[1] > 128 else 128), dtype=torch.uint8).pin_memory().numpy()
                buf[:, :pl.shape[1]] = pl
                views.append(buf[:, :pl.shape[1]])
            reqs.append((hip.REQ_YUV420, planes, 2, stream))
            strided.append((len(reqs) - 1, views))
    first_rgb = next(i for i, r in enumerate(reqs) if r[0] == hip.REQ_RGB and r[1][0].shape == (150, 200, 3))
    wide = torch.zeros((150, 250, 3), dtype=torch.uint8).pin_memory().numpy()
    wide[:, :200] = reqs[first_rgb][1][0]
    preqs = [(k, [pinned(p) for p in planes], f, s) for k, planes, f, s in reqs] + [(hip.REQ_RGB, [wide[:, :200]], 2, -1)]
    for i, views in strided:
        preqs[i] = (preqs[i][0], views, preqs[i][2], preqs[i][3])
    with hip.Server(net, "fp32") as srv:
        want = run(srv, reqs, together=True)
        outs = [[pinned(o) for o in out_planes(k, p, f)] for k, p, f, _ in preqs]
        tickets = [srv.submit(k, p, f, s, outs[i]) for i, (k, p, f, s) in enumerate(preqs)]
        for t in tickets:
            srv.wait(t, 60000)
    for i, (a, b) in enumerate(zip(want, outs)):
        for pa, pb in zip(a, b):
            assert np.array_equal(pa, pb), f"request {i}: page-locked planes differ from pageable ones"
    assert np.array_equal(outs[-1][0], want[first_rgb][0])
