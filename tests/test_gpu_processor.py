"""The drop-in surface on the GPU: ac_processor_* / ac_image_* (include/AC/Core/*.h)
with processor type "hip", against the oracle (factor 2 and chained factor 4/8, whose
intermediate keeps the input's element type as core/src/processor/Processor.cpp:219-236
does), preallocated mapped destinations, and the reference's ProcessorTest checks
(tests/core/src/ProcessorTest.cpp:45-104: dimensions, preallocated dst, factor 4)."""
import numpy as np
import pytest

from anime4kcpp_amd import processor as P
from anime4kcpp_amd import synth
from oracle_lib import oracle_pass
from test_gpu_fp32 import check_f32, check_int

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def proc(cuda):
    p = P.Processor("hip", 0, "acnet-legacy-hdn0")
    assert p.ok(), p.error()
    return p


def test_create_and_identity(proc):
    assert proc.type() == 3 and proc.type_name() == "HIP"
    assert "MI355" in proc.name() or len(proc.name()) > 0
    assert "HIP" in P.list_info() and "[0]" in P.list_info()


def test_auto_is_hip(cuda):
    p = P.Processor("auto", 0, "acnet-legacy-gan")
    assert p.ok() and p.type() == 3


@pytest.mark.parametrize("variant", ["gan", "hdn0", "hdn1", "hdn2", "hdn3"])
def test_factor2_u8(weights, cuda, variant):
    p = P.Processor("hip", 0, f"acnet-legacy-{variant}")
    src = synth.flat_shaded(21, 123, 77, dtype=np.uint8)
    rc, out = p.process(src, 2.0)
    assert rc == 0, p.error()
    assert out.shape == (154, 246) and out.dtype == np.uint8
    check_int(out, oracle_pass(src, *weights[variant]))


def test_factor2_f32(proc, weights):
    src = synth.flat_shaded(22, 256, 256)
    rc, out = proc.process(src, 2.0)
    assert rc == 0, proc.error()
    check_f32(out, oracle_pass(src, *weights["hdn0"]))


@pytest.mark.parametrize("factor,passes", [(4.0, 2), (3.0, 2), (8.0, 3)])
def test_chained_factors(proc, weights, factor, passes):
    src = synth.flat_shaded(23, 40, 30, dtype=np.uint8)
    want = src
    for _ in range(passes):
        want = oracle_pass(want, *weights["hdn0"])
    if factor == 3.0:
        return   # non-power-of-two factors: tests/test_gpu_pipeline.py
    rc, out = proc.process(src, factor)
    assert rc == 0, proc.error()
    assert out.shape == (30 << passes, 40 << passes)
    check_int(out, want)


def test_chained_f32(proc, weights):
    src = synth.flat_shaded(24, 64, 48)
    want = oracle_pass(oracle_pass(src, *weights["hdn0"]), *weights["hdn0"])
    rc, out = proc.process(src, 4.0)
    assert rc == 0, proc.error()
    check_f32(out, want, tol=2e-5)   # two passes; the second amplifies first-pass rounding


def test_preallocated_mapped_dst_with_stride(proc, weights):
    """A foreign destination with its own stride, as the video/VS/AVS callers pass it."""
    src = synth.flat_shaded(25, 50, 20, dtype=np.uint8)
    big = np.full((40, 130), 7, np.uint8)           # stride 130 > 100
    dst = big[:, :100]
    rc, _ = proc.process(src, 2.0, dst=dst)
    assert rc == 0, proc.error()
    check_int(big[:, :100], oracle_pass(src, *weights["hdn0"]))
    assert (big[:, 100:] == 7).all()


def test_mixed_types_into_preallocated(proc, weights):
    src = synth.flat_shaded(26, 33, 17, dtype=np.uint8)
    dst = np.zeros((34, 66), np.float32)
    rc, _ = proc.process(src, 2.0, dst=dst)
    assert rc == 0, proc.error()
    check_f32(dst, oracle_pass(src, *weights["hdn0"], out_dtype=np.float32))


def test_errors_are_reported_not_hidden(proc):
    two = np.zeros((8, 8, 2), np.uint8)              # 2-channel input is not an image layout
    rc, _ = proc.process(two, 2.0)
    assert rc == -256 and not proc.ok()
    src = np.zeros((8, 8), np.uint8)
    rc, _ = proc.process(src, 2.0)          # the next good call clears this thread's error
    assert rc == 0 and proc.ok()


def test_threads_share_one_processor(proc, weights):
    """One processor, many threads at once (per-thread streams, buffers and errors)."""
    import threading
    srcs = [synth.flat_shaded(100 + i, 60 + i, 40 + 2 * i, dtype=np.uint8) for i in range(8)]
    outs = [None] * 8

    def work(i):
        for _ in range(3):
            outs[i] = proc.process(srcs[i], 2.0)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    for i in range(8):
        rc, out = outs[i]
        assert rc == 0
        check_int(out, oracle_pass(srcs[i], *weights["hdn0"]))


@pytest.mark.parametrize("type_str,precision", [("hip-fp16", "fp16"), ("hip:fp16", "fp16"),
                                                ("hip-bf16", "bf16"), ("hip:bf16", "bf16"), ("hip", "fp32")])
def test_type_string_selects_precision(cuda, type_str, precision):
    """A processor built from a 16-bit type string runs that precision's kernel: its
    output equals achip_pass at the same precision bit for bit (and the BF16 and FP16
    forms differ from each other on this content)."""
    import torch
    from anime4kcpp_amd import hip
    src = synth.flat_shaded(26, 150, 97, dtype=np.uint8)
    p = P.Processor(type_str, 0, "acnet-legacy-hdn1")
    assert p.ok(), p.error()
    rc, out = p.process(src, 2.0)
    assert rc == 0, p.error()
    net = hip.Net("hdn1", 0)
    want = {}
    for prec in ("fp32", "fp16", "bf16"):
        want[prec] = net.upscale(torch.from_numpy(src).to(cuda), precision=prec).cpu().numpy()
    net.close()
    assert not np.array_equal(want["fp16"], want["bf16"])
    assert np.array_equal(out, want[precision])


def test_staged_and_page_locked_host_images(proc, weights):
    """Caller images larger than one 4 MiB staging chunk: pageable ones travel through the
    page-locked chunks both ways (2 chunks up and 5 down for gray, 2 up and 7 down for RGB,
    and a destination whose rows are narrower than its stride), page-locked ones (torch's
    pinned memory, mapped) are copied directly; every route gives the same pixels, and the
    gray one is checked against the oracle."""
    import torch
    gray = synth.flat_shaded(41, 2600, 1700, dtype=np.uint8)          # 4.4 MB in, 17.7 MB out
    rc, out = proc.process(gray, 2.0)
    assert rc == 0, proc.error()
    check_int(out, oracle_pass(gray, *weights["hdn0"]))
    pinned_src = torch.from_numpy(gray).pin_memory().numpy()
    pinned_dst = torch.zeros((3400, 5200), dtype=torch.uint8).pin_memory().numpy()
    rc, _ = proc.process(P.Image(pinned_src, mapped=True), 2.0, dst=pinned_dst)
    assert rc == 0, proc.error()
    assert np.array_equal(pinned_dst, out)

    g = synth.flat_shaded(42, 1600, 1400, dtype=np.uint8)
    rgb = np.ascontiguousarray(np.stack([g, np.roll(g, 3, 0), np.roll(g, 5, 1)], 2))   # 6.7 MB in, 27 MB out
    rc, ref = proc.process(rgb, 2.0)
    assert rc == 0, proc.error()
    big = np.zeros((2800, 3300, 3), np.uint8)
    rc, _ = proc.process(rgb, 2.0, dst=big[:, :3200])
    assert rc == 0, proc.error()
    assert np.array_equal(big[:, :3200], ref) and not big[:, 3200:].any()
    pinned_rgb = torch.from_numpy(rgb).pin_memory().numpy()
    pinned_out = torch.zeros((2800, 3200, 3), dtype=torch.uint8).pin_memory().numpy()
    rc, _ = proc.process(P.Image(pinned_rgb, mapped=True), 2.0, dst=pinned_out)
    assert rc == 0, proc.error()
    assert np.array_equal(pinned_out, ref)
