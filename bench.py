"""Benchmark: output Mpix/s of the ACNetLegacy 2x pass, 1920x1080 u8 luma -> 3840x2160.

One step = one 2x pass over a batch of synthetic 1080p u8 luma frames already resident in
HBM (two frames per weight set: GAN, HDN0, HDN1, HDN2, HDN3), written to resident 4K
outputs.  Prints one JSON line (rank 0):
  value      whole-job output Mpix/s of the FP32 pass the processor runs (achip_pass_ws with
             a workspace: at 1080p the layer-split form, four launches, the fused kernel's
             bits) over the timed steps, the step's frames alternating over two streams as
             two calling threads would drive the processor; the same on one stream, the fused
             kernel alone and the FP16-/BF16-MFMA forms are timed on the same frames and
             reported beside it, the 16-bit forms with their error against the FP64
             restatement
  roofline   achieved FP32 TFLOP/s of that pass (9424 algorithmic FLOP per input pixel x
             pixels per pass / mean pass time from HIP events on the launch stream) against
             the MI355X FP32 peak
  cpu_baseline  the reference CPU 2x pass (its own per-ISA kernels, oracle/_ref, or the
             oracle's restatement when that is absent) on a bounded sample of the same
             frames, timed on the host cores

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W_IN, H_IN = 1920, 1080
FLOP_PER_INPUT_PX = 9424          # 2 x 4712 MAC per input pixel per 2x pass
PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 peak
PEAK_MFMA16_TFLOPS = 2500.0       # MI355X published dense FP16/BF16 MFMA peak
PEAK_HBM_GBS = 8000.0             # MI355X published HBM3E bandwidth
PEAK_PCIE_GBS = 64.0              # PCIe 5.0 x16, one direction
FRAMES_PER_VARIANT = 2
# HBM bytes per 1080p u8 pass of the default configuration of each precision (fp32_split: the
# layer-split form's four launches together), measured with rocprofv3 --pmc FETCH_SIZE (x2.000)
# and WRITE_SIZE (x1.000) against a 256 MiB copy (profiles/acnet_fused_1080p_u8.md); counters
# cannot be read from inside this process.
MEASURED_TRAFFIC = {"fp32": 11.24e6, "fp16": 11.08e6, "bf16": 11.08e6, "fp32_split": 415.2e6}
# bytes per input pixel the layer-split form moves by design: the u8 source, a 32-byte f32 map
# written by the first launch, read and written by the two middle ones and read by the last,
# and the four u8 output pixels
SPLIT_FORM_BYTES_PER_PX = 1 + 32 * 6 + 4


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_reps(run, min_reps, min_seconds, max_reps):
    """One warm-up call, then at least min_reps timed calls (and at least min_seconds of
    them, up to max_reps); returns (median seconds per call, reps, total seconds)."""
    import statistics
    run()
    times = []
    while len(times) < min_reps or (sum(times) < min_seconds and len(times) < max_reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), len(times), sum(times)


def cpu_baseline(frames_u8, weights, min_seconds=10.0):
    """Bounded CPU sample: 2x passes of one 1080p u8 frame on all host cores, one warm-up then
    the median of >= 5 timed passes (repeated up to about min_seconds); output Mpix/s.  Where
    oracle/_ref is built, the passes run the reference's own per-ISA dot/conv kernels (OpImpl,
    compiled from its sources) inside the oracle's restatement of its per-pixel loops;
    otherwise the oracle's restatement alone."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    have_ref = OL.ref_available()
    arch = 5 if (have_ref and OL.ref().acref_has_arch(5)) else 4
    use_ref = have_ref and OL.ref().acref_has_arch(arch)
    k, b = weights["hdn0"]
    src = frames_u8[0]
    run = (lambda: OL.ref_pass(arch, src, k, b)) if use_ref else (lambda: OL.oracle_pass(src, k, b))
    out_mpix = 4 * W_IN * H_IN / 1e6
    med, n, el = _median_reps(run, 5, min_seconds, 200)
    what = (f"reference dot/conv kernels (OpImpl, arch {arch}) built from its sources, per-pixel loops restated"
            if use_ref else "oracle restatement (FMA order)")
    out = {"value": round(out_mpix / med, 2), "unit": "Mpix/s", "cores": threads,
           "kind": "reference" if use_ref else "port",
           "cpu_model": cpu_model(),
           "method": "1 warm-up, median of timed passes",
           "sample": f"median of {n} passes of one 1920x1080 u8 frame (HDN0), {el:.1f} s in all; {what}; "
                     f"OpenMP {threads} threads"}
    # the same passes on one thread (OpenMP's thread count of this thread's next parallel regions)
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
    except OSError:
        return out
    gomp.omp_set_num_threads(1)
    try:
        med1, n1, el1 = _median_reps(run, 5, min_seconds / 3, 50)
    finally:
        gomp.omp_set_num_threads(threads)
    out["one_thread"] = {"value": round(out_mpix / med1, 2), "unit": "Mpix/s",
                         "sample": f"median of {n1} passes of the same frame on one thread, {el1:.1f} s in all"}
    return out


def time_precision(nets, srcs, dsts, order, prec, steps, warmup, config, streams, workspaces=None):
    """Time `steps` steps of one precision; returns (wall seconds, seconds per pass).  Frame n
    of a step runs on streams[n % len(streams)], each stream with its own workspace, as the
    processor runs each calling thread's passes on that thread's own stream.  The time per
    pass is the step's span on the device (HIP events: after every stream's end, less the
    first stream's start) over its frames.  With workspaces the passes run through
    achip_pass_ws (the processor's pass: the faster form for the size), else through
    achip_pass (the fused kernel)."""
    import torch
    ns = len(streams)

    def step():
        for n, (i, v) in enumerate(order):
            k = n % ns
            nets[v].upscale(srcs[i], dsts[i], precision=prec, config=config, stream=streams[k],
                            workspace=workspaces[k] if workspaces else None)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(streams[0])
    for st in streams[1:]:
        st.wait_event(e0)
    for _ in range(steps):
        step()
    for st in streams[1:]:
        done = torch.cuda.Event()
        done.record(st)
        streams[0].wait_event(done)
    e1.record(streams[0])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) / 1e3 / (steps * len(srcs))


def layer_split_pass(net, src, dst, ws, stream):
    """One ACNetLegacy FP32 pass in the layer-split form (achip_pass_layers: layers 0-2, then
    layers 3..8 two per launch through f32 maps in `ws`)."""
    import ctypes
    from anime4kcpp_amd import _native, hip
    h, w = src.shape
    rc = _native.lib().achip_pass_layers(net.handle, ctypes.c_void_p(src.data_ptr()), src.stride(0) * src.element_size(),
                                         hip.elem_type(src.dtype), w, h, ctypes.c_void_p(dst.data_ptr()),
                                         dst.stride(0) * dst.element_size(), hip.elem_type(dst.dtype),
                                         ctypes.c_void_p(ws.data_ptr()), ws.numel(), ctypes.c_void_p(stream))
    if rc != 0:
        raise RuntimeError(f"achip_pass_layers: {_native.last_error()}")


def error_vs_fp64(nets, weights, prec):
    """Max/mean abs error and PSNR of one 1920x1080 f32 frame (HDN0, f32 output) against
    the FP64 restatement of the network."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from anime4kcpp_amd import synth
    src = synth.flat_shaded(7, W_IN, H_IN)
    out = nets["hdn0"].upscale(torch.from_numpy(src).cuda(), precision=prec, out_dtype=torch.float32)
    torch.cuda.synchronize()
    ref = OL.oracle_pass_f64(src, *weights["hdn0"])
    e = np.abs(out.cpu().numpy().astype(np.float64) - ref)
    mse = float((e ** 2).mean())
    return {"max_abs": float(e.max()), "mean_abs": float(e.mean()),
            "psnr_db": round(10 * np.log10(1.0 / mse), 2) if mse > 0 else None}


def batched_uneven(nets, prec, seed=2024, n=64, streams=None):
    """64 u8 images, sides uniform in [128, 2048] (seeded): one persistent launch with a tile
    work-queue per configuration (the autotuner's sweep) against 64 separate launches."""
    import numpy as np
    import torch
    from anime4kcpp_amd import hip, synth
    dev = torch.device("cuda:0")
    ws = synth.randint(seed, n, 128, 2048)
    hs = synth.randint(seed + 1, n, 128, 2048)
    srcs = [torch.from_numpy(synth.flat_shaded(700 + i, int(w), int(h), dtype=np.uint8)).to(dev) for i, (w, h) in enumerate(zip(ws, hs))]
    dsts = [torch.empty((2 * s.shape[0], 2 * s.shape[1]), dtype=torch.uint8, device=dev) for s in srcs]
    net = nets["hdn0"]
    out_mpix = sum(d.numel() for d in dsts) / 1e6
    from anime4kcpp_amd import _native
    ws_t = torch.empty(_native.lib().achip_batch_workspace_bytes(n), dtype=torch.uint8, device=dev)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    # three rounds over the configurations, the best round of each: a configuration measured
    # first in a single round read 5-15% slow
    sweep = {}
    for _ in range(3):
        for c in range(len(hip.configs(prec))):
            sweep[c] = min(sweep.get(c, float("inf")),
                           timed(lambda: net.upscale_batch(srcs, dsts, precision=prec, config=c, workspace=ws_t)))
    best = min(sweep, key=sweep.get)
    sep = timed(lambda: [net.upscale(s, d, precision=prec, config=best) for s, d in zip(srcs, dsts)])
    tb = sweep[best]
    names = {c: f"{hip.configs(prec)[c]['tw']}x{hip.configs(prec)[c]['th']}/{hip.configs(prec)[c]['nt']}thr"
                + ("/mfma" if hip.configs(prec)[c]["p"] == 0 else "/valu") + "/halo-recompute" for c in sweep}
    sweep_ms = {names[c]: round(t * 1e3, 3) for c, t in sweep.items()}
    if prec == "fp32":
        # the layer-split form (one image at a time: 4 launches, f32 maps through HBM), on one
        # stream and with the images alternating over two (a workspace each)
        nbytes = _native.lib().achip_pass8_workspace_bytes(2048, 2048)
        lws = [torch.empty(nbytes, dtype=torch.uint8, device=dev) for _ in range(2)]
        # the main section's two streams: a new stream here would change which hardware queues
        # the serving section's streams share (4 per process by default)
        sts = list(streams[:2]) if streams and len(streams) >= 2 else [torch.cuda.current_stream(dev), torch.cuda.Stream(dev)]
        st = sts[0].cuda_stream
        sweep_ms["62x16/256thr/mfma/layer-split (per image)"] = round(
            timed(lambda: [layer_split_pass(net, s, d, lws[0], st) for s, d in zip(srcs, dsts)]) * 1e3, 3)

        def two_streams():
            ev = torch.cuda.Event()
            ev.record(sts[0])
            sts[1].wait_event(ev)
            for i, (s, d) in enumerate(zip(srcs, dsts)):
                layer_split_pass(net, s, d, lws[i % 2], sts[i % 2].cuda_stream)
            ev2 = torch.cuda.Event()
            ev2.record(sts[1])
            sts[0].wait_event(ev2)
        sweep_ms["62x16/256thr/mfma/layer-split (per image, two streams)"] = round(timed(two_streams) * 1e3, 3)
        bws = torch.empty(net.batch_layers_bytes(srcs), dtype=torch.uint8, device=dev)
        sweep_ms["62x16/256thr/mfma/layer-split batched (4 launches over all images)"] = round(
            timed(lambda: net.upscale_batch_layers(srcs, dsts, workspace=bws)) * 1e3, 3)
        del bws
    out = {
        "images": n, "input_mpix": round(sum(s.numel() for s in srcs) / 1e6, 2), "output_mpix": round(out_mpix, 2),
        "one_launch": {"images_per_s": round(n / tb, 1), "mpix_per_s": round(out_mpix / tb, 1), "ms": round(tb * 1e3, 3)},
        "separate_launches": {"images_per_s": round(n / sep, 1), "mpix_per_s": round(out_mpix / sep, 1), "ms": round(sep * 1e3, 3)},
        "autotune_sweep_ms": sweep_ms,
        "chosen_config": best,
    }
    key = "62x16/256thr/mfma/layer-split batched (4 launches over all images)"
    if key in sweep_ms:
        tl = sweep_ms[key] / 1e3
        out["layer_split_batched"] = {"images_per_s": round(n / tl, 1), "mpix_per_s": round(out_mpix / tl, 1),
                                      "ms": round(tl * 1e3, 3), "vs_one_launch": round(tb / tl, 2)}
    out["fastest"] = min(sweep_ms, key=sweep_ms.get)
    return out


def precision_4k_x4(nets, weights, precs, crop=1024):
    """3840x2160 f32 -> 15360x8640 f32 (two chained 2x passes) per precision, each pass the
    processor's (achip_pass_ws with a workspace: in FP32 the layer-split form at these sizes):
    Mpix/s of the two-pass chain, and each pass's error against the FP64 restatement on a
    crop (interior only: the crop's own borders differ from the full image's)."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from anime4kcpp_amd import synth
    dev = torch.device("cuda:0")
    net = nets["hdn0"]
    k, b = weights["hdn0"]
    src_np = synth.flat_shaded(44, 3840, 2160)
    src = torch.from_numpy(src_np).to(dev)
    mid = torch.empty((4320, 7680), dtype=torch.float32, device=dev)
    out = torch.empty((8640, 15360), dtype=torch.float32, device=dev)
    out_mpix = out.numel() / 1e6
    res = {}
    for prec in precs:
        wsb = max(net.workspace_bytes(3840, 2160, precision=prec, src_dtype=torch.float32),
                  net.workspace_bytes(7680, 4320, precision=prec, src_dtype=torch.float32))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev) if wsb else None
        net.upscale(src, mid, precision=prec, workspace=ws)
        net.upscale(mid, out, precision=prec, workspace=ws)
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            net.upscale(src, mid, precision=prec, workspace=ws)
            net.upscale(mid, out, precision=prec, workspace=ws)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        errs = {}
        m = 12   # interior margin (input pixels) excluded from the crop comparison
        for pi, (inp_np, dev_out) in enumerate(((src_np, mid), (None, out))):
            if inp_np is None:
                inp_np = mid.cpu().numpy()
            y0, x0 = (inp_np.shape[0] - crop) // 2, (inp_np.shape[1] - crop) // 2
            c = np.ascontiguousarray(inp_np[y0:y0 + crop, x0:x0 + crop])
            ref = OL.oracle_pass_f64(c, k, b)[2 * m:-2 * m, 2 * m:-2 * m]
            got = dev_out[2 * y0 + 2 * m:2 * (y0 + crop) - 2 * m, 2 * x0 + 2 * m:2 * (x0 + crop) - 2 * m].cpu().numpy()
            e = np.abs(got.astype(np.float64) - ref)
            mse = float((e ** 2).mean())
            errs[f"pass{pi + 1}"] = {"max_abs": float(e.max()), "mean_abs": float(e.mean()),
                                     "psnr_db": round(10 * np.log10(1.0 / mse), 2) if mse > 0 else None}
        res[prec] = {"ms": round(t * 1e3, 3), "output_mpix_per_s": round(out_mpix / t, 1), "error_vs_fp64": errs,
                     "form": "layer-split" if wsb else "fused"}
        del ws
    base = res[precs[0]]["output_mpix_per_s"]
    for prec in precs[1:]:
        res[prec]["vs_" + precs[0]] = round(res[prec]["output_mpix_per_s"] / base, 2)
    res["workload"] = f"3840x2160 f32 -> 7680x4320 -> 15360x8640 f32 (HDN0); error on a {crop}x{crop} crop per pass"
    return res


@contextlib.contextmanager
def client_gc_paused():
    """Python's garbage collector run once, then off for the body, so that no collection pauses
    the submitting thread inside a measured region."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def serving(net, prec, loads, seconds, seed=5, bursts=5):
    """Poisson arrivals at each offered load (requests/s): 60% frames of 8 concurrent 1080p
    YUV420 u8 streams (Y through the network, UV bilinear 2x, hipGraph replays), 40% one-off
    RGB u8 images with sides uniform in [128, 4096] and factor 2 or 4.  Inputs are
    pre-generated in page-locked host memory and uploaded by the server; outputs stay in
    device memory and are released on completion."""
    import threading
    import numpy as np
    import torch
    from anime4kcpp_amd import hip, synth
    rng = np.random.default_rng(seed)
    frames = []
    for s in range(8):
        y = torch.from_numpy(synth.flat_shaded(3000 + s, W_IN, H_IN, dtype=np.uint8)).pin_memory()
        u = torch.from_numpy(synth.flat_shaded(3100 + s, W_IN // 2, H_IN // 2, dtype=np.uint8)).pin_memory()
        v = torch.from_numpy(synth.flat_shaded(3200 + s, W_IN // 2, H_IN // 2, dtype=np.uint8)).pin_memory()
        frames.append([y, u, v])
    sides = synth.randint(seed, 24, 128, 4096), synth.randint(seed + 1, 24, 128, 4096)
    pool = []
    for i, (w, h) in enumerate(zip(*sides)):
        g = synth.flat_shaded(3300 + i, int(w), int(h), dtype=np.uint8)
        pool.append(torch.from_numpy(np.ascontiguousarray(np.stack([g, np.roll(g, 1, 0), np.roll(g, 1, 1)], 2))).pin_memory())
    def pick():
        if rng.random() < 0.6:
            s = int(rng.integers(8))
            return (hip.REQ_YUV420, frames[s], 2, s), False
        return (hip.REQ_RGB, [pool[int(rng.integers(len(pool)))]], 2 if rng.random() < 0.5 else 4, -1), True

    out = {}
    srv = hip.Server(net, prec, hbm_budget=64 << 30)

    def burst(n):
        ts = [(srv.submit(*r), keep) for r, keep in (pick() for _ in range(n))]
        for t, keep in ts:
            srv.wait(t, 120000)
            if keep:
                srv.release(t)

    # one untimed burst first: a long-running server has its device pool, frame graphs and
    # kernels loaded, and the process its copy engines set up, before the loads below
    burst(120)
    srv.stats(reset=True)
    # capacity: the same mix submitted all at once, `bursts` times (each burst draws its own
    # 120 requests); the median burst is the figure, the spread beside it
    # The client's garbage collector stays out of the measured bursts and loads: a full
    # collection (38 ms, after the earlier sections' objects) paused this thread in the middle
    # of a burst while the server sat idle, and that burst read 72 against 92-96 Gpix/s.
    with client_gc_paused():
        rates, req_rates, steps = [], [], 0
        for _ in range(max(1, bursts)):
            burst(120)
            st = srv.stats(reset=True)
            rates.append(st["output_mpix"] / st["busy_s"])
            req_rates.append(120 / st["busy_s"])
            steps += st["steps"]
        import statistics
        out["saturated"] = {"requests": 120, "bursts": len(rates), "output_mpix_per_s": round(statistics.median(rates), 1),
                            "min_mpix_per_s": round(min(rates), 1), "max_mpix_per_s": round(max(rates), 1),
                            "per_burst_mpix_per_s": [round(r, 1) for r in rates],
                            "requests_per_s": round(statistics.median(req_rates), 1), "steps": steps}
        for load in loads:
            tickets = []
            lock = threading.Lock()
            done = threading.Event()

            def waiter():
                i = 0
                while True:
                    with lock:
                        t = tickets[i] if i < len(tickets) else None
                    if t is None:
                        if done.is_set():
                            with lock:
                                if i >= len(tickets):
                                    return
                        time.sleep(0.0005)
                        continue
                    srv.wait(t[0], 120000)
                    if t[1]:
                        srv.release(t[0])
                    i += 1

            th = threading.Thread(target=waiter)
            th.start()
            t_next = time.perf_counter()
            t_end = t_next + seconds
            n_frame = n_one = 0
            while True:
                t_next += rng.exponential(1.0 / load)
                if t_next > t_end:
                    break
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                r, keep = pick()
                t = srv.submit(*r)
                n_frame += 0 if keep else 1
                n_one += 1 if keep else 0
                with lock:
                    tickets.append((t, keep))
            done.set()
            th.join()
            st = srv.stats(reset=True)
            out[str(load)] = {
                "offered_req_per_s": load, "frames": n_frame, "oneoffs": n_one,
                "frame_p50_ms": round(st["frame_p50_ms"], 3), "frame_p99_ms": round(st["frame_p99_ms"], 3),
                "oneoff_p50_ms": round(st["oneoff_p50_ms"], 3), "oneoff_p99_ms": round(st["oneoff_p99_ms"], 3),
                "output_mpix_per_s": round(st["output_mpix"] / st["busy_s"], 1) if st["busy_s"] > 0 else None,
                "steps": st["steps"], "graph_replays": st["graph_replays"], "failed": st["failed"],
            }
    srv.close()
    out["method"] = ("one server per precision; an untimed burst of 120 requests of the mix, then the saturated bursts "
                     "(median), then each offered load")
    return out


def stages():
    """HBM-bound stages around the passes (colour split, chroma resize + merge, YUV420 chroma
    2x) and the PCIe copies of a frame, timed with HIP events on device-resident planes.
    Algorithmic bytes: every input element read once, every output element written once."""
    import torch
    from anime4kcpp_amd import hip
    dev = torch.device("cuda:0")

    def timeit(fn, n=20):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n / 1e3   # s

    def entry(sec, nbytes, peak, what):
        gbs = nbytes / sec / 1e9
        return {"us": round(sec * 1e6, 2), "bytes": nbytes, "gb_per_s": round(gbs, 1), "peak_gb_per_s": peak,
                "frac": round(gbs / peak, 4), "workload": what}

    out = {}
    W, H = 4096, 4096
    rgb = torch.randint(0, 256, (H, W, 3), dtype=torch.uint8, device=dev)
    y, uv = hip.rgb2yuv(rgb)
    out["colour_split_u8"] = entry(timeit(lambda: hip.rgb2yuv(rgb, y, uv)), 6 * W * H, PEAK_HBM_GBS,
                                   "RGB u8 4096x4096 -> Y + UV (ImageProcess.cpp rgb2yuv)")
    uvh = uv[: H // 2, : W // 2].contiguous()
    o = torch.empty_like(rgb)
    out["chroma_resize_merge_u8"] = entry(timeit(lambda: hip.resize_yuv2rgb(y, uvh, 1, o)),
                                          2 * (W // 2) * (H // 2) + 4 * W * H, PEAK_HBM_GBS,
                                          "UV 2048x2048 Catmull-Rom 2x + Y 4096x4096 -> RGB u8 (RGB one-off finish)")
    cu = torch.randint(0, 256, (540, 960), dtype=torch.uint8, device=dev)
    c2 = torch.empty((1080, 1920), dtype=torch.uint8, device=dev)
    out["yuv420_chroma_2x_u8"] = entry(timeit(lambda: hip.resize(cu, (1920, 1080), 16, c2)), 960 * 540 + 1920 * 1080,
                                       PEAK_HBM_GBS, "U or V 960x540 bilinear 2x (1080p YUV420 frame)")
    frame_in = 1920 * 1080 * 3 // 2
    frame_out = 3840 * 2160 * 3 // 2
    hin = torch.empty(frame_in, dtype=torch.uint8, pin_memory=True)
    din = torch.empty(frame_in, dtype=torch.uint8, device=dev)
    hout = torch.empty(frame_out, dtype=torch.uint8, pin_memory=True)
    dout = torch.empty(frame_out, dtype=torch.uint8, device=dev)
    out["upload_frame"] = entry(timeit(lambda: din.copy_(hin, non_blocking=True)), frame_in, PEAK_PCIE_GBS,
                                "1080p YUV420 u8 frame, page-locked host -> HBM")
    out["download_frame"] = entry(timeit(lambda: hout.copy_(dout, non_blocking=True)), frame_out, PEAK_PCIE_GBS,
                                  "4K YUV420 u8 frame, HBM -> page-locked host")
    return out


def acnet8(frames_np, weights=None, steps=25):
    """ACNet<8> (the reference's default model family) 1080p u8 -> 4K 2x passes for 4, 8 and 18
    blocks, in FP32 (FP32 MFMA, f32 maps) and in FP16 / BF16 (16-bit MFMA, 16-bit maps); blocks
    two per launch.  FLOP per input pixel: 2 x (72 + 576 blocks + 288) MAC.  HBM bytes per
    input pixel: the u8 source read twice (L0, the final residual), 4 output bytes, and an
    8-channel map (32 bytes FP32, 16 bytes 16-bit) written by the first launch, read and
    written once per map-to-map launch, and read by the last.  Error of each form against the FP64
    restatement on a 1080p f32 frame (B8 HDN)."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from anime4kcpp_amd import hip, synth
    dev = torch.device("cuda:0")
    srcs = [torch.from_numpy(f).to(dev) for f in frames_np[:4]]
    h, w = frames_np[0].shape
    dsts = [torch.empty((2 * h, 2 * w), dtype=torch.uint8, device=dev) for _ in srcs]
    out = {}
    for variant in ("b4_hdn", "b8_hdn", "b18_hdn"):
        net = hip.Net8(variant, 0)
        out[variant] = {"blocks": net.blocks}
        for prec in ("fp32", "fp16", "bf16"):
            for _ in range(3):   # warm-up: clocks, caches, first-launch costs
                for i in range(len(srcs)):
                    net.upscale(srcs[i], dsts[i], precision=prec)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(steps):
                for i in range(len(srcs)):
                    net.upscale(srcs[i], dsts[i], precision=prec)
            e1.record()
            torch.cuda.synchronize()
            sec = e0.elapsed_time(e1) / 1e3 / (steps * len(srcs))
            flop = 2 * (72 + 576 * net.blocks + 288) * w * h
            peak = PEAK_FP32_TFLOPS if prec == "fp32" else PEAK_MFMA16_TFLOPS
            mb = 32 if prec == "fp32" else 16
            # map-to-map launches: one per two blocks after L0 + block 1 (even counts) or L0
            # (odd), the last block running with the final layer
            maps = (net.blocks - 1) // 2
            entry = {"ms_per_frame": round(sec * 1e3, 4), "output_mpix_per_s": round(4 * w * h / sec / 1e6, 1),
                     "achieved_tflops": round(flop / sec / 1e12, 2), "peak_tflops": peak,
                     "frac": round(flop / sec / 1e12 / peak, 4),
                     "hbm_bytes_per_frame": (2 + 4 + mb + 2 * mb * maps + mb) * w * h}
            if variant == "b8_hdn":
                k, b, a, blocks = hip.load_weights8(variant)
                xf = synth.flat_shaded(7, w, h)
                if "ref64" not in out:
                    out["ref64"] = OL.oracle_pass8_f64(xf, k, b, a, blocks)
                got = net.upscale(torch.from_numpy(xf).to(dev), precision=prec, out_dtype=torch.float32).cpu().numpy()
                e = np.abs(got.astype(np.float64) - out["ref64"])
                mse = float((e ** 2).mean())
                entry["error_vs_fp64"] = {"max_abs": float(e.max()), "mean_abs": float(e.mean()),
                                          "psnr_db": round(10 * np.log10(1.0 / mse), 2) if mse > 0 else None}
            out[variant][prec] = entry
        net.close()
    out.pop("ref64", None)
    out["workload"] = ("1920x1080 u8 -> 3840x2160 u8, ACNet<8> HDN weights, layer by layer (8-channel maps in HBM: f32 "
                       "for fp32, 16-bit for fp16/bf16); error on a 1920x1080 f32 frame (B8 HDN) against FP64")
    return out


def layered(frames_np, family, steps=10):
    """ArtCNN (16 and 32 channels) or FSRCNNX (8 and 16 channels) 1080p u8 -> 4K 2x passes,
    layer by layer on FP32 MFMA (FSRCNNX F8: blocks two per launch).  FLOP per input pixel:
    ArtCNN 2 x (9 F + 9 F^2 (B + 1) + 36 F) MAC; FSRCNNX 2 x (25 F + 9 F^2 B + F^2 + 36 F) MAC
    (B = 4 blocks).  HBM bytes per input pixel: the u8 input and 4 outputs, plus 4F per map
    read or written."""
    import torch
    from anime4kcpp_amd import hip
    dev = torch.device("cuda:0")
    srcs = [torch.from_numpy(f).to(dev) for f in frames_np[:4]]
    h, w = frames_np[0].shape
    dsts = [torch.empty((2 * h, 2 * w), dtype=torch.uint8, device=dev) for _ in srcs]
    out = {}
    cls, variants = (hip.NetArt, ("f16", "f32")) if family == "artcnn" else (hip.NetFsr, ("f8", "f16"))
    for variant in variants:
        net = cls(variant, 0)
        for _ in range(3):   # warm-up: clocks, caches, first-launch costs
            for i in range(len(srcs)):
                net.upscale(srcs[i], dsts[i])
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            for i in range(len(srcs)):
                net.upscale(srcs[i], dsts[i])
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / (steps * len(srcs))
        f, b = net.channels, net.blocks
        if family == "artcnn":
            mac, maps = 9 * f + 9 * f * f * (b + 1) + 36 * f, 2 * (b + 2) + 1
        else:
            # head write; per launch of blocks a read and a write (the last block also reads
            # the head's map); the tail's read -- F8 runs the blocks two per launch
            mac, maps = 25 * f + 9 * f * f * b + f * f + 36 * f, (b + 3 if f == 8 else 2 * b + 3)
        flop = 2 * mac * w * h
        out[variant] = {"ms_per_frame": round(sec * 1e3, 4), "output_mpix_per_s": round(4 * w * h / sec / 1e6, 1),
                        "channels": f, "blocks": b, "achieved_tflops": round(flop / sec / 1e12, 2),
                        "frac_of_fp32_peak": round(flop / sec / 1e12 / PEAK_FP32_TFLOPS, 4),
                        "hbm_bytes_per_frame": (4 * f * maps + 1 + 4) * w * h}
        net.close()
    name = "ArtCNN C4" if family == "artcnn" else "FSRCNNX B4"
    out["workload"] = f"1920x1080 u8 -> 3840x2160 u8, {name} weights, layer by layer (F-channel f32 maps in HBM)"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-error", action="store_true", help="skip the FP64 error report")
    ap.add_argument("--precisions", default="fp32,fp16,bf16")
    ap.add_argument("--config", type=int, default=None, help="kernel tile configuration")
    ap.add_argument("--streams", type=int, default=2, help="streams the step's frames alternate over")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-image batched measurement")
    ap.add_argument("--no-4k", action="store_true", help="skip the 4K f32 x4 precision measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-stages", action="store_true", help="skip the colour / resize / PCIe stage measurement")
    ap.add_argument("--no-acnet8", action="store_true", help="skip the ACNet<8> measurement")
    ap.add_argument("--no-artcnn", action="store_true", help="skip the ArtCNN measurement")
    ap.add_argument("--no-fsrcnnx", action="store_true", help="skip the FSRCNNX measurement")
    ap.add_argument("--serving-loads", default="fp16:300,1200;fp32:60,240",
                    help="offered loads (requests/s) per precision")
    ap.add_argument("--serving-seconds", type=float, default=3.0)
    ap.add_argument("--serving-bursts", type=int, default=5, help="saturated bursts of 120 requests (median reported)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device copy of this size first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every image is independent: one process per GPU, each taking its own frames)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from anime4kcpp_amd import hip, synth

    dev = torch.device("cuda:0")
    if args.calib_copy_mib:
        a = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    weights = {v: hip.load_weights(v) for v in hip.VARIANTS}
    nets = {v: hip.Net(v, 0) for v in hip.VARIANTS}
    frames_np = [synth.flat_shaded(1000 + i, W_IN, H_IN, dtype=np.uint8)
                 for i in range(FRAMES_PER_VARIANT * len(hip.VARIANTS))]
    srcs = [torch.from_numpy(f).to(dev) for f in frames_np]
    dsts = [torch.empty((2 * H_IN, 2 * W_IN), dtype=torch.uint8, device=dev) for _ in srcs]
    order = [(i, hip.VARIANTS[i % len(hip.VARIANTS)]) for i in range(len(srcs))]
    stream = torch.cuda.current_stream(dev)
    out_px_per_step = len(srcs) * 4 * W_IN * H_IN

    precs = [p for p in args.precisions.split(",") if p]
    # the processor's pass (achip_pass_ws): each thread lends it a workspace, and at 1080p FP32
    # it runs the layer-split form (4 launches, the fused kernel's bits); an explicit --config
    # runs that fused configuration
    # The step's frames alternate over --streams streams (default 2), each with its own
    # workspace: the way a caller with several threads drives the processor (one stream per
    # thread), so one frame's launches fill the rounds another's leave idle.
    streams = [stream] + [torch.cuda.Stream(dev) for _ in range(max(1, args.streams) - 1)]
    ws_bytes = {p: (nets["hdn0"].workspace_bytes(W_IN, H_IN, precision=p) if args.config is None else 0) for p in precs}
    workspaces = {p: [torch.empty(n, dtype=torch.uint8, device=dev) for _ in streams] if n else None
                  for p, n in ws_bytes.items()}
    timings = {}
    for prec in precs:
        timings[prec] = time_precision(nets, srcs, dsts, order, prec, args.steps, args.warmup, args.config, streams,
                                       workspaces[prec])

    main_prec = precs[0]
    wall, mean_pass_s = timings[main_prec]
    value = out_px_per_step * args.steps / 1e6 / wall
    achieved = FLOP_PER_INPUT_PX * W_IN * H_IN / mean_pass_s / 1e12
    kcfg = hip.configs(main_prec)[args.config or 0]
    split = ws_bytes[main_prec] > 0
    if split:
        path = "achip_pass_ws: layer-split form, 4 launches (legacy_head, conv8_pair x2, conv8_pair + deconv), 8-channel f32 maps in HBM"
        engine = "FP32 MFMA (v_mfma_f32_16x16x4_f32), FP32 accumulate"
        traffic_key = main_prec + "_split"
    else:
        path = {"fp32": "fused FP32 kernel", "fp16": "FP16-MFMA fused kernel", "bf16": "BF16-MFMA fused kernel"}[main_prec]
        engine = (("FP32 MFMA (v_mfma_f32_16x16x4_f32), FP32 accumulate" if kcfg["p"] == 0 else
                   "FP32 VALU (packed FP32 disabled; non-packed FMA caps at half the FP32 peak)")
                  if main_prec == "fp32" else f"{main_prec.upper()} MFMA, FP32 accumulate")
        traffic_key = main_prec
    peak = PEAK_FP32_TFLOPS if main_prec == "fp32" else PEAK_MFMA16_TFLOPS

    result = {
        "metric": "output Mpix/s, 1080p->4K 2x (FP32 vs FP16-MFMA)",
        "value": round(value, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp16": "f16", "bf16": "bf16"}[main_prec],
        "data": "synthetic",
        "config": {
            "workload": "ACNetLegacy 2x pass, 1920x1080 u8 luma -> 3840x2160 u8, the processor's pass per frame",
            "path": path,
            "frames_per_step": len(srcs),
            "weights": "GAN, HDN0, HDN1, HDN2, HDN3 (2 frames each)",
            "content": "seeded flat-shaded polygons + Bezier strokes + noise (sigma 0.01), resident in HBM",
            "kernel_config": None if split else kcfg,
            "workspace_bytes": ws_bytes[main_prec],
            "streams": len(streams),
        },
        "roofline": {
            "bound": "mfma",
            "engine": engine,
            "unit_of_work": "one 2x pass of a 1080p frame" + (" (its 4 launches)" if split else " (one launch)")
                            + (f"; the step's frames alternate over {len(streams)} streams, so passes overlap and the time"
                               " per pass is the step's device time over its frames" if len(streams) > 1 else ""),
            "achieved": round(achieved, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": MEASURED_TRAFFIC.get(traffic_key) if args.config is None else None,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE per pass, profiles/acnet_fused_1080p_u8.md",
            "algorithmic_bytes": 5 * W_IN * H_IN,
            "form_bytes": SPLIT_FORM_BYTES_PER_PX * W_IN * H_IN if split else 5 * W_IN * H_IN,
            "mean_pass_us": round(mean_pass_s * 1e6, 2),
            "flop_per_pass": FLOP_PER_INPUT_PX * W_IN * H_IN,
        },
    }
    if len(streams) > 1:
        # the same pass with every frame on one stream (one calling thread)
        w1, m1 = time_precision(nets, srcs, dsts, order, main_prec, args.steps, args.warmup, args.config, [stream],
                                workspaces[main_prec][:1] if workspaces[main_prec] else None)
        a1 = FLOP_PER_INPUT_PX * W_IN * H_IN / m1 / 1e12
        result["one_stream"] = {"value": round(out_px_per_step * args.steps / 1e6 / w1, 1), "unit": "Mpix/s",
                                "ms_per_step": round(w1 * 1e3 / args.steps, 4), "mean_pass_us": round(m1 * 1e6, 2),
                                "achieved_tflops": round(a1, 2), "frac": round(a1 / peak, 4)}
    if split:
        # the fused kernel on the same frames (achip_pass, one launch per frame), same streams
        wf, mf = time_precision(nets, srcs, dsts, order, main_prec, args.steps, args.warmup, None, streams)
        achf = FLOP_PER_INPUT_PX * W_IN * H_IN / mf / 1e12
        result[main_prec + "_fused"] = {
            "value": round(out_px_per_step * args.steps / 1e6 / wf, 1), "unit": "Mpix/s",
            "ms_per_step": round(wf * 1e3 / args.steps, 4), "mean_launch_us": round(mf * 1e6, 2),
            "achieved_tflops": round(achf, 2), "frac": round(achf / peak, 4),
            "traffic": MEASURED_TRAFFIC.get(main_prec), "kernel_config": kcfg,
            "form": "one fused kernel per frame, 64x32 tiles with an 8-pixel recomputed halo"}
    for prec in precs[1:]:
        w2, ml2 = timings[prec]
        v2 = out_px_per_step * args.steps / 1e6 / w2
        ach2 = FLOP_PER_INPUT_PX * W_IN * H_IN / ml2 / 1e12
        result[f"{prec}_mfma"] = {
            "value": round(v2, 1), "unit": "Mpix/s", "vs_" + main_prec: round(v2 / value, 2),
            "ms_per_step": round(w2 * 1e3 / args.steps, 4), "mean_launch_us": round(ml2 * 1e6, 2),
            "achieved_tflops": round(ach2, 2), "peak_tflops": PEAK_MFMA16_TFLOPS,
            "frac": round(ach2 / PEAK_MFMA16_TFLOPS, 4),
            "kernel_config": hip.configs(prec)[args.config or 0],
        }
    if not args.no_error:
        errs = {p: error_vs_fp64(nets, weights, p) for p in precs}
        for p, e in errs.items():
            key = "fp32_error_vs_fp64" if p == main_prec else None
            if p == main_prec:
                result["error_vs_fp64"] = e
            else:
                result[f"{p}_mfma"]["error_vs_fp64"] = e
    if not args.no_batch:
        result["batched_uneven"] = {p: batched_uneven(nets, p, streams=streams) for p in precs[:2]}
    if not args.no_4k:
        result["precision_4k_x4"] = precision_4k_x4(nets, weights, precs)
    if not args.no_stages:
        result["stages"] = stages()
    if not args.no_acnet8:
        result["acnet8"] = acnet8(frames_np)
    if not args.no_artcnn:
        result["artcnn"] = layered(frames_np, "artcnn")
    if not args.no_fsrcnnx:
        result["fsrcnnx"] = layered(frames_np, "fsrcnnx")
    if not args.no_serving:
        result["serving"] = {}
        for spec in args.serving_loads.split(";"):
            p, loads = spec.split(":")
            result["serving"][p] = serving(nets["hdn0"], p, [float(x) for x in loads.split(",")], args.serving_seconds,
                                           bursts=args.serving_bursts)
        if not args.no_acnet8:   # the reference's default model (acnet-f8b8-hdn), FP32 and FP16
            from anime4kcpp_amd import hip as _hip
            n8 = _hip.Net8("b8_hdn", 0)
            result["serving"]["acnet8_b8_hdn_fp32"] = serving(n8, "fp32", [60.0, 240.0], args.serving_seconds,
                                                              bursts=args.serving_bursts)
            result["serving"]["acnet8_b8_hdn_fp16"] = serving(n8, "fp16", [300.0, 1200.0], args.serving_seconds,
                                                              bursts=args.serving_bursts)
            n8.close()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(frames_np, weights)
        result["cpu_baseline"]["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
    for n in nets.values():
        n.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
