"""Device-level Python mirror of the HIP ACNetLegacy path (torch tensors as device
memory; all compute goes through the C ABI of include/achip.h)."""
import ctypes
import os

import numpy as np

from . import _native
from ._native import check

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
VARIANTS = ("gan", "hdn0", "hdn1", "hdn2", "hdn3")
PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2}

U8, U16, F16, F32 = 0x001, 0x002, 0x202, 0x204


def load_weights(variant):
    """(kernels[4712], biases[72]) float32 of an ACNetLegacy variant (ACNet.hpp:34-60)."""
    v = variant.lower()
    if v not in VARIANTS:
        raise ValueError(f"unknown ACNetLegacy variant {variant!r}")
    k = np.fromfile(os.path.join(WEIGHTS_DIR, f"acnet_legacy_{v}.kernels.f32"), dtype="<f4")
    b = np.fromfile(os.path.join(WEIGHTS_DIR, f"acnet_legacy_{v}.biases.f32"), dtype="<f4")
    return np.ascontiguousarray(k), np.ascontiguousarray(b)


# ACNet<8> variants (core/include/AC/Core/Model/ACNet.hpp:67-72): blocks x {normal, hdn, box, box_hdn}
VARIANTS8 = tuple(f"b{b}{s}" for b in (4, 8, 18) for s in ("", "_hdn", "_box", "_box_hdn"))


def load_weights8(variant):
    """(kernels, biases, alphas, blocks) float32 of an ACNet<8> variant (ACNet.hpp:77-82)."""
    v = variant.lower()
    if v not in VARIANTS8:
        raise ValueError(f"unknown ACNet<8> variant {variant!r}")
    blocks = int(v[1:].split("_")[0])
    arrs = [np.ascontiguousarray(np.fromfile(os.path.join(WEIGHTS_DIR, f"acnet8_{v}.{kind}.f32"), dtype="<f4"))
            for kind in ("kernels", "biases", "alphas")]
    return arrs[0], arrs[1], arrs[2], blocks


# ArtCNN variants (core/include/AC/Core/Model/ArtCNN.hpp:20-23): channels x {normal, dn, ds}, 4 blocks
VARIANTS_ART = tuple(f"f{f}{s}" for f in (16, 32) for s in ("", "_dn", "_ds"))


def load_weights_artcnn(variant):
    """(kernels, biases, channels, blocks) float32 of an ArtCNN variant."""
    v = variant.lower()
    if v not in VARIANTS_ART:
        raise ValueError(f"unknown ArtCNN variant {variant!r}")
    k = np.ascontiguousarray(np.fromfile(os.path.join(WEIGHTS_DIR, f"artcnn_{v}.kernels.f32"), dtype="<f4"))
    b = np.ascontiguousarray(np.fromfile(os.path.join(WEIGHTS_DIR, f"artcnn_{v}.biases.f32"), dtype="<f4"))
    return k, b, int(v[1:3]), 4


# FSRCNNX variants (core/include/AC/Core/Model/FSRCNNX.hpp:21-24): channels x {normal,
# distort-plus}, 4 blocks
VARIANTS_FSR = ("f8", "f8_dp", "f16", "f16_dp")


def load_weights_fsrcnnx(variant):
    """(kernels, biases, alphas, channels, blocks) float32 of an FSRCNNX variant."""
    v = variant.lower()
    if v not in VARIANTS_FSR:
        raise ValueError(f"unknown FSRCNNX variant {variant!r}")
    k, b, a = (np.ascontiguousarray(np.fromfile(os.path.join(WEIGHTS_DIR, f"fsrcnnx_{v}.{kind}.f32"), dtype="<f4"))
               for kind in ("kernels", "biases", "alphas"))
    return k, b, a, int(v[1:].split("_")[0]), 4


def elem_type(dtype):
    import torch
    table = {torch.uint8: U8, torch.uint16: U16, torch.float16: F16, torch.float32: F32}
    if dtype not in table:
        raise TypeError(f"unsupported element type {dtype}")
    return table[dtype]


def _torch_stream(stream, dev):
    """A torch stream object for `stream`: None (torch's current stream), a torch stream, or a
    raw hipStream_t handle (int)."""
    import torch
    if stream is None:
        return torch.cuda.current_stream(dev)
    if hasattr(stream, "cuda_stream"):
        return stream
    return torch.cuda.ExternalStream(int(stream), device=dev)


_TORCH_FIRST = False


def _torch_first():
    """Initialise torch's HIP context before the library's first device call.  The objects of
    this module hand device tensors between torch and the library; when the library touched
    the device first, torch's own lazy initialisation reported no GPU on these machines."""
    global _TORCH_FIRST
    if _TORCH_FIRST:
        return
    _TORCH_FIRST = True
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:   # no torch, or no device: the library reports its own errors
        pass


def device_count():
    return _native.lib().achip_device_count()


class Net:
    """One ACNetLegacy weight set resident on one HIP device."""

    def __init__(self, variant="hdn0", device=0, kernels=None, biases=None):
        if kernels is None:
            kernels, biases = load_weights(variant)
        k = np.ascontiguousarray(kernels, dtype=np.float32)
        b = np.ascontiguousarray(biases, dtype=np.float32)
        if k.size != 4712 or b.size != 72:
            raise ValueError("ACNetLegacy needs 4712 kernel and 72 bias floats")
        _torch_first()
        h = ctypes.c_void_p()
        check(_native.lib().achip_net_create(device, k.ctypes.data, b.ctypes.data, ctypes.byref(h)), "achip_net_create")
        self._h = h
        self.variant = variant
        self.device = _native.lib().achip_net_device(h)

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().achip_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def workspace_bytes(self, w, h, precision="fp32", src_dtype=None, dst_dtype=None):
        """Device bytes achip_pass_ws wants for a w x h pass (0: the fused kernel runs)."""
        import torch
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        st = elem_type(src_dtype or torch.uint8)
        return _native.lib().achip_pass_ws_bytes(self._h, prec, st, elem_type(dst_dtype) if dst_dtype else st, w, h)

    def upscale(self, src, dst=None, precision="fp32", config=None, stream=None, out_dtype=None, workspace=None):
        """One 2x pass: src (H x W, 1 channel) -> dst (2H x 2W); launches on `stream`
        (a torch.cuda.Stream or raw hipStream_t int; default: torch's current stream).
        With a device tensor `workspace` (and no config) the pass runs through achip_pass_ws:
        the faster form for the size when the workspace holds workspace_bytes()."""
        import torch
        if src.dim() != 2 or not src.is_cuda:
            raise ValueError("src must be a 2-D device tensor")
        h, w = src.shape
        if dst is None:
            dst = torch.empty((2 * h, 2 * w), dtype=out_dtype or src.dtype, device=src.device)
        if dst.dim() != 2 or tuple(dst.shape) != (2 * h, 2 * w):
            raise ValueError(f"dst must be {2 * h} x {2 * w}")
        if src.stride(1) != 1 or dst.stride(1) != 1:
            raise ValueError("rows must be contiguous")
        if stream is None:
            stream = torch.cuda.current_stream(src.device).cuda_stream
        elif hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        args = (ctypes.c_void_p(src.data_ptr()), src.stride(0) * src.element_size(), elem_type(src.dtype), w, h,
                ctypes.c_void_p(dst.data_ptr()), dst.stride(0) * dst.element_size(), elem_type(dst.dtype),
                ctypes.c_void_p(stream))
        L = _native.lib()
        if workspace is not None and config is None:
            check(L.achip_pass_ws(self._h, prec, *args[:-1], ctypes.c_void_p(workspace.data_ptr()),
                                  workspace.numel() * workspace.element_size(), args[-1]), "achip_pass_ws")
        elif config is None:
            check(L.achip_pass(self._h, prec, *args), "achip_pass")
        else:
            check(L.achip_pass_config(self._h, prec, config, *args), "achip_pass_config")
        return dst


    def upscale_batch(self, srcs, dsts=None, precision="fp32", config=0, stream=None, workspace=None, out_dtype=None):
        """2x passes of many images (same element types) in ONE persistent launch with a tile
        work-queue (achip_pass_batch).  Returns the list of outputs."""
        import torch
        if not srcs:
            return []
        dev = srcs[0].device
        if dsts is None:
            dsts = [torch.empty((2 * s.shape[0], 2 * s.shape[1]), dtype=out_dtype or s.dtype, device=dev) for s in srcs]
        st = elem_type(srcs[0].dtype)
        dt = elem_type(dsts[0].dtype)
        arr = (_native.ACHipImage * len(srcs))()
        for i, (s, d) in enumerate(zip(srcs, dsts)):
            if elem_type(s.dtype) != st or elem_type(d.dtype) != dt:
                raise TypeError("a batch has one source and one destination element type")
            if s.stride(1) != 1 or d.stride(1) != 1 or tuple(d.shape) != (2 * s.shape[0], 2 * s.shape[1]):
                raise ValueError(f"image {i}: bad shape or layout")
            arr[i] = _native.ACHipImage(s.data_ptr(), s.stride(0) * s.element_size(), d.data_ptr(),
                                        d.stride(0) * d.element_size(), s.shape[1], s.shape[0])
        L = _native.lib()
        need = L.achip_batch_workspace_bytes(len(srcs))
        if workspace is None:
            workspace = torch.empty(need, dtype=torch.uint8, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        elif hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check(L.achip_pass_batch(self._h, prec, config, st, dt, arr, len(srcs), ctypes.c_void_p(workspace.data_ptr()),
                                 workspace.numel(), ctypes.c_void_p(stream)), "achip_pass_batch")
        self._keep = (workspace, arr)    # alive until the next call on this net
        return dsts

    def upscale_batch_layers(self, srcs, dsts=None, stream=None, workspace=None, out_dtype=None):
        """FP32 2x passes of many images (same element types) in the layer-split form, four
        launches over the tiles of all of them (achip_pass_batch_layers); outputs are the
        fused kernel's bits.  workspace: a device tensor of batch_layers_bytes(srcs) bytes, or
        None to allocate one.  Returns the list of outputs."""
        import torch
        if not srcs:
            return []
        dev = srcs[0].device
        if dsts is None:
            dsts = [torch.empty((2 * s.shape[0], 2 * s.shape[1]), dtype=out_dtype or s.dtype, device=dev) for s in srcs]
        st = elem_type(srcs[0].dtype)
        dt = elem_type(dsts[0].dtype)
        arr = (_native.ACHipImage * len(srcs))()
        for i, (s, d) in enumerate(zip(srcs, dsts)):
            if elem_type(s.dtype) != st or elem_type(d.dtype) != dt:
                raise TypeError("a batch has one source and one destination element type")
            if s.stride(1) != 1 or d.stride(1) != 1 or tuple(d.shape) != (2 * s.shape[0], 2 * s.shape[1]):
                raise ValueError(f"image {i}: bad shape or layout")
            arr[i] = _native.ACHipImage(s.data_ptr(), s.stride(0) * s.element_size(), d.data_ptr(),
                                        d.stride(0) * d.element_size(), s.shape[1], s.shape[0])
        L = _native.lib()
        need = L.achip_pass_batch_layers_bytes(arr, len(srcs))
        if workspace is None:
            workspace = torch.empty(need, dtype=torch.uint8, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        elif hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        check(L.achip_pass_batch_layers(self._h, st, dt, arr, len(srcs), ctypes.c_void_p(workspace.data_ptr()),
                                        workspace.numel(), ctypes.c_void_p(stream)), "achip_pass_batch_layers")
        self._keep = (workspace, arr)
        return dsts

    def batch_layers_bytes(self, srcs):
        """Workspace bytes of upscale_batch_layers for these images."""
        arr = (_native.ACHipImage * len(srcs))()
        for i, s in enumerate(srcs):
            arr[i] = _native.ACHipImage(0, 0, 0, 0, s.shape[1], s.shape[0])
        return _native.lib().achip_pass_batch_layers_bytes(arr, len(srcs))


    def autotune(self, w, h, precision="fp32", src_dtype=None, dst_dtype=None):
        """Time every kernel configuration of `precision` on w x h images (achip_autotune) and
        remember the fastest for that size class; returns (best, [ms per configuration])."""
        import torch
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        st = elem_type(src_dtype or torch.uint8)
        dt = elem_type(dst_dtype or src_dtype or torch.uint8)
        L = _native.lib()
        n = L.achip_config_count(prec)
        best = ctypes.c_int(-1)
        ms = (ctypes.c_float * max(n, 1))()
        check(L.achip_autotune(self._h, prec, st, dt, w, h, ctypes.byref(best), ms), "achip_autotune")
        return best.value, [ms[i] for i in range(n)]


class Net8:
    """One ACNet<8> weight set resident on one HIP device (achip_net8_create / achip_pass8)."""

    def __init__(self, variant="b8_hdn", device=0, weights=None):
        """weights: (kernels, biases, alphas, blocks) in the reference layout instead of a
        shipped variant's."""
        k, b, a, blocks = load_weights8(variant) if weights is None else weights
        k, b, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (k, b, a))
        _torch_first()
        h = ctypes.c_void_p()
        check(_native.lib().achip_net8_create(device, blocks, k.ctypes.data, b.ctypes.data, a.ctypes.data,
                                              ctypes.byref(h)), "achip_net8_create")
        self._h = h
        self.variant = variant
        self.blocks = blocks
        self.device = _native.lib().achip_net_device(h)
        self._ws = {}

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().achip_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _need(self, w, h):
        return _native.lib().achip_pass8_workspace_bytes(w, h)

    def workspace(self, w, h, dev, stream=None):
        """The feature-map workspace for w x h passes launched on `stream` (a torch stream;
        default: torch's current stream).  One per stream: passes on different streams never
        share feature maps, and each workspace is allocated on its own stream, so torch's
        allocator reuses a replaced one only after the work queued before it on that stream."""
        import torch
        ts = stream if stream is not None else torch.cuda.current_stream(dev)
        need = self._need(w, h)
        ws = self._ws.get(ts.cuda_stream)
        if ws is None or ws.numel() < need:
            with torch.cuda.stream(ts):
                ws = torch.empty(need, dtype=torch.uint8, device=dev)
            self._ws[ts.cuda_stream] = ws
        return ws

    def upscale(self, src, dst=None, stream=None, out_dtype=None, precision="fp32"):
        """One ACNet<8> 2x pass: src (H x W) -> dst (2H x 2W) on `stream` (default: torch's
        current stream) in `precision` (fp32, or the fp16/bf16 MFMA forms); the net keeps one
        feature-map workspace per stream between calls."""
        import torch
        if src.dim() != 2 or not src.is_cuda or src.stride(1) != 1:
            raise ValueError("src must be a 2-D device tensor with contiguous rows")
        h, w = src.shape
        if dst is None:
            dst = torch.empty((2 * h, 2 * w), dtype=out_dtype or src.dtype, device=src.device)
        ts = _torch_stream(stream, src.device)
        ws = self.workspace(w, h, src.device, ts)
        stream = ts.cuda_stream
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check(_native.lib().achip_pass8_precision(self._h, prec, ctypes.c_void_p(src.data_ptr()),
                                                  src.stride(0) * src.element_size(), elem_type(src.dtype), w, h,
                                                  ctypes.c_void_p(dst.data_ptr()), dst.stride(0) * dst.element_size(),
                                                  elem_type(dst.dtype), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                                  ctypes.c_void_p(stream)), "achip_pass8_precision")
        return dst


class NetArt(Net8):
    """One ArtCNN<16|32> weight set resident on one HIP device (achip_net_artcnn_create /
    achip_pass_net); same calls as Net8."""

    def __init__(self, variant="f16", device=0):
        k, b, f, blocks = load_weights_artcnn(variant)
        _torch_first()
        h = ctypes.c_void_p()
        check(_native.lib().achip_net_artcnn_create(device, f, blocks, k.ctypes.data, b.ctypes.data, ctypes.byref(h)),
              "achip_net_artcnn_create")
        self._h = h
        self.variant = variant
        self.channels = f
        self.blocks = blocks
        self.device = _native.lib().achip_net_device(h)
        self._ws = {}

    def _need(self, w, h):
        return _native.lib().achip_net_workspace_bytes(self._h, w, h)

    def upscale(self, src, dst=None, stream=None, out_dtype=None):
        """One ArtCNN 2x pass: src (H x W) -> dst (2H x 2W) on `stream` (default: torch's
        current stream); the net keeps one feature-map workspace per stream between calls."""
        import torch
        if src.dim() != 2 or not src.is_cuda or src.stride(1) != 1:
            raise ValueError("src must be a 2-D device tensor with contiguous rows")
        h, w = src.shape
        if dst is None:
            dst = torch.empty((2 * h, 2 * w), dtype=out_dtype or src.dtype, device=src.device)
        ts = _torch_stream(stream, src.device)
        ws = self.workspace(w, h, src.device, ts)
        stream = ts.cuda_stream
        check(_native.lib().achip_pass_net(self._h, ctypes.c_void_p(src.data_ptr()), src.stride(0) * src.element_size(),
                                           elem_type(src.dtype), w, h, ctypes.c_void_p(dst.data_ptr()),
                                           dst.stride(0) * dst.element_size(), elem_type(dst.dtype),
                                           ctypes.c_void_p(ws.data_ptr()), ws.numel(), ctypes.c_void_p(stream)),
              "achip_pass_net")
        return dst


class NetFsr(NetArt):
    """One FSRCNNX<8|16> weight set resident on one HIP device (achip_net_fsrcnnx_create /
    achip_pass_net); same calls as Net8."""

    def __init__(self, variant="f8", device=0):
        k, b, a, f, blocks = load_weights_fsrcnnx(variant)
        _torch_first()
        h = ctypes.c_void_p()
        check(_native.lib().achip_net_fsrcnnx_create(device, f, blocks, k.ctypes.data, b.ctypes.data, a.ctypes.data,
                                                     ctypes.byref(h)), "achip_net_fsrcnnx_create")
        self._h = h
        self.variant = variant
        self.channels = f
        self.blocks = blocks
        self.device = _native.lib().achip_net_device(h)
        self._ws = {}


def configs(precision="fp32"):
    L = _native.lib()
    prec = PRECISIONS[precision]
    out = []
    for i in range(L.achip_config_count(prec)):
        vals = [ctypes.c_int() for _ in range(5)]
        check(L.achip_config_info(prec, i, *[ctypes.byref(v) for v in vals]), "achip_config_info")
        out.append(dict(zip(("tw", "th", "nt", "p", "lds_bytes"), [v.value for v in vals])))   # p = 0: MFMA kernel
    return out


REQ_LUMA, REQ_YUV420, REQ_RGB, REQ_YUV = 0, 1, 2, 3
CHROMA_BILINEAR, CHROMA_CATMULL_ROM = 0, 1


class FrameFormat:
    """ACHipFrameFormat: plane layout of a REQ_YUV frame (the reference video module's formats)."""

    def __init__(self, planes, chroma_shift_x=1, chroma_shift_y=1, element_type=0x001, shift=0,
                 chroma_filter=CHROMA_BILINEAR):
        self.planes, self.chroma_shift_x, self.chroma_shift_y = planes, chroma_shift_x, chroma_shift_y
        self.element_type, self.shift, self.chroma_filter = element_type, shift, chroma_filter

    @property
    def dtype(self):
        return np.uint16 if self.element_type == 0x002 else np.uint8


# common layouts: (planes, chroma x shift, chroma y shift, element type, luma LSB shift)
FORMATS = {
    "gray8": (1, 0, 0, 0x001, 0), "gray10le": (1, 0, 0, 0x002, 6), "gray16": (1, 0, 0, 0x002, 0),
    "yuv420p": (3, 1, 1, 0x001, 0), "yuv422p": (3, 1, 0, 0x001, 0), "yuv444p": (3, 0, 0, 0x001, 0),
    "yuv420p10le": (3, 1, 1, 0x002, 6), "yuv420p12le": (3, 1, 1, 0x002, 4), "yuv444p16": (3, 0, 0, 0x002, 0),
    "nv12": (2, 1, 1, 0x001, 0), "nv16": (2, 1, 0, 0x001, 0), "nv24": (2, 0, 0, 0x001, 0),
    "p010": (2, 1, 1, 0x002, 0), "p016": (2, 1, 1, 0x002, 0), "p210": (2, 1, 0, 0x002, 0),
}


def frame_format(name, chroma_filter=CHROMA_BILINEAR):
    return FrameFormat(*FORMATS[name], chroma_filter=chroma_filter)


def _addr(a):
    return a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data


def _row_bytes(a):
    return a.stride(0) * a.element_size() if hasattr(a, "data_ptr") else a.strides[0]


class Server:
    """Serving loop on one Net (achip_server_*): requests of mixed kind, size and factor,
    admitted by an HBM budget, one-off passes batched step by step into persistent
    launches, video-stream frames replayed from captured hipGraphs.  Planes are u8 numpy
    arrays or torch tensors (host, page-locked preferred, or device)."""

    def __init__(self, net, precision="fp16", hbm_budget=0, step_tiles=0, max_workgroups=0):
        L = _native.lib()
        cfg = _native.ACHipServerConfig(PRECISIONS[precision], int(hbm_budget), int(step_tiles), int(max_workgroups))
        h = ctypes.c_void_p()
        check(L.achip_server_create(net.handle, ctypes.byref(cfg), ctypes.byref(h)), "achip_server_create")
        self._h = h
        self.net = net
        self.precision = precision
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().achip_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, kind, planes, factor=2, stream=-1, dst=None, fmt=None):
        """Queue one request; planes: [Y] (LUMA), [RGB H x W x 3] (RGB), [Y, U, V] (YUV420) or
        as `fmt` (a FrameFormat, kind REQ_YUV) lays them out; dst: matching output planes, or
        None to keep the output in the server.  Returns the ticket."""
        r = _native.ACHipRequest()
        r.kind, r.factor, r.stream = kind, factor, stream
        r.height, r.width = int(planes[0].shape[0]), int(planes[0].shape[1])
        for i, p in enumerate(planes):
            r.src[i] = _addr(p)
            r.src_stride[i] = _row_bytes(p)
        for i, p in enumerate(dst or []):
            r.dst[i] = _addr(p)
            r.dst_stride[i] = _row_bytes(p)
        if fmt is not None:
            r.format = _native.ACHipFrameFormat(fmt.planes, fmt.chroma_shift_x, fmt.chroma_shift_y, fmt.element_type,
                                                fmt.shift, fmt.chroma_filter)
        t = ctypes.c_ulonglong()
        check(_native.lib().achip_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)), "achip_server_submit")
        self._keep[t.value] = (planes, dst)
        return t.value

    def wait(self, ticket, timeout_ms=-1.0):
        """Block until done; returns the submit -> done latency in ms."""
        lat = ctypes.c_double()
        rc = _native.lib().achip_server_wait(self._h, ticket, float(timeout_ms), ctypes.byref(lat))
        if rc == 0 and self._keep.get(ticket, (None, None))[1] is not None:
            self._keep.pop(ticket, None)
        check(rc, "achip_server_wait")
        return lat.value

    def output(self, ticket, plane=0):
        ptr, stride = ctypes.c_void_p(), ctypes.c_long()
        check(_native.lib().achip_server_output(self._h, ticket, plane, ctypes.byref(ptr), ctypes.byref(stride)),
              "achip_server_output")
        return ptr.value, stride.value

    def fetch(self, ticket, plane, shape, dtype=np.uint8):
        """Copy a kept output plane (`shape`, `dtype`) from device memory into a new numpy array."""
        ptr, stride = self.output(ticket, plane)
        out = np.zeros(shape, dtype)
        row = out.strides[0]
        hip_rt = ctypes.CDLL("libamdhip64.so")
        hip_rt.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
        check(hip_rt.hipMemcpy2D(out.ctypes.data, row, ptr, stride, row, out.shape[0], 2), "hipMemcpy2D")   # D2H
        return out

    def release(self, ticket):
        check(_native.lib().achip_server_release(self._h, ticket), "achip_server_release")
        self._keep.pop(ticket, None)

    def stats(self, reset=False):
        s = _native.ACHipServerStats()
        check(_native.lib().achip_server_stats(self._h, ctypes.byref(s), 1 if reset else 0), "achip_server_stats")
        return {name: getattr(s, name) for name, _ in s._fields_}


# ---- device-memory colour and resize stages (achip_rgb2yuv / yuv2rgb / resize / resize_yuv2rgb) ----

def _stream_of(t, stream):
    import torch
    if stream is None:
        return torch.cuda.current_stream(t.device).cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream


def _check_hw(t, c):
    if not t.is_cuda or t.stride(-1) != 1 or (t.dim() == 3 and t.stride(1) != t.shape[2]):
        raise ValueError("planes must be CUDA tensors with packed pixels")
    if (t.dim() == 3 and t.shape[2] != c) or (t.dim() == 2 and c != 1):
        raise ValueError(f"expected {c} channels")


def rgb2yuv(rgb, y=None, uv=None, stream=None):
    """RGB(A) H x W x C tensor -> (Y H x W, UV(A) H x W x (C-1)) on the device."""
    import torch
    h, w, c = rgb.shape
    y = torch.empty((h, w), dtype=rgb.dtype, device=rgb.device) if y is None else y
    uv = torch.empty((h, w, c - 1), dtype=rgb.dtype, device=rgb.device) if uv is None else uv
    _check_hw(rgb, c), _check_hw(y, 1), _check_hw(uv, c - 1)
    check(_native.lib().achip_rgb2yuv(elem_type(rgb.dtype), c, rgb.data_ptr(), _row_bytes(rgb), w, h, y.data_ptr(),
                                      _row_bytes(y), uv.data_ptr(), _row_bytes(uv), _stream_of(rgb, stream)), "achip_rgb2yuv")
    return y, uv


def yuv2rgb(y, uv, out=None, stream=None):
    import torch
    h, w = y.shape
    c = uv.shape[2] + 1
    out = torch.empty((h, w, c), dtype=y.dtype, device=y.device) if out is None else out
    _check_hw(y, 1), _check_hw(uv, c - 1), _check_hw(out, c)
    check(_native.lib().achip_yuv2rgb(elem_type(y.dtype), c, y.data_ptr(), _row_bytes(y), uv.data_ptr(), _row_bytes(uv),
                                      w, h, out.data_ptr(), _row_bytes(out), _stream_of(y, stream)), "achip_yuv2rgb")
    return out


def resize(src, dsize, mode=1, out=None, stream=None):
    """Resize an H x W (x C) device tensor to dsize = (width, height); mode: ResizeModes."""
    import torch
    h, w = src.shape[:2]
    c = 1 if src.dim() == 2 else src.shape[2]
    dw, dh = dsize
    if out is None:
        out = torch.empty((dh, dw) + ((c,) if src.dim() == 3 else ()), dtype=src.dtype, device=src.device)
    _check_hw(src, c), _check_hw(out, c)
    check(_native.lib().achip_resize(elem_type(src.dtype), c, mode, src.data_ptr(), _row_bytes(src), w, h, out.data_ptr(),
                                     _row_bytes(out), dw, dh, _stream_of(src, stream)), "achip_resize")
    return out


def resize_yuv2rgb(y, uv, mode=1, out=None, stream=None):
    """Chroma UV(A) (any size) resized to Y's size and merged with Y into RGB(A) in one pass."""
    import torch
    h, w = y.shape
    uh, uw, cc = uv.shape
    out = torch.empty((h, w, cc + 1), dtype=y.dtype, device=y.device) if out is None else out
    _check_hw(y, 1), _check_hw(uv, cc), _check_hw(out, cc + 1)
    check(_native.lib().achip_resize_yuv2rgb(elem_type(y.dtype), cc + 1, mode, y.data_ptr(), _row_bytes(y), uv.data_ptr(),
                                             _row_bytes(uv), uw, uh, w, h, out.data_ptr(), _row_bytes(out),
                                             _stream_of(y, stream)), "achip_resize_yuv2rgb")
    return out
