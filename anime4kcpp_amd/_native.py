"""ctypes binding of the native library ``lib/libac_hip.so``.

Every entry point used here is declared in ``include/achip.h`` (the thin device
C ABI) or ``include/AC/Core/*.h`` (the reference-compatible C binding).  There is
no fallback: if the library is missing or fails to load, ``lib()`` raises.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libac_hip.so")

_lib = None

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_long = ctypes.c_long
c_double = ctypes.c_double
c_char_p = ctypes.c_char_p


class ACImage(ctypes.Structure):
    """binding/c/include/AC/Core/Image.h:52-61"""
    _fields_ = [("width", c_int), ("height", c_int), ("channels", c_int), ("stride", c_int),
                ("element_type", c_int), ("ptr", c_void_p), ("hptr", c_void_p)]


class ACHipImage(ctypes.Structure):
    """include/achip.h ACHipImage"""
    _fields_ = [("src", c_void_p), ("src_stride", c_long), ("dst", c_void_p), ("dst_stride", c_long),
                ("w", c_int), ("h", c_int)]


class ACHipFrameFormat(ctypes.Structure):
    """include/achip.h ACHipFrameFormat"""
    _fields_ = [("planes", c_int), ("chroma_shift_x", c_int), ("chroma_shift_y", c_int), ("element_type", c_int),
                ("shift", c_int), ("chroma_filter", c_int)]


class ACHipRequest(ctypes.Structure):
    """include/achip.h ACHipRequest"""
    _fields_ = [("kind", c_int), ("width", c_int), ("height", c_int), ("factor", c_int), ("stream", c_int),
                ("src", c_void_p * 3), ("src_stride", c_long * 3), ("dst", c_void_p * 3), ("dst_stride", c_long * 3),
                ("format", ACHipFrameFormat)]


class ACHipServerConfig(ctypes.Structure):
    """include/achip.h ACHipServerConfig"""
    _fields_ = [("precision", c_int), ("hbm_budget", ctypes.c_size_t), ("step_tiles", c_int), ("max_workgroups", c_int)]


class ACHipServerStats(ctypes.Structure):
    """include/achip.h ACHipServerStats"""
    _fields_ = [("completed_frames", c_long), ("completed_oneoffs", c_long), ("failed", c_long),
                ("frame_p50_ms", ctypes.c_double), ("frame_p99_ms", ctypes.c_double),
                ("oneoff_p50_ms", ctypes.c_double), ("oneoff_p99_ms", ctypes.c_double),
                ("output_mpix", ctypes.c_double), ("busy_s", ctypes.c_double),
                ("steps", c_long), ("graph_replays", c_long)]


class ACProcessor(ctypes.Structure):
    """binding/c/include/AC/Core/Processor.h:15-21"""
    _fields_ = [("device", c_int), ("type", c_char_p), ("model", c_char_p), ("hptr", c_void_p)]


# name -> (restype, argtypes); the C-ABI surface this package relies on
SIGNATURES = {
    "achip_net_artcnn_create": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
    "achip_net_fsrcnnx_create": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
    "achip_net_workspace_bytes": (ctypes.c_size_t, [c_void_p, c_int, c_int]),
    "achip_pass_net": (c_int, [c_void_p, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int, c_void_p,
                               ctypes.c_size_t, c_void_p]),
    "achip_net8_create": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
    "achip_net_blocks": (c_int, [c_void_p]),
    "achip_pass8_workspace_bytes": (ctypes.c_size_t, [c_int, c_int]),
    "achip_pass8": (c_int, [c_void_p, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int, c_void_p,
                            ctypes.c_size_t, c_void_p]),
    "achip_pass8_precision": (c_int, [c_void_p, c_int, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int,
                                      c_void_p, ctypes.c_size_t, c_void_p]),
    "achip_rgb2yuv": (c_int, [c_int, c_int, c_void_p, c_long, c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_void_p]),
    "achip_yuv2rgb": (c_int, [c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_int, c_int, c_void_p, c_long, c_void_p]),
    "achip_resize": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_int, c_int, c_void_p, c_long, c_int, c_int, c_void_p]),
    "achip_resize_yuv2rgb": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_int, c_int, c_int, c_int,
                                     c_void_p, c_long, c_void_p]),
    # include/achip.h
    "achip_device_count": (c_int, []),
    "achip_net_create": (c_int, [c_int, c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
    "achip_net_destroy": (None, [c_void_p]),
    "achip_net_device": (c_int, [c_void_p]),
    "achip_pass": (c_int, [c_void_p, c_int, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int, c_void_p]),
    "achip_pass_config": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int, c_void_p]),
    "achip_config_count": (c_int, [c_int]),
    "achip_config_info": (c_int, [c_int, c_int] + [ctypes.POINTER(c_int)] * 5),
    "achip_last_error": (c_char_p, []),
    "achip_batch_workspace_bytes": (ctypes.c_size_t, [c_int]),
    "achip_server_create": (c_int, [c_void_p, ctypes.POINTER(ACHipServerConfig), ctypes.POINTER(c_void_p)]),
    "achip_server_destroy": (None, [c_void_p]),
    "achip_server_submit": (c_int, [c_void_p, ctypes.POINTER(ACHipRequest), ctypes.POINTER(ctypes.c_ulonglong)]),
    "achip_server_wait": (c_int, [c_void_p, ctypes.c_ulonglong, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    "achip_server_output": (c_int, [c_void_p, ctypes.c_ulonglong, c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_long)]),
    "achip_server_release": (c_int, [c_void_p, ctypes.c_ulonglong]),
    "achip_server_stats": (c_int, [c_void_p, ctypes.POINTER(ACHipServerStats), c_int]),
    "achip_autotune": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(ctypes.c_float)]),
    "achip_pass_batch": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p, ctypes.c_size_t, c_void_p]),
    "achip_pass_layers": (c_int, [c_void_p, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int, c_void_p,
                                  ctypes.c_size_t, c_void_p]),
    "achip_pass_batch_layers_bytes": (ctypes.c_size_t, [c_void_p, c_int]),
    "achip_pass_batch_layers": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, ctypes.c_size_t, c_void_p]),
    "achip_pass_ws_bytes": (ctypes.c_size_t, [c_void_p, c_int, c_int, c_int, c_int, c_int]),
    "achip_pass_ws": (c_int, [c_void_p, c_int, c_void_p, c_long, c_int, c_int, c_int, c_void_p, c_long, c_int, c_void_p,
                              ctypes.c_size_t, c_void_p]),
    "achip_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "achip_pass_batch_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p]),
    # include/AC/Core/Image.h
    "ac_image_alloc": (ctypes.POINTER(ACImage), []),
    "ac_image_free": (None, [ctypes.POINTER(ctypes.POINTER(ACImage))]),
    "ac_image_ref": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage)]),
    "ac_image_unref": (None, [ctypes.POINTER(ACImage)]),
    "ac_image_create": (c_int, [ctypes.POINTER(ACImage)]),
    "ac_image_map": (c_int, [ctypes.POINTER(ACImage)]),
    "ac_image_from": (c_int, [ctypes.POINTER(ACImage), c_void_p]),
    "ac_image_view": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage), c_int, c_int, c_int, c_int]),
    "ac_image_clone": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage)]),
    "ac_image_to": (c_int, [ctypes.POINTER(ACImage), c_void_p, c_int]),
    "ac_resize": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage), ctypes.c_double, ctypes.c_double, c_int]),
    "ac_rgb2yuv": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage)]),
    "ac_rgba2yuva": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage)]),
    "ac_yuv2rgb": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage)]),
    "ac_yuva2rgba": (c_int, [ctypes.POINTER(ACImage), ctypes.POINTER(ACImage)]),
    # include/AC/Core/Processor.h
    "ac_processor_alloc": (ctypes.POINTER(ACProcessor), []),
    "ac_processor_free": (None, [ctypes.POINTER(ctypes.POINTER(ACProcessor))]),
    "ac_processor_ref": (c_int, [ctypes.POINTER(ACProcessor), ctypes.POINTER(ACProcessor)]),
    "ac_processor_unref": (None, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_create": (c_int, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_process": (c_int, [ctypes.POINTER(ACProcessor), ctypes.POINTER(ACImage), ctypes.POINTER(ACImage), c_double]),
    "ac_processor_ok": (c_int, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_error": (c_char_p, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_name": (c_char_p, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_type": (c_int, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_type_name": (c_char_p, [ctypes.POINTER(ACProcessor)]),
    "ac_processor_info": (c_char_p, [c_int]),
    "ac_processor_list_info": (c_char_p, []),
}


def lib():
    """Load the native library once; raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library {LIB_PATH} is not built; run anime4kcpp_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().achip_last_error().decode(errors="replace")


class AchipError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where} failed ({code}): {last_error()}")
        self.code = code


def check(code, where):
    if code != 0:
        raise AchipError(code, where)
    return code


def check_ac(code, where):
    """C API error codes (include/AC/Error.h): 0 success, negative errno otherwise."""
    if code != 0:
        raise RuntimeError(f"{where} failed ({code})")
    return code
