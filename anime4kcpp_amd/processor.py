"""Python view of the reference-compatible C API (include/AC/Core/{Image,Processor}.h):
``Processor("hip", 0, "acnet-legacy-hdn0").process(img, 2.0)`` as
core/include/AC/Core/Processor.hpp:25-29 and binding/c/src/Binding.cpp:190-203 define it."""
import ctypes

import numpy as np

from . import _native

ELEM = {np.dtype(np.uint8): 0x001, np.dtype(np.uint16): 0x002, np.dtype(np.float16): 0x202, np.dtype(np.float32): 0x204}
DTYPE = {v: k for k, v in ELEM.items()}


class Image:
    """An ACImage owned by the library (ac_image_from copies) or mapping a numpy array."""

    def __init__(self, array=None, mapped=False, _empty=False):
        L = _native.lib()
        self._p = L.ac_image_alloc()
        if not self._p:
            raise MemoryError("ac_image_alloc")
        self._keep = None
        if array is None:
            # an empty image still needs a handle: ac_image_create on a zeroed struct
            _native.check_ac(L.ac_image_create(self._p), "ac_image_create")
            return
        a = np.asarray(array)
        if a.ndim == 2:
            h, w, c = a.shape[0], a.shape[1], 1
        elif a.ndim == 3:
            h, w, c = a.shape
        else:
            raise ValueError("image must be H x W or H x W x C")
        if a.strides[-1] != a.itemsize or (a.ndim == 3 and a.strides[1] != c * a.itemsize):
            a = np.ascontiguousarray(a)
        im = self._p.contents
        im.width, im.height, im.channels = w, h, c
        im.element_type = ELEM[a.dtype]
        im.stride = a.strides[0]
        if mapped:
            self._keep = a
            im.ptr = a.ctypes.data
            _native.check_ac(L.ac_image_map(self._p), "ac_image_map")
        else:
            _native.check_ac(L.ac_image_from(self._p, ctypes.c_void_p(a.ctypes.data)), "ac_image_from")

    @property
    def c(self):
        return self._p.contents

    def empty(self):
        return not self.c.ptr

    def numpy(self):
        """Copy out (ac_image_to) as an H x W (x C) array."""
        im = self.c
        if not im.ptr:
            return None
        dt = DTYPE[im.element_type]
        shape = (im.height, im.width) if im.channels == 1 else (im.height, im.width, im.channels)
        out = np.empty(shape, dt)
        _native.check_ac(_native.lib().ac_image_to(self._p, ctypes.c_void_p(out.ctypes.data), out.strides[0]), "ac_image_to")
        return out

    def __del__(self):
        try:
            if self._p:
                _native.lib().ac_image_free(ctypes.byref(self._p))
        except Exception:
            pass


class Processor:
    def __init__(self, type="hip", device=0, model="acnet-legacy-hdn0"):
        L = _native.lib()
        self._p = L.ac_processor_alloc()
        self._type = type.encode() if type is not None else None
        self._model = model.encode() if model is not None else None
        pc = self._p.contents
        pc.device = device
        pc.type = self._type
        pc.model = self._model
        self.create_rc = L.ac_processor_create(self._p)

    def ok(self):
        return _native.lib().ac_processor_ok(self._p) == 0

    def error(self):
        e = _native.lib().ac_processor_error(self._p)
        return e.decode() if e else None

    def name(self):
        return _native.lib().ac_processor_name(self._p).decode()

    def type(self):
        return _native.lib().ac_processor_type(self._p)

    def type_name(self):
        return _native.lib().ac_processor_type_name(self._p).decode()

    def process(self, src, factor=2.0, dst=None):
        """src: numpy array or Image.  dst: None (allocated), or an Image/array mapped as a
        preallocated destination.  Returns (rc, output array)."""
        s = src if isinstance(src, Image) else Image(src)
        d = dst if isinstance(dst, Image) else (Image(dst, mapped=True) if dst is not None else Image())
        rc = _native.lib().ac_processor_process(self._p, s._p, d._p, float(factor))
        return rc, (d.numpy() if not d.empty() else None)

    def __del__(self):
        try:
            if self._p:
                _native.lib().ac_processor_free(ctypes.byref(self._p))
        except Exception:
            pass


def info(processor_type=3):
    return _native.lib().ac_processor_info(processor_type).decode()


def list_info():
    return _native.lib().ac_processor_list_info().decode()


# ResizeModes (core/include/AC/Core/Image.hpp:272-291), as pyac's core.ResizeModes exports them
(RESIZE_POINT, RESIZE_CATMULL_ROM, RESIZE_MITCHELL_NETRAVALI, RESIZE_BICUBIC_0_60, RESIZE_BICUBIC_0_75,
 RESIZE_BICUBIC_0_100, RESIZE_BICUBIC_20_50, RESIZE_SOFTCUBIC50, RESIZE_SOFTCUBIC75, RESIZE_SOFTCUBIC100,
 RESIZE_LANCZOS2, RESIZE_LANCZOS3, RESIZE_LANCZOS4, RESIZE_SPLINE16, RESIZE_SPLINE36, RESIZE_SPLINE64,
 RESIZE_BILINEAR) = range(17)


def resize(src, dsize=None, fx=0.0, fy=0.0, mode=RESIZE_BILINEAR):
    """pyac's core.resize (binding/python/src/Binding.cpp:98-133) through ac_resize: the output
    is dsize = (width, height) when given, else (int(w * fx), int(h * fy)); same dtype and
    channels as src.  Runs on the current HIP device."""
    a = np.asarray(src)
    if a.ndim not in (2, 3):
        raise ValueError("Incompatible dimension: expected 2 or 3.")
    if a.dtype not in ELEM:
        raise TypeError("Incompatible type: expected uint8, uint16, float16 or float32.")
    h, w = a.shape[:2]
    c = 1 if a.ndim == 2 else a.shape[2]
    dw, dh = int(w * fx), int(h * fy)
    if dsize is not None:
        if len(dsize) != 2:
            raise ValueError("dsize shouble be (width, height)")
        dw, dh = int(dsize[0]), int(dsize[1])
    out = np.empty((dh, dw) if c == 1 else (dh, dw, c), a.dtype)
    if dw == w and dh == h:
        out[...] = a   # ac_resize at the same size hands back src itself
        return out
    si = Image(a, mapped=True)
    di = Image(out, mapped=True)
    _native.check_ac(_native.lib().ac_resize(si._p, di._p, 0.0, 0.0, mode), "ac_resize")
    if di.c.ptr != out.ctypes.data:
        raise RuntimeError("ac_resize did not write the mapped output")
    return out
