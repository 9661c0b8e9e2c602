"""The C-ABI library loads and exports every function its public headers declare
(include/achip.h, include/AC/Core/Image.h, include/AC/Core/Processor.h).  No compute."""
import ctypes
import os
import re

import pytest

from anime4kcpp_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["include/achip.h", "include/AC/Core/Image.h", "include/AC/Core/Processor.h"]


def declared(path):
    text = open(os.path.join(ROOT, path)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"(?:ACHIP_API|AC_C_API)\s+[\w\s\*]+?\b(\w+)\s*\(", text)


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_symbol_is_exported(header):
    names = declared(header)
    assert len(names) >= 9, names
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_headers():
    names = set()
    for h in HEADERS:
        names.update(declared(h))
    assert names <= set(_native.SIGNATURES), names - set(_native.SIGNATURES)


def test_host_side_errors():
    L = _native.lib()
    h = ctypes.c_void_p()
    assert L.achip_net_create(0, None, None, ctypes.byref(h)) == -22
    assert "NULL weights" in _native.last_error()
    assert L.achip_pass(None, 0, None, 0, 1, 1, 1, None, 0, 1, None) == -22
    assert L.achip_config_count(0) >= 1
    assert L.achip_config_info(0, 999, None, None, None, None, None) == -22
    assert L.achip_pass_ws_bytes(None, 0, 1, 1, 1920, 1080) == 0
    assert L.achip_pass_ws(None, 0, None, 0, 1, 1, 1, None, 0, 1, None, 0, None) == -22
    assert L.achip_pass_batch_layers_bytes(None, 0) == 0
    assert L.achip_pass_batch_layers(None, 1, 1, None, 0, None, 0, None) == -22
    if L.achip_device_count() == 0:
        import numpy as np
        k = np.zeros(4712, np.float32)
        b = np.zeros(72, np.float32)
        assert L.achip_net_create(0, k.ctypes.data, b.ctypes.data, ctypes.byref(h)) == -19


def test_c_api_processor_without_device():
    from anime4kcpp_amd import processor as P
    if _native.lib().achip_device_count() > 0:
        pytest.skip("checks the no-device path")
    p = P.Processor("hip", 0, "acnet-legacy-hdn1")
    assert p.create_rc == -256 and not p.ok() and "no HIP device" in p.error()
    assert p.type() == 3 and p.type_name() == "HIP"
    for t in ("cpu", "opencl", "cuda", "nonsense"):
        q = P.Processor(t, 0, "acnet-legacy-hdn0")
        assert q.create_rc == -256 and "HIP processor only" in q.error()
    assert P.info(0) == "unsupported processor" and P.info(3).startswith("HIP:")


def test_c_api_image_semantics():
    import numpy as np
    from anime4kcpp_amd import processor as P
    a = np.arange(5 * 7, dtype=np.uint8).reshape(5, 7)
    im = P.Image(a)
    assert im.c.stride == 8 and np.array_equal(im.numpy(), a)      # ac_image_from aligns to 4
    m = P.Image(a, mapped=True)
    assert m.c.ptr == a.ctypes.data and m.c.stride == 7             # ac_image_map: no copy
    L = _native.lib()
    v = L.ac_image_alloc()
    assert L.ac_image_view(m._p, v, 2, 1, 10, 10) == 0
    assert v.contents.width == 5 and v.contents.height == 4 and v.contents.ptr == a.ctypes.data + 7 + 2
    c = L.ac_image_alloc()
    assert L.ac_image_clone(m._p, c) == 0 and c.contents.ptr != a.ctypes.data
    assert L.ac_image_ref(None, c) == -22
    L.ac_image_free(ctypes.byref(v))
    L.ac_image_free(ctypes.byref(c))
    assert not v and not c


def test_c_api_handle_lifecycle():
    """binding/c/src/Binding.cpp:33-105,161-189 semantics: ref shares pixels, unref zeroes
    the struct, free nulls the caller's pointer, every call on a struct without a handle
    (or a NULL struct) is AC_EINVAL / NULL, and processor ref copies device/type/model."""
    import numpy as np
    L = _native.lib()
    a = np.arange(6 * 9, dtype=np.uint8).reshape(6, 9)
    src = L.ac_image_alloc()
    s = src.contents
    s.width, s.height, s.channels, s.element_type, s.stride = 9, 6, 1, 0x001, 9
    s.ptr = a.ctypes.data
    assert L.ac_image_map(src) == 0 and s.ptr == a.ctypes.data
    r = L.ac_image_alloc()
    assert L.ac_image_ref(src, r) == 0
    assert (r.contents.ptr, r.contents.width, r.contents.height, r.contents.stride) == (a.ctypes.data, 9, 6, 9)
    # a second ref into an existing handle replaces what it held
    c = L.ac_image_alloc()
    assert L.ac_image_clone(src, c) == 0 and c.contents.ptr != a.ctypes.data
    assert L.ac_image_ref(src, c) == 0 and c.contents.ptr == a.ctypes.data
    # create on a zeroed struct gives an empty image with a handle
    e = L.ac_image_alloc()
    assert L.ac_image_create(e) == 0 and not e.contents.ptr and e.contents.hptr
    # unref zeroes every field; calls on a struct without a handle are rejected
    L.ac_image_unref(r)
    rc = r.contents
    assert (rc.width, rc.height, rc.channels, rc.stride, rc.element_type) == (0, 0, 0, 0, 0)
    assert not rc.ptr and not rc.hptr
    out = np.zeros(54, np.uint8)
    assert L.ac_image_to(r, out.ctypes.data, 9) == -22
    assert L.ac_image_view(r, c, 0, 0, 1, 1) == -22
    assert L.ac_image_to(src, None, 9) == -22
    assert L.ac_image_to(src, out.ctypes.data, 9) == 0 and np.array_equal(out.reshape(6, 9), a)
    assert L.ac_image_create(None) == -22 and L.ac_image_map(None) == -22 and L.ac_image_from(None, None) == -22
    assert L.ac_rgb2yuv(None, r) == -22 and L.ac_resize(src, r, 2.0, 2.0, 0) == -22
    L.ac_image_unref(None)
    L.ac_image_free(None)
    for p in (src, r, c, e):
        L.ac_image_free(ctypes.byref(p))
        assert not p

    # processors: queries without a created processor
    p = L.ac_processor_alloc()
    assert L.ac_processor_ok(p) == -22 and L.ac_processor_type(p) == -22
    assert L.ac_processor_error(p) is None and L.ac_processor_name(p) is None and L.ac_processor_type_name(p) is None
    assert L.ac_processor_create(None) == -22 and L.ac_processor_ref(None, p) == -22
    assert L.ac_processor_process(p, None, None, 2.0) == -22
    t, m = b"cpu", b"acnet-legacy-hdn0"
    p.contents.device, p.contents.type, p.contents.model = 3, t, m
    assert L.ac_processor_create(p) == -256 and L.ac_processor_type(p) == 0
    q = L.ac_processor_alloc()
    assert L.ac_processor_ref(p, q) == 0
    assert q.contents.device == 3 and q.contents.type == t and q.contents.model == m
    assert L.ac_processor_type(q) == 0 and L.ac_processor_type_name(q) == b"CPU"
    L.ac_processor_unref(q)
    assert not q.contents.hptr and q.contents.device == 0 and not q.contents.type
    for x in (p, q):
        L.ac_processor_free(ctypes.byref(x))
        assert not x
