"""MI355X-native (gfx950) ACNetLegacy upscaling path, a drop-in for the 2x pass of
TianZerL/Anime4KCPP's ``ac::core::Processor``.

The product is the native library ``lib/libac_hip.so`` (C ABI in ``include/``);
this package holds its build recipe, a ctypes binding and a device-level Python
mirror used by the tests and the benchmark.
"""
from ._native import LIB_PATH, lib  # noqa: F401
