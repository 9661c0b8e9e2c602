"""Builds tests/cpp/image_ops_test.cpp (the image helpers through the C++ API, read like the
reference's ImageProcessTest.cpp / ImageResizeTest.cpp) against include/, lib/libac_hip.so
and the oracle's conversions, and runs it: on a device the helpers' results, without one
that every helper fails loudly and leaves its outputs alone."""
import os
import subprocess

import pytest

from anime4kcpp_amd import _native

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _build(tmp_path):
    exe = str(tmp_path / "image_ops_test")
    libdir = os.path.dirname(_native.LIB_PATH)
    orc = os.path.join(ROOT, "oracle")
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "image_ops_test.cpp"),
           "-L" + libdir, "-lac_hip", "-Wl,-rpath," + libdir, os.path.join(orc, "liboracle.so"), "-Wl,-rpath," + orc,
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_image_helpers_without_device(tmp_path):
    if _native.lib().achip_device_count() > 0:
        pytest.skip("host-only variant; the gpu variant runs on a device")
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "no HIP device" in r.stderr


@pytest.mark.gpu
def test_image_helpers_gpu(tmp_path, cuda):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
