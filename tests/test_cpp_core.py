"""Builds tests/cpp/core_test.cpp against include/ and lib/libac_hip.so and runs it:
the C++ API read like the reference's own doctest suites."""
import os
import subprocess

import pytest

from anime4kcpp_amd import _native

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _build(tmp_path, name="core_test"):
    exe = str(tmp_path / name)
    libdir = os.path.dirname(_native.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", name + ".cpp"),
           "-L" + libdir, "-lac_hip", "-Wl,-rpath," + libdir, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_cpp_core_api_host(tmp_path):
    if _native.lib().achip_device_count() > 0:
        pytest.skip("host-only variant; the gpu variant runs on a device")
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_core_api_gpu(tmp_path, cuda):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_specs_listing_host(tmp_path):
    """AC/Specs.hpp compiled with the CLI's listing loops; counts and names checked."""
    if _native.lib().achip_device_count() > 0:
        pytest.skip("host-only variant; the gpu variant runs on a device")
    exe = _build(tmp_path, "specs_test")
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "acnet-legacy-hdn0:" in r.stdout and "hip-bf16" in r.stdout


@pytest.mark.gpu
def test_specs_listing_gpu(tmp_path, cuda):
    exe = _build(tmp_path, "specs_test")
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
