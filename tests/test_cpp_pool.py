"""Builds tests/cpp/pool_test.cpp (the serving loop's chunk pool, anime4kcpp_amd/csrc/chunk_pool.hpp,
over a fake host backend) with g++, plain and under AddressSanitizer/UBSan, and runs it: blocks never
overlap and stay inside backend chunks, freed ranges coalesce within a chunk and never across two
adjacent chunks, best fit reuses holes, a failed chunk allocation trims wholly free chunks and retries
the exact size, and every chunk goes back to the backend."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.parametrize("sanitize", [False, True], ids=["plain", "asan-ubsan"])
def test_chunk_pool(tmp_path, sanitize):
    exe = str(tmp_path / "pool_test")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "anime4kcpp_amd", "csrc"),
           os.path.join(HERE, "cpp", "pool_test.cpp"), "-o", exe]
    if sanitize:
        cmd[3:3] = ["-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    subprocess.run(cmd, check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
