"""The upload stage reads exactly the caller's bytes: tests/stage_check.cpp, built here with g++ (and once more with
AddressSanitizer on the host), puts source buffers that end at a PROT_NONE guard page. No device calls."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "stage_check.cpp")


@pytest.mark.parametrize("asan", [False, True])
def test_stage_reads_only_the_callers_bytes(asan):
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "stage_check")
        cmd = ["g++", "-O1", "-g", "-std=c++17", "-o", exe, SRC]
        if asan:
            cmd[1:1] = ["-fsanitize=address", "-fno-omit-frame-pointer"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if asan and r.returncode != 0 and "asan" in (r.stderr or "").lower():
            pytest.skip("AddressSanitizer runtime not available")
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
        assert "stage ok" in r.stdout
