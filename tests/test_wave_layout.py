"""The in-tile wavefront's LDS layout (messi_amd/csrc/wave_layout.h, the index
functions the inside kernel uses) checked on the host by enumeration: every read
of every lane and step stays in bounds and is a needed cell, a zero term or a
dead read (tests/wave_layout_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_wave_layout_reads(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "wave_layout_check")
    subprocess.run([gxx, "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "wave_layout_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for T in (16, 32, 8):
        assert f"ok T={T}" in r.stdout
