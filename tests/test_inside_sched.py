"""The single-launch (persistent) inside schedule (messi_amd/csrc/inside_sched.h,
the ticket table and decode the kernels use) checked on the host by enumeration:
every finalisation and bulk visit appears once, and each waits only for items of
smaller tickets -- its neighbouring finalisations, the previous diagonal, the
tile's previous visit -- so tickets handed out in order cannot deadlock
(tests/inside_sched_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_persistent_schedule_order(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "inside_sched_check")
    subprocess.run([gxx, "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "inside_sched_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
