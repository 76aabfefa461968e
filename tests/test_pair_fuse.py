"""The choice of pair nodes computed in their parent's launch
(messi_amd/csrc/pair_fuse.h, called by the dataset build) checked on the host:
on random trees it holds only candidates, never computes a node together with
its parent or both children of a node, and reaches the largest total saving,
found by enumeration on small trees (tests/pair_fuse_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_pair_fusion_choice(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "pair_fuse_check")
    subprocess.run([gxx, "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "pair_fuse_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
