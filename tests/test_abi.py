"""The C-ABI library loads and exports every entry point declared in include/messi.h.
No compute calls are made here (no GPU needed)."""
import ctypes
import os
import re

import pytest

from messi_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "messi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(messi_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ["messi_ctx_create", "messi_dataset_create", "messi_pair_loglik", "messi_site_loglik",
              "messi_inside", "messi_outside", "messi_bpp", "messi_eval", "messi_posterior"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lib.LIBPATH):
        pytest.skip("libmessi.so not built")
    so = ctypes.CDLL(lib.LIBPATH)
    for n in _declared():
        assert hasattr(so, n), n
        assert n in lib.SIGNATURES, n


def test_load_binds_signatures():
    if not os.path.exists(lib.LIBPATH):
        pytest.skip("libmessi.so not built")
    L = lib.load()
    assert L.messi_last_error.restype is ctypes.c_char_p
    assert L.messi_last_error(None) == b"null context"
