"""The C-ABI library loads and exports every symbol include/x2gpu.h declares
(no compute call: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "include", "x2gpu.h")


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(x2g_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared(HDR)
    for must in ("x2g_mcsor", "x2g_relax_batch", "x2g_putin", "x2g_wdot", "x2g_oneel", "x2g_moments"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from x2dhf_amd import gpu
    lib = C.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared(HDR) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from x2dhf_amd import gpu
    missing = [n for n in declared(HDR) if n not in gpu._PROTOS]
    assert not missing, missing


def test_last_error_and_version_without_device():
    from x2dhf_amd import gpu
    lib = gpu.load()
    assert lib.x2g_version() >= 100
    assert isinstance(lib.x2g_last_error(), bytes)


def test_bad_arguments_are_rejected_before_any_device_call():
    from x2dhf_amd import gpu
    lib = gpu.load()
    assert lib.x2g_ctx_create(0, None) == -1
    assert b"null" in lib.x2g_last_error()
    assert lib.x2g_set_sweep_engine(7, 0) == -1
    assert lib.x2g_set_sweep_engine(0, -1) == 0
