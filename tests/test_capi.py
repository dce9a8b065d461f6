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


def test_plan_graph_and_workspace_entry_points_check_their_arguments():
    """The sweep-plan, autotuner, graph and workspace entry points are host
    logic: their argument checks run without a device."""
    from x2dhf_amd import gpu
    lib = gpu.load()
    assert lib.x2g_set_sweep_plan(3, 0, 0) == -1 and b"block_rows" in lib.x2g_last_error()
    assert lib.x2g_set_sweep_plan(2, 17, 0) == -1 and b"wg_per_cu" in lib.x2g_last_error()
    assert lib.x2g_set_sweep_plan(2, 0, 2) == -1 and b"split" in lib.x2g_last_error()
    assert lib.x2g_set_sweep_plan(2, 1, 1) == 0 and lib.x2g_set_sweep_plan(0, 0, 0) == 0
    assert lib.x2g_set_graphs(None, 1) == -1 and b"null" in lib.x2g_last_error()
    c = C.c_longlong()
    assert lib.x2g_graph_counts(None, C.byref(c), C.byref(c)) == -1
    assert lib.x2g_tune_report(None, 10) == -1
    assert lib.x2g_set_autotune(1) == 0
    # workspace of one relaxation problem: FP64 two padded arrays + pairs,
    # FP32 two padded correction arrays + lhs/residual; bad sizes give -1
    w64, w32 = lib.x2g_relax_workspace_bytes(331, 601, 0), lib.x2g_relax_workspace_bytes(331, 601, 1)
    assert w64 > 2 * 8 * 339 * 609 and w32 > 2 * 4 * 339 * 609 and w64 > w32
    assert lib.x2g_relax_workspace_bytes(5, 601, 0) == -1 and lib.x2g_relax_workspace_bytes(331, 601, 2) == -1


def test_scf_control_record_layout_agrees():
    """x2g_scf_ctrl has one layout in the library, the ctypes mirror and the
    Fortran driver's bind(c) type (which checks it at job creation): 74
    doubles then 72 ints, 880 bytes, no padding; x2g_scf_control checks its arguments
    before any device call."""
    from x2dhf_amd import gpu
    lib = gpu.load()
    assert lib.x2g_scf_ctrl_bytes() == C.sizeof(gpu.ScfCtrl) == 880
    assert gpu.ScfCtrl.norb.offset == 74 * 8 and gpu.ScfCtrl.stop.offset == 880 - 4
    assert lib.x2g_scf_control(None, None, None, 1) == -1 and b"null" in lib.x2g_last_error()
