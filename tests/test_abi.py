"""The C-ABI library loads and exports every symbol include/cdbg_gpu.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT, cdbg_gpu as G


def _declared():
    txt = open(os.path.join(ROOT, "include", "cdbg_gpu.h")).read()
    return sorted(set(re.findall(r"CDBG_API\s+[\w\s\*]+?\b(cdbg_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("cdbg_gpu_job_create", "cdbg_gpu_job_phase1", "cdbg_gpu_job_phase2",
                 "cdbg_gpu_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared()) <= set(G._SIGS)


def test_errors_are_reported_not_thrown():
    G.lib()
    ctx = ctypes.c_void_p()
    rc = G.lib().cdbg_gpu_ctx_create(10_000, ctypes.byref(ctx))
    assert rc != 0
    assert G.lib().cdbg_gpu_last_error()


def test_library_carries_gfx950_code_only():
    """Every device code object in the library targets gfx950 (a HIP source
    built for another target has no kernel the MI355X can load)."""
    import re as _re
    data = open(G.LIB_PATH, "rb").read()
    targets = set(_re.findall(rb"amdhsa--(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}, targets
