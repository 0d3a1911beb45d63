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


def test_no_kernel_needs_scratch_memory():
    """No kernel of the library has a private segment.  The runtime
    allocates a queue's scratch at dispatch time, the first time a kernel
    needs it there; with device memory taken by job buffers that allocation
    fails and aborts the queue, so every kernel keeps its locals in
    registers and LDS."""
    import subprocess
    import tempfile
    llvm = "/opt/rocm/llvm/bin"
    tools = [os.path.join(llvm, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")]
    if not all(os.path.exists(t) for t in tools):
        import pytest
        pytest.skip("ROCm LLVM tools not present")
    objcopy, bundler, readelf = tools
    with tempfile.TemporaryDirectory() as td:
        fat = os.path.join(td, "fatbin")
        subprocess.run([objcopy, "--dump-section=.hip_fatbin=" + fat, G.LIB_PATH], check=True)
        blob = open(fat, "rb").read()
        magic = b"__CLANG_OFFLOAD_BUNDLE__"
        starts = [m.start() for m in re.finditer(re.escape(magic), blob)]
        assert starts, "no offload bundle in the library"
        sizes = []
        for i, s in enumerate(starts):
            part = os.path.join(td, f"b{i}")
            with open(part, "wb") as f:
                f.write(blob[s:starts[i + 1] if i + 1 < len(starts) else len(blob)])
            elf = part + ".elf"
            subprocess.run([bundler, "--unbundle", "--type=o", "--input=" + part,
                            "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + elf], check=True)
            notes = subprocess.run([readelf, "--notes", elf], check=True, capture_output=True, text=True).stdout
            sizes += [int(x) for x in re.findall(r"\.private_segment_fixed_size:\s*(\d+)", notes)]
    assert len(sizes) > 50  # every kernel was seen
    assert max(sizes) == 0, sorted(sizes)[-5:]


def test_chunk_budget_follows_key_width():
    """The automatic phase-2 chunk uses the same per-occurrence model as the
    server's admission (166 B with 64-bit k-mers, 236 B with 128-bit ones), so
    a k > 31 job plans proportionally smaller chunks."""
    usable = 200 << 30
    b31, b32, b63 = (G.chunk_budget(usable, k) for k in (31, 32, 63))
    assert b32 == b63 < b31
    assert abs(b31 / b63 - 236.0 / 166.0) < 1e-3
    assert G.chunk_budget(1 << 20, 31) == 1 << 20  # floor
    assert G.chunk_budget(1 << 50, 31) == (1 << 31) - 1  # ceiling
