"""The C-ABI library builds, loads without a GPU and exports every function
include/*.h declares; error paths return codes instead of exiting."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mc-mpi_amd", "libmcgpu.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+(mcgpu_[a-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"mcgpu_simulate", "mcgpu_simulate_device", "mcgpu_context_create"} <= names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libmcgpu.so first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared_functions() - exported
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    import mcgpu._lib as L
    assert declared_functions() <= set(L.SIGNATURES)


def test_library_loads_and_reports_errors_without_exiting():
    import mcgpu
    L = mcgpu.lib()
    assert L.mcgpu_version().startswith(b"mcgpu gfx950")
    assert L.mcgpu_status_string(-5) == b"history exceeded the event cap"
    # bad arguments return codes (no process exit, unlike gpu_errcheck)
    assert L.mcgpu_context_create(0, None) == -1
    assert L.mcgpu_simulate(None, 0, None, None, None, None, 0, 1, 0.1, 0.0, None, None, None) == -1
    assert L.mcgpu_simulate_device(None, 0, None, None, None, None, 0, 1, 0.1, 0.0, None, None,
                                   None, None, None, None) == -1


def test_context_create_without_device_returns_code():
    """On a machine without a HIP device the context constructor returns
    MCGPU_ERR_NO_DEVICE; with a device, an out-of-range ordinal does."""
    import mcgpu
    L = mcgpu.lib()
    h = C.c_void_p()
    rc = L.mcgpu_context_create(1000, C.byref(h))
    assert rc == -4 and not h.value
