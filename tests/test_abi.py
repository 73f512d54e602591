"""The C-ABI library loads and exports every function include/lattibol_amd.h
declares; struct layouts in the Python binding match the header (no GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lattibol_amd.h")


def declared_functions():
    text = open(HEADER).read()
    # the C declarations; the C++ inline overloads after them export nothing
    text = text.split("/* C++ callers")[0]
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, re.M):
        names.add(m.group(1))
    return names - {"if", "while", "for", "switch", "return", "sizeof"}


def test_header_declares_reference_entry_points():
    names = declared_functions()
    ref = {"streaming_d2q9", "streaming_d3q19", "collision_d2q9", "collision_d3q19", "get_border_d2q9",
           "get_border_d3q19", "copy_border_devices_d2q9", "copy_border_devices_d3q19", "apply_border_d2q9",
           "apply_border_d3q19", "read_value_2d", "read_value_3d"}
    assert ref <= names


def test_library_exports_every_declared_symbol():
    from lattibol_amd import _lib
    names = declared_functions()
    missing = [n for n in sorted(names) if not hasattr(_lib.lib, n)]
    assert not missing, missing
    # and the dynamic symbol table agrees (extern "C", unmangled)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert names <= exported, sorted(names - exported)


def test_binding_prototypes_cover_header():
    from lattibol_amd import _lib
    names = declared_functions()
    assert names <= set(_lib._PROTOS), sorted(names - set(_lib._PROTOS))


def test_struct_sizes_match_reference_layout():
    from lattibol_amd import _lib
    P = C.sizeof(C.c_void_p)
    # DataD2Q9: 5 + 9 + 6 + 6 pointers, 4 ints, 2 pointers, stream, 3 pointers
    assert C.sizeof(_lib.DataD2Q9) == 26 * P + 16 + 2 * P + P + 3 * P
    assert C.sizeof(_lib.DataD3Q19) == (6 + 19 + 10 + 10) * P + 20 + 4 + 3 * P
    assert C.sizeof(_lib.LbDim3) == 12


def test_error_strings_without_gpu():
    from lattibol_amd import _lib
    assert _lib.lib.lb_error_string(0) == b"success"
    assert b"grid*block" in _lib.lib.lb_error_string(-2)
    # the void reference launchers report through the sticky error
    _lib.lib.lb_last_error()
    _lib.lib.streaming_d2q9(None, _lib.LbDim3(1, 1, 1), _lib.LbDim3(1, 1, 1))
    assert _lib.lib.lb_last_error() == -1
    assert _lib.lib.lb_last_error() == 0


def test_create_fails_loudly_without_gpu():
    from lattibol_amd import _lib, make_config, Solver, LatticeError
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(LatticeError):
        Solver(make_config(9, (8, 8), tau=0.6))
