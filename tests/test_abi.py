"""The C ABI library loads and exports every entry point include/cubez_amd.h
declares (no GPU needed: nothing is called that touches a device)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cubez_amd.h")
LIB = os.path.join(ROOT, "cubez_amd", "libcubez_amd.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(cbz_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declares_the_step_api():
    names = declared()
    for must in ("cbz_create", "cbz_set_scene", "cbz_integrate", "cbz_narrowphase", "cbz_narrowphase_pairs",
                 "cbz_resolve", "cbz_step", "cbz_get_contacts"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (cbz_[a-z_0-9]+)$", out, re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_loads_and_python_binding_matches_header():
    from cubez_amd import _native

    L = ctypes.CDLL(LIB)
    for n in declared():
        assert hasattr(L, n), n
    assert sorted(_native.EXPORTS) == sorted(n for n in declared() if n in _native.EXPORTS)
    assert set(declared()) <= set(_native.EXPORTS)
    assert L.cbz_field_width(2) == 3 and L.cbz_field_width(17) == 1 and L.cbz_field_width(99) == -1


def test_field_table_matches_header():
    from cubez_amd import _native

    text = open(HEADER).read()
    ids = dict(re.findall(r"(CBZ_[A-Z_]+) = (\d+),", text))
    for name, (fid, width, kind) in _native.FIELDS.items():
        assert int(ids["CBZ_" + name.upper()]) == fid, name


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
@pytest.mark.parametrize("rows", [1, 31, 32, 33, 1000, 4096, 65536])
def test_surrogate_gemms_stay_inside_their_buffers(rows):
    """Every GEMM of a training step and an inference pass reads and writes
    inside the buffers the library allocates for them (host-only check: the
    operands get fake address ranges of the allocated sizes).  A layer-0
    weight-gradient GEMM once read Bt rows 48..63 of a 48-row buffer."""
    L = ctypes.CDLL(LIB)
    L.cbz_sur_check_gemms.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]
    msg = ctypes.create_string_buffer(512)
    rc = L.cbz_sur_check_gemms(rows, msg, 512)
    assert rc == 0, msg.value.decode()
