"""C ABI checks that need no GPU: libculsh.so loads, exports every function
include/culsh_c.h declares, and rejects invalid parameters before any device
work (cuLSH_HashTables.cu:70 parameter check)."""
import ctypes
import os
import re
import subprocess

import pytest

from culsh_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "culsh_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(culsh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_loads_and_reports_version():
    assert b"gfx950" in _lib.lib().culsh_version()


def test_invalid_parameters_fail_without_device():
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.culsh_ht_create(ctypes.byref(h)) == 0
    try:
        rc = L.culsh_ht_reset_random(h, 0, 16, 4, 6, 1.0, 1, None)
        assert rc == 1  # CULSH_E_ARG
        assert "positive" in _lib.last_error()
        rc = L.culsh_ht_reset_random(h, 100, 16, 4, 6, -1.0, 1, None)
        assert rc == 1
        rc = L.culsh_ht_reset_random(h, 100, 16, 4, 65, 1.0, 1, None)
        assert rc == 7  # CULSH_E_LIMIT: M > 64
        # index before a successful reset is a state error
        rc = L.culsh_ht_index(h, None, None)
        assert rc in (1, 2)
        assert L.culsh_ht_buckets(h, 0) == 0
        assert L.culsh_ht_A(h, 0) is None
    finally:
        L.culsh_ht_destroy(h)


def test_search_without_reset_is_a_state_error():
    L = _lib.lib()
    s = ctypes.c_void_p()
    assert L.culsh_st_create(ctypes.byref(s)) == 0
    try:
        rc = L.culsh_st_search(s, None, 10, None, None)
        assert rc == 2  # CULSH_E_STATE (cuLSH_SearchTables.cu:211-214)
        assert L.culsh_st_knn_ids(s) is None
        assert L.culsh_st_selectivity(s) == 0.0
    finally:
        L.culsh_st_destroy(s)


@pytest.mark.parametrize("header", ["culsh_c.h", "cuLSH.h"])
def test_headers_compile_as_plain_cpp(tmp_path, header):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include "{header}"\nint main() {{ return 0; }}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)
