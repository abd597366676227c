"""The C ABI library: it loads, exports every entry point include/axonexa.h
declares, and the ctypes mirror matches the header's struct layout.  No
compute call is made (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from axonexa_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "axonexa.h")
LIB = os.path.join(ROOT, "axonexa_amd", "libaxonexa.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(axx_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("axx_create", "axx_destroy", "axx_run_phase", "axx_plan_create", "axx_plan_walk",
              "axx_plan_reduce", "axx_plan_results", "axx_plan_get_state", "axx_plan_set_state",
              "axx_run_mag", "axx_mag_create", "axx_mag_walk", "axx_mag_results", "axx_mag_get_state"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(axx_[a-z_0-9]+)\b", out))
    assert set(declared_functions()) <= exported


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_error_strings_without_device():
    lib = C.CDLL(LIB)
    lib.axx_strerror.restype = C.c_char_p
    lib.axx_version.restype = C.c_char_p
    for code in range(-8, 1):
        assert lib.axx_strerror(code).decode()
    assert b"gfx950" in lib.axx_version()


def test_ctypes_structs_match_header_layout():
    # sizes/offsets of the descriptor structs as a C compiler lays them out
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "axonexa.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(axx_substrate), sizeof(axx_acq), sizeof(axx_run),
         sizeof(axx_result), sizeof(axx_tuning), offsetof(axx_run, bounce_cap), sizeof(axx_field),
         offsetof(axx_field, origin), offsetof(axx_field, T1), sizeof(axx_mag_result));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, "t.c")
        open(cf, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cf, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(abi.Substrate), C.sizeof(abi.Acq), C.sizeof(abi.Run), C.sizeof(abi.Result),
            C.sizeof(abi.Tuning), abi.Run.bounce_cap.offset, C.sizeof(abi.Field), abi.Field.origin.offset,
            abi.Field.T1.offset, C.sizeof(abi.MagResult)]
    assert got == want


def test_oracle_library_loads():
    from oracle import pyoracle as po
    assert po.lib() is not None
