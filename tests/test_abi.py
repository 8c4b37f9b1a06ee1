"""The C-ABI library without a GPU: it loads, exports exactly what include/onefft.h declares, and
its argument checks behave like the reference's (tests/test_onefft.c:170-182) before any device
work. No transform is run here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import onefft_amd

HEADER = os.path.join(ROOT, "include", "onefft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    c_part = text.split("#ifdef __cplusplus\n}")[0]
    return sorted(set(re.findall(r"\b(onefft_\w+)\s*\(", c_part)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(onefft_amd.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = onefft_amd.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", onefft_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = sorted(set(re.findall(r" T (\w+)", out)))
    assert exported == declared_functions()


def test_implementation_name():
    assert onefft_amd.implementation() == "hip-gfx950"


@pytest.mark.parametrize("n", [0, 3, 5, 6, 7, 12, 1000, 1023, 1025, 2**64 - 1, 2**31])
def test_invalid_sizes_rejected(n):
    lib = onefft_amd.lib()
    assert not lib.onefft_plan_create(n)
    a = np.zeros(6, np.float32)
    b = np.full(6, 7.0, np.float32)
    assert lib.onefft_forward(a.ctypes.data, b.ctypes.data, n) == -1
    assert lib.onefft_inverse(a.ctypes.data, b.ctypes.data, n) == -1
    assert (b == 7.0).all()
    assert lib.onefft_cache_prepare(n) == -1


def test_null_arguments_are_safe():
    lib = onefft_amd.lib()
    lib.onefft_plan_destroy(None)
    assert lib.onefft_plan_size(None) == 0
    lib.onefft_cache_clear()
    lib.onefft_cache_clear()
    assert lib.onefft_plan_forward_device(None, None, None, None) == -1
    assert lib.onefft_plan_inverse_batch(None, None, None, 1, 1, 1, None) == -1
    assert onefft_amd.last_error() != ""


def test_header_compiles_as_c99_and_cpp11(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "onefft.h"\nint main(void){ onefft_plan *p = onefft_plan_create(8); '
                 'onefft_plan_destroy(p); return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.dirname(HEADER),
                    "-c", str(c), "-o", str(tmp_path / "t.o")], check=True)
    cc = tmp_path / "t.cpp"
    cc.write_text('#include "onefft.h"\nint main(){ onefft::Plan p(8); onefft::Plan q(std::move(p)); '
                  'return (bool)p; }\n')
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-I", os.path.dirname(HEADER), "-c",
                    str(cc), "-o", str(tmp_path / "t2.o")], check=True)


def test_binding_raises_without_library(monkeypatch):
    monkeypatch.setattr(onefft_amd, "_lib", None)
    monkeypatch.setattr(onefft_amd, "LIB_PATH", "/nonexistent/libonefft_amd.so")
    with pytest.raises(onefft_amd.OneFFTError):
        onefft_amd.lib()
