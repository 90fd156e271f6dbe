"""The C-ABI library: it loads without a GPU, exports every function include/rabbitvar_amd.h
declares, the Python binding covers each of them, and opening an engine with no HIP device fails
with an error status instead of falling back to anything."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rabbitvar_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rv_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from rabbitvar_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built: run python -m rabbitvar_amd.build")
    return _lib


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 12
    so = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_header(lib):
    assert sorted(lib.SIGNATURES) == declared_functions()


def test_struct_sizes_match_header(lib):
    """ctypes mirrors of the ABI structs have the C sizes (a field added on one side only shows up
    here)."""
    import subprocess
    import tempfile

    src = r'''
#include <stdio.h>
#include "rabbitvar_amd.h"
int main(void){printf("%zu %zu %zu %zu %zu\n", sizeof(rv_config), sizeof(rv_reads), sizeof(rv_region),
 sizeof(rv_contig), sizeof(rv_stats)); return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(t) for t in (lib.RvConfig, lib.RvReads, lib.RvRegion, lib.RvContig, lib.RvStats)]
    assert got == want


def test_open_without_device_fails_loudly(lib):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    from rabbitvar_amd.engine import Engine, EngineError

    with pytest.raises(EngineError):
        Engine(sample="S1")
