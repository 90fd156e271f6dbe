"""rv_fisher_tables_noalloc (the allocation-free variant of rv_debug_fisher: its device memory is a
fixed file-scope array per GPU) against rv_debug_fisher through the C ABI: bit-identical p-values on
two calls in a row, the second with fewer tables, and the size limit refused."""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden", "fisher_bruteforce.json")


def noalloc(lib, tables):
    t = np.ascontiguousarray(tables, np.int32).reshape(-1, 4)
    out = np.full((len(t), 3), np.nan, np.float64)
    st = lib.rv_fisher_tables_noalloc(t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(t),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return st, out


def test_noalloc_matches_engine_entry(tools):
    from rabbitvar_amd import _lib
    from rabbitvar_amd.engine import Engine
    from test_gpu_fisher import deep_tables

    lib = _lib.load()
    with open(GOLDEN) as f:
        golden = [g["t"] for g in json.load(f)]
    rnd = np.random.default_rng(3)
    big = np.array(golden + deep_tables(9) + rnd.integers(0, 3000, size=(40000, 4)).tolist(), np.int32)
    small = np.array(deep_tables(4)[:300], np.int32)
    print(f"noalloc: {len(big)} then {len(small)} tables", flush=True)
    # the allocation-free calls first (in a fresh process they set the array up), then the engine's
    st1, got_big = noalloc(lib, big)
    st2, got_small = noalloc(lib, small)
    eng = Engine(device=0, sample="S1", fisher=1)
    want_big, want_small = eng.fisher(big), eng.fisher(small)
    eng.close()
    assert st1 == 0 and st2 == 0
    assert np.array_equal(got_big.view(np.uint64), want_big.view(np.uint64))
    assert np.array_equal(got_small.view(np.uint64), want_small.view(np.uint64))
    # beyond the fixed array: refused, nothing written
    st, got = noalloc(lib, np.zeros((65537, 4), np.int32))
    assert st == 5 and np.isnan(got).all()
    st, _ = noalloc(lib, np.zeros((0, 4), np.int32))
    assert st == 0
