"""The C-ABI library loads and exports every entry point include/macaw_amd.h
declares; host-side argument errors come back as messages, not crashes."""
import os
import re

import pytest

from macaw_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "macaw_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mcw_[a-z_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    names = declared()
    assert "mcw_run_batch" in names and "mcw_tally_get" in names
    assert set(names) == set(lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = lib.lib()
    for name in declared():
        assert hasattr(L, name), name


def test_struct_sizes_match_header():
    import ctypes as C
    assert C.sizeof(abi.Options) == 80
    assert C.sizeof(abi.Source) == 144
    assert C.sizeof(abi.History) == 72
    assert C.sizeof(abi.BatchStats) == 120
    assert C.sizeof(abi.TraceRecord) == 88 == abi.trace_dtype().itemsize


def test_default_options():
    o = lib.default_options()
    assert o.mode == abi.MODE_BATCHED and o.master_seed == 1 and o.precision == abi.PREC_FP64


def test_null_context_is_an_error():
    L = lib.lib()
    assert L.mcw_run_batch(None, 1, 10, None) != 0
    assert L.mcw_last_error(None) == b"null context"
