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
    assert C.sizeof(abi.Options) == 88
    assert C.sizeof(abi.Source) == 144
    assert C.sizeof(abi.History) == 72
    assert C.sizeof(abi.BatchStats) == 128
    assert C.sizeof(abi.TraceRecord) == 88 == abi.trace_dtype().itemsize
    assert C.sizeof(abi.Site) == 72 == abi.site_dtype().itemsize


STRUCTS = {"mcw_options": abi.Options, "mcw_nuclide": abi.Nuclide, "mcw_material": abi.Material,
           "mcw_source": abi.Source, "mcw_filter": abi.Filter, "mcw_batch_stats": abi.BatchStats,
           "mcw_history": abi.History, "mcw_trace_record": abi.TraceRecord, "mcw_site": abi.Site,
           "mcw_keff_result": abi.KeffResult, "mcw_request": abi.Request,
           "mcw_request_status": abi.RequestStatus}


def test_struct_layouts_match_c_compiler(tmp_path):
    """Every ctypes mirror has the C compiler's size and field offsets."""
    import ctypes as C
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "macaw_amd.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_default_options():
    o = lib.default_options()
    assert o.mode == abi.MODE_BATCHED and o.master_seed == 1 and o.precision == abi.PREC_FP64


def test_null_context_is_an_error():
    L = lib.lib()
    assert L.mcw_run_batch(None, 1, 10, None) != 0
    assert L.mcw_last_error(None) == b"null context"
