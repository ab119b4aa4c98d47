"""CPU checks of the C ABI: the library loads and exports every entry point the headers declare,
and the ctypes structs match the C layout."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in os.listdir(INCLUDE):
        if f.endswith(".h"):
            text = open(os.path.join(INCLUDE, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(vr_[a-z_0-9]+)\s*\(", text, re.M))
    return sorted(names)


def test_headers_declare_the_entry_points():
    names = declared_functions()
    for must in ("vr_agg_create", "vr_agg_update", "vr_agg_emit", "vr_join_probe", "vr_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from vecruntime_amd import vr
    lib = C.CDLL(vr.lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_c():
    """Compiles a probe of sizeof/offsetof for the ABI structs and compares with ctypes."""
    from vecruntime_amd import abi
    structs = {"vr_column": abi.VrColumn, "vr_out_column": abi.VrOutColumn, "vr_agg_desc": abi.VrAggDesc,
               "vr_agg_spec": abi.VrAggSpec, "vr_stats": abi.VrStats, "vr_batch_set": abi.VrBatchSet,
               "vr_launch_config": abi.VrLaunchConfig, "vr_join_spec": abi.VrJoinSpec,
               "vr_probe_output": abi.VrProbeOutput, "vr_server_config": abi.VrServerConfig,
               "vr_query": abi.VrQuery, "vr_query_result": abi.VrQueryResult,
               "vr_server_stats": abi.VrServerStats, "vr_scalar": abi.VrScalar,
               "vr_join_residual": abi.VrJoinResidual}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vr.h"', '#include "vr_serve.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0; }")
    tmp = os.path.join(ROOT, "_build_abi_probe")
    os.makedirs(tmp, exist_ok=True)
    src, exe = os.path.join(tmp, "probe.c"), os.path.join(tmp, "probe")
    open(src, "w").write("\n".join(lines))
    subprocess.run(["gcc", "-I", INCLUDE, src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(py, fname).offset, (cname, fname)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    """Without the built library the front end refuses to import (no CPU fallback)."""
    import importlib
    import sys
    import vecruntime_amd.vr as vrmod
    monkeypatch.setattr(os.path, "exists", lambda p: False if p.endswith("libvrgpu.so") else os.path.isfile(p) or os.path.isdir(p))
    sys.modules.pop("vecruntime_amd.vr", None)
    with pytest.raises(ImportError):
        importlib.import_module("vecruntime_amd.vr")
    sys.modules["vecruntime_amd.vr"] = vrmod


def test_prepared_batch_set_points_at_every_column():
    import numpy as np

    from vecruntime_amd import abi, vr
    from vecruntime_amd.columns import Column
    batches = []
    for b in range(3):
        n = 5 + b
        batches.append(([Column.longs(np.arange(n))], [Column.longs(np.arange(n)), Column.doubles(np.ones(n))], n))
    p = vr.GroupedAggregate.prepare_batches(batches, filter_col=1, threshold=7)
    bs = p.bs
    assert bs.num_batches == 3 and bs.filter_col == 1 and bs.filter_threshold == 7
    assert [bs.lengths[i] for i in range(3)] == [5, 6, 7]
    for b in range(3):
        assert bs.keys[b].length == 5 + b and bs.keys[b].type == abi.INT64
        assert bs.values[2 * b].type == abi.INT64 and bs.values[2 * b + 1].type == abi.FLOAT64
