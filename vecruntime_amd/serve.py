"""Python front end of the query server (include/vr_serve.h): queries submitted from any thread,
admitted against an HBM budget, run concurrently on per-worker HIP streams.

    srv = Server(dev, workers=4)
    t = srv.submit_agg(spec, [(keys, values, n), ...], expected_groups=g)
    res = srv.wait(t)                 # QueryResult: status, rows, latency
    cols = srv.fetch(t, res)          # host OutColumns: keys, then buffers
    srv.release(t)
"""
import ctypes as C

import numpy as np

from . import abi
from .columns import OutColumn, c_array
from .vr import _lib, _P, _sig, agg_key_dictionary, check, VrError

_sig("vr_server_create", [_P, C.POINTER(abi.VrServerConfig), C.POINTER(_P)])
_sig("vr_server_submit", [_P, C.POINTER(abi.VrQuery), C.POINTER(C.c_int64)])
_sig("vr_server_wait", [_P, C.c_int64, C.POINTER(abi.VrQueryResult)])
_sig("vr_server_copy_result", [_P, C.c_int64, C.c_int32, C.POINTER(abi.VrOutColumn)])
_sig("vr_server_release", [_P, C.c_int64])
_sig("vr_server_stats_get", [_P, C.POINTER(abi.VrServerStats)])
_sig("vr_server_destroy", [_P])

_I32P = C.POINTER(C.c_int32)


class QueryResult:
    def __init__(self, r):
        self.status = r.status
        self.rows = r.rows
        self.num_columns = r.num_columns
        self.types = [r.columns[i].type for i in range(r.num_columns)]
        self.queue_ms = r.queue_ms
        self.run_ms = r.run_ms
        self.latency_ms = r.latency_ms
        self.device_bytes = r.device_bytes
        self.agg = r.agg


def prewarm(device, specs, threads=4):
    """vr_agg_prewarm for each abi.AggSpec, `threads` compiles at a time."""
    from concurrent.futures import ThreadPoolExecutor
    specs = list(specs)
    with ThreadPoolExecutor(max(1, min(threads, len(specs) or 1))) as ex:
        for st in ex.map(lambda sp: _lib.vr_agg_prewarm(device.h, C.byref(sp.c)), specs):
            check(st)


class Server:
    def __init__(self, device, workers=4, hbm_budget=0, admission=abi.ADMIT_SIZE_AWARE, aging_ms=0):
        self.dev = device
        self.h = _P()
        cfg = abi.VrServerConfig(int(workers), int(admission), int(hbm_budget), int(aging_ms))
        check(_lib.vr_server_create(device.h, C.byref(cfg), C.byref(self.h)))
        self._keep = {}

    def submit_agg(self, spec, batches, expected_groups=0, filter_col=-1, threshold=0):
        """spec: abi.AggSpec; batches: [(key Columns, value Columns, n)] of device columns."""
        ka = c_array([c for b in batches for c in b[0]])
        va = c_array([c for b in batches for c in b[1]])
        ln = np.asarray([b[2] for b in batches], np.int64)
        bs = abi.VrBatchSet(len(batches), C.cast(ka, C.POINTER(abi.VrColumn)), C.cast(va, C.POINTER(abi.VrColumn)),
                            ln.ctypes.data_as(C.POINTER(C.c_int64)), filter_col, threshold)
        q = abi.VrQuery()
        q.kind = abi.QUERY_AGG
        q.agg = C.pointer(spec.c)
        q.batches = C.pointer(bs)
        q.expected_groups = int(expected_groups)
        t = C.c_int64()
        check(_lib.vr_server_submit(self.h, C.byref(q), C.byref(t)))
        self._keep[t.value] = ("agg", spec, batches)
        return t.value

    def submit_join(self, spec, build_cols, build_rows, probe_cols, probe_keys, probe_rows, join_type,
                    probe_out=(), build_out=(), expected_rows=0):
        """spec: abi.JoinSpec; build_cols / probe_cols: device Columns."""
        bc, pc = c_array(build_cols), c_array(probe_cols)
        pk, po, bo = abi.int_array(probe_keys), abi.int_array(list(probe_out)), abi.int_array(list(build_out))
        q = abi.VrQuery()
        q.kind = abi.QUERY_JOIN
        q.join_type = join_type
        q.join = C.pointer(spec.c)
        q.build_columns = C.cast(bc, C.POINTER(abi.VrColumn))
        q.build_rows = int(build_rows)
        q.probe_columns = C.cast(pc, C.POINTER(abi.VrColumn))
        q.num_probe_columns = len(probe_cols)
        q.num_probe_out = len(probe_out)
        q.probe_key_columns = C.cast(pk, _I32P)
        q.probe_out_columns = C.cast(po, _I32P)
        q.probe_rows = int(probe_rows)
        q.num_build_out = len(build_out)
        q.build_out_columns = C.cast(bo, _I32P)
        q.expected_rows = int(expected_rows)
        t = C.c_int64()
        check(_lib.vr_server_submit(self.h, C.byref(q), C.byref(t)))
        self._keep[t.value] = ("join", spec, build_cols, probe_cols)
        return t.value

    def wait(self, ticket):
        r = abi.VrQueryResult()
        check(_lib.vr_server_wait(self.h, ticket, C.byref(r)))
        return QueryResult(r)

    def fetch(self, ticket, res):
        """The result columns copied to host OutColumns."""
        if res.status != abi.VR_OK:
            raise VrError(res.status, "the query failed")
        cols = []
        for i, t in enumerate(res.types):
            oc = OutColumn(t, max(1, res.rows))
            check(_lib.vr_server_copy_result(self.h, ticket, i, C.byref(oc.c())))
            cols.append(oc)
        return cols

    def key_dictionary(self, res, c):
        """The strings behind the ids of UTF8 key column `c` of a finished group-by (valid until
        the query is released)."""
        if not res.agg:
            raise ValueError("the query has no UTF8 key dictionaries")
        return agg_key_dictionary(res.agg, c)

    def release(self, ticket):
        check(_lib.vr_server_release(self.h, ticket))
        self._keep.pop(ticket, None)

    def stats(self):
        s = abi.VrServerStats()
        check(_lib.vr_server_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.VrServerStats._fields_}

    def close(self):
        if self.h:
            _lib.vr_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
