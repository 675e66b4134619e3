"""ctypes bindings of libolapcore.so (include/olapcore.h).

Every call goes through the C ABI; device buffers are torch tensors on a ROCm
device (torch only provides memory and streams). There is no CPU fallback:
if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libolapcore.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "olapcore.h")

MAX_DIMS = 4
DIM_NULL_I8 = 127
DIM_NULL_I32 = -1
GID_NULL = -1
STARJOIN_NULL = 32767

OK, E_INVALID, E_HIP, E_NOMEM, E_CAPACITY, E_NODEVICE, E_BUSY = range(7)
ROWWISE, VWM_DV, VWM_SV, CWM_DV, CWM_SV = range(5)
STRATEGIES = {"rowwise": ROWWISE, "vwm_dv": VWM_DV, "vwm_sv": VWM_SV,
              "cwm_dv": CWM_DV, "cwm_sv": CWM_SV}
AGGREGATIONS = ["global atomics", "two-pass staged", "hot-group cache", "one-pass staged",
                "LDS table"]
# table slots as the reference's main() indexes them
TABLES = {"customer": 0, "supplier": 1, "part": 2, "date": 3}


class OlapError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"olapcore status {status}: {msg}")
        self.status = status


class Plan(C.Structure):
    _fields_ = [("dimvec_nums", C.c_int32),
                ("orders", C.c_int32 * MAX_DIMS),
                ("factor", C.c_int32 * MAX_DIMS),
                ("group_nums", C.c_int64)]

    def as_dict(self):
        return {"dimvec_nums": self.dimvec_nums, "orders": list(self.orders),
                "factor": list(self.factor), "group_nums": self.group_nums}


class Columns(C.Structure):
    _fields_ = [("dimvec", C.c_void_p * MAX_DIMS),
                ("dim_rows", C.c_int64 * MAX_DIMS),
                ("code_bytes", C.c_int32 * MAX_DIMS),
                ("fk", C.c_void_p * MAX_DIMS),
                ("m1", C.c_void_p),
                ("m2", C.c_void_p),
                ("rows", C.c_int64)]


MAX_TERMS = 8
MAX_IN_VALUES = 4096
TERM_RANGE, TERM_IN, TERM_HOP = range(3)
AND, OR = range(2)
GROUP_NONE, GROUP_COLUMN, GROUP_HOP = range(3)


class Term(C.Structure):
    _fields_ = [("kind", C.c_int32), ("combine", C.c_int32), ("negate", C.c_int32),
                ("lo", C.c_int32), ("hi", C.c_int32), ("nvalues", C.c_int32),
                ("parent_code_bytes", C.c_int32), ("key_bias", C.c_int32),
                ("col", C.c_void_p), ("values", C.c_void_p), ("parent", C.c_void_p),
                ("parent_rows", C.c_int64)]


class DimProgram(C.Structure):
    """olap_dim_program. Build with DimProgram.of(terms, group=...): terms are
    ("range", col, lo, hi), ("in", col, values), ("hop", col, parent_vector,
    key_bias) tuples, each optionally preceded by "and"/"or" (combine with the
    terms before, default "and") and "not" (negate)."""
    _fields_ = [("nterms", C.c_int32), ("group", C.c_int32), ("group_mod", C.c_int32),
                ("group_term", C.c_int32), ("group_col", C.c_void_p),
                ("term", Term * MAX_TERMS)]

    @classmethod
    def of(cls, terms, group=GROUP_NONE, group_col=None, group_mod=0, group_term=0):
        import numpy as np
        p = cls()
        p.nterms = len(terms)
        p.group, p.group_mod, p.group_term = group, group_mod, group_term
        p.group_col = _ptr(group_col)
        keep = [group_col]
        for i, t in enumerate(terms):
            t = list(t)
            comb, neg = AND, 0
            while t[0] in ("and", "or", "not"):
                w = t.pop(0)
                if w == "not":
                    neg = 1
                else:
                    comb = AND if w == "and" else OR
            x = p.term[i]
            x.combine, x.negate = comb, neg
            kind = t[0]
            x.col = _ptr(t[1])
            keep.append(t[1])
            if kind == "range":
                x.kind, x.lo, x.hi = TERM_RANGE, int(t[2]), int(t[3])
            elif kind == "in":
                v = np.ascontiguousarray(t[2], dtype=np.int32)
                keep.append(v)
                x.kind, x.values, x.nvalues = TERM_IN, v.ctypes.data, v.size
            elif kind == "hop":
                parent = t[2]
                keep.append(parent)
                x.kind, x.parent = TERM_HOP, _ptr(parent)
                x.parent_rows = parent.numel()
                x.parent_code_bytes = parent.element_size()
                x.key_bias = int(t[3]) if len(t) > 3 else 0
            else:
                raise ValueError(f"unknown term kind {kind}")
        p._keep = keep
        return p


class TriJoin(C.Structure):
    _fields_ = [("rows", C.c_int64), ("key", C.c_void_p * 2), ("key_bias", C.c_int32 * 2),
                ("dimvec", C.c_void_p * 2), ("dim_rows", C.c_int64 * 2), ("code_bytes", C.c_int32 * 2),
                ("price", C.c_void_p), ("discount", C.c_void_p), ("group_nums", C.c_int32)]


class BatchQuery(C.Structure):
    _fields_ = [("ndims", C.c_int32),
                ("table", C.c_int32 * MAX_DIMS),
                ("lo", C.c_int32 * MAX_DIMS),
                ("hi", C.c_int32 * MAX_DIMS),
                ("group_mod", C.c_int32 * MAX_DIMS),
                ("factor", C.c_int32 * MAX_DIMS),
                ("group_nums", C.c_int64),
                ("program", C.POINTER(DimProgram) * MAX_DIMS)]

    @classmethod
    def make(cls, steps, factor, group_nums, programs=None):
        """steps: [(table, lo, hi, group_mod)] in probe order; programs:
        optional DimProgram (or None) per step, ANDed with the range."""
        q = cls()
        q.ndims = len(steps)
        for i, (t, lo, hi, g) in enumerate(steps):
            q.table[i], q.lo[i], q.hi[i], q.group_mod[i] = t, lo, hi, g
            q.factor[i] = factor[i]
        q.group_nums = group_nums
        if programs:
            q._programs = list(programs)
            for i, pr in enumerate(programs):
                if pr is not None:
                    q.program[i] = C.pointer(pr)
        return q

    def as_dict(self):
        n = self.ndims
        return {"steps": [(self.table[i], self.lo[i], self.hi[i], self.group_mod[i])
                          for i in range(n)],
                "factor": list(self.factor)[:n], "group_nums": self.group_nums}


class Star(C.Structure):
    _fields_ = [("filter", C.c_void_p * MAX_DIMS),
                ("attr", C.c_void_p * MAX_DIMS),
                ("dim_rows", C.c_int64 * MAX_DIMS),
                ("fk", C.c_void_p * MAX_DIMS),
                ("m1", C.c_void_p),
                ("m2", C.c_void_p),
                ("rows", C.c_int64)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int32),
                ("streams", C.c_int32),
                ("memory_budget", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("rejected", C.c_int64),
                ("cancelled", C.c_int64), ("batches", C.c_int64),
                ("mean_batch", C.c_double),
                ("p50_us", C.c_double), ("p99_us", C.c_double), ("max_us", C.c_double),
                ("mean_us", C.c_double),
                ("bytes_in_use", C.c_int64), ("bytes_peak", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)

# olap_server_exec_fn / olap_server_alloc_fn / olap_server_free_fn
SERVER_EXEC_FN = C.CFUNCTYPE(C.c_int, _P, C.c_int32, C.POINTER(BatchQuery), C.c_int32, _I64P, _I64P,
                             _I64P)
SERVER_ALLOC_FN = C.CFUNCTYPE(_P, _P, C.c_size_t)
SERVER_FREE_FN = C.CFUNCTYPE(None, _P, _P)

# name -> (restype, argtypes); the export test checks this list against the header
SIGNATURES = {
    "olap_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "olap_ctx_destroy": (C.c_int, [_P]),
    "olap_last_error": (C.c_char_p, [_P]),
    "olap_device_info": (C.c_int, [_P, C.POINTER(C.c_int32), _I64P, C.c_char_p, C.c_size_t]),
    "olap_dev_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "olap_dev_free": (C.c_int, [_P, _P]),
    "olap_copy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t, _P]),
    "olap_copy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t, _P]),
    "olap_stream_sync": (C.c_int, [_P, _P]),
    "olap_gen_uniform_i32": (C.c_int, [_P, _P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64,
                                       C.c_int32, C.c_int32, _P]),
    "olap_zipf_table": (C.c_int, [C.c_int64, C.c_double, _P]),
    "olap_gen_zipf_i32": (C.c_int, [_P, _P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, _P,
                                    C.c_int64, _P]),
    "olap_make_plan": (C.c_int, [_P, _P, _P, C.POINTER(Plan)]),
    "olap_build_dimvec": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int32, _P, _P]),
    "olap_build_dimvec_program": (C.c_int, [_P, C.c_int64, C.POINTER(DimProgram), C.c_int32, _P,
                                            _P]),
    "olap_dict_encode": (C.c_int, [_P, C.c_int64, _P, _I64P]),
    "olap_dict_code_of": (C.c_int32, [_P, _P, C.c_int64, C.c_char_p]),
    "olap_query_create": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), C.POINTER(_P)]),
    "olap_query_run": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "olap_query_destroy": (C.c_int, [_P]),
    "olap_query_device_bytes": (C.c_int64, [_P]),
    "olap_query_launch_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]),
    "olap_num_shapes": (C.c_int32, [C.c_int]),
    "olap_shape_info": (C.c_int, [C.c_int, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "olap_shape_pipeline": (C.c_int32, [C.c_int, C.c_int32]),
    "olap_query_pipeline": (C.c_int32, [C.c_void_p]),
    "olap_query_set_shape": (C.c_int, [_P, C.c_int32]),
    "olap_query_set_aggregation": (C.c_int, [_P, C.c_int32]),
    "olap_query_autotune": (C.c_int, [_P, C.c_int, _P, _P, _P, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double)]),
    "olap_rowwise": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), _P, _P, _I64P, _P]),
    "olap_vwm_dv": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), _P, _P, _I64P, _P]),
    "olap_vwm_sv": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), _P, _P, _I64P, _P]),
    "olap_cwm_dv": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), _P, _P, _I64P, _P]),
    "olap_cwm_sv": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), _P, _P, _I64P, _P]),
    "olap_rowwise_noalloc": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(Columns), _P, _P, _I64P,
                                       _P]),
    "olap_join_cwm_sv": (C.c_int, [_P, _P, _P]),
    "olap_agg_cwm_sv": (C.c_int, [_P, _P, _P, _P, _P]),
    "olap_join_cwm_dv": (C.c_int, [_P, _P]),
    "olap_agg_cwm_dv": (C.c_int, [_P, _P, _P, _P]),
    "olap_batch_create": (C.c_int, [_P, _P, _P, C.c_int32, C.POINTER(_P)]),
    "olap_batch_offsets": (C.c_int, [_P, _P]),
    "olap_batch_run": (C.c_int, [_P, _P, _P, _P]),
    "olap_batch_num_configs": (C.c_int32, []),
    "olap_batch_config_form": (C.c_int32, [C.c_int32]),
    "olap_batch_config_info": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "olap_batch_set_config": (C.c_int, [_P, C.c_int32]),
    "olap_batch_autotune": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "olap_batch_graph_capture": (C.c_int, [_P, _P, _P, _P]),
    "olap_batch_graph_launch": (C.c_int, [_P, _P]),
    "olap_batch_expected_events": (C.c_double, [_P]),
    "olap_batch_destroy": (C.c_int, [_P]),
    "olap_server_create": (C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "olap_server_create_host": (C.c_int, [C.POINTER(ServerConfig), SERVER_EXEC_FN, _P,
                                          C.POINTER(_P)]),
    "olap_server_set_allocator": (C.c_int, [_P, SERVER_ALLOC_FN, SERVER_FREE_FN, _P]),
    "olap_server_last_error": (C.c_char_p, [_P]),
    "olap_server_add_dataset": (C.c_int, [_P, C.POINTER(Star), C.POINTER(C.c_int32)]),
    "olap_server_submit": (C.c_int, [_P, C.c_int32, C.POINTER(BatchQuery), _P, _P,
                                     C.POINTER(C.c_uint64)]),
    "olap_server_wait": (C.c_int, [_P, C.c_uint64, C.c_int64, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double)]),
    "olap_server_pause": (C.c_int, [_P, C.c_int32]),
    "olap_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "olap_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "olap_server_reset_stats": (C.c_int, [_P]),
    "olap_server_destroy": (C.c_int, [_P]),
    "olap_server_request_bytes": (C.c_int64, [C.POINTER(Star), C.POINTER(BatchQuery)]),
    "olap_server_plan": (C.c_int32, [_P, _P, C.c_int32, C.c_int64, C.c_int32, _P]),
    "olap_legacy_count": (C.c_int32, [_P, C.c_int64]),
    "olap_legacy_u32": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "olap_starjoin": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P, _P, _P, _I64P, _P]),
    "olap_vector_group": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, C.c_int64, _I64P, _I64P,
                                    _P]),
    "olap_agg_expr": (C.c_int, [_P, C.c_int64, _P, _P, _P, _I64P, _P]),
    "olap_tri_table_join": (C.c_int, [_P, C.POINTER(TriJoin), _P, _P, _P]),
    "olap_join_build": (C.c_int, [_P, C.c_int64, _P, _P, _P, C.c_int64, _I64P, _P]),
    "olap_join_probe": (C.c_int, [_P, C.c_int64, _P, _P, C.c_int64, _I64P, _I64P, _P]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libolapcore.so and declare signatures. Raises when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: run `python -m memory_database_amd.build` first")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(t) -> Optional[int]:
    """Address of a torch tensor's or numpy array's data."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def make_plan(groups: Sequence[int], bitmap: Sequence[int], sele: Sequence[float]) -> Plan:
    """sort_by_rate + factor plan (host code in the library)."""
    lib = load()
    g = (C.c_int32 * 4)(*groups)
    b = (C.c_int32 * 4)(*bitmap)
    s = (C.c_double * 4)(*sele)
    p = Plan()
    st = lib.olap_make_plan(g, b, s, C.byref(p))
    if st != OK:
        raise OlapError(st, "make_plan rejected its arguments")
    return p


def plan_from(orders: Sequence[int], factor: Sequence[int], group_nums: int) -> Plan:
    p = Plan()
    p.dimvec_nums = len(orders)
    for i, o in enumerate(orders):
        p.orders[i] = o
        p.factor[i] = factor[i]
    p.group_nums = group_nums
    return p


def server_request_bytes(star: Star, query: BatchQuery) -> int:
    """Device bytes the server charges a request against its memory budget."""
    return int(load().olap_server_request_bytes(C.byref(star), C.byref(query)))


def server_plan(datasets, request_bytes, budget_left: int, max_batch: int):
    """The batch an idle server worker admits from pending requests (arrival
    order): indices into the pending list."""
    import numpy as np
    ds = np.ascontiguousarray(datasets, dtype=np.int32)
    by = np.ascontiguousarray(request_bytes, dtype=np.int64)
    out = np.zeros(max(len(ds), 1), dtype=np.int32)
    n = load().olap_server_plan(ds.ctypes.data, by.ctypes.data, len(ds), budget_left, max_batch,
                                out.ctypes.data)
    return [int(i) for i in out[:n]]


def legacy_count(sums) -> int:
    """The reference driver's int return value from int64 sums (host array)."""
    import numpy as np
    a = np.ascontiguousarray(np.asarray(sums, dtype=np.int64))
    return int(load().olap_legacy_count(a.ctypes.data, a.size))


def config_info(cfg: int):
    """(threads, rows per thread, LDS KiB for query tables, aggregation: 0 atomics,
    1 two-pass staged, 2 one-pass staged) of a batch config."""
    th, rp, kib, st = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    load().olap_batch_config_info(cfg, C.byref(th), C.byref(rp), C.byref(kib), C.byref(st))
    return th.value, rp.value, kib.value, st.value


def config_kernel(cfg: int) -> str:
    """Name of the scan kernel a batch config runs (k_batch, k_batch_ev or
    k_batch_seg)."""
    form = int(load().olap_batch_config_form(cfg))
    if form < 0:
        raise ValueError(f"no batch config {cfg}")
    return ("k_batch", "k_batch_ev", "k_batch_seg")[form]


def dict_encode(strings):
    """olap_dict_encode: int32 codes (rank among the distinct strings in byte
    order) and the number of distinct strings."""
    import numpy as np
    enc = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
    arr = (C.c_char_p * max(len(enc), 1))(*enc)
    codes = np.zeros(max(len(enc), 1), dtype=np.int32)
    nd = C.c_int64()
    st = load().olap_dict_encode(arr, len(enc), codes.ctypes.data, C.byref(nd))
    if st != OK:
        raise OlapError(st, "dict_encode")
    return codes[: len(enc)], nd.value


def dict_code_of(strings, codes, value: str) -> int:
    import numpy as np
    enc = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
    arr = (C.c_char_p * max(len(enc), 1))(*enc)
    c = np.ascontiguousarray(codes, dtype=np.int32)
    return int(load().olap_dict_code_of(arr, c.ctypes.data, len(enc), value.encode()))


def zipf_table(d: int, theta: float):
    import numpy as np
    thr = np.zeros(d, dtype=np.uint64)
    st = load().olap_zipf_table(d, theta, thr.ctypes.data)
    if st != OK:
        raise OlapError(st, "zipf_table")
    return thr


class Context:
    """One olap_ctx on a device. Not thread-safe (one per host thread)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = _P()
        st = self.lib.olap_ctx_create(device, C.byref(h))
        if st != OK:
            raise OlapError(st, "olap_ctx_create failed (no GPU?)")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.olap_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int, what: str):
        if st != OK:
            msg = self.lib.olap_last_error(self.h)
            raise OlapError(st, f"{what}: {msg.decode() if msg else ''}")

    def info(self):
        n = C.c_int32()
        b = C.c_int64()
        arch = C.create_string_buffer(64)
        self.check(self.lib.olap_device_info(self.h, C.byref(n), C.byref(b), arch, 64), "info")
        return {"num_cus": n.value, "hbm_bytes": b.value, "arch": arch.value.decode()}

    # ---- data --------------------------------------------------------
    def gen_uniform(self, out, seed: int, stream_id: int, lo: int, hi: int,
                    first_row: int = 0, stream=None):
        self.check(self.lib.olap_gen_uniform_i32(self.h, _ptr(out), out.numel(), first_row,
                                                 seed, stream_id, lo, hi, stream),
                   "gen_uniform")
        return out

    def gen_zipf(self, out, seed: int, stream_id: int, thr_dev, first_row: int = 0,
                 stream=None):
        self.check(self.lib.olap_gen_zipf_i32(self.h, _ptr(out), out.numel(), first_row, seed,
                                              stream_id, _ptr(thr_dev), thr_dev.numel(), stream),
                   "gen_zipf")
        return out

    def build_dimvec(self, filter_col, group_col, lo: int, hi: int, bitmap: int, out,
                     group_mod: int = 0, stream=None):
        cb = out.element_size()
        self.check(self.lib.olap_build_dimvec(self.h, _ptr(filter_col), _ptr(group_col),
                                              filter_col.numel(), lo, hi, bitmap, group_mod, cb,
                                              _ptr(out), stream), "build_dimvec")
        return out

    def build_dimvec_program(self, prog: "DimProgram", rows: int, out, stream=None):
        """Dimension vector of a general predicate program (int8 or int32 out)."""
        self.check(self.lib.olap_build_dimvec_program(self.h, rows, C.byref(prog),
                                                      out.element_size(), _ptr(out), stream),
                   "build_dimvec_program")
        return out

    # ---- star join ---------------------------------------------------
    @staticmethod
    def columns(dimvecs, fks, m1, m2=None, rows: Optional[int] = None) -> Columns:
        """dimvecs/fks: sequences of 4 (tensor or None) indexed by table slot."""
        c = Columns()
        for t in range(MAX_DIMS):
            dv = dimvecs[t]
            if dv is not None:
                c.dimvec[t] = dv.data_ptr()
                c.dim_rows[t] = dv.numel()
                c.code_bytes[t] = dv.element_size()
            else:
                c.code_bytes[t] = 1
            if fks[t] is not None:
                c.fk[t] = fks[t].data_ptr()
                if rows is None:
                    rows = fks[t].numel()
        c.m1 = _ptr(m1)
        c.m2 = _ptr(m2)
        c.rows = rows if rows is not None else (m1.numel() if m1 is not None else 0)
        # the struct holds raw device pointers: keep the tensors alive with it
        c._keep = (list(dimvecs), list(fks), m1, m2)
        return c

    def query(self, plan: Plan, cols: Columns) -> "Query":
        return Query(self, plan, cols)

    @staticmethod
    def star(filters, attrs, fks, m1, m2=None, rows: Optional[int] = None) -> Star:
        s = Star()
        for t in range(MAX_DIMS):
            if filters[t] is not None:
                s.filter[t] = _ptr(filters[t])
                s.dim_rows[t] = filters[t].shape[0]
            if attrs[t] is not None:
                s.attr[t] = _ptr(attrs[t])
            if fks[t] is not None:
                s.fk[t] = _ptr(fks[t])
                if rows is None:
                    rows = fks[t].shape[0]
        s.m1 = _ptr(m1)
        s.m2 = _ptr(m2)
        s.rows = rows if rows is not None else 0
        s._keep = (list(filters), list(attrs), list(fks), m1, m2)
        return s

    def server(self, max_batch: int = 64, streams: int = 2,
               memory_budget: int = 8 << 30) -> "Server":
        return Server(self, max_batch, streams, memory_budget)

    def batch(self, star: Star, queries) -> "Batch":
        return Batch(self, star, queries)

    def run(self, strategy: str, plan: Plan, cols: Columns, out_sum, out_cnt, stream=None):
        """One-shot reference-named entry point (or "rowwise_noalloc", the
        allocation-free row-wise form); returns the total count."""
        fn = getattr(self.lib, "olap_" + strategy)
        tot = C.c_int64()
        self.check(fn(self.h, C.byref(plan), C.byref(cols), _ptr(out_sum), _ptr(out_cnt),
                      C.byref(tot), stream), strategy)
        return tot.value

    # ---- sub-steps ---------------------------------------------------
    def starjoin(self, strategy: str, dimvecs, fks, result=None, stream=None) -> int:
        dv = (_P * 4)(*[d.data_ptr() for d in dimvecs])
        dr = (C.c_int64 * 4)(*[d.numel() for d in dimvecs])
        fk = (_P * 4)(*[f.data_ptr() for f in fks])
        cnt = C.c_int64()
        self.check(self.lib.olap_starjoin(self.h, STRATEGIES[strategy], fks[0].numel(), dv, dr,
                                          fk, _ptr(result), C.byref(cnt), stream), "starjoin")
        return cnt.value

    def vector_group(self, key1, m1, m2, group_vector, stream=None):
        tot = C.c_int64()
        bad = C.c_int64()
        self.check(self.lib.olap_vector_group(self.h, key1.numel(), _ptr(key1), _ptr(m1),
                                              _ptr(m2), _ptr(group_vector), group_vector.numel(),
                                              C.byref(tot), C.byref(bad), stream),
                   "vector_group")
        return tot.value, bad.value

    def join_build(self, key, payload, vector, stream=None) -> int:
        """vecjoin build into the int32 device vector; returns the keys
        outside it."""
        bad = C.c_int64()
        self.check(self.lib.olap_join_build(self.h, key.numel(), _ptr(key), _ptr(payload),
                                            _ptr(vector), vector.numel(), C.byref(bad), stream),
                   "join_build")
        return bad.value

    def join_probe(self, key, vector, stream=None):
        """(int64 sum of vector[key], keys outside the vector)."""
        tot = C.c_int64()
        bad = C.c_int64()
        self.check(self.lib.olap_join_probe(self.h, key.numel(), _ptr(key), _ptr(vector),
                                            vector.numel(), C.byref(tot), C.byref(bad), stream),
                   "join_probe")
        return tot.value, bad.value

    def tri_table_join(self, supp_key, supp_vec, order_key, order_vec, price, discount,
                       group_nums, out_sum, out_cnt, supp_bias=-1, order_bias=0, stream=None):
        """TPC-H Q5's fused snowflake probe into out_sum / out_cnt (int64,
        group_nums each): revenue in cents x percent."""
        j = TriJoin()
        j.rows = supp_key.numel()
        j.key[0], j.key[1] = _ptr(supp_key), _ptr(order_key)
        j.key_bias[0], j.key_bias[1] = supp_bias, order_bias
        j.dimvec[0], j.dimvec[1] = _ptr(supp_vec), _ptr(order_vec)
        j.dim_rows[0], j.dim_rows[1] = supp_vec.numel(), order_vec.numel()
        j.code_bytes[0], j.code_bytes[1] = supp_vec.element_size(), order_vec.element_size()
        j.price, j.discount = _ptr(price), _ptr(discount)
        j.group_nums = group_nums
        self.check(self.lib.olap_tri_table_join(self.h, C.byref(j), _ptr(out_sum), _ptr(out_cnt),
                                                stream), "tri_table_join")

    def agg_expr(self, p, q, t, stream=None) -> int:
        r = C.c_int64()
        self.check(self.lib.olap_agg_expr(self.h, p.numel(), _ptr(p), _ptr(q), _ptr(t),
                                          C.byref(r), stream), "agg_expr")
        return r.value

    def legacy_u32(self, sums, out, stream=None):
        """The reference's uint32 group vector from int64 sums (device)."""
        self.check(self.lib.olap_legacy_u32(self.h, _ptr(sums), sums.numel(), _ptr(out), stream),
                   "legacy_u32")
        return out

    def sync(self, stream=None):
        self.check(self.lib.olap_stream_sync(self.h, stream), "sync")


class Query:
    """olap_query: validated plan + packed dimension vectors."""

    def __init__(self, ctx: Context, plan: Plan, cols: Columns):
        self.ctx = ctx
        self.plan = plan
        self.cols = cols
        h = _P()
        self.h = None
        ctx.check(ctx.lib.olap_query_create(ctx.h, C.byref(plan), C.byref(cols), C.byref(h)),
                  "query_create")
        self.h = h

    def run(self, strategy: str, out_sum, out_cnt, stream=None):
        self.ctx.check(self.ctx.lib.olap_query_run(self.h, STRATEGIES[strategy], _ptr(out_sum),
                                                   _ptr(out_cnt), stream), "query_run")

    def join_cwm_sv(self, gid_vec, stream=None):
        self.ctx.check(self.ctx.lib.olap_join_cwm_sv(self.h, _ptr(gid_vec), stream), "join_cwm_sv")

    def agg_cwm_sv(self, gid_vec, out_sum, out_cnt, stream=None):
        self.ctx.check(self.ctx.lib.olap_agg_cwm_sv(self.h, _ptr(gid_vec), _ptr(out_sum),
                                                    _ptr(out_cnt), stream), "agg_cwm_sv")

    def join_cwm_dv(self, stream=None):
        self.ctx.check(self.ctx.lib.olap_join_cwm_dv(self.h, stream), "join_cwm_dv")

    def agg_cwm_dv(self, out_sum, out_cnt, stream=None):
        self.ctx.check(self.ctx.lib.olap_agg_cwm_dv(self.h, _ptr(out_sum), _ptr(out_cnt), stream),
                       "agg_cwm_dv")

    def autotune(self, strategy, out_sum, out_cnt, stream=None):
        """Returns (shape index, threads, rows per thread, median ms)."""
        sh = C.c_int32()
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.olap_query_autotune(self.h, STRATEGIES[strategy], _ptr(out_sum),
                                                        _ptr(out_cnt), stream, C.byref(sh),
                                                        C.byref(ms)), "autotune")
        th, rp = C.c_int32(), C.c_int32()
        self.ctx.lib.olap_shape_info(STRATEGIES[strategy], sh.value, C.byref(th), C.byref(rp))
        return sh.value, th.value, rp.value, ms.value

    def set_shape(self, shape: int):
        self.ctx.check(self.ctx.lib.olap_query_set_shape(self.h, shape), "set_shape")

    def set_aggregation(self, layout: int):
        """-1 automatic, 0 global atomics, 1 staged (two passes), 2 global
        atomics behind an LDS cache of hot groups, 3 one-pass staged (group
        tables beyond LDS)."""
        self.ctx.check(self.ctx.lib.olap_query_set_aggregation(self.h, layout), "set_aggregation")

    def launch_info(self):
        """(threads, rows per thread, aggregation name) of the last run."""
        th, rp, ag = C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx.check(self.ctx.lib.olap_query_launch_info(self.h, C.byref(th), C.byref(rp),
                                                           C.byref(ag)), "launch_info")
        return th.value, rp.value, AGGREGATIONS[ag.value]

    def pipeline(self):
        """How the last row-wise run fetched: dependent fetches one tile
        ahead, and the second step's keys streamed with the first's."""
        p = int(self.ctx.lib.olap_query_pipeline(self.h))
        if p < 0:
            raise OlapError(E_INVALID, "pipeline: the query has not run")
        return {"fetch_ahead": bool(p & 1), "survivor_queue": bool(p & 4),
                "second_step_streamed": bool(p & 2)}

    def device_bytes(self) -> int:
        return int(self.ctx.lib.olap_query_device_bytes(self.h))

    def close(self):
        if self.h:
            self.ctx.lib.olap_query_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """olap_batch: up to 64 queries evaluated by one fused scan."""

    def __init__(self, ctx: Context, star: Star, queries):
        self.ctx = ctx
        self.star = star
        self.queries = list(queries)
        arr = (BatchQuery * len(self.queries))(*self.queries)
        self._arr = arr
        h = _P()
        self.h = None
        ctx.check(ctx.lib.olap_batch_create(ctx.h, C.byref(star), arr, len(self.queries),
                                            C.byref(h)), "batch_create")
        self.h = h
        import numpy as np
        off = np.zeros(len(self.queries) + 1, dtype=np.int64)
        ctx.check(ctx.lib.olap_batch_offsets(self.h, off.ctypes.data), "batch_offsets")
        self.offsets = off

    @property
    def total_slots(self) -> int:
        return int(self.offsets[-1])

    def expected_events(self) -> float:
        return float(self.ctx.lib.olap_batch_expected_events(self.h))

    def run(self, out_sum, out_cnt, stream=None):
        self.ctx.check(self.ctx.lib.olap_batch_run(self.h, _ptr(out_sum), _ptr(out_cnt), stream),
                       "batch_run")

    def set_config(self, cfg: int):
        self.ctx.check(self.ctx.lib.olap_batch_set_config(self.h, cfg), "batch_set_config")

    def autotune(self, out_sum, out_cnt, stream=None):
        """Returns (config, threads, rows per thread, LDS KiB for tables,
        staged, ms)."""
        cfg = C.c_int32()
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.olap_batch_autotune(self.h, _ptr(out_sum), _ptr(out_cnt), stream,
                                                        C.byref(cfg), C.byref(ms)), "batch_autotune")
        return (cfg.value, *config_info(cfg.value), ms.value)

    def capture(self, out_sum, out_cnt, stream):
        self.ctx.check(self.ctx.lib.olap_batch_graph_capture(self.h, _ptr(out_sum), _ptr(out_cnt),
                                                             stream), "batch_graph_capture")

    def launch(self, stream):
        self.ctx.check(self.ctx.lib.olap_batch_graph_launch(self.h, stream), "batch_graph_launch")

    def close(self):
        if self.h:
            self.ctx.lib.olap_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """olap_server: continuous batching of submitted queries over resident
    star schemas. submit() returns a ticket; wait() returns (sum, count) as
    int64 numpy arrays of group_nums entries.

    Server(ctx, ...) runs batches with the library's fused GPU scan;
    Server.host(executor, ...) runs them with a Python executor
    `executor(dataset, queries, offsets, sum, count) -> status` over numpy
    arrays (no device involved)."""

    def __init__(self, ctx: Optional[Context], max_batch: int = 64, streams: int = 2,
                 memory_budget: int = 8 << 30, _executor=None):
        self.lib = load()
        self.ctx = ctx
        self.h = None
        self._stars = []
        self._out = {}
        self._keep = []
        cfg = ServerConfig(max_batch, streams, memory_budget)
        h = _P()
        if _executor is None:
            ctx.check(self.lib.olap_server_create(ctx.h, C.byref(cfg), C.byref(h)), "server_create")
        else:
            cb = SERVER_EXEC_FN(_executor)
            self._keep.append(cb)
            st = self.lib.olap_server_create_host(C.byref(cfg), cb, None, C.byref(h))
            if st != OK:
                raise OlapError(st, "server_create_host: max_batch 1..64, streams 1..8, budget > 0")
        self.h = h

    @classmethod
    def host(cls, executor, max_batch: int = 64, streams: int = 2, memory_budget: int = 8 << 30):
        import numpy as np

        def trampoline(user, dataset, queries, n, offsets, osum, ocnt):
            try:
                off = np.ctypeslib.as_array(offsets, shape=(n + 1,)).copy()
                total = max(int(off[-1]), 1)
                s = np.ctypeslib.as_array(osum, shape=(total,))
                c = np.ctypeslib.as_array(ocnt, shape=(total,))
                qs = [queries[i] for i in range(n)]
                return int(executor(int(dataset), qs, off, s, c))
            except Exception:  # an executor exception fails its batch
                return E_INVALID

        return cls(None, max_batch, streams, memory_budget, _executor=trampoline)

    def check(self, st: int, what: str):
        if st != OK:
            msg = self.lib.olap_server_last_error(self.h) if self.h else b""
            raise OlapError(st, f"{what}: {msg.decode() if msg else ''}")

    def set_allocator(self, alloc, free):
        """alloc(bytes) -> address or 0; free(address)."""
        a = SERVER_ALLOC_FN(lambda user, n: alloc(int(n)) or None)
        f = SERVER_FREE_FN(lambda user, p: free(p))
        self._keep += [a, f]
        self.check(self.lib.olap_server_set_allocator(self.h, a, f, None), "server_set_allocator")

    def add_dataset(self, star: Star) -> int:
        i = C.c_int32()
        self.check(self.lib.olap_server_add_dataset(self.h, C.byref(star), C.byref(i)),
                   "server_add_dataset")
        self._stars.append(star)
        return i.value

    def submit(self, dataset: int, query: BatchQuery) -> int:
        import numpy as np
        g = max(int(query.group_nums), 1)
        s = np.empty(g, dtype=np.int64)
        c = np.empty(g, dtype=np.int64)
        t = C.c_uint64()
        self.check(self.lib.olap_server_submit(self.h, dataset, C.byref(query), s.ctypes.data,
                                               c.ctypes.data, C.byref(t)), "server_submit")
        self._out[t.value] = (s, c, query)
        return t.value

    def wait(self, ticket: int, timeout_us: int = -1, with_latency: bool = False):
        """(sum, count), plus the submit-to-result latency in us when asked.
        OlapError(E_BUSY) on timeout; the request's own error otherwise."""
        st = C.c_int32()
        lat = C.c_double()
        self.check(self.lib.olap_server_wait(self.h, ticket, timeout_us, C.byref(st), C.byref(lat)),
                   "server_wait")
        s, c, _ = self._out.pop(ticket)
        if st.value != OK:
            raise OlapError(st.value, f"request {ticket} failed in its batch")
        return (s, c, lat.value) if with_latency else (s, c)

    def pause(self, paused: bool = True):
        self.check(self.lib.olap_server_pause(self.h, int(paused)), "server_pause")

    def cancel(self, ticket: int):
        self.check(self.lib.olap_server_cancel(self.h, ticket), "server_cancel")
        self._out.pop(ticket, None)

    def stats(self) -> dict:
        st = ServerStats()
        self.check(self.lib.olap_server_get_stats(self.h, C.byref(st)), "server_stats")
        return st.as_dict()

    def reset_stats(self):
        self.check(self.lib.olap_server_reset_stats(self.h), "server_reset_stats")

    def close(self):
        if self.h:
            self.lib.olap_server_destroy(self.h)
            self.h = None
            self._out.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
