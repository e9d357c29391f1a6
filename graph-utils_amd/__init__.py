"""Python host binding of the MI355X multi-source search library.

A thin ctypes layer over ``include/graph_utils_amd.h`` (the C ABI built into
``libgraph_utils_amd.so``).  It mirrors the reference's host objects:

* :class:`HostGraph` -- ``host_graph`` (parse.h:15-35): R/C/F int32 arrays,
  ``parse_metis`` (parse.cpp:115-208) and the seeded generators;
* :class:`DeviceGraph` -- ``device_graph`` (device_graph.h:7-22) plus the
  multi-search entry points (``multi_search_*_setup``,
  ``count_shortest_paths_setup``, ``betweenness_centrality_setup``,
  ``diameter_sampling_setup``);
* :func:`lbs` -- ``extract_edges_block`` (load_balanced_search.cu:199-202).

Every call goes to the HIP library; there is no CPU fallback.  Importing works
without a GPU (host-only functions such as the parser and generators run);
device calls raise :class:`GuError` when no GPU is present.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraph_utils_amd.so")

OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_RANGE, ERR_PARSE = -1, -2, -3, -4, -5
ERR_WEIGHTED, ERR_IO, ERR_NODEVICE, ERR_CAPACITY = -6, -7, -8, -9

VARIANT_LINEAR = 0
VARIANT_EDGE_PARALLEL = 1
VARIANT_WARP = 2
VARIANT_SCAN = 3
VARIANT_CTA_WARP = 4
VARIANT_SHUFFLE = 5
VARIANT_AUTO = 6
VARIANTS = {
    "linear": VARIANT_LINEAR,
    "edge_parallel": VARIANT_EDGE_PARALLEL,
    "warp": VARIANT_WARP,
    "scan": VARIANT_SCAN,
    "cta_warp": VARIANT_CTA_WARP,
    "shuffle": VARIANT_SHUFFLE,
    "auto": VARIANT_AUTO,
}
DIR_AUTO, DIR_TOPDOWN, DIR_BOTTOMUP = 0, 1, 2
ENGINE_AUTO, ENGINE_BATCHED, ENGINE_PER_SOURCE = 0, 1, 2

INT_MAX = 2147483647

# Names the C ABI must export (checked by the CPU tests against include/*.h).
EXPORTS = (
    "gu_choose_device", "gu_device_count", "gu_timer_start", "gu_timer_stop",
    "gu_malloc", "gu_free", "gu_memcpy_h2d", "gu_memcpy_d2h", "gu_memcpy_d2d", "gu_memset_d",
    "gu_synchronize", "gu_strerror", "gu_graph_create", "gu_graph_destroy",
    "gu_graph_arrays", "gu_graph_device_bytes", "gu_multi_search",
    "gu_betweenness", "gu_diameter", "gu_lbs", "gu_parse_metis",
    "gu_write_metis", "gu_gen_er", "gu_gen_rmat", "gu_gen_grid",
    "gu_gen_zipf_counts", "gu_sample_sources", "gu_host_free", "gu_graph_share",
    "gu_multi_search_rows", "gu_betweenness_rows", "gu_server_create", "gu_server_add_graph",
    "gu_server_submit", "gu_server_wait", "gu_server_next_done", "gu_server_poll",
    "gu_server_release", "gu_server_now_ms", "gu_server_get_stats", "gu_server_destroy",
    "gu_autotune", "gu_parse_market", "gu_parse_snap", "gu_write_edgelist", "gu_reverse",
    "gu_lbs_noalloc", "gu_power_start", "gu_power_stop", "gu_trace_push", "gu_trace_pop",
)

QUERY_BFS, QUERY_BC, QUERY_DIAMETER = 0, 1, 2
QF_SIGMA, QF_DELTA, QF_HOST = 1, 2, 4
QR_SIGMA_WRAPPED = 1  # gu_query_result.flags: a path count wrapped mod 2^64
ERR_CAPACITY, ERR_TIMEOUT, ERR_UNSUPPORTED = -9, -10, -11


class GuError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__("%s: %s (%d)" % (what, _strerror(status), status))


class Config(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int32),
        ("direction", ctypes.c_int32),
        ("batch_width", ctypes.c_int32),
        ("wg_size", ctypes.c_int32),
        ("alpha", ctypes.c_int32),
        ("beta", ctypes.c_int32),
        ("sync_every", ctypes.c_int32),
        ("engine", ctypes.c_int32),
        ("level_graphs", ctypes.c_int32),
        ("source_queue", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("teps_units", ctypes.c_int64),
        ("reached", ctypes.c_int64),
        ("dag_edges", ctypes.c_int64),
        ("max_level", ctypes.c_int32),
        ("batches", ctypes.c_int32),
        ("topdown_levels", ctypes.c_int32),
        ("bottomup_levels", ctypes.c_int32),
        ("sigma_overflow", ctypes.c_int32),
        ("variant_used", ctypes.c_int32),
        ("batch_width_used", ctypes.c_int32),
        ("engine_used", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceProps(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("device", ctypes.c_int32),
        ("compute_units", ctypes.c_int32),
        ("max_threads_per_block", ctypes.c_int32),
        ("wave_size", ctypes.c_int32),
        ("total_mem", ctypes.c_int64),
        ("free_mem", ctypes.c_int64),
        ("mem_bandwidth_gbs", ctypes.c_double),
        ("l2_bytes", ctypes.c_int32),
        ("clock_khz", ctypes.c_int32),
    ]


_lib = None
_P = ctypes.c_void_p
class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("slots", ctypes.c_int32),
                ("max_pack", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("engine", ctypes.c_int32), ("host_budget", ctypes.c_int64)]


class Query(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_int32), ("kind", ctypes.c_int32), ("nsrc", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("sources", ctypes.c_void_p)]


class QueryResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("kind", ctypes.c_int32), ("nsrc", ctypes.c_int32),
                ("diameter", ctypes.c_int32), ("launches", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("dist_dev", ctypes.c_void_p),
                ("sigma_dev", ctypes.c_void_p), ("delta_dev", ctypes.c_void_p),
                ("bc_dev", ctypes.c_void_p), ("submit_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("first_launch_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("hbm_bytes", ctypes.c_int64),
                ("dist_host", ctypes.c_void_p), ("sigma_host", ctypes.c_void_p),
                ("delta_host", ctypes.c_void_p), ("bc_host", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("rejected", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("sources", ctypes.c_int64), ("hbm_in_use", ctypes.c_int64),
                ("hbm_peak", ctypes.c_int64), ("hbm_budget", ctypes.c_int64),
                ("waiting", ctypes.c_int32), ("active", ctypes.c_int32),
                ("host_in_use", ctypes.c_int64), ("host_peak", ctypes.c_int64),
                ("host_budget", ctypes.c_int64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TuneResult(ctypes.Structure):
    _fields_ = [("engine", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("batch_width", ctypes.c_int32), ("wg_size", ctypes.c_int32),
                ("alpha", ctypes.c_int32), ("trials", ctypes.c_int32), ("sources", ctypes.c_int32),
                ("beta", ctypes.c_int32), ("ms_default", ctypes.c_double),
                ("ms_tuned", ctypes.c_double)]

    def as_dict(self):
        d = {f[0]: getattr(self, f[0]) for f in self._fields_}
        d["ms_default"] = round(d["ms_default"], 4)
        d["ms_tuned"] = round(d["ms_tuned"], 4)
        return d


_I32P = ctypes.POINTER(ctypes.c_int32)
_I32PP = ctypes.POINTER(_I32P)
_I64P = ctypes.POINTER(ctypes.c_int64)


def lib():
    """Load libgraph_utils_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libgraph_utils_amd.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "gu_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "gu_device_count": (ctypes.c_int, [_I32P]),
        "gu_choose_device": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(DeviceProps)]),
        "gu_timer_start": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
        "gu_timer_stop": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
        "gu_power_start": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P)]),
        "gu_power_stop": (ctypes.c_int, [_P, ctypes.POINTER(PowerResult)]),
        "gu_trace_push": (ctypes.c_int, [ctypes.c_char_p]),
        "gu_trace_pop": (ctypes.c_int, []),
        "gu_malloc": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_size_t]),
        "gu_free": (ctypes.c_int, [_P]),
        "gu_memcpy_h2d": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
        "gu_memcpy_d2h": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
        "gu_memset_d": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_size_t]),
        "gu_memcpy_d2d": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
        "gu_synchronize": (ctypes.c_int, []),
        "gu_graph_create": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int64,
                                           ctypes.c_int32, ctypes.POINTER(_P)]),
        "gu_graph_destroy": (ctypes.c_int, [_P]),
        "gu_graph_arrays": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                           ctypes.POINTER(_P), _I32P, _I64P]),
        "gu_graph_device_bytes": (ctypes.c_int, [_P, _I64P]),
        "gu_multi_search": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P,
                                           ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_betweenness": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P,
                                          ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_diameter": (ctypes.c_int, [_P, _P, ctypes.c_int32, _I32P,
                                       ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_lbs": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, ctypes.c_int64, _I64P,
                                  ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_lbs_noalloc": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, ctypes.c_int64, _I64P,
                                          ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_parse_metis": (ctypes.c_int, [ctypes.c_char_p, _I32PP, _I32PP, _I32PP, _I32P, _I64P]),
        "gu_write_metis": (ctypes.c_int, [ctypes.c_char_p, _P, _P, ctypes.c_int32, ctypes.c_int64]),
        "gu_parse_market": (ctypes.c_int, [ctypes.c_char_p, _I32PP, _I32PP, _I32PP, _I32P, _I64P]),
        "gu_parse_snap": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, _I32PP, _I32PP, _I32PP,
                                         _I32P, _I64P]),
        "gu_write_edgelist": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, _P, _P, ctypes.c_int32,
                                             ctypes.c_int64]),
        "gu_reverse": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int64, _I32PP, _I32PP, _I32PP]),
        "gu_gen_er": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                                     _I32PP, _I32PP, _I32PP, _I32P, _I64P]),
        "gu_gen_rmat": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                       _I32PP, _I32PP, _I32PP, _I32P, _I64P]),
        "gu_gen_grid": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.c_uint64,
                                       ctypes.c_int32, _I32PP, _I32PP, _I32PP, _I32P, _I64P]),
        "gu_gen_zipf_counts": (ctypes.c_int, [ctypes.c_double, ctypes.c_int32, ctypes.c_int64,
                                              ctypes.c_uint64, _I32PP, _I64P]),
        "gu_sample_sources": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_uint64, _P]),
        "gu_host_free": (None, [_P]),
        "gu_graph_share": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
        "gu_multi_search_rows": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P, _P,
                                                ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_betweenness_rows": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P,
                                               ctypes.POINTER(Config), ctypes.POINTER(Stats)]),
        "gu_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
        "gu_server_add_graph": (ctypes.c_int, [_P, _P, _I32P]),
        "gu_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Query), _I64P]),
        "gu_server_wait": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.POINTER(QueryResult)]),
        "gu_server_next_done": (ctypes.c_int, [_P, ctypes.c_int32, _I64P,
                                               ctypes.POINTER(QueryResult)]),
        "gu_server_poll": (ctypes.c_int, [_P, ctypes.c_int64, _I32P]),
        "gu_server_release": (ctypes.c_int, [_P, ctypes.c_int64]),
        "gu_server_now_ms": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
        "gu_server_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
        "gu_server_destroy": (ctypes.c_int, [_P]),
        "gu_autotune": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(TuneResult)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _strerror(status):
    try:
        return lib().gu_strerror(status).decode()
    except Exception:  # pragma: no cover - only when the library is missing
        return "status"


def _check(rc, what):
    if rc != OK:
        raise GuError(rc, what)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _take(p, count, dtype=np.int32):
    """Copy a malloc'd host array from the library into numpy and free it."""
    if not p:
        return np.zeros(0, dtype=dtype)
    out = np.ctypeslib.as_array(p, shape=(max(count, 0),)).astype(dtype, copy=True) \
        if count > 0 else np.zeros(0, dtype=dtype)
    lib().gu_host_free(ctypes.cast(p, ctypes.c_void_p))
    return out


def make_config(variant="auto", direction=DIR_AUTO, batch_width=0, wg_size=0, alpha=0,
                beta=0, sync_every=0, engine=ENGINE_AUTO, level_graphs=0, source_queue=0):
    v = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    return Config(v, int(direction), int(batch_width), int(wg_size), int(alpha), int(beta),
                  int(sync_every), int(engine), int(level_graphs), int(source_queue), None)


# --------------------------------------------------------------------------
# util_device
# --------------------------------------------------------------------------
def device_count():
    c = ctypes.c_int32(0)
    rc = lib().gu_device_count(ctypes.byref(c))
    return c.value if rc == OK else 0


def choose_device(requested=-1):
    """choose_device (util_device.cu:23-76); returns the device properties."""
    p = DeviceProps()
    _check(lib().gu_choose_device(requested, ctypes.byref(p)), "choose_device")
    return {
        "name": p.name.decode(), "arch": p.arch.decode(), "device": p.device,
        "compute_units": p.compute_units, "max_threads_per_block": p.max_threads_per_block,
        "wave_size": p.wave_size, "total_mem": p.total_mem, "free_mem": p.free_mem,
        "mem_bandwidth_gbs": p.mem_bandwidth_gbs, "l2_bytes": p.l2_bytes,
        "clock_khz": p.clock_khz,
    }


def synchronize():
    _check(lib().gu_synchronize(), "synchronize")


class DeviceBuffer:
    """Device (HBM) allocation owned by Python; freed on close / GC."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().gu_malloc(ctypes.byref(p), max(self.nbytes, 1)), "gu_malloc")
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _check(lib().gu_memcpy_h2d(b.ptr, _ptr(a), a.nbytes), "gu_memcpy_h2d")
        return b

    def to_host(self, dtype, count=None, offset_items=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize - offset_items
        out = np.empty(count, dtype=dtype)
        if count:
            src = ctypes.c_void_p(self.ptr.value + offset_items * dtype.itemsize)
            _check(lib().gu_memcpy_d2h(_ptr(out), src, out.nbytes), "gu_memcpy_d2h")
        return out

    def copy_from(self, other, nbytes=None):
        nb = self.nbytes if nbytes is None else int(nbytes)
        _check(lib().gu_memcpy_d2d(self.ptr, other.ptr, nb), "gu_memcpy_d2d")

    def zero(self):
        _check(lib().gu_memset_d(self.ptr, 0, self.nbytes), "gu_memset_d")

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().gu_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PowerResult(ctypes.Structure):
    _fields_ = [("avg_w", ctypes.c_double), ("max_w", ctypes.c_double),
                ("energy_avg_w", ctypes.c_double), ("energy_j", ctypes.c_double),
                ("seconds", ctypes.c_double), ("samples", ctypes.c_int32)]


class PowerMeasurer:
    """power_measurer (util_device.cuh:47-62, util_device.cu:79-130): the
    device's socket power sampled through AMD SMI every `period_ms` between
    start() and end().  end() returns the mean of the samples in W, or -1 when
    the host cannot read the device's power (the reference's answer on
    boards NVML cannot read); `result` then holds the energy-counter average
    and the sample count, or None."""

    def __init__(self, period_ms, device=0):
        self.period_ms = int(period_ms)
        self.device = int(device)
        self._h = None
        self.result = None
        self.status = None

    def start(self):
        h = ctypes.c_void_p()
        self.status = lib().gu_power_start(self.device, self.period_ms, ctypes.byref(h))
        if self.status == ERR_UNSUPPORTED:
            self._h = None
            return False
        _check(self.status, "gu_power_start")
        self._h = h
        return True

    def end(self):
        if self._h is None:
            self.result = None
            return -1.0
        r = PowerResult()
        h, self._h = self._h, None
        _check(lib().gu_power_stop(h, ctypes.byref(r)), "gu_power_stop")
        self.result = r
        return r.avg_w

    start_power_sample = start
    end_power_sample = end


class TraceRange:
    """A roctx range (gu_trace_push / gu_trace_pop) around a with-block."""

    def __init__(self, message):
        self.message = message.encode()
        self.on = False

    def __enter__(self):
        self.on = lib().gu_trace_push(self.message) == 0
        return self

    def __exit__(self, *exc):
        if self.on:
            lib().gu_trace_pop()
        return False


class Timer:
    """start_clock / end_clock (util_device.cu:4-21) in milliseconds."""

    def __enter__(self):
        self._t = ctypes.c_void_p()
        _check(lib().gu_timer_start(None, ctypes.byref(self._t)), "gu_timer_start")
        return self

    def __exit__(self, *exc):
        ms = ctypes.c_double()
        _check(lib().gu_timer_stop(self._t, ctypes.byref(ms)), "gu_timer_stop")
        self.ms = ms.value
        return False


# --------------------------------------------------------------------------
# host_graph
# --------------------------------------------------------------------------
class HostGraph:
    """host_graph (parse.h:15-35): hybrid CSR (R, C) / COO (F) in int32."""

    def __init__(self, R, C, F=None, directed=False, m=None):
        self.R = np.ascontiguousarray(R, dtype=np.int32)
        self.C = np.ascontiguousarray(C, dtype=np.int32)
        self.n = int(len(self.R) - 1)
        self.m = int(len(self.C)) if m is None else int(m)
        if F is None:
            F = np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.R))
        self.F = np.ascontiguousarray(F, dtype=np.int32)
        self.directed = bool(directed)

    @staticmethod
    def _from_out(R, C, F, n, m):
        Ra = _take(R, n.value + 1)
        Ca = _take(C, m.value)
        Fa = _take(F, m.value)
        return HostGraph(Ra, Ca, Fa, m=m.value)

    @classmethod
    def parse_metis(cls, path):
        R, C, F = _I32P(), _I32P(), _I32P()
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gu_parse_metis(os.fsencode(path), ctypes.byref(R), ctypes.byref(C),
                                    ctypes.byref(F), ctypes.byref(n), ctypes.byref(m)),
               "parse_metis(%s)" % path)
        return cls._from_out(R, C, F, n, m)

    @classmethod
    def parse_market(cls, path):
        """parse_market (parse.cpp:210-332): symmetrised pattern Matrix Market."""
        R, C, F = _I32P(), _I32P(), _I32P()
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gu_parse_market(os.fsencode(path), ctypes.byref(R), ctypes.byref(C),
                                     ctypes.byref(F), ctypes.byref(n), ctypes.byref(m)),
               "parse_market(%s)" % path)
        return cls._from_out(R, C, F, n, m)

    @classmethod
    def parse_snap(cls, path, header=False):
        """parse_snap (parse.cpp:335-434): directed edge list."""
        R, C, F = _I32P(), _I32P(), _I32P()
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gu_parse_snap(os.fsencode(path), 1 if header else 0, ctypes.byref(R),
                                   ctypes.byref(C), ctypes.byref(F), ctypes.byref(n), ctypes.byref(m)),
               "parse_snap(%s)" % path)
        g = cls._from_out(R, C, F, n, m)
        g.directed = True
        return g

    def reverse(self):
        """reverse (graph-utils/reverse.cpp:3-44): the transposed graph."""
        R, C, F = _I32P(), _I32P(), _I32P()
        _check(lib().gu_reverse(_ptr(self.R), _ptr(self.C), self.n, self.m, ctypes.byref(R),
                                ctypes.byref(C), ctypes.byref(F)), "reverse")
        g = HostGraph(_take(R, self.n + 1), _take(C, self.m), _take(F, self.m), directed=self.directed,
                      m=self.m)
        return g

    def write_edgelist(self, path, header=False):
        """host_graph::write_edgelist_to_file (parse.cpp:436-456)."""
        _check(lib().gu_write_edgelist(os.fsencode(path), 1 if header else 0, _ptr(self.F), _ptr(self.C),
                                       self.n, self.m), "write_edgelist")

    @classmethod
    def er(cls, n, m_undirected, seed=0):
        R, C, F = _I32P(), _I32P(), _I32P()
        nn, m = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gu_gen_er(n, m_undirected, seed, ctypes.byref(R), ctypes.byref(C),
                               ctypes.byref(F), ctypes.byref(nn), ctypes.byref(m)), "gen_er")
        return cls._from_out(R, C, F, nn, m)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0):
        R, C, F = _I32P(), _I32P(), _I32P()
        nn, m = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gu_gen_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(R),
                                 ctypes.byref(C), ctypes.byref(F), ctypes.byref(nn),
                                 ctypes.byref(m)), "gen_rmat")
        return cls._from_out(R, C, F, nn, m)

    @classmethod
    def grid(cls, side, p_delete=0.10, seed=0, lcc_only=True):
        R, C, F = _I32P(), _I32P(), _I32P()
        nn, m = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gu_gen_grid(side, p_delete, seed, 1 if lcc_only else 0, ctypes.byref(R),
                                 ctypes.byref(C), ctypes.byref(F), ctypes.byref(nn),
                                 ctypes.byref(m)), "gen_grid")
        return cls._from_out(R, C, F, nn, m)

    def write_metis(self, path):
        _check(lib().gu_write_metis(os.fsencode(path), _ptr(self.R), _ptr(self.C), self.n,
                                    self.m), "write_metis")

    def degrees(self):
        return np.diff(self.R)

    def sample_sources(self, k, seed=0):
        out = np.zeros(k, dtype=np.int32)
        _check(lib().gu_sample_sources(_ptr(self.R), self.n, k, seed, _ptr(out)),
               "sample_sources")
        return out


def zipf_counts(s=1.1, cap=1 << 20, W=1 << 26, seed=0):
    p = _I32P()
    S = ctypes.c_int64()
    _check(lib().gu_gen_zipf_counts(s, cap, W, seed, ctypes.byref(p), ctypes.byref(S)),
           "zipf_counts")
    return _take(p, S.value)


# --------------------------------------------------------------------------
# device_graph and the search entry points
# --------------------------------------------------------------------------
class DeviceGraph:
    """device_graph (device_graph.h:7-22) plus the batched search engine."""

    def __init__(self, hg):
        self.n, self.m = hg.n, hg.m
        h = ctypes.c_void_p()
        _check(lib().gu_graph_create(_ptr(hg.R), _ptr(hg.C), _ptr(hg.F), hg.n, hg.m,
                                     1 if hg.directed else 0, ctypes.byref(h)),
               "gu_graph_create")
        self.h = h

    def close(self):
        if self.h is not None and self.h.value:
            lib().gu_graph_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        b = ctypes.c_int64()
        _check(lib().gu_graph_device_bytes(self.h, ctypes.byref(b)), "device_bytes")
        return b.value

    # -- device-resident forms (outputs stay in HBM) -------------------------
    def multi_search_device(self, sources, dist_buf=None, sigma_buf=None, config=None):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = Stats()
        cfg = config if config is not None else make_config()
        _check(lib().gu_multi_search(self.h, _ptr(src), len(src),
                                     dist_buf.ptr if dist_buf is not None else None,
                                     sigma_buf.ptr if sigma_buf is not None else None,
                                     ctypes.byref(cfg), ctypes.byref(st)), "gu_multi_search")
        return st

    def betweenness_device(self, sources, delta_buf=None, bc_buf=None, config=None):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = Stats()
        cfg = config if config is not None else make_config()
        _check(lib().gu_betweenness(self.h, _ptr(src), len(src),
                                    delta_buf.ptr if delta_buf is not None else None,
                                    bc_buf.ptr if bc_buf is not None else None,
                                    ctypes.byref(cfg), ctypes.byref(st)), "gu_betweenness")
        return st

    # -- host-returning forms ------------------------------------------------
    def multi_search(self, sources, dist=True, sigma=False, config=None):
        """Distances [k x n] int32 (INT_MAX unreachable) and/or path counts
        [k x n] uint64 for the given sources; returns (dist, sigma, stats)."""
        k = len(sources)
        db = DeviceBuffer(4 * k * self.n) if dist else None
        sb = DeviceBuffer(8 * k * self.n) if sigma else None
        try:
            st = self.multi_search_device(sources, db, sb, config)
            d = db.to_host(np.int32).reshape(k, self.n) if dist else None
            s = sb.to_host(np.uint64).reshape(k, self.n) if sigma else None
        finally:
            for b in (db, sb):
                if b is not None:
                    b.close()
        return d, s, st

    def betweenness(self, sources, delta=True, bc=True, config=None):
        """Dependency rows [k x n] float32 and BC [n] float64 (sum over the
        sources, no 1/2 factor); returns (delta, bc, stats)."""
        k = len(sources)
        dbuf = DeviceBuffer(4 * k * self.n) if delta else None
        bbuf = DeviceBuffer(8 * self.n) if bc else None
        try:
            if bbuf is not None:
                bbuf.zero()
            st = self.betweenness_device(sources, dbuf, bbuf, config)
            d = dbuf.to_host(np.float32).reshape(k, self.n) if delta else None
            b = bbuf.to_host(np.float64) if bc else None
        finally:
            for x in (dbuf, bbuf):
                if x is not None:
                    x.close()
        return d, b, st

    def diameter(self, sources, config=None):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        out = ctypes.c_int32()
        st = Stats()
        cfg = config if config is not None else make_config()
        _check(lib().gu_diameter(self.h, _ptr(src), len(src), ctypes.byref(out),
                                 ctypes.byref(cfg), ctypes.byref(st)), "gu_diameter")
        return out.value, st

    # -- per-source output rows (gu_multi_search_rows / gu_betweenness_rows) --
    def multi_search_rows(self, sources, dist_rows=None, sigma_rows=None, want_ecc=False,
                          config=None):
        """dist_rows / sigma_rows: sequences of device addresses (int, 0 = skip)
        of n-entry rows, one per source.  Returns (ecc int32[k] or None, stats)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        k = len(src)
        dr = np.ascontiguousarray(dist_rows, dtype=np.uint64) if dist_rows is not None else None
        sr = np.ascontiguousarray(sigma_rows, dtype=np.uint64) if sigma_rows is not None else None
        ecc = np.zeros(k, dtype=np.int32) if want_ecc else None
        st = Stats()
        cfg = config if config is not None else make_config()
        _check(lib().gu_multi_search_rows(self.h, _ptr(src), k,
                                          _ptr(dr) if dr is not None else None,
                                          _ptr(sr) if sr is not None else None,
                                          _ptr(ecc) if ecc is not None else None,
                                          ctypes.byref(cfg), ctypes.byref(st)),
               "gu_multi_search_rows")
        return ecc, st

    def betweenness_rows(self, sources, delta_rows=None, bc_targets=None, config=None):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        xr = np.ascontiguousarray(delta_rows, dtype=np.uint64) if delta_rows is not None else None
        bt = np.ascontiguousarray(bc_targets, dtype=np.uint64) if bc_targets is not None else None
        st = Stats()
        cfg = config if config is not None else make_config()
        _check(lib().gu_betweenness_rows(self.h, _ptr(src), len(src),
                                         _ptr(xr) if xr is not None else None,
                                         _ptr(bt) if bt is not None else None,
                                         ctypes.byref(cfg), ctypes.byref(st)),
               "gu_betweenness_rows")
        return st

    def autotune(self, sources, kind=0):
        """gu_autotune for kind QUERY_BFS (searches) or QUERY_BC (Brandes);
        returns the chosen shape as a dict."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        r = TuneResult()
        _check(lib().gu_autotune(self.h, _ptr(src), len(src), kind, ctypes.byref(r)), "gu_autotune")
        return r.as_dict()

    def share(self):
        """A second handle on this resident graph with its own workspace
        (gu_graph_share); close it before this one."""
        h = ctypes.c_void_p()
        _check(lib().gu_graph_share(self.h, ctypes.byref(h)), "gu_graph_share")
        v = DeviceGraph.__new__(DeviceGraph)
        v.n, v.m, v.h = self.n, self.m, h
        return v


class Server:
    """Continuous-batching query server (gu_server_*): resident graphs,
    HBM-budget admission, shared multi-source launches on `slots` streams."""

    def __init__(self, hbm_budget=0, slots=0, max_pack=0, variant="auto", engine=ENGINE_AUTO, host_budget=0):
        c = ServerConfig(hbm_budget, slots, max_pack,
                         VARIANTS[variant] if isinstance(variant, str) else int(variant), engine, host_budget)
        h = ctypes.c_void_p()
        _check(lib().gu_server_create(ctypes.byref(c), ctypes.byref(h)), "gu_server_create")
        self.h = h
        self.n = {}

    def add_graph(self, dg):
        gid = ctypes.c_int32()
        _check(lib().gu_server_add_graph(self.h, dg.h, ctypes.byref(gid)), "gu_server_add_graph")
        self.n[gid.value] = dg.n
        return gid.value

    def submit(self, graph, kind, sources, flags=0):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        q = Query(graph, kind, len(src), flags, src.ctypes.data)
        t = ctypes.c_int64()
        _check(lib().gu_server_submit(self.h, ctypes.byref(q), ctypes.byref(t)), "gu_server_submit")
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        r = QueryResult()
        _check(lib().gu_server_wait(self.h, ticket, timeout_ms, ctypes.byref(r)), "gu_server_wait")
        return r

    def next_done(self, timeout_ms=-1):
        r = QueryResult()
        t = ctypes.c_int64()
        _check(lib().gu_server_next_done(self.h, timeout_ms, ctypes.byref(t), ctypes.byref(r)),
               "gu_server_next_done")
        return t.value, r

    def release(self, ticket):
        _check(lib().gu_server_release(self.h, ticket), "gu_server_release")

    def now_ms(self):
        x = ctypes.c_double()
        _check(lib().gu_server_now_ms(self.h, ctypes.byref(x)), "gu_server_now_ms")
        return x.value

    def stats(self):
        s = ServerStats()
        _check(lib().gu_server_get_stats(self.h, ctypes.byref(s)), "gu_server_get_stats")
        return s.as_dict()

    def fetch(self, res, graph):
        """Copies of a finished query's outputs: dict of arrays (from the
        device block, or from the host arena for GU_QF_HOST queries)."""
        n, k = self.n[graph], res.nsrc
        out = {"status": res.status, "diameter": res.diameter, "flags": res.flags}

        def get(addr, dtype, count):
            a = np.empty(count, dtype=dtype)
            if count:
                _check(lib().gu_memcpy_d2h(_ptr(a), ctypes.c_void_p(addr), a.nbytes), "d2h")
            return a

        def host(addr, dtype, count):
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(addr)
            return np.frombuffer(buf, dtype=dtype, count=count).copy()

        for key, dt, cnt in (("dist", np.int32, k * n), ("sigma", np.uint64, k * n),
                             ("delta", np.float32, k * n), ("bc", np.float64, n)):
            h, d = getattr(res, key + "_host"), getattr(res, key + "_dev")
            if h:
                out[key] = host(h, dt, cnt)
            elif d:
                out[key] = get(d, dt, cnt)
            if key in out and key != "bc":
                out[key] = out[key].reshape(k, n)
        return out

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().gu_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def lbs_device(counts_buf, S, scan_buf, result_buf, result_cap):
    W = ctypes.c_int64()
    st = Stats()
    cfg = make_config()
    rc = lib().gu_lbs(counts_buf.ptr, S, scan_buf.ptr,
                      result_buf.ptr if result_buf is not None else None, result_cap,
                      ctypes.byref(W), ctypes.byref(cfg), ctypes.byref(st))
    _check(rc, "gu_lbs")
    return W.value, st


# Largest inputs of gu_lbs_noalloc (GU_LBS_NOALLOC_MAX_SEGMENTS / _ITEMS).
LBS_NOALLOC_MAX_SEGMENTS = 1 << 27
LBS_NOALLOC_MAX_ITEMS = 1 << 28


def lbs_device_noalloc(counts_buf, S, scan_buf, result_buf, result_cap):
    """gu_lbs_noalloc: lbs_device without any allocation (a fixed scratch
    per GPU); returns (W, stats) or raises GuError."""
    W = ctypes.c_int64()
    st = Stats()
    cfg = make_config()
    rc = lib().gu_lbs_noalloc(counts_buf.ptr, S, scan_buf.ptr,
                              result_buf.ptr if result_buf is not None else None, result_cap,
                              ctypes.byref(W), ctypes.byref(cfg), ctypes.byref(st))
    _check(rc, "gu_lbs_noalloc")
    return W.value, st


def lbs(counts):
    """Exclusive scan and load-balanced search of the counts; returns
    (scan int32[S], result int32[W], W)."""
    c = np.ascontiguousarray(counts, dtype=np.int32)
    S = len(c)
    W = int(c.astype(np.int64).sum())
    cb = DeviceBuffer.from_array(c)
    sb = DeviceBuffer(4 * max(S, 1))
    rb = DeviceBuffer(4 * max(W, 1))
    try:
        Wd, _ = lbs_device(cb, S, sb, rb, W)
        scan = sb.to_host(np.int32, S)
        res = rb.to_host(np.int32, Wd)
    finally:
        for b in (cb, sb, rb):
            b.close()
    return scan, res, Wd


def row_sources(start, end, grid):
    """Sources held by the rows of a [start,end) result (common.cu:17-25,
    multi_search_test.cu:43-63)."""
    total = end - start
    if total <= 0:
        return np.zeros(0, dtype=np.int32)
    r = min(total, grid)
    out = np.zeros(r, dtype=np.int32)
    if total <= grid:
        out[:] = np.arange(start, start + r)
    else:
        for i in range(r):
            out[(total + i) % r] = end - r + i
    return out
