"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU restatement of the reference's sequential checkers (see oracle.h for
the file:line each function follows).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg load this module; the product never does.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_PARSE = os.path.join(HERE, "_ref", "libref_parse.so")
INT_MAX = 2147483647

_lib = None
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        L.or_bfs.argtypes = [_P, _P, i32, i32, _P]
        L.or_sigma.argtypes = [_P, _P, i32, i32, _P, _P]
        L.or_brandes_f32.argtypes = [_P, _P, i32, i32, _P, _P, _P]
        L.or_brandes_f64.argtypes = [_P, _P, i32, i32, _P, _P, _P]
        L.or_bc_f64.argtypes = [_P, _P, i32, _P, i32, _P]
        L.or_bc_f64_mt.argtypes = [_P, _P, i32, _P, i32, i32, _P]
        L.or_eccentricity.argtypes = [_P, _P, i32, i32]
        L.or_eccentricity.restype = i32
        L.or_exclusive_scan.argtypes = [_P, i64, _P]
        L.or_exclusive_scan.restype = i64
        L.or_lbs.argtypes = [i64, _P, i64, _P]
        L.or_rows_for_range.argtypes = [i32, i32, i32, _P]
        L.or_rows_for_range.restype = i32
        L.or_teps_units.argtypes = [_P, i32, _P]
        L.or_teps_units.restype = i64
        L.or_sigma_batch_timed.argtypes = [_P, _P, i32, _P, i32, i32, ctypes.POINTER(ctypes.c_double)]
        L.or_sigma_batch_timed.restype = i64
        L.or_lbs_timed.argtypes = [_P, i64, _P, _P, i64]
        L.or_lbs_timed.restype = ctypes.c_double
        L.or_rows_mt.argtypes = [_P, _P, i32, _P, i32, i32, _P, _P, _P]
        L.or_check_search_rows.argtypes = [_P, _P, i32, _P, i32, _P, _P, i32,
                                           ctypes.POINTER(ctypes.c_int64)]
        L.or_check_search_rows.restype = i64
        L.or_sigma_wraps.argtypes = [_P, _P, i32, i32]
        L.or_sigma_wraps.restype = i32
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _csr(R, C):
    return np.ascontiguousarray(R, dtype=np.int32), np.ascontiguousarray(C, dtype=np.int32)


def bfs(R, C, src):
    R, C = _csr(R, C)
    n = len(R) - 1
    d = np.empty(n, dtype=np.int32)
    lib().or_bfs(_p(R), _p(C), n, int(src), _p(d))
    return d


def sigma(R, C, src):
    R, C = _csr(R, C)
    n = len(R) - 1
    d = np.empty(n, dtype=np.int32)
    s = np.empty(n, dtype=np.uint64)
    lib().or_sigma(_p(R), _p(C), n, int(src), _p(d), _p(s))
    return d, s


def brandes_f32(R, C, src):
    R, C = _csr(R, C)
    n = len(R) - 1
    delta = np.empty(n, dtype=np.float32)
    lib().or_brandes_f32(_p(R), _p(C), n, int(src), None, None, _p(delta))
    return delta


def brandes_f64(R, C, src):
    R, C = _csr(R, C)
    n = len(R) - 1
    delta = np.empty(n, dtype=np.float64)
    lib().or_brandes_f64(_p(R), _p(C), n, int(src), None, None, _p(delta))
    return delta


def bc_f64(R, C, sources):
    R, C = _csr(R, C)
    n = len(R) - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    bc = np.empty(n, dtype=np.float64)
    lib().or_bc_f64(_p(R), _p(C), n, _p(src), len(src), _p(bc))
    return bc


def bc_f64_mt(R, C, sources, nthreads):
    R, C = _csr(R, C)
    n = len(R) - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    bc = np.empty(n, dtype=np.float64)
    lib().or_bc_f64_mt(_p(R), _p(C), n, _p(src), len(src), int(nthreads), _p(bc))
    return bc


def eccentricity(R, C, src):
    R, C = _csr(R, C)
    return int(lib().or_eccentricity(_p(R), _p(C), len(R) - 1, int(src)))


def exclusive_scan(counts):
    c = np.ascontiguousarray(counts, dtype=np.int32)
    out = np.empty(len(c), dtype=np.int32)
    W = lib().or_exclusive_scan(_p(c), len(c), _p(out))
    return out, int(W)


def lbs(W, scan):
    s = np.ascontiguousarray(scan, dtype=np.int32)
    out = np.empty(int(W), dtype=np.int32)
    lib().or_lbs(int(W), _p(s), len(s), _p(out))
    return out


def rows_for_range(start, end, G):
    out = np.zeros(max(min(end - start, G), 1), dtype=np.int32)
    r = lib().or_rows_for_range(start, end, G, _p(out))
    return out[:r]


def teps_units(R, dist):
    R = np.ascontiguousarray(R, dtype=np.int32)
    d = np.ascontiguousarray(dist, dtype=np.int32)
    return int(lib().or_teps_units(_p(R), len(R) - 1, _p(d)))


def sigma_batch_timed(R, C, sources, nthreads=1):
    R, C = _csr(R, C)
    src = np.ascontiguousarray(sources, dtype=np.int32)
    secs = ctypes.c_double()
    units = lib().or_sigma_batch_timed(_p(R), _p(C), len(R) - 1, _p(src), len(src),
                                       int(nthreads), ctypes.byref(secs))
    return int(units), secs.value


def brandes_timed(R, C, sources):
    """Times the reference's sequential FP32 Brandes (bc.cu:11-57) source
    after source on one core; returns (TEPS units = edges scanned from the
    reached vertices, summed over sources; seconds inside the calls)."""
    import time

    R, C = _csr(R, C)
    n = len(R) - 1
    d = np.empty(n, dtype=np.int32)
    s = np.empty(n, dtype=np.uint64)
    delta = np.empty(n, dtype=np.float32)
    units, secs = 0, 0.0
    for src in np.asarray(sources, dtype=np.int32):
        t0 = time.perf_counter()
        lib().or_brandes_f32(_p(R), _p(C), n, int(src), _p(d), _p(s), _p(delta))
        secs += time.perf_counter() - t0
        units += teps_units(R, d)
    return units, secs


def lbs_timed(counts):
    c = np.ascontiguousarray(counts, dtype=np.int32)
    W = int(c.astype(np.int64).sum())
    scan = np.empty(len(c), dtype=np.int32)
    res = np.empty(W, dtype=np.int32)
    secs = lib().or_lbs_timed(_p(c), len(c), _p(scan), _p(res), W)
    return scan, res, secs


def default_threads():
    """Threads for the checker: the cores this process may use, at most 16
    (the GPU box's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def rows_mt(R, C, sources, sigma=True, ecc=False, nthreads=None):
    """Oracle rows for many sources on threads: (dist, sigma or None, ecc or
    None), rows [len(sources)][n]."""
    R, C = _csr(R, C)
    n = len(R) - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    k = len(src)
    d = np.empty((k, n), dtype=np.int32)
    s = np.empty((k, n), dtype=np.uint64) if sigma else None
    e = np.empty(k, dtype=np.int32) if ecc else None
    lib().or_rows_mt(_p(R), _p(C), n, _p(src), k, int(nthreads or default_threads()), _p(d),
                     _p(s) if sigma else None, _p(e) if ecc else None)
    return d, s, e


def eccentricities_mt(R, C, sources, nthreads=None):
    R, C = _csr(R, C)
    src = np.ascontiguousarray(sources, dtype=np.int32)
    e = np.empty(len(src), dtype=np.int32)
    lib().or_rows_mt(_p(R), _p(C), len(R) - 1, _p(src), len(src), int(nthreads or default_threads()),
                     None, None, _p(e))
    return e


def sigma_wraps(R, C, src):
    """True when some vertex reached from src has at least 2^64 shortest
    paths (its path count wraps modulo 2^64)."""
    R, C = _csr(R, C)
    return bool(lib().or_sigma_wraps(_p(R), _p(C), len(R) - 1, int(src)))


def check_search_rows(RT, CT, sources, dist, sigma=None, nthreads=None):
    """Checks search rows against the BFS equations (see oracle.h).  Returns
    (violations, (row, vertex) of one violation or None)."""
    RT, CT = _csr(RT, CT)
    n = len(RT) - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    d = np.ascontiguousarray(dist, dtype=np.int32)
    assert d.shape == (len(src), n)
    if sigma is not None:
        sigma = np.ascontiguousarray(sigma, dtype=np.uint64)
        assert sigma.shape == d.shape
    bad = ctypes.c_int64(-1)
    v = lib().or_check_search_rows(_p(RT), _p(CT), n, _p(src), len(src), _p(d),
                                   _p(sigma) if sigma is not None else None,
                                   int(nthreads or default_threads()), ctypes.byref(bad))
    where = None if bad.value < 0 else divmod(int(bad.value), n)
    return int(v), where


def ref_parse_available():
    return os.path.exists(REF_PARSE)


def ref_parse_metis(path):
    """The reference's own parse_metis (parse.cpp:115-208), compiled from the
    reference checkout into oracle/_ref.  Returns (R, C, F, n, m, directed)."""
    L = ctypes.CDLL(REF_PARSE)
    I32P = ctypes.POINTER(ctypes.c_int32)
    L.ref_parse_metis.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(I32P),
                                  ctypes.POINTER(I32P), ctypes.POINTER(I32P),
                                  ctypes.POINTER(ctypes.c_int32)]
    L.ref_free.argtypes = [ctypes.c_void_p]
    n, m, dr = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
    R, C, F = I32P(), I32P(), I32P()
    L.ref_parse_metis(os.fsencode(path), ctypes.byref(n), ctypes.byref(m), ctypes.byref(R),
                      ctypes.byref(C), ctypes.byref(F), ctypes.byref(dr))

    def take(p, k):
        a = np.ctypeslib.as_array(p, shape=(k,)).copy() if k > 0 else np.zeros(0, np.int32)
        L.ref_free(ctypes.cast(p, ctypes.c_void_p))
        return a

    return take(R, n.value + 1), take(C, m.value), take(F, m.value), n.value, m.value, bool(dr.value)


def _ref_lib():
    L = ctypes.CDLL(REF_PARSE)
    I32P = ctypes.POINTER(ctypes.c_int32)
    L.ref_parse_market.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(I32P),
                                   ctypes.POINTER(I32P), ctypes.POINTER(I32P),
                                   ctypes.POINTER(ctypes.c_int64)]
    L.ref_parse_snap.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(I32P),
                                 ctypes.POINTER(I32P), ctypes.POINTER(I32P),
                                 ctypes.POINTER(ctypes.c_int32)]
    L.ref_free.argtypes = [ctypes.c_void_p]
    return L, I32P


def _take_ref(L, p, k):
    a = np.ctypeslib.as_array(p, shape=(k,)).copy() if k > 0 else np.zeros(0, np.int32)
    L.ref_free(ctypes.cast(p, ctypes.c_void_p))
    return a


def ref_parse_market(path):
    """The reference's parse_market (parse.cpp:210-332).  Returns
    (R, C, F, n, m) with C and F cut to m entries."""
    L, I32P = _ref_lib()
    n, m, cl = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
    R, C, F = I32P(), I32P(), I32P()
    L.ref_parse_market(os.fsencode(path), ctypes.byref(n), ctypes.byref(m), ctypes.byref(R),
                       ctypes.byref(C), ctypes.byref(F), ctypes.byref(cl))
    Ra = _take_ref(L, R, n.value + 1)
    Ca = _take_ref(L, C, cl.value)[:m.value]
    Fa = _take_ref(L, F, cl.value)[:m.value]
    return Ra, Ca, Fa, n.value, m.value


def ref_parse_snap(path, header=False):
    """The reference's parse_snap (parse.cpp:335-434).  Returns
    (R, C, F, n, m, directed)."""
    L, I32P = _ref_lib()
    n, m, dr = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
    R, C, F = I32P(), I32P(), I32P()
    L.ref_parse_snap(os.fsencode(path), 1 if header else 0, ctypes.byref(n), ctypes.byref(m),
                     ctypes.byref(R), ctypes.byref(C), ctypes.byref(F), ctypes.byref(dr))
    return (_take_ref(L, R, n.value + 1), _take_ref(L, C, m.value), _take_ref(L, F, m.value),
            n.value, m.value, bool(dr.value))


def ref_reverse(R, C, F):
    """The reference's reverse (graph-utils/reverse.cpp:3-44)."""
    L, I32P = _ref_lib()
    P = ctypes.c_void_p
    L.ref_reverse.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(I32P),
                              ctypes.POINTER(I32P), ctypes.POINTER(I32P)]
    R, C, F = (np.ascontiguousarray(x, dtype=np.int32) for x in (R, C, F))
    n, m = len(R) - 1, len(C)
    R2, C2, F2 = I32P(), I32P(), I32P()
    L.ref_reverse(_p(R), _p(C), _p(F), n, m, ctypes.byref(R2), ctypes.byref(C2), ctypes.byref(F2))
    return _take_ref(L, R2, n + 1), _take_ref(L, C2, m), _take_ref(L, F2, m)


def ref_reach_dag(R, C, F):
    """The reference's get_reachability_dag (graph-utils/dag_reachability.cpp:31-65)
    on a directed acyclic graph: list of sorted member arrays."""
    L, I32P = _ref_lib()
    P = ctypes.c_void_p
    I64P = ctypes.POINTER(ctypes.c_int64)
    L.ref_reach_dag.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(I64P),
                                ctypes.POINTER(I32P)]
    R, C, F = (np.ascontiguousarray(x, dtype=np.int32) for x in (R, C, F))
    n, m = len(R) - 1, len(C)
    off, mem = I64P(), I32P()
    L.ref_reach_dag(_p(R), _p(C), _p(F), n, m, ctypes.byref(off), ctypes.byref(mem))
    o = np.ctypeslib.as_array(off, shape=(n + 1,)).copy()
    L.ref_free(ctypes.cast(off, ctypes.c_void_p))
    members = _take_ref(L, mem, int(o[-1]))
    return [members[o[i]:o[i + 1]] for i in range(n)]
