"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() (as the checker) and bench.py's
cpu_baseline leg use this module.  The product path never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class OracleResult(ctypes.Structure):
    _fields_ = [
        ("score", ctypes.c_float), ("best_row", ctypes.c_int32), ("best_col", ctypes.c_int32),
        ("start1", ctypes.c_int32), ("start2", ctypes.c_int32), ("columns", ctypes.c_int32),
        ("identity", ctypes.c_int32), ("similarity", ctypes.c_int32), ("gaps", ctypes.c_int32),
        ("transitions", ctypes.c_int32), ("transversions", ctypes.c_int32),
        ("nw_region_len", ctypes.c_int32), ("nw_region_identity", ctypes.c_int32), ("nw_similarity", ctypes.c_int32),
        ("kimura_gaps", ctypes.c_int32), ("pos_similarity", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def cpu_info():
    """Host CPUs this process may run on: nproc (os.cpu_count), the affinity
    mask, the cgroup CPU quota (cpu.max, v2; cfs_quota_us, v1) if one is set,
    and `usable` = the affinity count capped by the quota."""
    nproc = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        try:
            q = int(open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us").read())
            period = int(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
            if q > 0:
                quota = q / period
        except (OSError, ValueError):
            quota = None
    usable = affinity if quota is None else max(1, min(affinity, int(quota + 0.5)))
    return {"nproc": nproc, "affinity": affinity, "cgroup_quota_cpus": quota, "usable": usable}


def cpu_threads():
    """Oracle worker threads: one per usable host CPU (cpu_info)."""
    return cpu_info()["usable"]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        L.oracle_sw_literal.argtypes = [vp, i32, vp, i32, vp, f32, f32, ctypes.POINTER(OracleResult), vp, vp]
        L.oracle_sw_carry.argtypes = [vp, i32, vp, i32, vp, i32, i32, ctypes.POINTER(OracleResult)]
        L.oracle_nw_literal.argtypes = [vp, i32, vp, i32, vp, f32, f32, ctypes.POINTER(OracleResult), vp, vp]
        L.oracle_dist_i16_fraction.argtypes = [i32, i32]
        L.oracle_dist_i16_fraction.restype = ctypes.c_int16
        L.oracle_dist_f64_fraction.argtypes = [i32, i32]
        L.oracle_dist_f64_fraction.restype = ctypes.c_double
        L.oracle_kimura2.argtypes = [i32, i32, i32, i32]
        L.oracle_kimura2.restype = ctypes.c_double
        L.oracle_jukes_cantor.argtypes = [i32, i32, i32, ctypes.POINTER(i32)]
        L.oracle_jukes_cantor.restype = ctypes.c_double
        L.oracle_f64_to_i16.argtypes = [ctypes.c_double]
        L.oracle_f64_to_i16.restype = ctypes.c_int16
        L.oracle_sw_pairs_mt.argtypes = [vp, vp, vp, vp, vp, i64, vp, f32, f32, i32, i32, i32, vp, vp, i32]
        L.oracle_pairs_mt.argtypes = [vp, vp, vp, vp, vp, i64, vp, f32, f32, i32, i32, i32, i32, i32, vp, vp, i32]
        _lib = L
    return _lib


def _buf(b):
    b = bytes(b)
    arr = np.frombuffer(b, dtype=np.uint8) if b else np.zeros(1, np.uint8)
    return arr, len(b)


def sw_literal(s1, s2, scores, gap_open, gap_ext, want_alignment=False):
    a1, n1 = _buf(s1)
    a2, n2 = _buf(s2)
    sc = np.ascontiguousarray(scores, dtype=np.float32)
    r = OracleResult()
    o1 = ctypes.create_string_buffer(n1 + n2 + 1) if want_alignment else None
    o2 = ctypes.create_string_buffer(n1 + n2 + 1) if want_alignment else None
    rc = lib().oracle_sw_literal(a1.ctypes.data, n1, a2.ctypes.data, n2, sc.ctypes.data, gap_open, gap_ext,
                                 ctypes.byref(r), o1, o2)
    if rc != 0:
        raise IndexError("symbol outside the 127-entry scoring matrix")
    d = r.as_dict()
    if want_alignment:
        d["aln1"] = o1.raw[:r.columns].decode("latin1")
        d["aln2"] = o2.raw[:r.columns].decode("latin1")
    return d


def sw_carry(s1, s2, scores, gap_open, gap_ext):
    a1, n1 = _buf(s1)
    a2, n2 = _buf(s2)
    isc = np.ascontiguousarray(np.asarray(scores).astype(np.int32))
    r = OracleResult()
    rc = lib().oracle_sw_carry(a1.ctypes.data, n1, a2.ctypes.data, n2, isc.ctypes.data, int(gap_open), int(gap_ext),
                               ctypes.byref(r))
    if rc != 0:
        raise IndexError("symbol outside the 127-entry scoring matrix")
    return r.as_dict()


def nw_literal(s1, s2, scores, gap_open, gap_ext, want_alignment=False):
    a1, n1 = _buf(s1)
    a2, n2 = _buf(s2)
    sc = np.ascontiguousarray(scores, dtype=np.float32)
    r = OracleResult()
    o1 = ctypes.create_string_buffer(n1 + n2 + 1) if want_alignment else None
    o2 = ctypes.create_string_buffer(n1 + n2 + 1) if want_alignment else None
    rc = lib().oracle_nw_literal(a1.ctypes.data, n1, a2.ctypes.data, n2, sc.ctypes.data, gap_open, gap_ext,
                                 ctypes.byref(r), o1, o2)
    if rc != 0:
        raise IndexError("symbol outside the 127-entry scoring matrix")
    d = r.as_dict()
    if want_alignment:
        d["aln1"] = o1.raw[:r.columns].decode("latin1")
        d["aln2"] = o2.raw[:r.columns].decode("latin1")
    return d


def dist_i16(count, columns):
    return int(lib().oracle_dist_i16_fraction(count, columns))


def dist_f64(count, columns):
    return float(lib().oracle_dist_f64_fraction(count, columns))


def kimura2(columns, gaps, ts, tv):
    return float(lib().oracle_kimura2(columns, gaps, ts, tv))


def f64_to_i16(x):
    return int(lib().oracle_f64_to_i16(x))


def jukes_cantor(columns, gaps, differences):
    undefined = ctypes.c_int32(0)
    v = float(lib().oracle_jukes_cantor(columns, gaps, differences, ctypes.byref(undefined)))
    if undefined.value:
        raise ArithmeticError("Jukes and Cantor Distance Undefined - Log(Zero)")
    return v


def sw_pairs_mt(residues, offsets, a, b, scores, gap_open, gap_ext, nucleic=True, dist_fn=0, tdist_double=False,
                threads=1, want_stats=False, program=0, global_=None):
    """Multi-threaded literal aligner over pairs (a[k], b[k]); returns
    (distances, stats[k, (score, columns, identity, similarity)] or None).
    program 1 = the SWMS ComputeDistance helpers; dist_fn -1 (or global_)
    = the global aligner and the NeedlemanWunschTPL distances."""
    residues = np.ascontiguousarray(residues, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    lengths = np.diff(offsets).astype(np.int32)
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    sc = np.ascontiguousarray(scores, dtype=np.float32)
    n = len(a)
    out = np.zeros(n, dtype=np.float64 if tdist_double else np.int16)
    st = np.zeros((n, 4), dtype=np.int32) if want_stats else None
    res_ptr = residues.ctypes.data if residues.size else np.zeros(1, np.uint8).ctypes.data
    glob = (dist_fn < 0) if global_ is None else bool(global_)
    rc = lib().oracle_pairs_mt(res_ptr, offsets.ctypes.data, lengths.ctypes.data, a.ctypes.data, b.ctypes.data,
                               n, sc.ctypes.data, gap_open, gap_ext, int(glob), int(program), int(nucleic),
                               max(dist_fn, 0), int(tdist_double), out.ctypes.data,
                               st.ctypes.data if want_stats else None, threads)
    if rc == -3:
        raise ArithmeticError("Jukes and Cantor Distance Undefined - Log(Zero)")
    if rc != 0:
        raise IndexError("symbol outside the 127-entry scoring matrix")
    return out, st


def nw_pairs_mt(residues, offsets, a, b, scores, gap_open, gap_ext, nucleic=True, tdist_double=False, threads=1,
                want_stats=False, dist_fn=0):
    """Global-alignment counterpart of sw_pairs_mt (oracle_nw_literal and the
    distance of Salsa.NeedlemanWunschTPL/Program.cs:269-300); stats rows are
    (score, columns, count, region length).  dist_fn 3 = MinMaxNormScore."""
    return sw_pairs_mt(residues, offsets, a, b, scores, gap_open, gap_ext, nucleic=nucleic, dist_fn=dist_fn,
                       tdist_double=tdist_double, threads=threads, want_stats=want_stats, global_=True)
