"""ctypes binding of oracle/liboracle.so -- the CPU checker used by the tests.

The oracle is test infrastructure: this module is imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.environ.get("MYERS_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")

MODE_ALIGNED = 0
MODE_EXACT = 1

CAND_DTYPE = np.dtype([("query", "<u4"), ("pad", "<u4"), ("position", "<u8")])

_lib = None


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    lib = ctypes.CDLL(ORACLE_SO)
    vp, u32, u64, f32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_int
    lib.oracle_window_len.argtypes = [u32, f32]
    lib.oracle_window_len.restype = u32
    lib.oracle_myers_batch.argtypes = [vp, u64, u64, vp, vp, vp, u32, f32, i32, u32, vp]
    lib.oracle_myers_batch.restype = None
    lib.oracle_myers_trace.argtypes = [vp, u32, vp, u32, vp, vp]
    lib.oracle_dp_trace.argtypes = [vp, u32, vp, u32, vp, vp]
    lib.oracle_pack_reference.argtypes = [vp, u64, vp]
    lib.oracle_build_peq.argtypes = [vp, u32, vp]
    lib.oracle_bpm64_batch.argtypes = [vp, u64, u64, vp, vp, vp, u32, f32, i32, vp]
    lib.oracle_bpm64_batch.restype = None
    lib.oracle_threshold.argtypes = [u32, f32]
    lib.oracle_threshold.restype = u32
    lib.oracle_apply_threshold.argtypes = [vp, vp, u64, f32, vp]
    lib.oracle_apply_threshold.restype = None
    _lib = lib
    return lib


def window_len(m, distance):
    return load().oracle_window_len(m, distance)


def make_cands(query, position):
    c = np.zeros(len(query), dtype=CAND_DTYPE)
    c["query"] = query
    c["position"] = position
    return c


def myers_batch(ref_words, ref_size, peq, qinfo, cands, distance, mode=MODE_ALIGNED, max_query_len=0):
    """Returns an (n, 2) uint32 array of [column, score] per candidate."""
    ref_words = np.ascontiguousarray(ref_words, dtype=np.uint32)
    peq = np.ascontiguousarray(peq, dtype=np.uint32)
    qinfo = np.ascontiguousarray(qinfo, dtype=np.uint32)
    cands = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    out = np.zeros((len(cands), 2), dtype=np.uint32)
    load().oracle_myers_batch(_ptr(ref_words), len(ref_words), int(ref_size), _ptr(peq), _ptr(qinfo),
                              _ptr(cands), len(cands), float(distance), mode, max_query_len, _ptr(out))
    return out


def bpm64_batch(ref_words, ref_size, peq, qinfo, cands, distance, mode=MODE_ALIGNED):
    """GEMTools-style 64-bit block BPM over the same inputs: (n, 2) [column, score]."""
    ref_words = np.ascontiguousarray(ref_words, dtype=np.uint32)
    peq = np.ascontiguousarray(peq, dtype=np.uint32)
    qinfo = np.ascontiguousarray(qinfo, dtype=np.uint32)
    cands = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    out = np.zeros((len(cands), 2), dtype=np.uint32)
    load().oracle_bpm64_batch(_ptr(ref_words), len(ref_words), int(ref_size), _ptr(peq), _ptr(qinfo),
                              _ptr(cands), len(cands), float(distance), mode, _ptr(out))
    return out


def threshold(m, e):
    return int(load().oracle_threshold(int(m), float(e)))


def apply_threshold(qinfo, cands, results, e):
    """Copy of results with the threshold mode's no-match sentinel applied."""
    qinfo = np.ascontiguousarray(qinfo, dtype=np.uint32)
    cands = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    out = np.ascontiguousarray(results, dtype=np.uint32).copy()
    load().oracle_apply_threshold(_ptr(qinfo), _ptr(cands), len(cands), float(e), _ptr(out))
    return out


def myers_trace(query_codes, text_codes):
    q = np.ascontiguousarray(query_codes, dtype=np.uint8)
    t = np.ascontiguousarray(text_codes, dtype=np.uint8)
    scores = np.zeros(len(t), dtype=np.int32)
    res = np.zeros(2, dtype=np.uint32)
    load().oracle_myers_trace(_ptr(q), len(q), _ptr(t), len(t), _ptr(scores), _ptr(res))
    return scores, (int(res[1]), int(res[0]))


def dp_trace(query_codes, text_codes):
    q = np.ascontiguousarray(query_codes, dtype=np.uint8)
    t = np.ascontiguousarray(text_codes, dtype=np.uint8)
    scores = np.zeros(len(t), dtype=np.int32)
    res = np.zeros(2, dtype=np.uint32)
    load().oracle_dp_trace(_ptr(q), len(q), _ptr(t), len(t), _ptr(scores), _ptr(res))
    return scores, (int(res[1]), int(res[0]))


def pack_reference(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros((len(codes) + 7) // 8, dtype=np.uint32)
    load().oracle_pack_reference(_ptr(codes), len(codes), _ptr(out))
    return out


def build_peq(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros(((len(codes) + 127) // 128, 20), dtype=np.uint32)
    load().oracle_build_peq(_ptr(codes), len(codes), _ptr(out))
    return out
