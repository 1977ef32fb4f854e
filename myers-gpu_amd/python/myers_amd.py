"""ctypes binding of the MI355X candidate-verification library.

Loads lib/libmyers_amd.so (C host library + gfx950 kernels) and
lib/libmyers_tools.so (synthetic data, packing).  There is no fallback: if a
library is missing, import fails with an error that says how to build it.

The record layouts mirror include/myers-interface.h:
  qryEntry_t  -> uint32[20]            (bitmap[5][4])
  qryInfo_t   -> uint32[2]             {posEntry, size}
  candInfo_t  -> {u32 query, u32 pad, u64 position}
  resEntry_t  -> uint32[2]             {column, score}
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
# MYERS_AMD_LIB / MYERS_TOOLS_LIB select other builds of the two libraries
# (the CPU sanitizer builds of `make -C myers-gpu_amd san`).
LIB_PATH = os.environ.get("MYERS_AMD_LIB") or os.path.join(LIB_DIR, "libmyers_amd.so")
TOOLS_PATH = os.environ.get("MYERS_TOOLS_LIB") or os.path.join(LIB_DIR, "libmyers_tools.so")

CAND_DTYPE = np.dtype([("query", "<u4"), ("pad", "<u4"), ("position", "<u8")])
SKIP_SCORE = 0xFFFFFFFF

WINDOW_ALIGNED = 0
WINDOW_EXACT = 1
ENGINE_AUTO = 0
ENGINE_PLAIN = 1 << 4
ENGINE_PERSISTENT = 2 << 4
ENGINE_BINNED = 3 << 4
USE_GRAPH = 1 << 6
NO_LONG_LAUNCH = 1 << 7
SERVE_RESIDENT = 1 << 8              # myersServeConfig_t.flags: one resident kernel instead of batches
NOALLOC_MAX_CANDIDATES = 1 << 21     # myersVerifyDeviceNoAlloc's static workspace

ERRORS = {
    0: "SUCCESS", 1: "E_OPENING_FILE", 2: "E_READING_FILE", 3: "E_INSUFFICIENT_MEM_GPU",
    4: "E_ALLOCATE_MEM", 5: "E_INCOMPATIBLE_GPU", 16: "E_INVALID_ARGUMENT", 17: "E_CAPACITY",
    18: "E_HIP", 19: "E_QUERY_TOO_LONG", 20: "E_NO_DEVICE", 21: "E_REFERENCE_TOO_LARGE",
}
globals().update({name: code for code, name in ERRORS.items()})


class MyersError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({code})")


class Config(ctypes.Structure):
    _fields_ = [("structSize", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("exitOnError", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("bufferBytes", ctypes.c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [("structSize", ctypes.c_uint32), ("launches", ctypes.c_uint32),
                ("variantsUsed", ctypes.c_uint32), ("binned", ctypes.c_uint32),
                ("kernelMs", ctypes.c_double), ("batchMs", ctypes.c_double),
                ("cells", ctypes.c_uint64), ("candidates", ctypes.c_uint64),
                ("kernelMsTotal", ctypes.c_double), ("kernelLaunchesTotal", ctypes.c_uint64),
                ("batchesTotal", ctypes.c_uint64), ("graphCaptures", ctypes.c_uint64),
                ("verifyMs", ctypes.c_double), ("verifyMsTotal", ctypes.c_double)]


class VariantInfo(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_uint32), ("words", ctypes.c_uint32),
                ("maxQuerySize", ctypes.c_uint32), ("candidatesPerWave", ctypes.c_uint32),
                ("ldsBytes", ctypes.c_uint32), ("workgroupsPerCU", ctypes.c_int32)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("structSize", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("distance", ctypes.c_float),
                ("flags", ctypes.c_uint32), ("hbmBudgetBytes", ctypes.c_uint64), ("pollMicros", ctypes.c_uint32),
                ("maxBatchCandidates", ctypes.c_uint32), ("maxBatchCells", ctypes.c_uint64),
                ("arenaBytes", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [("peq", ctypes.c_void_p), ("qinfo", ctypes.c_void_p), ("cands", ctypes.c_void_p),
                ("results", ctypes.c_void_p), ("numPEQEntries", ctypes.c_uint32),
                ("numQueries", ctypes.c_uint32), ("numCandidates", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("structSize", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("requests", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("candidates", ctypes.c_uint64),
                ("maxBatchRequests", ctypes.c_uint64), ("busyMs", ctypes.c_double),
                ("admitMs", ctypes.c_double), ("completeMs", ctypes.c_double),
                ("residentWorkgroups", ctypes.c_uint32), ("residentLdsBytes", ctypes.c_uint32),
                ("residentItems", ctypes.c_uint64), ("liveRequests", ctypes.c_uint64),
                ("drainBatches", ctypes.c_uint64)]


class SynthSpec(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("numQueries", ctypes.c_uint32),
                ("lengthMode", ctypes.c_uint32), ("length", ctypes.c_uint32),
                ("numChoices", ctypes.c_uint32), ("choices", ctypes.c_uint32 * 16),
                ("minLength", ctypes.c_uint32), ("maxLength", ctypes.c_uint32),
                ("subRate", ctypes.c_double), ("insRate", ctypes.c_double),
                ("delRate", ctypes.c_double), ("errMin", ctypes.c_double),
                ("errMax", ctypes.c_double), ("subFrac", ctypes.c_double),
                ("indelGeomP", ctypes.c_double), ("candMode", ctypes.c_uint32),
                ("candidates", ctypes.c_uint32), ("candMean", ctypes.c_double),
                ("candCap", ctypes.c_uint32), ("distance", ctypes.c_float),
                ("keepCodes", ctypes.c_uint32)]


class SynthBatch(ctypes.Structure):
    _fields_ = [("peq", ctypes.c_void_p), ("qinfo", ctypes.c_void_p), ("cands", ctypes.c_void_p),
                ("codes", ctypes.c_void_p), ("codeOffset", ctypes.c_void_p),
                ("trueStart", ctypes.c_void_p), ("numEntries", ctypes.c_uint64),
                ("numQueries", ctypes.c_uint64), ("numCandidates", ctypes.c_uint64),
                ("totalBases", ctypes.c_uint64), ("cells", ctypes.c_uint64)]


def _load(path):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `make -C myers-gpu_amd`")
    return ctypes.CDLL(path)


_lib = None
_tools = None


def lib():
    global _lib
    if _lib is None:
        L = _load(LIB_PATH)
        vp, u32, u64, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float
        pvp = ctypes.POINTER(ctypes.c_void_p)
        ppvp = ctypes.POINTER(pvp)
        for name in ("getPEQBuffer", "getCandidatesBuffer", "getPEQInfoBuffer", "getResultsBuffer"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = vp
        for name in ("getMaxPEQEntries", "getMaxCandidates", "getMaxQueries", "myersGetNumBuffers",
                     "myersNumVariants"):
            getattr(L, name).restype = u32
        for name in ("getMaxPEQEntries", "getMaxCandidates", "getMaxQueries", "myersGetNumBuffers"):
            getattr(L, name).argtypes = [vp]
        L.myersNumVariants.argtypes = []
        L.initMyers.argtypes = [ppvp, u32, ctypes.c_char_p, i32, i32]
        L.initMyers.restype = ctypes.c_int
        L.myersInitEx.argtypes = [ppvp, u32, ctypes.c_char_p, i32, i32, ctypes.POINTER(Config)]
        L.myersInitFromMemory.argtypes = [ppvp, u32, vp, u64, u64, i32, i32, ctypes.POINTER(Config)]
        L.sendMyersBuffer.argtypes = [vp, f32, u32, u32, u32]
        L.sendMyersBuffer.restype = None
        L.receiveMyersBuffer.argtypes = [vp]
        L.receiveMyersBuffer.restype = None
        L.endMyers.argtypes = [ppvp]
        L.endMyers.restype = None
        L.myersSendEx.argtypes = [vp, f32, u32, u32, u32, u32]
        L.myersReceiveEx.argtypes = [vp]
        L.myersEnd.argtypes = [ppvp]
        L.myersGetErrorString.argtypes = [ctypes.c_int]
        L.myersGetErrorString.restype = ctypes.c_char_p
        L.myersGetLastError.argtypes = [vp]
        L.myersGetStats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.myersResetStats.argtypes = [vp]
        L.myersResetStats.restype = None
        L.myersGetReferenceSize.argtypes = [vp]
        L.myersGetReferenceSize.restype = u64
        L.myersVerifyDevice.argtypes = [vp, vp, u32, vp, u32, vp, u32, f32, u32, vp, vp]
        L.myersVerifyDeviceNoAlloc.argtypes = [vp, vp, u32, vp, u32, vp, u32, f32, u32, vp, vp]
        L.myersSynchronize.argtypes = [vp]
        L.myersGetStream.argtypes = [vp]
        L.myersGetStream.restype = vp
        L.myersDeviceMalloc.argtypes = [pvp, ctypes.c_size_t]
        L.myersDeviceFree.argtypes = [vp]
        L.myersMemcpyHtoD.argtypes = [vp, vp, ctypes.c_size_t]
        L.myersMemcpyDtoH.argtypes = [vp, vp, ctypes.c_size_t]
        L.myersGetVariantInfo.argtypes = [vp, u32, ctypes.POINTER(VariantInfo)]
        L.myersGetBinTable.argtypes = [vp, vp, u32]
        L.myersSetBinTable.argtypes = [vp, vp, u32]
        L.myersSetChunkFactor.argtypes = [vp, u32, u32]
        L.myersSetThreshold.argtypes = [vp, ctypes.c_float]
        L.myersGetThreshold.argtypes = [vp]
        L.myersGetThreshold.restype = ctypes.c_float
        L.myersGetChunkFactor.argtypes = [vp, u32, ctypes.POINTER(ctypes.c_uint32)]
        L.myersLoadTuning.argtypes = [vp, ctypes.c_char_p]
        L.myersServeDefaultConfig.argtypes = [ctypes.POINTER(ServeConfig)]
        L.myersServeDefaultConfig.restype = None
        L.myersServeStart.argtypes = [pvp, pvp, u32, ctypes.POINTER(ServeConfig)]
        L.myersServeSubmit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
        L.myersServeWait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        L.myersServeGetStats.argtypes = [vp, ctypes.POINTER(ServeStats)]
        L.myersServeStop.argtypes = [vp]
        L.myersServeRunOpenLoop.argtypes = [vp, u32, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.myersServeRunOpenLoopEx.argtypes = [vp, u32, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double)]
        L.myersDefaultConfig.argtypes = [ctypes.POINTER(Config)]
        L.myersDefaultConfig.restype = None
        _lib = L
    return _lib


def tools():
    global _tools
    if _tools is None:
        T = _load(TOOLS_PATH)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        T.myersSynthReference.argtypes = [u64, u64, ctypes.c_double, vp]
        T.myersSynthReference.restype = None
        T.myersSynthBatch.argtypes = [ctypes.POINTER(SynthSpec), vp, u64, ctypes.POINTER(SynthBatch)]
        T.myersSynthFreeBatch.argtypes = [ctypes.POINTER(SynthBatch)]
        T.myersSynthFreeBatch.restype = None
        T.myersBuildPeq.argtypes = [vp, u32, vp]
        T.myersBuildPeq.restype = None
        T.myersPackCodes.argtypes = [vp, u64, vp]
        T.myersPackCodes.restype = None
        T.myersPackFasta.argtypes = [ctypes.c_char_p, u64, vp, u64]
        T.myersPackFasta.restype = u64
        T.myersWriteReferenceFile.argtypes = [ctypes.c_char_p, vp, u32, u32]
        T.myersReadReferenceFile.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                             ctypes.POINTER(u32), ctypes.POINTER(u32)]
        T.myersReadGemQry.argtypes = [ctypes.c_char_p, ctypes.POINTER(GemQry)]
        T.myersWriteGemQry.argtypes = [ctypes.c_char_p, ctypes.POINTER(GemQry)]
        T.myersFreeGemQry.argtypes = [ctypes.POINTER(GemQry)]
        T.myersFreeGemQry.restype = None
        T.myersParseProf.argtypes = [ctypes.c_char_p, ctypes.POINTER(GemQry)]
        _tools = T
    return _tools


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class GemQry(ctypes.Structure):
    _fields_ = [("totalSizeQueries", ctypes.c_uint32), ("numEntries", ctypes.c_uint32),
                ("numQueries", ctypes.c_uint32), ("numCandidates", ctypes.c_uint32),
                ("peq", ctypes.c_void_p), ("cands", ctypes.c_void_p), ("qinfo", ctypes.c_void_p),
                ("gemScores", ctypes.c_void_p), ("bases", ctypes.c_void_p)]


def check(rc, where):
    if rc != 0:
        raise MyersError(rc, where)


# ------------------------------------------------------------------ tools

N_RUN_PROB = 0.0082   # one N run per ~122 chunks of 64 bases: ~0.1% of bases are N


def synth_reference(seed, size, n_run_prob=N_RUN_PROB):
    words = np.empty((size + 7) // 8, dtype=np.uint32)
    tools().myersSynthReference(seed, size, n_run_prob, _ptr(words))
    return words


def pack_codes(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.empty((len(codes) + 7) // 8, dtype=np.uint32)
    tools().myersPackCodes(_ptr(codes), len(codes), _ptr(out))
    return out


def build_peq(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros(((len(codes) + 127) // 128, 20), dtype=np.uint32)
    tools().myersBuildPeq(_ptr(codes), len(codes), _ptr(out))
    return out


def _copy_out(ptr, count, dtype, shape=None):
    if count == 0 or not ptr:
        return np.zeros(shape or (0,), dtype=dtype)
    buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
    a = np.frombuffer(buf, dtype=dtype).copy()
    return a.reshape(shape) if shape else a


def _gem_to_dict(g):
    return {
        "peq": _copy_out(g.peq, g.numEntries * 20, np.uint32, (g.numEntries, 20)),
        "qinfo": _copy_out(g.qinfo, g.numQueries * 2, np.uint32, (g.numQueries, 2)),
        "cands": _copy_out(g.cands, g.numCandidates, CAND_DTYPE),
        "gem_scores": _copy_out(g.gemScores, g.numCandidates, np.uint32),
        "bases": _copy_out(g.bases, g.totalSizeQueries, np.uint8).tobytes(),
    }


def read_reference_file(path):
    """A .4bits.ref file -> (words, size)."""
    w, ne, sz = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
    if tools().myersReadReferenceFile(os.fsencode(path), ctypes.byref(w), ctypes.byref(ne), ctypes.byref(sz)) != 0:
        raise OSError(f"cannot read reference file {path}")
    try:
        return _copy_out(w.value, ne.value, np.uint32), sz.value
    finally:
        ctypes.CDLL(None).free(w)


def write_reference_file(path, words, size):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    if tools().myersWriteReferenceFile(os.fsencode(path), _ptr(words), len(words), size) != 0:
        raise OSError(f"cannot write reference file {path}")


def parse_prof(path):
    """A GEM .prof file -> dict(peq, qinfo, cands, gem_scores, bases)."""
    g = GemQry()
    if tools().myersParseProf(os.fsencode(path), ctypes.byref(g)) != 0:
        raise ValueError(f"cannot parse .prof file {path}")
    try:
        return _gem_to_dict(g)
    finally:
        tools().myersFreeGemQry(ctypes.byref(g))


def read_gem_qry(path):
    g = GemQry()
    if tools().myersReadGemQry(os.fsencode(path), ctypes.byref(g)) != 0:
        raise ValueError(f"cannot read .gem.qry file {path}")
    try:
        return _gem_to_dict(g)
    finally:
        tools().myersFreeGemQry(ctypes.byref(g))


def write_gem_qry(path, d):
    peq = np.ascontiguousarray(d["peq"], dtype=np.uint32)
    qinfo = np.ascontiguousarray(d["qinfo"], dtype=np.uint32)
    cands = np.ascontiguousarray(d["cands"], dtype=CAND_DTYPE)
    scores = np.ascontiguousarray(d.get("gem_scores", np.zeros(len(cands), np.uint32)), dtype=np.uint32)
    bases = np.frombuffer(d.get("bases", b""), dtype=np.uint8).copy()
    g = GemQry(len(bases), len(peq), len(qinfo), len(cands), peq.ctypes.data, cands.ctypes.data,
               qinfo.ctypes.data, scores.ctypes.data, bases.ctypes.data if len(bases) else None)
    if tools().myersWriteGemQry(os.fsencode(path), ctypes.byref(g)) != 0:
        raise OSError(f"cannot write .gem.qry file {path}")


def make_spec(seed, num_queries, *, length=None, choices=None, log_uniform=None, sub=0.0, ins=0.0,
              dele=0.0, err_range=None, sub_frac=0.8, geom_p=0.7, candidates=None, poisson=None,
              cap=64, distance=0.1, keep_codes=False):
    s = SynthSpec()
    s.seed = seed
    s.numQueries = num_queries
    if choices is not None:
        s.lengthMode = 1
        s.numChoices = len(choices)
        for i, c in enumerate(choices):
            s.choices[i] = c
    elif log_uniform is not None:
        s.lengthMode = 2
        s.minLength, s.maxLength = log_uniform
    else:
        s.lengthMode = 0
        s.length = length
    s.subRate, s.insRate, s.delRate = sub, ins, dele
    if err_range is not None:
        s.errMin, s.errMax = err_range
        s.subFrac = sub_frac
    s.indelGeomP = geom_p
    if poisson is not None:
        s.candMode = 1
        s.candMean = poisson
        s.candCap = cap
    else:
        s.candMode = 0
        s.candidates = candidates or 1
    s.distance = distance
    s.keepCodes = 1 if keep_codes else 0
    return s


def synth_batch(spec, ref_words, ref_size):
    """Returns a dict of numpy arrays (copied out of the C batch)."""
    b = SynthBatch()
    rc = tools().myersSynthBatch(ctypes.byref(spec), _ptr(ref_words), ref_size, ctypes.byref(b))
    if rc != 0:
        raise ValueError("myersSynthBatch rejected the specification")
    try:
        def arr(ptr, count, dtype, shape=None):
            if count == 0:
                return np.zeros(shape or (0,), dtype=dtype)
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            a = np.frombuffer(buf, dtype=dtype).copy()
            return a.reshape(shape) if shape else a
        out = {
            "peq": arr(b.peq, b.numEntries * 20, np.uint32, (b.numEntries, 20)),
            "qinfo": arr(b.qinfo, b.numQueries * 2, np.uint32, (b.numQueries, 2)),
            "cands": arr(b.cands, b.numCandidates, CAND_DTYPE),
            "true_start": arr(b.trueStart, b.numQueries, np.uint32),
            "cells": int(b.cells),
        }
        if spec.keepCodes:
            out["codes"] = arr(b.codes, b.totalBases, np.uint8)
            out["code_offset"] = arr(b.codeOffset, b.numQueries + 1, np.uint64)
        return out
    finally:
        tools().myersSynthFreeBatch(ctypes.byref(b))


# --------------------------------------------------------------- library

class Myers:
    """The buffer API (initMyers / send / receive / endMyers) over a packed
    reference held in host memory."""

    def __init__(self, ref_words, ref_size, num_buffers=1, avg_query=1024, cands_per_query=1,
                 buffer_bytes=0, exit_on_error=False, ref_file=None):
        L = lib()
        cfg = Config()
        L.myersDefaultConfig(ctypes.byref(cfg))
        cfg.exitOnError = 1 if exit_on_error else 0
        cfg.bufferBytes = buffer_bytes
        self._handles = ctypes.POINTER(ctypes.c_void_p)()
        if ref_file is not None:
            rc = L.myersInitEx(ctypes.byref(self._handles), num_buffers, ref_file.encode(), avg_query,
                               cands_per_query, ctypes.byref(cfg))
        else:
            ref_words = np.ascontiguousarray(ref_words, dtype=np.uint32)
            rc = L.myersInitFromMemory(ctypes.byref(self._handles), num_buffers, _ptr(ref_words),
                                       len(ref_words), ref_size, avg_query, cands_per_query,
                                       ctypes.byref(cfg))
        check(rc, "myersInit")
        self.num_buffers = num_buffers
        self._last = [0] * num_buffers       # candidates of each buffer's last send

    def buffer(self, i=0):
        return self._handles[i]

    def capacity(self, i=0):
        L, b = lib(), self.buffer(i)
        return L.getMaxPEQEntries(b), L.getMaxQueries(b), L.getMaxCandidates(b)

    def views(self, i=0):
        """numpy views of the buffer's pinned host arrays."""
        L, b = lib(), self.buffer(i)
        ne, nq, nc = self.capacity(i)

        def view(ptr, count, dtype):
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dtype)
        return {
            "peq": view(L.getPEQBuffer(b), ne * 20, np.uint32).reshape(ne, 20),
            "qinfo": view(L.getPEQInfoBuffer(b), nq * 2, np.uint32).reshape(nq, 2),
            "cands": view(L.getCandidatesBuffer(b), nc, CAND_DTYPE),
            "results": view(L.getResultsBuffer(b), nc * 2, np.uint32).reshape(nc, 2),
        }

    def send(self, peq, qinfo, cands, distance, flags=0, i=0):
        v = self.views(i)
        v["peq"][:len(peq)] = peq
        v["qinfo"][:len(qinfo)] = qinfo
        v["cands"][:len(cands)] = cands
        check(lib().myersSendEx(self.buffer(i), distance, len(peq), len(qinfo), len(cands), flags),
              "myersSendEx")
        self._last[i] = len(cands)

    def receive(self, i=0, count=None):
        rc = lib().myersReceiveEx(self.buffer(i))
        res = self.views(i)["results"][: (count if count is not None else self._last[i])].copy()
        check(rc, "myersReceiveEx")
        return res

    def verify(self, peq, qinfo, cands, distance, flags=0, i=0):
        self.send(peq, qinfo, cands, distance, flags, i)
        return self.receive(i, len(cands))

    def stats(self, i=0):
        s = Stats()
        check(lib().myersGetStats(self.buffer(i), ctypes.byref(s)), "myersGetStats")
        return s

    def variants(self):
        out = []
        for v in range(lib().myersNumVariants()):
            info = VariantInfo()
            check(lib().myersGetVariantInfo(self.buffer(0), v, ctypes.byref(info)), "myersGetVariantInfo")
            out.append({k: getattr(info, k) for k, _ in VariantInfo._fields_})
        return out

    def bin_table(self, max_words=768):
        t = np.zeros(max_words + 1, dtype=np.uint8)
        check(lib().myersGetBinTable(self.buffer(0), _ptr(t), max_words), "myersGetBinTable")
        return t

    def set_threshold(self, error_rate, i=None):
        """Threshold mode on buffer i (all buffers if None); error_rate < 0 switches it off."""
        for j in (range(self.num_buffers) if i is None else [i]):
            check(lib().myersSetThreshold(self.buffer(j), error_rate), "myersSetThreshold")

    def threshold(self, i=0):
        return lib().myersGetThreshold(self.buffer(i))

    def set_chunk_factor(self, variant, chunks):
        check(lib().myersSetChunkFactor(self.buffer(0), variant, chunks), "myersSetChunkFactor")

    def chunk_factor(self, variant):
        c = ctypes.c_uint32()
        check(lib().myersGetChunkFactor(self.buffer(0), variant, ctypes.byref(c)), "myersGetChunkFactor")
        return c.value

    def load_tuning(self, path):
        check(lib().myersLoadTuning(self.buffer(0), path.encode()), "myersLoadTuning")

    def variant_index(self, lanes, words):
        for i, d in enumerate(self.variants()):
            if d["lanes"] == lanes and d["words"] == words:
                return i
        raise KeyError((lanes, words))

    def set_bin_table(self, table):
        t = np.ascontiguousarray(table, dtype=np.uint8)
        check(lib().myersSetBinTable(self.buffer(0), _ptr(t), len(t) - 1), "myersSetBinTable")

    def close(self):
        if self._handles:
            check(lib().myersEnd(ctypes.byref(self._handles)), "myersEnd")
            self._handles = ctypes.POINTER(ctypes.c_void_p)()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceArray:
    """A device allocation filled from / read into numpy (no torch needed)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.nbytes = max(host.nbytes, 1)
        self.dtype, self.shape = host.dtype, host.shape
        self.ptr = ctypes.c_void_p()
        check(lib().myersDeviceMalloc(ctypes.byref(self.ptr), self.nbytes), "myersDeviceMalloc")
        if host.nbytes:
            check(lib().myersMemcpyHtoD(self.ptr, _ptr(host), host.nbytes), "myersMemcpyHtoD")

    @classmethod
    def empty(cls, shape, dtype):
        return cls(np.zeros(shape, dtype=dtype))

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            check(lib().myersMemcpyDtoH(_ptr(out), self.ptr, out.nbytes), "myersMemcpyDtoH")
        return out

    def free(self):
        if self.ptr:
            lib().myersDeviceFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def verify_device(m, peq_d, num_entries, qinfo_d, num_queries, cands_d, num_cands, distance, results_d,
                  flags=0, h_qinfo=None, i=0):
    hq_arr = np.ascontiguousarray(h_qinfo, dtype=np.uint32) if h_qinfo is not None else None
    hq = _ptr(hq_arr) if hq_arr is not None else None
    check(lib().myersVerifyDevice(m.buffer(i), peq_d.ptr, num_entries, qinfo_d.ptr, num_queries, cands_d.ptr,
                                  num_cands, distance, flags, results_d.ptr, hq), "myersVerifyDevice")


def synchronize(m, i=0):
    check(lib().myersSynchronize(m.buffer(i)), "myersSynchronize")


# ----------------------------------------------------------------- serving

class Server:
    """Continuous-batching server over all buffers of a Myers instance."""

    def __init__(self, myers, distance=0.1, flags=0, hbm_budget=0, poll_us=0, max_batch_candidates=0,
                 max_batch_cells=0, arena_bytes=0):
        cfg = ServeConfig()
        lib().myersServeDefaultConfig(ctypes.byref(cfg))
        cfg.distance = distance
        cfg.flags = flags
        if hbm_budget:
            cfg.hbmBudgetBytes = hbm_budget
        if poll_us:
            cfg.pollMicros = poll_us
        cfg.maxBatchCandidates = max_batch_candidates
        if max_batch_cells:
            cfg.maxBatchCells = max_batch_cells
        cfg.arenaBytes = arena_bytes
        self._h = ctypes.c_void_p()
        check(lib().myersServeStart(ctypes.byref(self._h), myers._handles, myers.num_buffers, ctypes.byref(cfg)),
              "myersServeStart")
        self._keep = []

    @staticmethod
    def request(peq, qinfo, cands):
        """A Request over numpy arrays (kept alive by the returned tuple)."""
        peq = np.ascontiguousarray(peq, dtype=np.uint32)
        qinfo = np.ascontiguousarray(qinfo, dtype=np.uint32)
        cands = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
        res = np.zeros((len(cands), 2), dtype=np.uint32)
        r = Request(peq.ctypes.data, qinfo.ctypes.data, cands.ctypes.data, res.ctypes.data,
                    len(peq), len(qinfo), len(cands), 0)
        return r, (peq, qinfo, cands, res)

    def submit(self, req):
        t = ctypes.c_uint64()
        check(lib().myersServeSubmit(self._h, ctypes.byref(req), ctypes.byref(t)), "myersServeSubmit")
        return t.value

    def wait(self, ticket):
        lat = ctypes.c_double()
        check(lib().myersServeWait(self._h, ticket, ctypes.byref(lat)), "myersServeWait")
        return lat.value

    def run_open_loop(self, reqs, arrivals_ns, queue=False):
        """Latencies (ms) and wall time (ms); with queue=True also each
        request's wait from arrival to admission into a batch (ms)."""
        n = len(reqs)
        arr = (Request * n)(*reqs)
        times = np.ascontiguousarray(arrivals_ns, dtype=np.uint64)
        lat = np.zeros(n, dtype=np.float64)
        qms = np.zeros(n, dtype=np.float64)
        wall = ctypes.c_double()
        dp = ctypes.POINTER(ctypes.c_double)
        check(lib().myersServeRunOpenLoopEx(self._h, n, arr, times.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                            lat.ctypes.data_as(dp), qms.ctypes.data_as(dp) if queue else None,
                                            ctypes.byref(wall)),
              "myersServeRunOpenLoopEx")
        return (lat, wall.value, qms) if queue else (lat, wall.value)

    def stats(self):
        s = ServeStats()
        s.structSize = ctypes.sizeof(s)
        check(lib().myersServeGetStats(self._h, ctypes.byref(s)), "myersServeGetStats")
        return s

    def stop(self):
        if self._h:
            check(lib().myersServeStop(self._h), "myersServeStop")
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
