"""The C-ABI boundary: headers, exported symbols, record layouts, and a C
caller written against the reference API (src-api/test.c style) that builds
with plain -std=c99 and links against libmyers_amd.so."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import myers_amd as M

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(REPO, "include")
CDIR = os.path.join(REPO, "tests", "c")

HEADERS = {
    "myers-interface.h": M.LIB_PATH,
    "myers-amd.h": M.LIB_PATH,
    "myers-tools.h": M.TOOLS_PATH,
}


def declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//.*", "", text)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)       # preprocessor lines
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name in ("sizeof", "if", "while", "return"):
            continue
        names.add(name)
    return sorted(names)


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(header):
    lib = ctypes.CDLL(HEADERS[header])
    names = declared_functions(header)
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_api_names_present():
    """The ten entry points of src-api/myers-interface.h:31-46."""
    want = ["getPEQBuffer", "getCandidatesBuffer", "getPEQInfoBuffer", "getResultsBuffer", "getMaxPEQEntries",
            "getMaxCandidates", "getMaxQueries", "initMyers", "sendMyersBuffer", "receiveMyersBuffer", "endMyers"]
    assert set(want) <= set(declared_functions("myers-interface.h"))
    lib = ctypes.CDLL(M.LIB_PATH)
    for n in want:
        getattr(lib, n)


def test_record_layouts_compile():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "sizes")
        subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, os.path.join(CDIR, "record_sizes.c"), "-o", exe],
                       check=True)
        subprocess.run([exe], check=True)


def build_api_caller(outdir):
    exe = os.path.join(outdir, "api_caller")
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", INCLUDE, os.path.join(CDIR, "api_caller.c"),
                    "-o", exe, "-L", M.LIB_DIR, "-lmyers_amd", f"-Wl,-rpath,{M.LIB_DIR}",
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


def test_c_caller_links_with_plain_c99():
    with tempfile.TemporaryDirectory() as d:
        assert os.path.exists(build_api_caller(d))


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_init_without_device_reports_error():
    ref = np.zeros(16, dtype=np.uint32)
    with pytest.raises(M.MyersError) as e:
        M.Myers(ref, 128)
    assert e.value.code == 20   # MYERS_E_NO_DEVICE


def test_error_strings():
    lib = M.lib()
    assert lib.myersGetErrorString(1) == b"Myers - Error: opening file"
    assert lib.myersGetErrorString(3).startswith(b"Myers - Error: there aren't enough GPU memory")
    assert lib.myersGetErrorString(12345) == b"Myers - Unknown error"


def test_missing_reference_file():
    h = ctypes.POINTER(ctypes.c_void_p)()
    rc = M.lib().myersInitEx(ctypes.byref(h), 1, b"/nonexistent/ref.4bits.ref", 100, 1, None)
    # no device here -> E_NO_DEVICE may come first only after the file is read
    assert rc == 1


def test_truncated_reference_file():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "bad.4bits.ref")
        with open(p, "wb") as f:
            f.write(np.array([10, 80], dtype=np.uint32).tobytes() + b"\x00" * 12)
        h = ctypes.POINTER(ctypes.c_void_p)()
        assert M.lib().myersInitEx(ctypes.byref(h), 1, p.encode(), 100, 1, None) == 2


@pytest.mark.gpu
def test_c_caller_end_to_end():
    """The C caller over a .4bits.ref file and a batch file, checked against the oracle."""
    import oracle_lib as O
    ref_size = 500_000
    ref = M.synth_reference(13, ref_size)
    b = M.synth_batch(M.make_spec(14, 300, choices=[100, 300, 1000], sub=0.03, ins=0.005, dele=0.005,
                                  candidates=9, distance=0.1), ref, ref_size)
    with tempfile.TemporaryDirectory() as d:
        exe = build_api_caller(d)
        rf = os.path.join(d, "ref.4bits.ref")
        assert M.tools().myersWriteReferenceFile(rf.encode(), M._ptr(ref), len(ref), ref_size) == 0
        bf = os.path.join(d, "batch.bin")
        with open(bf, "wb") as f:
            f.write(np.array([len(b["peq"]), len(b["qinfo"]), len(b["cands"])], dtype=np.uint32).tobytes())
            f.write(b["peq"].tobytes())
            f.write(b["qinfo"].tobytes())
            f.write(b["cands"].tobytes())
        out = os.path.join(d, "res.txt")
        print("[gpu-case] c caller", flush=True)
        subprocess.run([exe, "0.1", rf, bf, out], check=True, timeout=120)
        lines = open(out).read().split("\n")
        n = int(lines[0])
        got = np.zeros((n, 2), dtype=np.uint32)
        for line in lines[1:1 + n]:
            i, s, c = (int(x) for x in line.split())
            got[i] = (c, s)
    exp = O.myers_batch(ref, ref_size, b["peq"], b["qinfo"], b["cands"], 0.1, 0)
    np.testing.assert_array_equal(got, exp)


def test_reference_size_limit_before_device_work():
    """The kernels index reference words with signed 32-bit offsets, so
    myersInitFromMemory refuses 2^31 - 64 words or more (with the 64 guard
    words) before it touches the device; the same call runs here, without a
    GPU, and the refused sizes never read the word array."""
    words = np.zeros(4, dtype=np.uint32)
    handles = ctypes.POINTER(ctypes.c_void_p)()
    L = M.lib()
    for nw in (2**31 - 64, 2**31, 2**32 - 1):
        rc = L.myersInitFromMemory(ctypes.byref(handles), 1, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                   nw, 8 * nw, 1000, 11, None)
        assert rc == M.E_REFERENCE_TOO_LARGE, (nw, rc)
        assert not handles
    # a size larger than the words hold is an argument error, also before device work
    rc = L.myersInitFromMemory(ctypes.byref(handles), 1, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                               4, 33, 1000, 11, None)
    assert rc == M.E_INVALID_ARGUMENT
