"""myersVerifyDeviceNoAlloc: the allocation-free variant of
myersVerifyDevice (its pipeline workspace is the library's static device
array).  Called through the C-ABI and compared bit-exactly with
myersVerifyDevice on the same device batches -- two calls in a row, the
second smaller, so the second runs over what the first left in the shared
array -- for every engine and both window modes, and in threshold mode; the
oracle checks both."""
import ctypes

import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu

REF_SIZE = 3_000_000


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def make_batch(ref, seed, nq):
    # short, mid and long reads: one-lane bins, multi-lane bins and the long-read launch
    return M.synth_batch(M.make_spec(seed, nq, choices=[100, 250, 1000, 3000, 9000], sub=0.03, ins=0.005,
                                     dele=0.005, candidates=5, distance=0.1), ref, REF_SIZE)


def run(m, fn, b, flags):
    dp, dq, dc = M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"])
    dr = M.DeviceArray.empty((len(b["cands"]), 2), np.uint32)
    hq = np.ascontiguousarray(b["qinfo"], dtype=np.uint32)
    try:
        rc = fn(m.buffer(0), dp.ptr, len(b["peq"]), dq.ptr, len(b["qinfo"]), dc.ptr, len(b["cands"]),
                ctypes.c_float(0.1), flags, dr.ptr, hq.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0, M.ERRORS.get(rc, rc)
        M.synchronize(m)
        return dr.get()
    finally:
        for a in (dp, dq, dc, dr):
            a.free()


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_no_alloc_matches_verify_device(mode):
    ref = M.synth_reference(77, REF_SIZE)
    big, small = make_batch(ref, 78, 400), make_batch(ref, 79, 60)
    L = M.lib()
    with M.Myers(ref, REF_SIZE, avg_query=1000, cands_per_query=5, buffer_bytes=64 << 20) as m:
        for engine in (M.ENGINE_AUTO, M.ENGINE_PLAIN, M.ENGINE_PERSISTENT):
            for name, b in (("first", big), ("second, smaller", small)):
                announce(f"no-alloc mode={mode} engine={engine} {name}: {len(b['cands'])} candidates")
                want = run(m, L.myersVerifyDevice, b, mode | engine)
                got = run(m, L.myersVerifyDeviceNoAlloc, b, mode | engine)
                np.testing.assert_array_equal(got, want)
                exp = O.myers_batch(ref, REF_SIZE, b["peq"], b["qinfo"], b["cands"], 0.1, mode)
                np.testing.assert_array_equal(got, exp)
        # threshold mode: the band probe and the span order run in the same pipeline
        m.set_threshold(0.2)
        for name, b in (("first", big), ("second, smaller", small)):
            announce(f"no-alloc threshold mode={mode} {name}: {len(b['cands'])} candidates")
            want = run(m, L.myersVerifyDevice, b, mode | M.ENGINE_PERSISTENT)
            got = run(m, L.myersVerifyDeviceNoAlloc, b, mode | M.ENGINE_PERSISTENT)
            np.testing.assert_array_equal(got, want)
            full = O.myers_batch(ref, REF_SIZE, b["peq"], b["qinfo"], b["cands"], 0.1, mode)
            np.testing.assert_array_equal(got, O.apply_threshold(b["qinfo"], b["cands"], full, 0.2))
        m.set_threshold(-1.0)


def test_no_alloc_refuses_oversize_and_graphs():
    ref = M.synth_reference(80, REF_SIZE)
    L = M.lib()
    with M.Myers(ref, REF_SIZE, avg_query=1000, cands_per_query=5, buffer_bytes=64 << 20) as m:
        dummy = M.DeviceArray(np.zeros(16, dtype=np.uint32))
        try:
            # both are refused before any device work (the pointers are never read)
            rc = L.myersVerifyDeviceNoAlloc(m.buffer(0), dummy.ptr, 1, dummy.ptr, 1, dummy.ptr,
                                            M.NOALLOC_MAX_CANDIDATES + 1, ctypes.c_float(0.1), 0, dummy.ptr, None)
            assert rc == M.E_CAPACITY
            rc = L.myersVerifyDeviceNoAlloc(m.buffer(0), dummy.ptr, 1, dummy.ptr, 1, dummy.ptr, 1,
                                            ctypes.c_float(0.1), M.USE_GRAPH, dummy.ptr, None)
            assert rc == M.E_INVALID_ARGUMENT
        finally:
            dummy.free()
