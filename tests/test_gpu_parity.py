"""Bit-exact parity of the HIP path (through the C-ABI library) with the oracle.

Every comparison is exact: (column, score) per candidate.  Inputs are the
golden fixtures written by the reference's own programs, seeded synthetic
batches, and hand-made edge cases.
"""
import glob
import os
import threading

import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O
from test_oracle import golden_batch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def synth_case(seed, ref_size, nq, **kw):
    ref = M.synth_reference(seed, ref_size)
    b = M.synth_batch(M.make_spec(seed + 1, nq, **kw), ref, ref_size)
    return ref, b


def concat_batches(parts):
    """One batch from several: Peq entries, reads and candidates appended in
    order, with posEntry and the candidates' read ids shifted."""
    peq, qinfo, cands, e_off, q_off = [], [], [], 0, 0
    for b in parts:
        qi = b["qinfo"].copy()
        qi[:, 0] += np.uint32(e_off)
        c = b["cands"].copy()
        c["query"] += np.uint32(q_off)
        peq.append(b["peq"])
        qinfo.append(qi)
        cands.append(c)
        e_off += len(b["peq"])
        q_off += len(b["qinfo"])
    return {"peq": np.concatenate(peq), "qinfo": np.concatenate(qinfo), "cands": np.concatenate(cands)}


def oracle_for(ref, ref_size, b, distance, mode, max_query_len=0):
    return O.myers_batch(ref, ref_size, b["peq"], b["qinfo"], b["cands"], distance, mode, max_query_len)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_golden_fixtures(path):
    announce(f"golden {os.path.basename(path)}")
    z, peq, qinfo, cands = golden_batch(path)
    mode = int(z["mode"])
    with M.Myers(z["ref_words"], int(z["ref_size"]), avg_query=1024, cands_per_query=8,
                 buffer_bytes=64 << 20) as m:
        got = m.verify(peq, qinfo, cands, float(z["distance"]), flags=mode)
    keep = ~z["reads_past_end"]
    if int(z["max_query_len"]):
        keep &= qinfo[z["cand_query"], 1] <= int(z["max_query_len"])
    np.testing.assert_array_equal(got[keep], z["results"][keep])
    # everything (incl. > 4096 and the past-the-end windows, which read the
    # all-N guard) against the oracle, which has the same guard rule
    ref_o = O.myers_batch(z["ref_words"], z["ref_size"], peq, qinfo, cands, float(z["distance"]), mode, 0)
    np.testing.assert_array_equal(got, ref_o)


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_short_reads_full(mode):
    """4 Mbase seed-42 reference, 10K x 100 bp, 11 candidates each (110K)."""
    announce(f"10K x 100-base reads mode={mode}")
    ref, b = synth_case(42, 4_000_000, 10_000, length=100, sub=0.02, ins=0.0025, dele=0.0025,
                        candidates=11, distance=0.1)
    with M.Myers(ref, 4_000_000, avg_query=100, cands_per_query=11, buffer_bytes=256 << 20) as m:
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode)
        st = m.stats()
    assert st.launches == 1 and st.binned == 0
    np.testing.assert_array_equal(got, oracle_for(ref, 4_000_000, b, 0.1, mode))


def test_mixed_lengths_binned():
    """Reads of 100-1000 bases mixed in one batch through device-side binning (several variants)."""
    announce("mixed lengths")
    ref, b = synth_case(5, 2_000_000, 600, choices=[100, 200, 400, 600, 800, 1000], err_range=(0.01, 0.05),
                        sub_frac=0.8, poisson=11, cap=64, distance=0.1)
    with M.Myers(ref, 2_000_000, avg_query=550, cands_per_query=11, buffer_bytes=256 << 20) as m:
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
        st = m.stats()
    assert st.binned == 1 and st.launches == 1          # one queue launch over the six bins
    np.testing.assert_array_equal(got, oracle_for(ref, 2_000_000, b, 0.1, O.MODE_ALIGNED))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_long_reads(mode):
    """Long noisy reads: 2-20 kb, 12% error, over the multi-lane variants."""
    announce(f"long reads mode={mode}")
    ref, b = synth_case(11, 3_000_000, 40, log_uniform=(2000, 20000), err_range=(0.12, 0.12), sub_frac=0.5,
                        candidates=4, distance=0.1)
    with M.Myers(ref, 3_000_000, avg_query=8000, cands_per_query=4, buffer_bytes=256 << 20) as m:
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode)
    np.testing.assert_array_equal(got, oracle_for(ref, 3_000_000, b, 0.1, mode))


def test_every_variant():
    """Force each compiled (lanes, words) variant over queries up to its capacity."""
    ref_size = 1_000_000
    ref = M.synth_reference(21, ref_size)
    with M.Myers(ref, ref_size, avg_query=2000, cands_per_query=6, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        default = m.bin_table()
        for v, info in enumerate(variants):
            cap = info["maxQuerySize"]
            announce(f"variant {v} lanes={info['lanes']} words={info['words']}")
            lengths = sorted({max(1, cap - 32 * info["words"] * info["lanes"] + 1), cap // 2 + 1, cap - 17, cap})
            spec = M.make_spec(100 + v, 8, choices=[x for x in lengths if x >= 1], sub=0.04, ins=0.01, dele=0.01,
                               candidates=6, distance=0.1)
            b = M.synth_batch(spec, ref, ref_size)
            table = default.copy()
            table[:] = v
            table[(np.arange(len(table)) * 32) > cap] = default[(np.arange(len(table)) * 32) > cap]
            m.set_bin_table(table)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
            np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, O.MODE_ALIGNED),
                                          err_msg=f"variant {info}")
        m.set_bin_table(default)


def test_edge_cases():
    """Skip rule boundaries, m = 1, N-only text, text nibbles 5..15, windows at the reference end."""
    announce("edge cases")
    rng = np.random.default_rng(99)
    size = 50_003
    words = rng.integers(0, 2**32, size=(size + 7) // 8, dtype=np.uint64).astype(np.uint32)   # any nibble 0..15
    words[100:140] = 0x44444444                                                                # an N run
    qlens = [1, 2, 31, 32, 33, 64, 100, 128, 129, 500, 1024, 1025, 4096, 4097]
    peqs, qinfo, qc, off = [], [], [], 0
    for m_ in qlens:
        q = rng.integers(0, 5, size=m_).astype(np.uint8)
        p = O.build_peq(q)
        peqs.append(p)
        qinfo.append([off, m_])
        off += len(p)
    peq = np.concatenate(peqs)
    qinfo = np.array(qinfo, dtype=np.uint32)
    cq, cp = [], []
    for i, m_ in enumerate(qlens):
        n = O.window_len(m_, 0.2)
        for p in [0, 1, 7, 8, 803, 805, size - n - 1, size - n, size - n - 9, size - 1, size, 2**40,
                  int(rng.integers(0, size - n - 1))]:
            cq.append(i)
            cp.append(p)
    cands = O.make_cands(cq, cp)
    with M.Myers(words, size, avg_query=512, cands_per_query=13, buffer_bytes=64 << 20) as m:
        for mode in (M.WINDOW_ALIGNED, M.WINDOW_EXACT):
            got = m.verify(peq, qinfo, cands, 0.2, flags=mode)
            np.testing.assert_array_equal(got, O.myers_batch(words, size, peq, qinfo, cands, 0.2, mode))


def test_query_too_long_gets_sentinel():
    announce("too long")
    ref_size = 200_000
    ref = M.synth_reference(3, ref_size)
    rng = np.random.default_rng(1)
    lens = [100, 24577, 300]
    peqs, qinfo, off = [], [], 0
    for m_ in lens:
        p = O.build_peq(rng.integers(0, 4, size=m_).astype(np.uint8))
        peqs.append(p)
        qinfo.append([off, m_])
        off += len(p)
    cands = O.make_cands([0, 1, 2, 1], [10, 20, 30, 40])
    with M.Myers(ref, ref_size, avg_query=8192, cands_per_query=1, buffer_bytes=64 << 20) as m:
        peq, qinfo = np.concatenate(peqs), np.array(qinfo, dtype=np.uint32)
        m.send(peq, qinfo, cands, 0.1)
        with pytest.raises(M.MyersError) as e:
            m.receive()
        assert e.value.code == 19
        got = m.views()["results"][:4]
        assert tuple(got[1]) == (0, 0xFFFFFFFF) and tuple(got[3]) == (0, 0xFFFFFFFF)
        exp = O.myers_batch(ref, ref_size, peq, qinfo, cands, 0.1, 0)
        np.testing.assert_array_equal(got[[0, 2]], exp[[0, 2]])


def test_device_resident_equals_buffer_api():
    announce("device resident")
    ref, b = synth_case(8, 1_000_000, 500, length=250, sub=0.03, ins=0.005, dele=0.005, candidates=7,
                        distance=0.1)
    with M.Myers(ref, 1_000_000, avg_query=250, cands_per_query=7, buffer_bytes=64 << 20) as m:
        host = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
        dp, dq, dc = M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"])
        dr = M.DeviceArray.empty((len(b["cands"]), 2), np.uint32)
        for hint in (b["qinfo"], None):
            M.verify_device(m, dp, len(b["peq"]), dq, len(b["qinfo"]), dc, len(b["cands"]), 0.1, dr, h_qinfo=hint)
            M.synchronize(m)
            np.testing.assert_array_equal(dr.get(), host)
        for a in (dp, dq, dc, dr):
            a.free()
    np.testing.assert_array_equal(host, oracle_for(ref, 1_000_000, b, 0.1, 0))


def test_buffers_from_threads():
    """Several buffers driven concurrently from host threads (src-api/test.c:184-193 pattern)."""
    announce("threads")
    ref_size = 2_000_000
    ref = M.synth_reference(77, ref_size)
    batches = [M.synth_batch(M.make_spec(200 + i, 300, length=150 + 50 * i, sub=0.03, ins=0.005, dele=0.005,
                                         candidates=11, distance=0.1), ref, ref_size) for i in range(4)]
    out = [None] * 4
    with M.Myers(ref, ref_size, num_buffers=4, avg_query=300, cands_per_query=11, buffer_bytes=64 << 20) as m:
        def work(i):
            b = batches[i]
            for _ in range(3):
                out[i] = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, i=i)
        th = [threading.Thread(target=work, args=(i,)) for i in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    for i, b in enumerate(batches):
        np.testing.assert_array_equal(out[i], oracle_for(ref, ref_size, b, 0.1, 0))


def test_repeat_sends_are_independent():
    """Reusing a buffer: results depend only on the batch (the reference's numWarps grew per send)."""
    announce("repeat sends")
    ref, b = synth_case(9, 1_000_000, 400, choices=[100, 1000], sub=0.03, ins=0.005, dele=0.005, candidates=5,
                        distance=0.1)
    with M.Myers(ref, 1_000_000, avg_query=550, cands_per_query=5, buffer_bytes=64 << 20) as m:
        first = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
        half = len(b["cands"]) // 2
        m.verify(b["peq"], b["qinfo"], b["cands"][:half], 0.1)
        again = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
    np.testing.assert_array_equal(first, again)


@pytest.mark.parametrize("engine", [M.ENGINE_PLAIN, M.ENGINE_PERSISTENT, M.ENGINE_BINNED,
                                    M.ENGINE_PERSISTENT | M.NO_LONG_LAUNCH])
def test_engines_mixed_lengths(engine):
    """Plain per-bin launches, the queue launches, per-bin resident grids and
    the queue launches without the long-read launch give identical results."""
    announce(f"engine {engine}")
    ref, b = synth_case(31, 3_000_000, 900, choices=[64, 100, 200, 333, 600, 1000, 1500, 3000],
                        err_range=(0.01, 0.08), sub_frac=0.7, poisson=9, cap=40, distance=0.1)
    with M.Myers(ref, 3_000_000, avg_query=800, cands_per_query=9, buffer_bytes=256 << 20) as m:
        for mode in (M.WINDOW_ALIGNED, M.WINDOW_EXACT):
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | engine)
            assert m.stats().binned == 1
            np.testing.assert_array_equal(got, oracle_for(ref, 3_000_000, b, 0.1, mode))


def test_persistent_single_length():
    """The persistent engine on a one-bin batch (binned, one queue)."""
    announce("persistent single bin")
    ref, b = synth_case(32, 1_000_000, 700, length=1000, sub=0.03, ins=0.005, dele=0.005, candidates=11,
                        distance=0.1)
    with M.Myers(ref, 1_000_000, avg_query=1000, cands_per_query=11, buffer_bytes=256 << 20) as m:
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=M.ENGINE_PERSISTENT)
    np.testing.assert_array_equal(got, oracle_for(ref, 1_000_000, b, 0.1, 0))


def test_auto_engine_by_batch_size():
    """The auto engine runs a one-length batch on the plain engine (no
    binning) up to 2048 candidates per CU and on the queue launch (binned)
    above that (256 CUs: 524,288 candidates); both equal the oracle."""
    ref_size = 2_000_000
    ref = M.synth_reference(33, ref_size)
    announce("auto engine by batch size")
    with M.Myers(ref, ref_size, avg_query=100, cands_per_query=11, buffer_bytes=256 << 20) as m:
        for reads, binned in ((10_000, 0), (50_000, 1)):
            b = M.synth_batch(M.make_spec(34 + reads, reads, length=100, sub=0.03, ins=0.005, dele=0.005,
                                          candidates=11, distance=0.1), ref, ref_size)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
            st = m.stats()
            assert st.binned == binned, f"{len(b['cands'])} candidates"
            assert st.launches == 1
            np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, 0))


@pytest.mark.parametrize("engine", [M.ENGINE_PLAIN, M.ENGINE_PERSISTENT, M.ENGINE_BINNED])
def test_graph_replays_match_eager(engine):
    """A captured pipeline, replayed several times with eager work and
    synchronisations in between, matches the eager result every time."""
    announce(f"graph replays engine {engine}")
    ref, b = synth_case(33, 2_000_000, 800, choices=[100, 400, 1000], err_range=(0.01, 0.05), poisson=11,
                        distance=0.1)
    ref2, b2 = synth_case(34, 2_000_000, 300, length=150, sub=0.03, ins=0.005, dele=0.005, candidates=5,
                          distance=0.1)
    with M.Myers(ref, 2_000_000, avg_query=500, cands_per_query=11, buffer_bytes=256 << 20) as m:
        dp, dq, dc = M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"])
        dr = M.DeviceArray.empty((len(b["cands"]), 2), np.uint32)
        nq, nc, ne = len(b["qinfo"]), len(b["cands"]), len(b["peq"])
        M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=engine, h_qinfo=b["qinfo"])
        M.synchronize(m)
        eager = dr.get()
        np.testing.assert_array_equal(eager, oracle_for(ref, 2_000_000, b, 0.1, 0))
        for rep in range(4):
            dr_clear = np.zeros((nc, 2), np.uint32)
            M.check(M.lib().myersMemcpyHtoD(dr.ptr, M._ptr(dr_clear), dr_clear.nbytes), "clear")
            M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=engine | M.USE_GRAPH, h_qinfo=b["qinfo"])
            M.synchronize(m)
            np.testing.assert_array_equal(dr.get(), eager, err_msg=f"replay {rep}")
            # eager work of another shape on the same buffer between replays
            other = m.verify(b2["peq"], b2["qinfo"], b2["cands"], 0.1)
            assert len(other) == len(b2["cands"])
        assert m.stats().graphCaptures == 1
        for x in (dp, dq, dc, dr):
            x.free()


@pytest.mark.parametrize("engine", [M.ENGINE_PLAIN, M.ENGINE_PERSISTENT, M.ENGINE_BINNED])
def test_multiwave_variants_in_mixed_batches(engine):
    """Long reads on the multi-wave variants (one candidate per workgroup of
    2 or 4 waves) next to one-wave variants in one binned batch."""
    announce(f"multi-wave engine {engine}")
    ref, b = synth_case(35, 3_000_000, 120, choices=[300, 1000, 3000, 7000, 12000, 16000, 20000, 24000],
                        err_range=(0.01, 0.12), sub_frac=0.5, candidates=3, distance=0.1)
    with M.Myers(ref, 3_000_000, avg_query=6000, cands_per_query=3, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        wide = {(v["lanes"], v["words"]): i for i, v in enumerate(variants) if v["lanes"] > 64}
        assert set(wide) >= {(128, 4), (128, 6), (256, 2), (256, 3)}
        table = m.bin_table()
        w = np.arange(len(table))
        table[(w > 128) & (w <= 512)] = wide[(256, 2)]
        table[(w > 512) & (w <= 640)] = wide[(128, 6)]
        table[w > 640] = wide[(256, 3)]
        m.set_bin_table(table)
        for mode in (M.WINDOW_ALIGNED, M.WINDOW_EXACT):
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | engine)
            np.testing.assert_array_equal(got, oracle_for(ref, 3_000_000, b, 0.1, mode))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_long_read_bins_share_one_launch(mode):
    """Every one-wave long-read variant (4..64 lanes) holds one bin of the same
    batch, next to short-read bins: the long bins run in the one long-read
    launch, whose waves switch variants between chunks."""
    announce(f"long-read launch mode={mode}")
    ref_size = 3_000_000
    ref = M.synth_reference(36, ref_size)
    with M.Myers(ref, ref_size, avg_query=6000, cands_per_query=3, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        longs = sorted((v["lanes"] * v["words"], i) for i, v in enumerate(variants) if 4 <= v["lanes"] <= 64)
        table = m.bin_table()
        used, lengths = set(), []
        for cap, i in longs:
            w = cap
            while w in used:
                w -= 1
            used.add(w)
            table[w] = i
            lengths.append(32 * w - 5)
        lengths += [100, 400, 1000]                       # short-read bins beside them
        m.set_bin_table(table)
        b = concat_batches([M.synth_batch(M.make_spec(37 + k, 1, length=n, sub=0.06, ins=0.02, dele=0.02,
                                                      candidates=2, distance=0.1), ref, ref_size)
                            for k, n in enumerate(lengths)])
        assert sorted(b["qinfo"][:, 1].tolist()) == sorted(lengths)
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | M.ENGINE_PERSISTENT)
        st = m.stats()
        wq = (b["qinfo"][:, 1].astype(np.int64) + 31) // 32
        bins = set(int(table[x]) for x in wq)
        nlong = sum(1 for v in bins if 4 <= variants[v]["lanes"] <= 64)
        assert nlong >= 10
        assert st.launches == 2                # the long-read launch and one queue launch for the short bins
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, mode))


def test_long_read_launch_wide_rounds():
    """Long reads with far more candidates than the batch average: the host's
    estimate puts their bin on whole workgroups, and the device finds more
    candidates than workgroups, so later rounds pull from the wide counter."""
    announce("long-read launch, wide rounds")
    ref_size = 3_000_000
    ref = M.synth_reference(38, ref_size)
    parts = [M.synth_batch(M.make_spec(39, 1500, length=100, sub=0.02, ins=0.005, dele=0.005, candidates=1,
                                       distance=0.1), ref, ref_size),
             M.synth_batch(M.make_spec(40, 20, length=9000, sub=0.06, ins=0.02, dele=0.02, candidates=20,
                                       distance=0.1), ref, ref_size),
             M.synth_batch(M.make_spec(41, 20, length=4000, sub=0.06, ins=0.02, dele=0.02, candidates=4,
                                       distance=0.1), ref, ref_size)]
    b = concat_batches(parts)
    with M.Myers(ref, ref_size, avg_query=400, cands_per_query=2, buffer_bytes=256 << 20) as m:
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=M.ENGINE_PERSISTENT)
        st = m.stats()
        table = m.bin_table()
        bins = set(int(table[x]) for x in (b["qinfo"][:, 1].astype(np.int64) + 31) // 32)
        assert len(bins) == 3 and st.launches == 2       # the two long bins share one launch
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, O.MODE_ALIGNED))


def test_graph_replays_long_read_launch():
    """The long-read launch (wide phase and one-wave phase) inside a captured
    graph: every replay, with eager work in between, matches the eager run."""
    announce("graph replays, long-read launch")
    ref_size = 3_000_000
    ref = M.synth_reference(42, ref_size)
    b = concat_batches([M.synth_batch(M.make_spec(43 + k, n, length=length, sub=0.05, ins=0.02, dele=0.02,
                                                  candidates=c, distance=0.1), ref, ref_size)
                        for k, (n, length, c) in enumerate([(200, 300, 6), (6, 3000, 3), (4, 9000, 3),
                                                            (3, 17000, 2)])])
    with M.Myers(ref, ref_size, avg_query=600, cands_per_query=5, buffer_bytes=256 << 20) as m:
        dp, dq, dc = M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"])
        nq, nc, ne = len(b["qinfo"]), len(b["cands"]), len(b["peq"])
        dr = M.DeviceArray.empty((nc, 2), np.uint32)
        M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=M.ENGINE_PERSISTENT, h_qinfo=b["qinfo"])
        M.synchronize(m)
        eager = dr.get()
        assert m.stats().launches < len(set(m.bin_table()[(b["qinfo"][:, 1].astype(np.int64) + 31) // 32]))
        np.testing.assert_array_equal(eager, oracle_for(ref, ref_size, b, 0.1, O.MODE_ALIGNED))
        for rep in range(3):
            clear = np.zeros((nc, 2), np.uint32)
            M.check(M.lib().myersMemcpyHtoD(dr.ptr, M._ptr(clear), clear.nbytes), "clear")
            M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=M.ENGINE_PERSISTENT | M.USE_GRAPH,
                            h_qinfo=b["qinfo"])
            M.synchronize(m)
            np.testing.assert_array_equal(dr.get(), eager, err_msg=f"replay {rep}")
            m.verify(b["peq"][:8], b["qinfo"][:1], b["cands"][:3], 0.1)     # eager work in between
        assert m.stats().graphCaptures == 1
        for x in (dp, dq, dc, dr):
            x.free()


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_multiwave_windows_over_65536_columns(mode):
    """Distance 1.0 gives 20-24 kb reads windows of 60,000-73,700 columns: the
    multi-wave top wave keeps its first minimum as (score, column) pairs there
    instead of one 16-bit-column key."""
    announce(f"multi-wave wide windows mode={mode}")
    ref_size = 2_000_000
    ref = M.synth_reference(44, ref_size)
    b = M.synth_batch(M.make_spec(45, 3, choices=[20000, 22000, 24500], sub=0.05, ins=0.02, dele=0.02,
                                  candidates=2, distance=1.0), ref, ref_size)
    with M.Myers(ref, ref_size, avg_query=22000, cands_per_query=2, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        wide = {(v["lanes"], v["words"]): i for i, v in enumerate(variants) if v["lanes"] > 64}
        table = m.bin_table()
        w = np.arange(len(table))
        table[w > 512] = wide[(256, 3)]
        m.set_bin_table(table)
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 1.0, flags=mode | M.ENGINE_PLAIN)
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 1.0, mode))


def test_long_read_launch_edge_cases():
    """Long-read bins in the one launch with nothing to do: reads without
    candidates, candidates that all take the skip rule, and a batch of long
    reads only (no short-read bin beside the launch)."""
    announce("long-read launch edge cases")
    ref_size = 3_000_000
    ref = M.synth_reference(46, ref_size)
    parts = [M.synth_batch(M.make_spec(47 + k, n, length=length, sub=0.05, ins=0.02, dele=0.02,
                                       candidates=c, distance=0.1), ref, ref_size)
             for k, (n, length, c) in enumerate([(50, 300, 4), (3, 6000, 2), (3, 12000, 2), (2, 19000, 2)])]
    b = concat_batches(parts)
    with M.Myers(ref, ref_size, avg_query=800, cands_per_query=3, buffer_bytes=256 << 20) as m:
        # (1) the 12 kb reads keep their Peq and qinfo but lose their candidates
        keep = b["qinfo"][b["cands"]["query"], 1] != 12000
        b1 = dict(b, cands=b["cands"][keep])
        got = m.verify(b1["peq"], b1["qinfo"], b1["cands"], 0.1, flags=M.ENGINE_PERSISTENT)
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b1, 0.1, O.MODE_ALIGNED))
        # (2) every long candidate past the end of the reference
        c2 = b["cands"].copy()
        long_c = b["qinfo"][c2["query"], 1] > 1000
        c2["position"][long_c] = ref_size - 10
        b2 = dict(b, cands=c2)
        got = m.verify(b2["peq"], b2["qinfo"], b2["cands"], 0.1, flags=M.ENGINE_PERSISTENT)
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b2, 0.1, O.MODE_ALIGNED))
        assert np.all(got[long_c] == [0, 0xFFFFFFFF])
        # (3) long reads only
        b3 = concat_batches(parts[1:])
        got = m.verify(b3["peq"], b3["qinfo"], b3["cands"], 0.1, flags=M.ENGINE_PERSISTENT)
        assert m.stats().launches == 1
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b3, 0.1, O.MODE_ALIGNED))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_queue_launch_mixed_short_reads(mode):
    """Reads of 100-1000 bases (six one-lane bins) through the persistent engine: one
    queue launch over every bin, bit-exact against the oracle."""
    announce(f"queue launch mixed 100-1000-base reads mode={mode}")
    ref, b = synth_case(41, 3_000_000, 1500, choices=[100, 200, 400, 600, 800, 1000], err_range=(0.01, 0.05),
                        sub_frac=0.8, poisson=11, cap=64, distance=0.1)
    with M.Myers(ref, 3_000_000, avg_query=550, cands_per_query=11, buffer_bytes=256 << 20) as m:
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | M.ENGINE_PERSISTENT)
        st = m.stats()
    assert st.binned == 1 and st.launches == 1 and st.variantsUsed == 6
    np.testing.assert_array_equal(got, oracle_for(ref, 3_000_000, b, 0.1, mode))


def test_queue_launch_every_variant():
    """Every variant of a queue class, forced for every read length up to its
    capacity, through the queue launch (the long-read launch switched off),
    and a batch holding a bin of every one-lane and every multi-lane queue
    variant at once (two queue launches)."""
    ref_size = 1_000_000
    ref = M.synth_reference(42, ref_size)
    with M.Myers(ref, ref_size, avg_query=2000, cands_per_query=6, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        default = m.bin_table()
        queue = [v for v, info in enumerate(variants) if info["lanes"] == 1 or
                 (info["lanes"] <= 64 and info["words"] <= 12)]
        assert len(queue) >= 30
        flags = M.ENGINE_PERSISTENT | M.NO_LONG_LAUNCH
        for v in queue:
            info = variants[v]
            cap = info["maxQuerySize"]
            announce(f"queue variant {v} lanes={info['lanes']} words={info['words']}")
            lengths = sorted({max(1, cap - 32 * info["words"] * info["lanes"] + 1), cap // 2 + 1, cap - 17, cap})
            b = M.synth_batch(M.make_spec(300 + v, 8, choices=[x for x in lengths if x >= 1], sub=0.04,
                                          ins=0.01, dele=0.01, candidates=6, distance=0.1), ref, ref_size)
            table = default.copy()
            w = np.arange(len(table))
            table[(w >= 1) & (w * 32 <= cap)] = v
            m.set_bin_table(table)
            mode = M.WINDOW_EXACT if v % 2 else M.WINDOW_ALIGNED
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | flags)
            assert m.stats().launches == 1
            np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, mode), err_msg=f"variant {info}")
        m.set_bin_table(default)
        # one bin per queue variant: lengths spread so that each word count maps to its own variant
        table = default.copy()
        lens, used = [], []
        for w in range(1, min(len(table), 24576 // 32 + 1)):
            if table[w] in queue and table[w] not in used:
                used.append(int(table[w]))
                lens.append(32 * w - 5)
        announce(f"queue launch with {len(used)} bins")
        b = concat_batches([M.synth_batch(M.make_spec(399 + k, 6 * len(lens[k:k + 8]), choices=lens[k:k + 8],
                                                      sub=0.04, ins=0.01, dele=0.01, candidates=3, distance=0.1),
                                          ref, ref_size) for k in range(0, len(lens), 8)])
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=flags)
        assert m.stats().launches == 2
        np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, O.MODE_ALIGNED))


def test_queue_launch_shared_staging():
    """The multi-lane queue kernel's two-wave workgroups share one staged Peq:
    every multi-lane queue variant with one candidate per read (more distinct
    reads in a chunk pair than query slots: several passes), with three per
    read (reads straddling the two waves), and with chunk counts that leave
    the pair's second wave empty, in both window modes, against the oracle."""
    ref_size = 1_000_000
    ref = M.synth_reference(43, ref_size)
    with M.Myers(ref, ref_size, avg_query=4000, cands_per_query=3, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        default = m.bin_table()
        multi = [v for v, info in enumerate(variants) if 2 <= info["lanes"] <= 64 and info["words"] <= 12]
        assert len(multi) >= 15
        flags = M.ENGINE_PERSISTENT | M.NO_LONG_LAUNCH
        for v in multi:
            info = variants[v]
            cap = info["maxQuerySize"]
            slots = 64 // info["lanes"]
            table = default.copy()
            w = np.arange(len(table))
            table[(w >= 1) & (w * 32 <= cap)] = v
            m.set_bin_table(table)
            lengths = sorted({max(1, cap - 32 * info["words"] * info["lanes"] + 1), cap - 9, cap})
            for k, (per_read, reads) in enumerate([(1, 2 * slots + 1), (3, (2 * slots) // 3 + 2)]):
                announce(f"shared staging variant {v} lanes={info['lanes']} words={info['words']} "
                         f"candidates/read={per_read} reads={reads}")
                b = M.synth_batch(M.make_spec(700 + 4 * v + k, reads, choices=lengths, sub=0.04, ins=0.01,
                                              dele=0.01, candidates=per_read, distance=0.1), ref, ref_size)
                mode = M.WINDOW_EXACT if (v + k) % 2 else M.WINDOW_ALIGNED
                got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | flags)
                assert m.stats().launches == 1
                np.testing.assert_array_equal(got, oracle_for(ref, ref_size, b, 0.1, mode),
                                              err_msg=f"variant {info} candidates/read {per_read}")
        m.set_bin_table(default)


def test_reference_beyond_32_bit_positions():
    """A 4.5 Gbase reference (loaded from memory with 64-bit sizes): random
    loci over the whole reference, candidates placed around 2^31 and 2^32,
    and the skip-rule edges at the end, through the plain, persistent and
    per-bin engines in both window modes, against the oracle."""
    size = 4_500_000_000
    announce(f"reference of {size} bases")
    ref = M.synth_reference(91, size)
    assert len(ref) * 8 >= size
    b = concat_batches([M.synth_batch(M.make_spec(92 + k, 40, length=n, sub=0.03, ins=0.005, dele=0.005,
                                                  candidates=4, distance=0.1), ref, size)
                        for k, n in enumerate([100, 1000, 5000])])
    cands = b["cands"].copy()
    m_of = b["qinfo"][cands["query"], 1].astype(np.int64)
    n_of = np.array([O.load().oracle_window_len(int(x), 0.1) for x in m_of], dtype=np.int64)
    rng = np.random.default_rng(93)
    picks = rng.permutation(len(cands))
    edges = []
    for x in (2**31 - 3, 2**31, 2**31 + 5, 2**32 - 7, 2**32, 2**32 + 3):
        edges += [x, x + 1]
    k = 0
    for x in edges:                               # around the 32-bit boundaries
        cands["position"][picks[k]] = x
        k += 1
    for off in (1, 2, 0, -1, -5):                  # size - pos vs n: live for pos <= size - n - 1
        i = picks[k]
        cands["position"][i] = size - n_of[i] - off
        k += 1
    for pos in (size - 1, size, size + 10):        # beyond the end: skipped
        cands["position"][picks[k]] = pos
        k += 1
    b = dict(b, cands=cands)
    assert int(cands["position"].max()) >= 2**32 and (cands["position"] > 2**31).sum() > 20
    with M.Myers(ref, size, avg_query=2000, cands_per_query=4, buffer_bytes=256 << 20) as m:
        for mode in (M.WINDOW_ALIGNED, M.WINDOW_EXACT):
            want = oracle_for(ref, size, b, 0.1, mode)
            assert (want[:, 1] == 0xFFFFFFFF).sum() >= 5
            for engine in (M.ENGINE_PLAIN, M.ENGINE_PERSISTENT, M.ENGINE_BINNED):
                announce(f"beyond 32 bits mode={mode} engine={engine}")
                got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | engine)
                np.testing.assert_array_equal(got, want)
