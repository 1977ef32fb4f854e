"""Threshold mode (myersSetThreshold): bit-exact results for candidates whose
score is <= (uint32)((float)m * e), {0xFFFFFFFF, 0xFFFFFFFF} for the other
verified candidates, the skip sentinel unchanged.  The oracle computes every
candidate in full and applies the threshold afterwards; the kernels for
queries of up to 1024 bases skip rows above the threshold (Ukkonen cut-off),
so equality here checks that the cut-off loses nothing under the threshold.
"""
import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu

RATES = (0.0, 0.05, 0.1, 0.2, 0.35, 1.0)


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def expected(ref, size, b, distance, mode, e):
    full = O.myers_batch(ref, size, b["peq"], b["qinfo"], b["cands"], distance, mode)
    return O.apply_threshold(b["qinfo"], b["cands"], full, e)


def test_threshold_every_single_lane_variant():
    """Every one-lane variant (cut-off kernels for W >= 8, among them the odd
    W whose top group holds fewer than four words; post-filtered full kernels
    for W < 8), lengths across the variant's top four words in one wave,
    error rates around the thresholds."""
    size = 2_000_000
    ref = M.synth_reference(61, size)
    with M.Myers(ref, size, avg_query=600, cands_per_query=6, buffer_bytes=256 << 20) as m:
        variants = m.variants()
        default = m.bin_table()
        for v, info in enumerate(variants):
            if info["lanes"] != 1:
                continue
            cap = info["maxQuerySize"]
            lo = max(1, cap - 32 * 4 + 1)
            spec = M.make_spec(300 + v, 40, choices=sorted({lo, (lo + cap) // 2, cap - 3, cap}),
                               err_range=(0.0, 0.4), sub_frac=0.6, candidates=6, distance=0.1)
            b = M.synth_batch(spec, ref, size)
            table = default.copy()
            table[:] = v
            table[(np.arange(len(table)) * 32) > cap] = default[(np.arange(len(table)) * 32) > cap]
            m.set_bin_table(table)
            for e in RATES:
                announce(f"threshold variant W={info['words']} e={e}")
                m.set_threshold(e)
                got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
                np.testing.assert_array_equal(got, expected(ref, size, b, 0.1, 0, e),
                                              err_msg=f"variant {info} e={e}")
        m.set_threshold(-1.0)
        m.set_bin_table(default)


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_threshold_1000_base_reads(mode):
    """1000-base reads, 11 candidates each (one true locus), e = 0.2, both engines."""
    size = 3_000_000
    ref = M.synth_reference(62, size)
    b = M.synth_batch(M.make_spec(63, 3000, length=1000, sub=0.03, ins=0.005, dele=0.005, candidates=11,
                                  distance=0.1), ref, size)
    want = expected(ref, size, b, 0.1, mode, 0.2)
    under = np.mean(want[:, 1].reshape(-1, 11)[:, 0] != 0xFFFFFFFF)
    assert under > 0.99                       # the true loci stay under the threshold
    with M.Myers(ref, size, avg_query=1000, cands_per_query=11, buffer_bytes=256 << 20) as m:
        m.set_threshold(0.2)
        assert abs(m.threshold() - 0.2) < 1e-7
        for engine in (M.ENGINE_PLAIN, M.ENGINE_PERSISTENT):
            announce(f"threshold 1000-base reads mode={mode} engine={engine}")
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | engine)
            np.testing.assert_array_equal(got, want)
        m.set_threshold(-1.0)
        got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode)
        np.testing.assert_array_equal(got, O.myers_batch(ref, size, b["peq"], b["qinfo"], b["cands"], 0.1, mode))


def test_threshold_mixed_lengths_and_long_reads():
    """Mixed short reads and long reads (multi-lane kernels) in one batch,
    with skip-rule candidates at the reference end."""
    size = 3_000_000
    ref = M.synth_reference(64, size)
    b = M.synth_batch(M.make_spec(65, 900, choices=[50, 100, 200, 333, 600, 1000, 1500, 3000],
                                  err_range=(0.01, 0.3), sub_frac=0.7, poisson=9, cap=40, distance=0.1),
                      ref, size)
    cands = b["cands"].copy()
    cands["position"][::97] = size - 5                    # skipped: window past the end
    b = dict(b, cands=cands)
    with M.Myers(ref, size, avg_query=800, cands_per_query=9, buffer_bytes=256 << 20) as m:
        for e in (0.02, 0.1, 0.25):
            announce(f"threshold mixed e={e}")
            m.set_threshold(e)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1)
            want = expected(ref, size, b, 0.1, 0, e)
            np.testing.assert_array_equal(got, want)
            assert np.all(got[::97] == [0, 0xFFFFFFFF])


def test_threshold_graph_recaptures_on_change():
    """A captured pipeline is keyed by the threshold: changing it captures anew."""
    announce("threshold graph")
    size = 2_000_000
    ref = M.synth_reference(66, size)
    b = M.synth_batch(M.make_spec(67, 600, choices=[100, 400, 1000], err_range=(0.01, 0.2), poisson=11,
                                  distance=0.1), ref, size)
    with M.Myers(ref, size, avg_query=500, cands_per_query=11, buffer_bytes=256 << 20) as m:
        dp, dq, dc = M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"])
        dr = M.DeviceArray.empty((len(b["cands"]), 2), np.uint32)
        nq, nc, ne = len(b["qinfo"]), len(b["cands"]), len(b["peq"])
        for i, e in enumerate((0.1, 0.1, 0.05, -1.0)):
            m.set_threshold(e)
            M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=M.USE_GRAPH, h_qinfo=b["qinfo"])
            M.synchronize(m)
            full = O.myers_batch(ref, size, b["peq"], b["qinfo"], b["cands"], 0.1, 0)
            want = full if e < 0 else O.apply_threshold(b["qinfo"], b["cands"], full, e)
            np.testing.assert_array_equal(dr.get(), want, err_msg=f"call {i} e={e}")
        assert m.stats().graphCaptures == 3
        for x in (dp, dq, dc, dr):
            x.free()


def test_threshold_long_read_launch():
    """Threshold mode through the long-read launch: reads over 8 kb run on whole
    workgroups, 3-5 kb reads on one wave, next to short-read bins."""
    from test_gpu_parity import concat_batches
    size = 3_000_000
    ref = M.synth_reference(67, size)
    b = concat_batches([M.synth_batch(M.make_spec(68 + k, n, length=length, sub=0.04, ins=0.02, dele=0.02,
                                                  candidates=c, distance=0.1), ref, size)
                        for k, (n, length, c) in enumerate([(150, 200, 6), (5, 3500, 3), (4, 9000, 3),
                                                            (3, 17000, 2)])])
    with M.Myers(ref, size, avg_query=600, cands_per_query=5, buffer_bytes=256 << 20) as m:
        for e in (0.05, 0.15, 0.3):
            announce(f"threshold long-read launch e={e}")
            m.set_threshold(e)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=M.ENGINE_PERSISTENT)
            np.testing.assert_array_equal(got, expected(ref, size, b, 0.1, 0, e))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_threshold_long_reads_early_stop(mode):
    """Long noisy reads (1 true and 3 random candidates each) on the per-bin
    multi-lane kernels in threshold mode, where a wave stops once its
    candidates cannot get under the threshold; rates on both sides of the
    true loci's error."""
    size = 3_000_000
    ref = M.synth_reference(72, size)
    b = M.synth_batch(M.make_spec(73, 24, log_uniform=(1500, 20000), err_range=(0.12, 0.12), sub_frac=0.5,
                                  candidates=4, distance=0.1), ref, size)
    with M.Myers(ref, size, avg_query=8000, cands_per_query=4, buffer_bytes=256 << 20) as m:
        for e in (0.05, 0.15, 0.25):
            announce(f"threshold long reads early stop mode={mode} e={e}")
            m.set_threshold(e)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | M.ENGINE_PLAIN)
            np.testing.assert_array_equal(got, expected(ref, size, b, 0.1, mode, e))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_threshold_queue_launches(mode):
    """Threshold mode through the queue launches of both classes (the cut-off
    kernels of the one-lane class, the early stop of the multi-lane class)."""
    size = 3_000_000
    ref = M.synth_reference(66, size)
    b = M.synth_batch(M.make_spec(67, 400, choices=[100, 333, 600, 1000, 1500, 3000, 6000, 12000],
                                  err_range=(0.01, 0.3), sub_frac=0.7, poisson=5, cap=20, distance=0.1),
                      ref, size)
    with M.Myers(ref, size, avg_query=3000, cands_per_query=5, buffer_bytes=256 << 20) as m:
        for e in (0.05, 0.2):
            announce(f"threshold queue launches mode={mode} e={e}")
            m.set_threshold(e)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1,
                           flags=mode | M.ENGINE_PERSISTENT | M.NO_LONG_LAUNCH)
            assert m.stats().launches == 2
            np.testing.assert_array_equal(got, expected(ref, size, b, 0.1, mode, e))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_threshold_probe_order(mode):
    """The band probe and the reordering of probed bins only change the
    order of the work.  Batches whose 128-slot spans hold no class-0
    candidate (random windows only), a few (one true locus per read,
    Poisson candidate counts, so spans cut reads apart), or more than 64
    (every candidate at its true locus), lengths over one-lane cut-off
    variants with even and odd W, skipped candidates, every engine: all
    equal the oracle's filtered results."""
    size = 3_000_000
    ref = M.synth_reference(66, size)
    b = M.synth_batch(M.make_spec(67, 2400, choices=[300, 520, 777, 900, 1000], err_range=(0.01, 0.08),
                                  sub_frac=0.7, poisson=11, cap=40, distance=0.1), ref, size)
    rnd = b["cands"].copy()
    first = np.r_[0, np.cumsum(np.bincount(rnd["query"], minlength=len(b["qinfo"])))[:-1]]
    rng = np.random.default_rng(68)
    rnd["position"][first] = rng.integers(0, size - 1300, len(first))     # no true locus at all
    true_all = b["cands"].copy()
    true_all["position"] = true_all["position"][first[true_all["query"]]]  # every candidate at the locus
    mixed = b["cands"].copy()
    mixed["position"][::53] = size - 7                                      # skipped: past the end
    with M.Myers(ref, size, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        for name, cands in (("random", rnd), ("true-all", true_all), ("mixed", mixed)):
            bb = dict(b, cands=cands)
            want = expected(ref, size, bb, 0.1, mode, 0.2)
            m.set_threshold(0.2)
            for engine in (M.ENGINE_PLAIN, M.ENGINE_PERSISTENT, M.ENGINE_BINNED):
                announce(f"threshold probe order {name} mode={mode} engine={engine}")
                got = m.verify(bb["peq"], bb["qinfo"], bb["cands"], 0.1, flags=mode | engine)
                np.testing.assert_array_equal(got, want, err_msg=f"{name} engine={engine}")
        m.set_threshold(-1.0)
