"""Serving: the continuous-batching server and the resident-kernel server.
Every request's results are the same alone and in the mix, and equal to the
oracle; admission honours the HBM budget."""
import time

import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu

REF_SIZE = 3_000_000


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def carve_requests(b, sizes):
    """Split a synthetic batch into requests of `sizes` reads each (rebased arrays)."""
    reqs, q0 = [], 0
    per_read = np.bincount(b["cands"]["query"], minlength=len(b["qinfo"]))
    cand_start = np.concatenate([[0], np.cumsum(per_read)])
    for n in sizes:
        q1 = min(q0 + n, len(b["qinfo"]))
        qi = b["qinfo"][q0:q1].copy()
        e0 = qi[0, 0]
        e1 = qi[-1, 0] + (qi[-1, 1] + 127) // 128
        peq = b["peq"][e0:e1]
        qi[:, 0] -= e0
        c = b["cands"][cand_start[q0]:cand_start[q1]].copy()
        c["query"] -= q0
        reqs.append((peq, qi, c))
        q0 = q1
    return reqs


def make_pool(seed):
    ref = M.synth_reference(seed, REF_SIZE)
    short = M.synth_batch(M.make_spec(seed + 1, 1200, choices=[100, 200, 400, 600, 800, 1000],
                                      err_range=(0.01, 0.05), sub_frac=0.8, poisson=11, cap=64, distance=0.1),
                          ref, REF_SIZE)
    long = M.synth_batch(M.make_spec(seed + 2, 60, log_uniform=(2000, 20000), err_range=(0.12, 0.12),
                                     sub_frac=0.5, candidates=4, distance=0.1), ref, REF_SIZE)
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in np.exp(rng.uniform(0, np.log(200), size=24)).round()]
    reqs = carve_requests(short, sizes)
    reqs += carve_requests(long, [1, 3, 5, 2, 10])
    order = rng.permutation(len(reqs))
    return ref, [reqs[i] for i in order]


def oracle_req(ref, req):
    peq, qi, c = req
    return O.myers_batch(ref, REF_SIZE, peq, qi, c, 0.1, O.MODE_ALIGNED)


MODES = [0, M.SERVE_RESIDENT]
MODE_IDS = ["batches", "resident"]


@pytest.mark.parametrize("serve", MODES, ids=MODE_IDS)
def test_results_same_alone_and_in_the_mix(serve):
    announce(f"serve mix invariance flags={serve}")
    ref, reqs = make_pool(41)
    with M.Myers(ref, REF_SIZE, num_buffers=3, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        alone = [m.verify(*r, 0.1) for r in reqs]
        with M.Server(m, distance=0.1, flags=serve) as sv:
            made = [M.Server.request(*r) for r in reqs]
            tickets = [sv.submit(x[0]) for x in made]
            for t in tickets:
                assert sv.wait(t) >= 0.0
            st = sv.stats()
            assert st.requests == len(reqs)
            if not serve:
                assert st.maxBatchRequests >= 2      # requests really shared batches
            else:
                assert st.batches >= 1               # resident kernel launches
        for (req, keep), a, r in zip(made, alone, reqs):
            np.testing.assert_array_equal(keep[3], a)
            np.testing.assert_array_equal(a, oracle_req(ref, r))


@pytest.mark.parametrize("mode", [M.WINDOW_ALIGNED, M.WINDOW_EXACT])
def test_resident_window_modes_and_threshold(mode):
    """The resident kernel in both window modes, and in threshold mode (the
    cut-off kernels), against the oracle; the mix is submitted twice, the
    second time after the kernel has left the device when idle.  The arena
    (1 MB) wraps several times, so host memory the kernel read for earlier
    requests is rewritten under it and read again."""
    announce(f"serve resident mode={mode}")
    ref, reqs = make_pool(45)
    with M.Myers(ref, REF_SIZE, num_buffers=1, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        for e in (-1.0, 0.1):
            m.set_threshold(e)
            with M.Server(m, distance=0.1, flags=mode | M.SERVE_RESIDENT, arena_bytes=1 << 20) as sv:
                for rnd in range(2):
                    made = [M.Server.request(*r) for r in reqs]
                    tickets = [sv.submit(x[0]) for x in made]
                    for t in tickets:
                        sv.wait(t)
                    for (req, keep), r in zip(made, reqs):
                        peq, qi, c = r
                        full = O.myers_batch(ref, REF_SIZE, peq, qi, c, 0.1, mode)
                        want = full if e < 0 else O.apply_threshold(qi, c, full, e)
                        np.testing.assert_array_equal(keep[3], want, err_msg=f"round {rnd} e={e}")
                    if rnd == 0:
                        time.sleep(0.2)              # the kernel leaves after ~50 ms without work
                launches = sv.stats().batches
            assert launches >= 2
        m.set_threshold(-1.0)


@pytest.mark.parametrize("serve", MODES, ids=MODE_IDS)
def test_open_loop_poisson(serve):
    announce(f"serve open loop flags={serve}")
    ref, reqs = make_pool(42)
    rng = np.random.default_rng(5)
    with M.Myers(ref, REF_SIZE, num_buffers=3, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        with M.Server(m, distance=0.1, flags=serve) as sv:
            made = [M.Server.request(*r) for r in reqs * 3]
            gaps = rng.exponential(1e6 / 2000.0, size=len(made))          # 2000 requests/s
            arrivals = (np.cumsum(gaps) * 1e3).astype(np.uint64)
            lat, wall, queue = sv.run_open_loop([x[0] for x in made], arrivals, queue=True)
        assert np.all(lat >= 0) and wall > 0
        assert np.all(queue >= 0) and np.all(queue <= lat)      # admitted before completion
        for (req, keep), r in zip(made, reqs * 3):
            np.testing.assert_array_equal(keep[3], oracle_req(ref, r))


@pytest.mark.parametrize("serve", MODES, ids=MODE_IDS)
def test_budget_limits_admission(serve):
    announce(f"serve budget flags={serve}")
    ref, reqs = make_pool(43)
    with M.Myers(ref, REF_SIZE, num_buffers=2, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        with M.Server(m, distance=0.1, hbm_budget=1, flags=serve) as sv:     # one request in flight at a time
            made = [M.Server.request(*r) for r in reqs[:10]]
            tickets = [sv.submit(x[0]) for x in made]
            for t in tickets:
                sv.wait(t)
            if not serve:
                assert sv.stats().maxBatchRequests == 1
        for (req, keep), r in zip(made, reqs[:10]):
            np.testing.assert_array_equal(keep[3], oracle_req(ref, r))


def test_oversized_request_rejected():
    announce("serve oversized")
    ref, reqs = make_pool(44)
    with M.Myers(ref, REF_SIZE, num_buffers=1, avg_query=100, cands_per_query=1,
                 buffer_bytes=1 << 16) as m:
        cap = m.capacity()[2]
        big = max(reqs, key=lambda r: len(r[2]))
        assert len(big[2]) > cap
        with M.Server(m, distance=0.1) as sv:
            req, keep = M.Server.request(*big)
            t = sv.submit(req)
            with pytest.raises(M.MyersError) as e:
                sv.wait(t)
            assert e.value.code == 17


def test_resident_oversized_request_rejected():
    """A request larger than the resident server's arena fails with
    MYERS_E_CAPACITY; the requests after it are served."""
    announce("serve resident oversized")
    ref, reqs = make_pool(46)
    with M.Myers(ref, REF_SIZE, num_buffers=1, avg_query=800, cands_per_query=11, buffer_bytes=64 << 20) as m:
        big = max(reqs, key=lambda r: len(r[2]))
        small = min(reqs, key=lambda r: len(r[2]))
        arena = 64 * 1024
        need = sum(x.nbytes for x in big) + len(big[2]) * 12
        assert need > arena
        with M.Server(m, distance=0.1, flags=M.SERVE_RESIDENT, arena_bytes=arena) as sv:
            req, keep = M.Server.request(*big)
            t = sv.submit(req)
            req2, keep2 = M.Server.request(*small)
            t2 = sv.submit(req2)
            with pytest.raises(M.MyersError) as e:
                sv.wait(t)
            assert e.value.code == 17
            sv.wait(t2)
        np.testing.assert_array_equal(keep2[3], oracle_req(ref, small))


def test_resident_variant_boundaries():
    """Reads at the capacities of the long-read variants the resident kernel
    switches between (448, 512, 640 and 768 words), the longest supported
    read, and short reads, in one request and in both window modes, against
    the oracle."""
    announce("serve resident boundaries")
    ref = M.synth_reference(47, REF_SIZE)
    lengths = [14336, 14337, 16384, 16385, 20480, 20481, 24575, 24576, 1000, 100]
    parts = [M.synth_batch(M.make_spec(48 + k, 1, length=n, sub=0.04, ins=0.01, dele=0.01, candidates=3,
                                       distance=0.1), ref, REF_SIZE) for k, n in enumerate(lengths)]
    peq, qinfo, cands, e_off = [], [], [], 0
    for k, b in enumerate(parts):
        qi = b["qinfo"].copy()
        qi[:, 0] += np.uint32(e_off)
        c = b["cands"].copy()
        c["query"] += np.uint32(k)
        peq.append(b["peq"])
        qinfo.append(qi)
        cands.append(c)
        e_off += len(b["peq"])
    req = (np.concatenate(peq), np.concatenate(qinfo), np.concatenate(cands))
    assert sorted(req[1][:, 1].tolist()) == sorted(lengths)
    with M.Myers(ref, REF_SIZE, num_buffers=1, avg_query=8000, cands_per_query=3, buffer_bytes=256 << 20) as m:
        for mode in (M.WINDOW_ALIGNED, M.WINDOW_EXACT):
            with M.Server(m, distance=0.1, flags=mode | M.SERVE_RESIDENT) as sv:
                r, keep = M.Server.request(*req)
                sv.wait(sv.submit(r))
            want = O.myers_batch(ref, REF_SIZE, req[0], req[1], req[2], 0.1, mode)
            np.testing.assert_array_equal(keep[3], want)


@pytest.mark.parametrize("serve", MODES, ids=MODE_IDS)
def test_server_releases_requests(serve):
    """A wait releases its ticket: after 1,600 small requests (submitted and
    waited in rounds of 16) the server holds none, and waiting again on a
    released ticket is an error.  Results stay equal to the oracle."""
    announce(f"serve release flags={serve}")
    ref, reqs = make_pool(48)
    small = sorted(reqs, key=lambda r: len(r[2]))[:8]
    want = [oracle_req(ref, r) for r in small]
    with M.Myers(ref, REF_SIZE, num_buffers=2, avg_query=800, cands_per_query=11, buffer_bytes=64 << 20) as m:
        with M.Server(m, distance=0.1, flags=serve) as sv:
            peak = 0
            for rnd in range(100):
                made = [M.Server.request(*small[i % len(small)]) for i in range(16)]
                tickets = [sv.submit(x[0]) for x in made]
                peak = max(peak, sv.stats().liveRequests)
                for t in tickets:
                    sv.wait(t)
                if rnd % 25 == 0:
                    for i, (req, keep) in enumerate(made):
                        np.testing.assert_array_equal(keep[3], want[i % len(small)])
            st = sv.stats()
            assert st.requests == 1600
            assert st.liveRequests == 0
            assert peak <= 16
            with pytest.raises(M.MyersError):
                sv.wait(tickets[-1])


def test_backlog_drain_batches():
    """A backlog of more than three capped batches is drained in batches of
    up to four times the cap (which then take the throughput path for their
    long reads); every request's results stay those it gets alone."""
    announce("serve backlog drain")
    ref, reqs = make_pool(47)
    cells = sum(int((r[1][r[2]["query"], 1].astype(np.int64) *
                     np.floor(r[1][r[2]["query"], 1].astype(np.float32) * np.float32(1.2))).sum())
                for r in reqs)
    cap = max(1, cells // 40)
    with M.Myers(ref, REF_SIZE, num_buffers=3, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        alone = [m.verify(*r, 0.1) for r in reqs]
        with M.Server(m, distance=0.1, max_batch_cells=cap) as sv:
            made = [M.Server.request(*r) for r in reqs]
            tickets = [sv.submit(x[0]) for x in made]
            for t in tickets:
                assert sv.wait(t) >= 0.0
            st = sv.stats()
            assert st.requests == len(reqs)
            assert st.drainBatches >= 1, (st.batches, cap, cells)
        for (req, keep), a, r in zip(made, alone, reqs):
            np.testing.assert_array_equal(keep[3], a)
            np.testing.assert_array_equal(a, oracle_req(ref, r))
