"""Pins the CPU oracle (oracle/myers_oracle.c) before it is trusted as the checker.

Each function of the oracle is checked against something it does not share
code with: golden vectors written by the reference's own compiled C programs
(tests/golden/*.npz), the debug/ known-answer trace, a brute-force DP, and
closed forms.
"""
import glob
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ENC = {c: i for i, c in enumerate("ACGT")}


def encode(s):
    return np.array([ENC.get(c, 4) for c in s], dtype=np.uint8)


def load_kat():
    with open(os.path.join(GOLDEN, "kat_candidate_15415.json")) as f:
        return json.load(f)


def golden_batch(path):
    """Inputs of a golden fixture in the library's layouts."""
    z = np.load(path)
    if "peq" in z.files:
        peq, qinfo = z["peq"], z["qinfo"]
    else:  # fixed-length fixture: rebuild 128-base entries from the read codes
        m = int(z["query_len"])
        codes = z["query_codes"]
        nq = len(codes) // m
        e = (m + 127) // 128
        peq = np.concatenate([O.build_peq(codes[i * m:(i + 1) * m]) for i in range(nq)])
        qinfo = np.array([[i * e, m] for i in range(nq)], dtype=np.uint32)
    cands = O.make_cands(z["cand_query"], z["cand_pos"])
    return z, peq, qinfo, cands


FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def test_fixtures_present():
    assert len(FIXTURES) >= 6


def test_kat_per_column_trace():
    """debug/SCORES-Myers-caos-code.txt: 140 per-column scores, min 19 at column 88."""
    kat = load_kat()
    scores, res = O.myers_trace(encode(kat["query"]), encode(kat["text"]))
    assert scores.tolist() == kat["scores"]
    assert res == (kat["final_score"], kat["final_column"])


def test_kat_dp_trace():
    kat = load_kat()
    scores, res = O.dp_trace(encode(kat["query"]), encode(kat["text"]))
    assert scores.tolist() == kat["scores"]
    assert res == (19, 88)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_matches_reference_programs(path):
    z, peq, qinfo, cands = golden_batch(path)
    got = O.myers_batch(z["ref_words"], z["ref_size"], peq, qinfo, cands, float(z["distance"]),
                        int(z["mode"]), int(z["max_query_len"]))
    keep = ~z["reads_past_end"]
    assert keep.sum() > 100
    np.testing.assert_array_equal(got[keep], z["results"][keep])
    # the fixtures exercise the skip rule
    assert (z["results"][:, 1] == 0xFFFFFFFF).sum() > 10


@pytest.mark.parametrize("name", ["api_aligned_d01.npz", "api_aligned_d02_mixed.npz"])
def test_packer_and_peq_match_reference_programs(name):
    """greference-4bits / gcandidates-reorder outputs vs the oracle's restatements."""
    z = np.load(os.path.join(GOLDEN, name))
    np.testing.assert_array_equal(O.pack_reference(z["ref_codes"]), z["ref_words"])
    codes, off = z["query_codes"], 0
    for pe, m in z["qinfo"]:
        p = O.build_peq(codes[off:off + m])
        np.testing.assert_array_equal(p, z["peq"][pe:pe + len(p)])
        off += m


def test_over_4096_skipped_like_the_api_twin():
    z = np.load(os.path.join(GOLDEN, "api_aligned_d01.npz"))
    long_q = np.flatnonzero(z["qinfo"][:, 1] > 4096)
    assert len(long_q) >= 2
    sel = np.isin(z["cand_query"], long_q)
    assert (z["results"][sel, 1] == 0xFFFFFFFF).all()


@pytest.mark.parametrize("m,d,n", [(100, 0.1, 120), (1000, 0.1, 1200), (333, 0.1, 399), (4096, 0.1, 4915),
                                   (20000, 0.1, 24000), (100, 0.2, 140), (1, 0.1, 1), (0, 0.1, 0),
                                   (7, 0.2, 9), (2500, 0.2, 3500)])
def test_window_len(m, d, n):
    assert O.window_len(m, d) == n
    assert O.window_len(m, d) == int(np.float32(np.float32(m) * (np.float32(1) + np.float32(2) * np.float32(d))))


def _window_text(ref_codes, pos, m, d, mode):
    n = O.window_len(m, d)
    if mode == O.MODE_ALIGNED:
        e = n // 8 + (2 if n % 8 else 1)
        start, cols = pos // 8 * 8, 8 * e
    else:
        start, cols = pos, n
    t = np.full(cols, 4, dtype=np.uint8)
    avail = max(0, min(cols, len(ref_codes) - start))
    t[:avail] = ref_codes[start:start + avail]
    return t


@pytest.mark.parametrize("mode", [O.MODE_ALIGNED, O.MODE_EXACT])
def test_oracle_equals_bruteforce_dp(mode):
    """Random tiny cases: block recurrence + window rules == plain DP on the window text."""
    rng = np.random.default_rng(7 + mode)
    size = 3000
    ref_codes = rng.integers(0, 4, size=size).astype(np.uint8)
    ref_codes[rng.integers(0, size, 60)] = 4
    words = O.pack_reference(ref_codes)
    lengths = [1, 2, 3, 5, 17, 31, 32, 33, 40, 63, 64, 65, 96, 97, 127, 128, 129, 150]
    peqs, qinfo, qcodes, off = [], [], [], 0
    for m in lengths:
        q = rng.integers(0, 5, size=m).astype(np.uint8)
        qcodes.append(q)
        p = O.build_peq(q)
        peqs.append(p)
        qinfo.append([off, m])
        off += len(p)
    peq = np.concatenate(peqs)
    qinfo = np.array(qinfo, dtype=np.uint32)
    cq, cp = [], []
    for qi, m in enumerate(lengths):
        n = O.window_len(m, 0.1)
        for _ in range(12):
            cq.append(qi)
            cp.append(int(rng.integers(0, size - n - 1)))
        cq += [qi, qi, qi]
        cp += [size - n - 1, size - n, size + 5]   # last, first skipped, out of range
    cands = O.make_cands(cq, cp)
    got = O.myers_batch(words, size, peq, qinfo, cands, 0.1, mode)
    for i, (qi, p) in enumerate(zip(cq, cp)):
        m = lengths[qi]
        n = O.window_len(m, 0.1)
        if not (p < size and size - p > n):
            assert tuple(got[i]) == (0, 0xFFFFFFFF)
            continue
        _, (score, col) = O.dp_trace(qcodes[qi], _window_text(ref_codes, p, m, 0.1, mode))
        assert (int(got[i, 1]), int(got[i, 0])) == (score, col), (m, p, mode)


def test_dp_closed_forms():
    q = np.array([0, 1, 2, 3, 0, 1], dtype=np.uint8)
    # the query itself inside the text: distance 0 at its last base
    t = np.concatenate([np.full(5, 4, np.uint8), q, np.full(4, 4, np.uint8)])
    s, res = O.dp_trace(q, t)
    assert res == (0, 10)
    # nothing matches: D[m][j] = m for every column, first minimum at column 0
    s, res = O.dp_trace(q, np.full(9, 4, np.uint8))
    assert s.tolist() == [6] * 9 and res == (6, 0)
    # one substitution
    t2 = t.copy()
    t2[7] = (t2[7] + 1) % 4
    assert O.dp_trace(q, t2)[1][0] == 1
    # text codes 5..7 behave as A (0); 4 (N) matches only a query N
    qa = np.array([0, 0, 0], np.uint8)
    assert O.dp_trace(qa, np.array([5, 6, 7], np.uint8))[1] == (0, 2)
    assert O.dp_trace(qa, np.array([4, 4, 4], np.uint8))[1] == (3, 0)
    assert O.dp_trace(np.array([4, 4], np.uint8), np.array([4, 4], np.uint8))[1] == (0, 1)


def test_myers_trace_equals_dp_random():
    rng = np.random.default_rng(3)
    for m in [1, 8, 31, 32, 33, 64, 65, 100, 129, 250]:
        q = rng.integers(0, 5, size=m).astype(np.uint8)
        t = rng.integers(0, 8, size=int(1.3 * m) + 5).astype(np.uint8)
        a = O.myers_trace(q, t)
        b = O.dp_trace(q, t)
        assert a[0].tolist() == b[0].tolist() and a[1] == b[1]


def test_threshold_matches_reference_checksum():
    """(uint32)((float)m * e) as src/checksum-2.c computes it, read back from
    the reference program's own under-threshold counts; the set holds pairs
    where float32, double and decimal products disagree."""
    kat = json.load(open(os.path.join(GOLDEN, "threshold_kat.json")))
    assert len(kat) >= 120
    for x in kat:
        assert O.threshold(x["m"], float(x["e"])) == x["threshold"], x
    # a double product of the float32 rate would give 6 here, the reference gives 7
    assert {(x["e"], x["m"]): x["threshold"] for x in kat}[("0.7", 10)] == 7


def test_apply_threshold_sentinels():
    qinfo = np.array([[0, 100], [1, 10]], dtype=np.uint32)
    cands = np.zeros(5, O.CAND_DTYPE)
    cands["query"] = [0, 0, 0, 1, 1]
    res = np.array([[5, 20], [6, 21], [0, 0xFFFFFFFF], [3, 7], [4, 8]], dtype=np.uint32)
    out = O.apply_threshold(qinfo, cands, res, 0.2)
    np.testing.assert_array_equal(out, [[5, 20], [0xFFFFFFFF, 0xFFFFFFFF], [0, 0xFFFFFFFF],
                                        [0xFFFFFFFF, 0xFFFFFFFF], [0xFFFFFFFF, 0xFFFFFFFF]])
    out = O.apply_threshold(qinfo, cands, res, 0.7)
    np.testing.assert_array_equal(out[3:], [[3, 7], [0xFFFFFFFF, 0xFFFFFFFF]])


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_bpm64_baseline_matches_reference_programs(path):
    """The 64-bit block BPM baseline (GEMTools formulation) gives the reference's results."""
    z, peq, qinfo, cands = golden_batch(path)
    got = O.bpm64_batch(z["ref_words"], z["ref_size"], peq, qinfo, cands, float(z["distance"]), int(z["mode"]))
    keep = ~z["reads_past_end"]
    limit = int(z["max_query_len"])
    if limit:                                      # the API twin skips queries over 4096 bases
        keep &= qinfo[cands["query"], 1] <= limit
    np.testing.assert_array_equal(got[keep], z["results"][keep])


@pytest.mark.parametrize("mode", [O.MODE_ALIGNED, O.MODE_EXACT])
def test_bpm64_equals_oracle_random(mode):
    """Two formulations of one DP (u32 blocks with int8 carry, u64 blocks with
    1-bit hin) agree on random batches with every length residue mod 64."""
    rng = np.random.default_rng(70 + mode)
    size = 200_000
    codes = rng.integers(0, 5, size=size).astype(np.uint8)
    words = O.pack_reference(codes)
    lengths = list(range(1, 200)) + [255, 256, 257, 511, 512, 513, 1000, 1023, 1024, 1025, 3000]
    peqs, qinfo, off = [], [], 0
    for m in lengths:
        start = int(rng.integers(0, size - m))
        q = codes[start:start + m].copy()
        q[rng.random(m) < 0.1] = rng.integers(0, 5)
        p = O.build_peq(q)
        peqs.append(p)
        qinfo.append([off, m])
        off += len(p)
    peq = np.concatenate(peqs)
    qinfo = np.array(qinfo, dtype=np.uint32)
    cq = np.repeat(np.arange(len(lengths)), 3)
    cp = rng.integers(0, size, size=len(cq))
    cands = O.make_cands(cq, cp)
    a = O.myers_batch(words, size, peq, qinfo, cands, 0.1, mode)
    b = O.bpm64_batch(words, size, peq, qinfo, cands, 0.1, mode)
    np.testing.assert_array_equal(a, b)
