"""Data preparation (libmyers_tools.so): packing and Peq layouts against the
reference programs' outputs and the oracle, and the synthetic generator's
determinism and statistics."""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pack_fasta_matches_greference():
    """FASTA text -> words equals src/genReference-4bits.c's output."""
    z = np.load(os.path.join(GOLDEN, "api_aligned_d01.npz"))
    chars = z["ref_chars"].tobytes()
    fasta = b">chrSynthetic test\n" + b"\n".join(chars[i:i + 60] for i in range(0, len(chars), 60)) + b"\n"
    out = np.zeros(len(z["ref_words"]) + 1, dtype=np.uint32)
    n = M.tools().myersPackFasta(fasta, len(fasta), M._ptr(out), len(out))
    assert n == int(z["ref_size"])
    np.testing.assert_array_equal(out[:len(z["ref_words"])], z["ref_words"])


def test_pack_codes_matches_oracle_and_golden():
    z = np.load(os.path.join(GOLDEN, "api_aligned_d02_mixed.npz"))
    np.testing.assert_array_equal(M.pack_codes(z["ref_codes"]), z["ref_words"])
    rng = np.random.default_rng(0)
    for n in (1, 7, 8, 9, 1001):
        c = rng.integers(0, 9, size=n).astype(np.uint8)
        np.testing.assert_array_equal(M.pack_codes(c), O.pack_reference(c))


def test_build_peq_matches_golden():
    z = np.load(os.path.join(GOLDEN, "api_aligned_d01.npz"))
    codes, off = z["query_codes"], 0
    for pe, m in z["qinfo"]:
        p = M.build_peq(codes[off:off + m])
        np.testing.assert_array_equal(p, z["peq"][pe:pe + len(p)])
        off += m


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_synth_reference_statistics():
    ref = M.synth_reference(42, 4_000_000)
    codes = ((ref[:, None] >> (4 * np.arange(8, dtype=np.uint32))) & 15).astype(np.uint8).ravel()
    counts = np.bincount(codes, minlength=5)
    assert counts[5:].sum() == 0 if len(counts) > 5 else True
    frac = counts / counts.sum()
    assert 0.0005 < frac[4] < 0.002            # ~0.1% N
    assert np.all(np.abs(frac[:4] - 0.25) < 0.002)
    assert _digest(ref) == _digest(M.synth_reference(42, 4_000_000))
    assert _digest(ref) != _digest(M.synth_reference(43, 4_000_000))


def test_synth_independent_of_thread_count():
    code = ("import sys; sys.path.insert(0, %r); import myers_amd as M, hashlib; "
            "r = M.synth_reference(7, 3_000_001); "
            "b = M.synth_batch(M.make_spec(8, 2000, choices=[100, 1000], err_range=(0.01, 0.05), poisson=11, "
            "distance=0.1), r, 3_000_001); "
            "print(hashlib.sha256(r.tobytes() + b['peq'].tobytes() + b['cands'].tobytes()).hexdigest())"
            % os.path.dirname(M.__file__))
    outs = []
    for t in ("1", "5"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, check=True, capture_output=True,
                                   text=True).stdout.strip())
    assert outs[0] == outs[1]


def test_synth_batch_layout():
    size = 2_000_000
    ref = M.synth_reference(1, size)
    spec = M.make_spec(2, 500, choices=[100, 200, 400, 600, 800, 1000], err_range=(0.01, 0.05), sub_frac=0.8,
                       poisson=11, cap=64, distance=0.1, keep_codes=True)
    b = M.synth_batch(spec, ref, size)
    q = b["qinfo"]
    entries = (q[:, 1] + 127) // 128
    np.testing.assert_array_equal(q[:, 0], np.concatenate([[0], np.cumsum(entries)[:-1]]))
    assert set(np.unique(q[:, 1])) <= {100, 200, 400, 600, 800, 1000}
    c = b["cands"]
    assert np.all(np.diff(c["query"].astype(np.int64)) >= 0)
    per_q = np.bincount(c["query"], minlength=len(q))
    assert per_q.min() >= 1 and per_q.max() <= 64 and 9 < per_q.mean() < 13
    first = np.concatenate([[0], np.cumsum(per_q)[:-1]])
    margin = q[:, 1] // 10
    np.testing.assert_array_equal(c["position"][first], np.maximum(b["true_start"].astype(np.int64) - margin, 0))
    n = np.array([O.window_len(int(m), 0.1) for m in q[c["query"], 1]], dtype=np.uint64)
    assert np.all(c["position"] + n < size)
    assert b["cells"] == int((q[c["query"], 1].astype(np.uint64) * n).sum())
    off = b["code_offset"]
    for i in range(0, 500, 71):
        np.testing.assert_array_equal(M.build_peq(b["codes"][off[i]:off[i + 1]]), b["peq"][q[i, 0]:q[i, 0] + entries[i]])


def test_synth_true_locus_scores_low():
    """Reads sampled with ~2.5% error score far below random loci (oracle)."""
    size = 4_000_000
    ref = M.synth_reference(42, size)
    b = M.synth_batch(M.make_spec(43, 2000, length=100, sub=0.02, ins=0.0025, dele=0.0025, candidates=11,
                                  distance=0.1), ref, size)
    r = O.myers_batch(ref, size, b["peq"], b["qinfo"], b["cands"], 0.1, 0)
    sc = r[:, 1].reshape(-1, 11).astype(np.int64)
    assert np.mean(sc[:, 0] <= 20) > 0.99
    assert np.mean(sc[:, 1:] <= 20) < 0.01


def test_reference_file_roundtrip():
    ref = M.synth_reference(5, 1001)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "r.4bits.ref")
        assert M.tools().myersWriteReferenceFile(p.encode(), M._ptr(ref), len(ref), 1001) == 0
        raw = np.fromfile(p, dtype=np.uint32)
    assert raw[0] == len(ref) and raw[1] == 1001
    np.testing.assert_array_equal(raw[2:], ref)
