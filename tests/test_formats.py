"""On-disk formats of the tools library: .prof -> Peq converter, .gem.qry and
.4bits.ref readers/writers.

The converter is pinned by the golden fixtures written with the reference's own
gcandidates-reorder program (src-api/genCandidate-reorder.c): the fixture holds
the reads and candidates that went into its .prof file and the Peq entries and
qinfo records it wrote; the .prof text is rebuilt here by the same rules the
fixture script used (lowercase first half of every read with id % 4 == 1, GEM
score id % 7).
"""
import os
import struct
import tempfile

import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CODE2CHAR = np.frombuffer(b"ACGTN", dtype=np.uint8)


def prof_text_from_fixture(z):
    lengths = z["qinfo"][:, 1].astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    lines = []
    for qid, m in enumerate(lengths):
        chars = CODE2CHAR[z["query_codes"][offs[qid]:offs[qid + 1]]].copy()
        if qid % 4 == 1:
            chars[: m // 2] += 32
        pos = z["cand_pos"][z["cand_query"] == qid]
        lines.append(chars.tobytes().decode() + "".join(f"\tchr1:+:{int(p)}:{qid % 7}" for p in pos) + "\n")
    return "".join(lines)


@pytest.mark.parametrize("name", ["api_aligned_d01", "api_aligned_d02_mixed"])
def test_parse_prof_matches_reference_converter(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "reads.prof")
        with open(path, "w") as f:
            f.write(prof_text_from_fixture(z))
        g = M.parse_prof(path)
    np.testing.assert_array_equal(g["peq"], z["peq"])
    np.testing.assert_array_equal(g["qinfo"], z["qinfo"])
    np.testing.assert_array_equal(g["cands"]["query"], z["cand_query"])
    np.testing.assert_array_equal(g["cands"]["position"], z["cand_pos"])
    np.testing.assert_array_equal(g["gem_scores"], z["cand_query"] % 7)
    assert len(g["bases"]) == int(z["qinfo"][:, 1].sum())


def test_parse_prof_whitespace_and_empty_lines():
    text = "ACGTN\tx:+:10:3 y:-:20:4\n\nacgt\n" + "G" * 300 + "\tz:+:4294967295:1\r\n"
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "r.prof")
        with open(path, "w") as f:
            f.write(text)
        g = M.parse_prof(path)
    np.testing.assert_array_equal(g["qinfo"], [[0, 5], [1, 4], [2, 300]])
    np.testing.assert_array_equal(g["cands"]["query"], [0, 0, 2])
    np.testing.assert_array_equal(g["cands"]["position"], [10, 20, 4294967295])
    np.testing.assert_array_equal(g["gem_scores"], [3, 4, 1])
    np.testing.assert_array_equal(g["peq"][0], M.build_peq(np.array([0, 1, 2, 3, 4], np.uint8))[0])
    np.testing.assert_array_equal(g["peq"][1], M.build_peq(np.array([0, 1, 2, 3], np.uint8))[0])
    np.testing.assert_array_equal(g["peq"][2:5], M.build_peq(np.full(300, 2, np.uint8)))
    assert g["bases"] == b"ACGTNacgt" + b"G" * 300


def test_parse_prof_rejects_bad_candidate():
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "bad.prof")
        with open(path, "w") as f:
            f.write("ACGT\tchr1:+:12\n")
        with pytest.raises(ValueError):
            M.parse_prof(path)
        with pytest.raises(ValueError):
            M.parse_prof(os.path.join(d, "missing.prof"))


def test_gem_qry_layout_and_roundtrip():
    z = np.load(os.path.join(GOLDEN, "api_aligned_d01.npz"))
    nq, nc, ne = len(z["qinfo"]), len(z["cand_pos"]), len(z["peq"])
    cands = np.zeros(nc, M.CAND_DTYPE)
    cands["query"] = z["cand_query"]
    cands["position"] = z["cand_pos"]
    d = {"peq": z["peq"], "qinfo": z["qinfo"], "cands": cands,
         "gem_scores": np.arange(nc, dtype=np.uint32), "bases": bytes(CODE2CHAR[z["query_codes"]])}
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "x.gem.qry")
        M.write_gem_qry(path, d)
        raw = open(path, "rb").read()
        back = M.read_gem_qry(path)
        # a file that stops after qinfo (what src-api/test.c reads) is accepted too
        cut = os.path.join(t, "cut.gem.qry")
        with open(cut, "wb") as f:
            f.write(raw[: 20 + 80 * ne + 16 * nc + 8 * nq])
        short = M.read_gem_qry(cut)
        with open(cut, "wb") as f:
            f.write(raw[: 20 + 80 * ne + 16 * nc + 8 * nq - 1])
        with pytest.raises(ValueError):
            M.read_gem_qry(cut)
    total = int(z["qinfo"][:, 1].sum())
    assert struct.unpack("<5I", raw[:20]) == (total, ne, 0, nq, nc)
    assert len(raw) == 20 + 80 * ne + 16 * nc + 8 * nq + 4 * nc + total
    off = 20 + 80 * ne
    c16 = np.frombuffer(raw[off:off + 16 * nc], dtype=M.CAND_DTYPE)
    np.testing.assert_array_equal(c16["position"], z["cand_pos"])
    for k in ("peq", "qinfo", "gem_scores"):
        np.testing.assert_array_equal(back[k], d[k])
    np.testing.assert_array_equal(back["cands"]["position"], cands["position"])
    assert back["bases"] == d["bases"]
    np.testing.assert_array_equal(short["peq"], d["peq"])
    np.testing.assert_array_equal(short["qinfo"], d["qinfo"])
    assert not short["gem_scores"].any()


def test_reference_file_read_write():
    words = M.synth_reference(11, 12345)
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "r.4bits.ref")
        M.write_reference_file(path, words, 12345)
        back, size = M.read_reference_file(path)
        with pytest.raises(OSError):
            M.read_reference_file(os.path.join(t, "none.ref"))
    assert size == 12345
    np.testing.assert_array_equal(back, words)


# ------------------------------------------------------------ programs

BIN = os.path.join(M.PKG_DIR, "bin")


def write_fixture_inputs(z, d):
    """The fixture's reference as FASTA and its reads as .prof, in directory d."""
    fasta = os.path.join(d, "ref.fa")
    with open(fasta, "wb") as f:
        f.write(b">chrSynthetic test\n")
        chars = z["ref_chars"]
        for i in range(0, len(chars), 60):
            f.write(chars[i:i + 60].tobytes() + b"\n")
    prof = os.path.join(d, "reads.prof")
    with open(prof, "w") as f:
        f.write(prof_text_from_fixture(z))
    return fasta, prof


def run(args, cwd):
    import subprocess
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def test_prep_program_writes_reference_formats():
    z = np.load(os.path.join(GOLDEN, "api_aligned_d02_mixed.npz"))
    size = int(z["ref_size"])
    with tempfile.TemporaryDirectory() as d:
        fasta, prof = write_fixture_inputs(z, d)
        run([os.path.join(BIN, "myers-prep"), "fasta2ref", fasta], d)
        words, got_size = M.read_reference_file(f"{fasta}.{size}.4bits.ref")
        run([os.path.join(BIN, "myers-prep"), "prof2qry", prof, os.path.join(d, "x.gem.qry")], d)
        g = M.read_gem_qry(os.path.join(d, "x.gem.qry"))
    assert got_size == size
    np.testing.assert_array_equal(words, z["ref_words"])
    np.testing.assert_array_equal(g["peq"], z["peq"])
    np.testing.assert_array_equal(g["qinfo"], z["qinfo"])
    np.testing.assert_array_equal(g["cands"]["position"], z["cand_pos"])


def read_results_text(path):
    with open(path) as f:
        lines = f.read().split("\n")
    count = int(lines[0])
    res = np.zeros((count, 2), dtype=np.uint32)          # [column, score]
    for line in lines[1:1 + count]:
        idx, score, col = (int(x) for x in line.split())
        res[idx] = (col, score)
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["api_aligned_d01", "api_aligned_d02_mixed"])
def test_run_program_matches_golden(name):
    """The test.c-equivalent harness, end to end from the reference's file
    formats, against the results of the reference's own CPU program."""
    print(f"[gpu-case] myers-run {name}", flush=True)
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    size = int(z["ref_size"])
    keep = ~z["reads_past_end"].astype(bool)
    with tempfile.TemporaryDirectory() as d:
        fasta, prof = write_fixture_inputs(z, d)
        run([os.path.join(BIN, "myers-prep"), "fasta2ref", fasta], d)
        ref = f"{fasta}.{size}.4bits.ref"
        run([os.path.join(BIN, "myers-prep"), "prof2qry", prof, os.path.join(d, "a.gem.qry")], d)
        dist = repr(float(z["distance"]))
        # two buffers (one .gem.qry, one .prof), batches of at most 50 candidates
        out = run([os.path.join(BIN, "myers-run"), "-t", "0.2", "-n", "50", dist, ref,
                   os.path.join(d, "a.gem.qry"), prof], d)
        res = read_results_text(os.path.join(d, "a.gem.qry.gem-int.pad.gpu"))
    nc = len(z["cand_pos"])
    assert len(res) == 2 * nc
    # the reference's API twin skips reads over 4096 bases; the library verifies
    # them (an extension), so those are checked against the oracle instead
    m = z["qinfo"][z["cand_query"], 1]
    limit = int(z["max_query_len"])
    long_reads = (m > limit) if limit else np.zeros(nc, bool)
    cands = O.make_cands(z["cand_query"], z["cand_pos"])
    full = O.myers_batch(z["ref_words"], size, z["peq"], z["qinfo"], cands, float(z["distance"]), O.MODE_ALIGNED)
    for half in (res[:nc], res[nc:]):
        np.testing.assert_array_equal(half[keep & ~long_reads], z["results"][keep & ~long_reads])
        np.testing.assert_array_equal(half[long_reads], full[long_reads])
    assert "under threshold" in out
