#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ with the reference's own CPU programs.

Runs only in a container that holds the reference sources: it builds them
with ``make -C oracle ref`` (outputs in oracle/_ref/), writes small seeded
inputs in the reference's file formats, runs the programs, and stores the
inputs together with the programs' outputs as ``.npz`` fixtures (numbers
only, no reference source).

Programs and what their outputs pin:
  greference-4bits      (src/genReference-4bits.c)      FASTA -> packed words
  gcandidates-reorder   (src-api/genCandidate-reorder.c) .prof -> Peq entries + qinfo
  myersCPU-GEM-padding  (src-integration/myers-reorderBuffer-padding.c)
                        aligned windows, variable lengths, 128-base Peq entries
  myers4bitsCPU         (src/myers-4bits.c)            exact-start windows, fixed length
  myers4bitsCPU-padding (src/myers-4bits-padding.c)    aligned windows, fixed length
  checksum-2            (src/checksum-2.c)              under-threshold counts, read back
                        as the threshold (uint32)((float)m * e) per (e, m)

Usage: python tests/golden/make_golden.py [threshold]
       (writes tests/golden/*.npz and threshold_kat.json; "threshold" only the latter)
"""
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")
CODE2CHAR = np.frombuffer(b"ACGTN", dtype=np.uint8)


def build_ref_programs():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True,
                   stdout=subprocess.DEVNULL)


def random_reference(rng, size):
    """ACGT with N runs, some lowercase letters and a few IUPAC letters (all non-ACGT -> N)."""
    codes = rng.integers(0, 4, size=size).astype(np.uint8)
    for _ in range(max(1, size // 2000)):
        start = int(rng.integers(0, size))
        length = int(rng.geometric(0.1))
        codes[start:start + length] = 4
    chars = CODE2CHAR[codes].copy()
    lower = rng.random(size) < 0.05
    chars[lower] = chars[lower] + 32          # a c g t n
    iupac = np.flatnonzero(codes == 4)[::7]
    chars[iupac] = ord("R")
    return codes, chars


def mutate(rng, seq, sub, ins, dele):
    out = []
    i = 0
    n = len(seq)
    while i < n:
        r = rng.random()
        if r < sub:
            out.append((int(seq[i]) + int(rng.integers(1, 4))) % 4)
            i += 1
        elif r < sub + ins:
            out.extend(int(x) for x in rng.integers(0, 4, size=int(rng.geometric(0.7))))
        elif r < sub + ins + dele:
            i += int(rng.geometric(0.7))
        else:
            out.append(int(seq[i]))
            i += 1
    return np.array(out, dtype=np.uint8)


def window_len(m, d):
    return int(np.float32(np.float32(m) * (np.float32(1.0) + np.float32(2.0) * np.float32(d))))


def reads_past_end(cand_query, cand_pos, lengths, distance, size, num_words, mode):
    """Aligned windows of candidates at the very end of the reference read
    the word after the last stored one; the reference programs then read
    past their buffer (undefined), so such results pin nothing."""
    out = np.zeros(len(cand_query), dtype=bool)
    if mode != 0:
        return out
    for i, (q, p) in enumerate(zip(cand_query, cand_pos)):
        n = window_len(lengths[q], distance)
        if p < size and size - p > n:
            entries = n // 8 + (2 if n % 8 else 1)
            out[i] = p // 8 + entries > num_words
    return out


def write_fasta(path, chars):
    with open(path, "wb") as f:
        f.write(b">chrSynthetic test\n")
        for i in range(0, len(chars), 60):
            f.write(chars[i:i + 60].tobytes() + b"\n")


def read_4bits_ref(path):
    with open(path, "rb") as f:
        num_entries, size = struct.unpack("<II", f.read(8))
        words = np.frombuffer(f.read(4 * num_entries), dtype="<u4").copy()
    return size, words


def read_results(path, count):
    with open(path) as f:
        lines = f.read().split("\n")
    assert int(lines[0]) == count, (lines[0], count)
    res = np.zeros((count, 2), dtype=np.uint32)          # [column, score]
    for line in lines[1:1 + count]:
        idx, score, col = (int(x) for x in line.split())
        res[idx, 0] = col
        res[idx, 1] = score
    return res


def candidate_positions(rng, true_pos, m, n, size, extra):
    """True locus (window centred on the read) plus random and edge-case loci."""
    pos = [max(0, true_pos - m // 10)]
    pos += [int(x) for x in rng.integers(0, size - n - 1, size=extra)]
    return pos


def make_variable_length(tmp, rng, name, size, lengths, distance, extra_cands):
    codes, chars = random_reference(rng, size)
    fasta = os.path.join(tmp, "ref.fa")
    write_fasta(fasta, chars)
    subprocess.run([os.path.join(REF_BIN, "greference-4bits"), fasta], check=True,
                   stdout=subprocess.DEVNULL, cwd=tmp)
    ref_size, ref_words = read_4bits_ref(f"{fasta}.{size}.4bits.ref")
    assert ref_size == size

    # Queries as a GEM .prof file (src-api/genCandidate-reorder.c:114-235).
    queries, cand_q, cand_pos = [], [], []
    lines = []
    for qid, m in enumerate(lengths):
        n = window_len(m, distance)
        start = int(rng.integers(0, size - m - 1))
        q = codes[start:start + m].copy()
        q[q == 4] = rng.integers(0, 4, size=int((q == 4).sum())).astype(np.uint8)
        q = mutate(rng, q, 0.03, 0.005, 0.005)[:m]
        if len(q) < m:
            q = np.concatenate([q, rng.integers(0, 4, size=m - len(q)).astype(np.uint8)])
        if m > 20 and qid % 3 == 0:
            q[int(rng.integers(0, m))] = 4       # an N in the read
        qchars = CODE2CHAR[q].copy()
        if qid % 4 == 1:
            qchars[: m // 2] += 32               # lowercase half
        pos = candidate_positions(rng, start, m, n, size, extra_cands)
        r = pos[-1] % 8
        pos[-1] += (qid % 8) - r                 # cover every pos % 8 residue
        pos[-1] = max(0, pos[-1])
        if qid % 5 == 0:                          # edge cases of the skip rule
            pos += [size - n - 1, size - n, size - n + 1, size - 1, size, size + 17, 0]
        queries.append(q)
        for p in pos:
            cand_q.append(qid)
            cand_pos.append(p)
        lines.append(qchars.tobytes().decode() + "".join(
            f"\tchr1:+:{p}:{qid % 7}" for p in pos) + "\n")
    prof = os.path.join(tmp, "reads.prof")
    with open(prof, "w") as f:
        f.writelines(lines)
    subprocess.run([os.path.join(REF_BIN, "gcandidates-reorder"), prof], check=True,
                   stdout=subprocess.DEVNULL, cwd=tmp)
    nq, nc = len(lengths), len(cand_q)
    gem = f"{prof}.{nq}.{nc // nq}.0.gem.qry"
    with open(gem, "rb") as f:
        total_size, total_entries, _zero, num_q, num_c = struct.unpack("<5I", f.read(20))
        peq = np.frombuffer(f.read(80 * total_entries), dtype="<u4").reshape(total_entries, 20).copy()
        c16 = np.frombuffer(f.read(16 * num_c), dtype=np.dtype([("q", "<u4"), ("pad", "<u4"), ("pos", "<u8")])).copy()
        qinfo = np.frombuffer(f.read(8 * num_q), dtype="<u4").reshape(num_q, 2).copy()
    assert num_q == nq and num_c == nc and total_size == sum(lengths)
    assert np.array_equal(c16["q"], np.array(cand_q, dtype=np.uint32))
    assert np.array_equal(c16["pos"], np.array(cand_pos, dtype=np.uint64))
    expect_pe = np.concatenate([[0], np.cumsum([(m + 127) // 128 for m in lengths])[:-1]])
    assert np.array_equal(qinfo[:, 0], expect_pe), "genCandidate-reorder posEntry drifted"
    assert np.array_equal(qinfo[:, 1], np.array(lengths, dtype=np.uint32))

    # The API twin reads 8-byte candidates {u32 query; u32 position}.
    qry = os.path.join(tmp, "reads.bin.qry")
    with open(qry, "wb") as f:
        f.write(struct.pack("<5I", total_size, total_entries, 0, nq, nc))
        f.write(peq.astype("<u4").tobytes())
        c8 = np.zeros((nc, 2), dtype="<u4")
        c8[:, 0] = cand_q
        c8[:, 1] = cand_pos
        f.write(c8.tobytes())
        f.write(qinfo.astype("<u4").tobytes())
    subprocess.run([os.path.join(REF_BIN, "myersCPU-GEM-padding"), repr(distance), f"{fasta}.{size}.4bits.ref", qry],
                   check=True, stdout=subprocess.DEVNULL, cwd=tmp, env=dict(os.environ, OMP_NUM_THREADS="4"))
    res = read_results(qry + ".gemres.pad.cpu", nc)

    qcodes = np.concatenate(queries).astype(np.uint8)
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        ref_codes=codes, ref_chars=chars, ref_words=ref_words, ref_size=np.uint64(size),
        query_codes=qcodes, peq=peq, qinfo=qinfo,
        cand_query=np.array(cand_q, dtype=np.uint32), cand_pos=np.array(cand_pos, dtype=np.uint64),
        distance=np.float32(distance), mode=np.int32(0), max_query_len=np.uint32(4096),
        results=res,
        reads_past_end=reads_past_end(cand_q, cand_pos, lengths, distance, size, len(ref_words), 0))
    return nc


def make_fixed_length(tmp, rng, name, size, m, nq, distance, program, suffix, mode):
    codes, chars = random_reference(rng, size)
    fasta = os.path.join(tmp, "ref.fa")
    write_fasta(fasta, chars)
    subprocess.run([os.path.join(REF_BIN, "greference-4bits"), fasta], check=True,
                   stdout=subprocess.DEVNULL, cwd=tmp)
    ref_size, ref_words = read_4bits_ref(f"{fasta}.{size}.4bits.ref")
    n = window_len(m, distance)
    words_per_q = (m + 31) // 32
    queries, cand_q, cand_pos = [], [], []
    for qid in range(nq):
        start = int(rng.integers(0, size - m - 1))
        q = codes[start:start + m].copy()
        q[q == 4] = 0
        q = mutate(rng, q, 0.04, 0.01, 0.01)[:m]
        if len(q) < m:
            q = np.concatenate([q, rng.integers(0, 4, size=m - len(q)).astype(np.uint8)])
        if qid % 3 == 0:
            q[int(rng.integers(0, m))] = 4
        queries.append(q)
        pos = candidate_positions(rng, start, m, n, size, 3)
        pos.append(int(rng.integers(0, size - n - 1)) // 8 * 8 + qid % 8)
        if qid % 7 == 0:
            pos += [size - n - 1, size - n, size - 1, size]
        for p in pos:
            cand_q.append(qid)
            cand_pos.append(p)
    # Fixed-length .bin.qry: 20-byte Peq per 32 bases (src/genCandidates.c:303-329).
    peq32 = np.zeros((nq * words_per_q, 5), dtype=np.uint32)
    for qid, q in enumerate(queries):
        for i, c in enumerate(q):
            peq32[qid * words_per_q + i // 32, min(int(c), 4)] |= np.uint32(1 << (i % 32))
    qry = os.path.join(tmp, f"{name}.bin.qry")
    nc = len(cand_q)
    with open(qry, "wb") as f:
        f.write(struct.pack("<5I", m * nq, nq * words_per_q, m, nq, nc))
        f.write(peq32.astype("<u4").tobytes())
        c8 = np.zeros((nc, 2), dtype="<u4")
        c8[:, 0] = cand_q
        c8[:, 1] = cand_pos
        f.write(c8.tobytes())
    subprocess.run([os.path.join(REF_BIN, program), repr(distance), f"{fasta}.{size}.4bits.ref", qry],
                   check=True, stdout=subprocess.DEVNULL, cwd=tmp, env=dict(os.environ, OMP_NUM_THREADS="4"))
    res = read_results(qry + suffix, nc)
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        ref_words=ref_words, ref_size=np.uint64(size), query_codes=np.concatenate(queries),
        query_len=np.uint32(m), cand_query=np.array(cand_q, dtype=np.uint32),
        cand_pos=np.array(cand_pos, dtype=np.uint64), distance=np.float32(distance),
        mode=np.int32(mode), max_query_len=np.uint32(0), results=res,
        reads_past_end=reads_past_end(cand_q, cand_pos, [m] * nq, distance, size, len(ref_words), mode))
    return nc


def make_threshold_kat(tmp):
    """For each (e, m), results with scores 0..top go through checksum-2 twice
    (old = new); its count of scores <= threshold is threshold + 1."""
    import json
    import re
    pairs = [(e, m) for e in ("0", "0.02", "0.05", "0.1", "0.15", "0.2", "0.3", "0.7", "1", "1.5")
             for m in (1, 7, 10, 30, 64, 100, 150, 333, 1000, 1023, 4096, 20000)]
    # pairs where a float32 product, a double product of the decimal rate and a
    # double product of the float32 rate disagree
    pairs += [("0.7", 90), ("0.7", 170), ("0.12", 25), ("0.08", 25), ("0.35", 20), ("0.45", 20),
              ("0.9", 10), ("0.9", 20)]
    out = []
    for e, m in pairs:
        path = os.path.join(tmp, f"res_{e}_{m}.txt")
        top = int(m * max(float(e), 1.0)) + 2
        with open(path, "w") as f:
            f.write(f"{top + 1}\n")
            for i in range(top + 1):
                f.write(f"{i} {i} 0\n")
        r = subprocess.run([os.path.join(REF_BIN, "checksum-2"), e, str(m), path, path], check=True,
                           capture_output=True, text=True)
        shown = int(re.search(r"with threshold (\d+)", r.stdout).group(1))
        under = int(re.search(r"New Matching Candidates: (\d+)", r.stdout).group(1))
        assert under == shown + 1, (e, m, r.stdout)
        out.append({"e": e, "m": m, "threshold": shown})
    with open(os.path.join(HERE, "threshold_kat.json"), "w") as f:
        json.dump(out, f, indent=0)
    return len(out)


def main():
    if not os.path.isdir("/root/reference"):
        sys.exit("the reference sources are not present here; fixtures are already committed")
    build_ref_programs()
    tmp = tempfile.mkdtemp(prefix="golden_")
    if sys.argv[1:] == ["threshold"]:
        try:
            print("threshold pairs written:", make_threshold_kat(tmp))
        finally:
            shutil.rmtree(tmp)
        return
    try:
        make_threshold_kat(tmp)
        rng = np.random.default_rng(20140501)
        edge = [1, 2, 7, 8, 31, 32, 33, 63, 64, 65, 96, 100, 127, 128, 129, 150, 200, 255,
                256, 257, 300, 511, 512, 513, 1000, 1023, 1024, 1025, 2000, 4095, 4096, 4097, 5000]
        n1 = make_variable_length(tmp, rng, "api_aligned_d01", 48000, edge, 0.1, 4)
        mixed = [int(x) for x in rng.choice([100, 200, 400, 600, 800, 1000], size=60)]
        n2 = make_variable_length(tmp, rng, "api_aligned_d02_mixed", 48000, mixed, 0.2, 6)
        n3 = make_fixed_length(tmp, rng, "exact_m100_d02", 32000, 100, 120, 0.2,
                               "myers4bitsCPU", ".res-4bits-rev2.cpu", 1)
        n4 = make_fixed_length(tmp, rng, "exact_m64_d01", 32000, 64, 60, 0.1,
                               "myers4bitsCPU", ".res-4bits-rev2.cpu", 1)
        n5 = make_fixed_length(tmp, rng, "exact_m150_d01", 32000, 150, 60, 0.1,
                               "myers4bitsCPU", ".res-4bits-rev2.cpu", 1)
        n6 = make_fixed_length(tmp, rng, "aligned_m100_d01", 32000, 100, 120, 0.1,
                               "myers4bitsCPU-padding", ".res-4bits-padding.cpu", 0)
        print("fixtures written:", n1, n2, n3, n4, n5, n6, "candidates")
    finally:
        shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
