#!/usr/bin/env python3
"""Autotuner for the length-bin table of libmyers_amd.so.

For each representative read length m (query words w = ceil(m/32)) it times
every kernel variant whose bit-vector covers w without more than ~2x padding
(lanes per read L, u32 words per lane W), each with several chunk factors
(chunks of 64/L candidates a persistent workgroup takes per queue pull, i.e.
candidates per workgroup pull = chunk_factor * 64 / L), on a synthetic batch
of reads of that length, and keeps the fastest.  The result is written as the
library's tuning table (lines "maxWords lanes words chunkFactor"), which
initMyers loads from tuning/gfx950.tsv next to the library directory.

  python myers-gpu_amd/python/autotune.py [--out myers-gpu_amd/tuning/gfx950.tsv]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import myers_amd as M  # noqa: E402

SHORT = [64, 100, 128, 160, 200, 256, 320, 400, 480, 600, 700, 800, 900, 1000]
# Long reads: one length at each capacity (32 * lanes * words bases) of the
# multi-lane variants, so that the variant chosen at a capacity serves exactly
# the reads between the previous capacity and it.
LONG = [1536, 2048, 3072, 4096, 5120, 6144, 7168, 8192, 10240, 12288, 14336, 16384, 20480, 24576]
CHUNK_FACTORS = (1, 2, 4)
# A variant of the queue launch (one lane, or at most 64 lanes of at most 12
# words) keeps its bin in the batch's single launch; another variant needs a
# launch of its own in mixed batches, so it is chosen only when it is this
# much faster alone.
QUEUE_PREFERENCE = 1.05
REF_SIZE = 200_000_000
# Each measurement must run in the throughput regime: enough candidates to
# keep every SIMD of the 256 CUs busy for many rounds (a batch that fills the
# GPU only once favours variants that spread a read over more lanes).
TARGET_CELLS = 1.5e12
MIN_CANDIDATES, MAX_CANDIDATES = 20_000, 2_000_000


def log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def candidates_for(variants, w):
    out = []
    for i, v in enumerate(variants):
        t = v["lanes"] * v["words"]
        if t >= w and t <= max(2 * w, w + 4):
            out.append(i)
    return out


def in_queue_launch(info):
    return info["lanes"] == 1 or (info["lanes"] <= 64 and info["words"] <= 12)


def read_table(path):
    """Rows (maxWords, lanes, words, chunkFactor, comment) of a tuning table."""
    rows = []
    with open(path) as f:
        for line in f:
            body, _, comment = line.partition("#")
            parts = body.split()
            if len(parts) == 4:
                rows.append((int(parts[0]), int(parts[1]), int(parts[2]), int(parts[3]), comment.strip()))
    return rows


def time_variant(m, bufs, b, v, cf, reps):
    table = m.bin_table()
    w = (int(b["qinfo"][0, 1]) + 31) // 32
    table[w] = v
    m.set_bin_table(table)
    m.set_chunk_factor(v, cf)
    dp, dq, dc, dr = bufs
    nq, nc, ne = len(b["qinfo"]), len(b["cands"]), len(b["peq"])
    M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=M.ENGINE_PERSISTENT, h_qinfo=b["qinfo"])
    M.synchronize(m)
    best = float("inf")
    for _ in range(reps):
        M.verify_device(m, dp, ne, dq, nq, dc, nc, 0.1, dr, flags=M.ENGINE_PERSISTENT, h_qinfo=b["qinfo"])
        M.synchronize(m)
        best = min(best, m.stats().kernelMs)
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(M.PKG_DIR, "tuning", "gfx950.tsv"))
    ap.add_argument("--json", default=None, help="write every measurement here")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--lengths", default=None, help="comma-separated read lengths (default: built-in set)")
    ap.add_argument("--merge", default=None,
                    help="keep this table's rows for the word counts the run does not time")
    args = ap.parse_args()
    lengths = [int(x) for x in args.lengths.split(",")] if args.lengths else SHORT + LONG

    ref = M.synth_reference(99, REF_SIZE)
    # start from the built-in table (no tuning file) so earlier tables do not leak in
    os.environ["MYERS_TUNING_FILE"] = os.devnull
    m = M.Myers(ref, REF_SIZE, avg_query=1000, cands_per_query=11, buffer_bytes=3 << 30)
    variants = m.variants()
    default_table = m.bin_table()
    records, chosen = [], []
    for length in lengths:
        w = (length + 31) // 32
        cands_per_read = 11 if length <= 1024 else 4
        n_cands = int(min(MAX_CANDIDATES, max(MIN_CANDIDATES, TARGET_CELLS / (length * 1.2 * length))))
        n_reads = -(-n_cands // cands_per_read)
        spec = M.make_spec(1000 + length, n_reads, length=length, sub=0.03, ins=0.005, dele=0.005,
                           candidates=cands_per_read, distance=0.1)
        b = M.synth_batch(spec, ref, REF_SIZE)
        bufs = (M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"]),
                M.DeviceArray.empty((len(b["cands"]), 2), np.uint32))
        best = best_queue = None
        for v in candidates_for(variants, w):
            for cf in CHUNK_FACTORS:
                ms = time_variant(m, bufs, b, v, cf, args.reps)
                gcups = b["cells"] / (ms / 1e3) / 1e9
                rec = {"length": length, "words": w, "lanes": variants[v]["lanes"], "wpl": variants[v]["words"],
                       "chunk_factor": cf, "kernel_ms": ms, "gcups": gcups, "candidates": len(b["cands"])}
                records.append(rec)
                if best is None or ms < best[0]:
                    best = (ms, v, cf, gcups)
                if in_queue_launch(variants[v]) and (best_queue is None or ms < best_queue[0]):
                    best_queue = (ms, v, cf, gcups)
        if best_queue is not None and best_queue[0] <= best[0] * QUEUE_PREFERENCE:
            best = best_queue
        for x in bufs:
            x.free()
        m.set_bin_table(default_table)
        ms, v, cf, gcups = best
        chosen.append((w, variants[v]["lanes"], variants[v]["words"], cf, length, gcups))
        log(f"m={length:5d} w={w:3d}: L{variants[v]['lanes']}xW{variants[v]['words']} cf={cf} "
            f"{gcups:9.0f} GCUPS (default L{variants[default_table[w]]['lanes']}xW{variants[default_table[w]]['words']})")
    m.close()

    rows = [(w, lanes, words, cf, f"m={length}: {gcups:.0f} GCUPS")
            for w, lanes, words, cf, length, gcups in chosen]
    if args.merge:
        timed = {r[0] for r in rows}
        rows += [r for r in read_table(args.merge) if r[0] not in timed]
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("# maxWords lanes words chunkFactor  -- written by autotune.py on gfx950\n")
        f.write(f"# {time.strftime('%Y-%m-%d')}: best persistent-kernel time per read length "
                "(reads of m bases use the first line with maxWords >= ceil(m/32)); a variant outside "
                f"the queue launch only when {QUEUE_PREFERENCE:.2f}x faster\n")
        prev = 0
        for w, lanes, words, cf, note in sorted(rows):
            if w <= prev:
                continue
            f.write(f"{w} {lanes} {words} {cf}    # {note}\n")
            prev = w
    log(f"wrote {args.out}")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(records, f)


if __name__ == "__main__":
    main()
