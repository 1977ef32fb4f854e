#!/usr/bin/env python3
"""Throughput of the candidate-verification path on one MI355X.

Headline workload (config 3, the mixed-length batch): seed-7 synthetic
reference of 3.1 Gbases (1.55 GB packed, resident in HBM), 1M reads with
lengths uniform on {100, 200, 400, 600, 800, 1000} bp, error uniform in
[1%, 5%] (80% substitutions, 20% indels), Poisson(11) candidates per read
capped at 64 with 1 true locus, window 1.2 m (distance 0.1), aligned windows
(the API kernel's semantics), persistent engine (one queue launch over all
length bins).  It is compared on the same batch with the plain engine (one
launch per length bin, in turn), per-bin resident grids, and one launch of
the kernel sized for the longest read.

Beside the headline, by default: config 2 (100K x 1000 bp, plain one-length
kernel) and config 4 (50K long noisy reads, 2-20 kb), each with its own
roofline and a CPU check on a sample; threshold mode on each; the
serving sweep (config 5).

A step is one pass of the pipeline (binning + verification) over the whole
batch, with the inputs already in HBM.  The script prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--config 1|2|3|4] [--legs 2,4]
                  [--engine auto|plain|persistent|binned]
"""
import argparse
import json
import os
import struct
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "myers-gpu_amd", "python"))
# The serving leg runs several buffers at once, each with a main stream and
# side streams; HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
# default), and kernels of streams that share a queue run one after another.
# Give the process enough queues (at least 16) before the runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np  # noqa: E402

import myers_amd as M  # noqa: E402

# MI355X published figures: 256 CUs, 2.4 GHz peak engine clock, 8 TB/s HBM3E,
# 157.3 TFLOP/s FP32 vector = 256 CUs x 128 lanes x 2 flops (FMA) x 2.4 GHz.
CUS, CLOCK_HZ = 256, 2.4e9
# VALU issue peak: each SIMD issues a wave64 VALU instruction per quad-cycle,
# two of the double-rate kinds (and/or/xor/bitop3/add without carry), so a
# CU retires at most 128 lane-ops per clock: 78.6e12 lane-ops/s.
PEAK_VALU_OPS = CUS * 128 * CLOCK_HZ
PEAK_SINGLE_ISSUE_OPS = CUS * 64 * CLOCK_HZ                 # one instruction per quad-cycle
PEAK_HBM_GBS = 8000.0
# Algorithmic VALU ops per 32 cells (one u32 word of one column):
#   K_GFX950 = 10: and, add-with-carry, 3 three-input logic ops (Xh, Ph, Pv),
#   Mh and, Mv and, Xv or, 2 funnel shifts -- the recurrence's minimum on gfx950.
#   7 of them issue at the double rate, the carry add and the two funnel
#   shifts at the single rate: 6.5 single-issue slots per word, which caps
#   this op mix at 10/6.5 x PEAK_SINGLE_ISSUE_OPS = 60.5e12 ops/s (194 TCUPS).
#   K_REF = 18: the reference's u32 formulation (src/myers-4bits.c:117-134).
K_GFX950 = 10
K_REF = 18
OP_MIX_SLOTS = 7 * 0.5 + 3 * 1.0
PEAK_OP_MIX_OPS = PEAK_SINGLE_ISSUE_OPS * K_GFX950 / OP_MIX_SLOTS

CONFIGS = {
    1: dict(name="config1", ref_seed=42, ref_size=4_000_000, reads=10_000, spec=dict(
        length=100, sub=0.02, ins=0.0025, dele=0.0025, candidates=11)),
    2: dict(name="config2", ref_seed=7, ref_size=3_100_000_000, reads=100_000, spec=dict(
        length=1000, sub=0.03, ins=0.005, dele=0.005, candidates=11)),
    3: dict(name="config3", ref_seed=7, ref_size=3_100_000_000, reads=1_000_000, spec=dict(
        choices=[100, 200, 400, 600, 800, 1000], err_range=(0.01, 0.05), sub_frac=0.8, poisson=11, cap=64)),
    4: dict(name="config4", ref_seed=7, ref_size=3_100_000_000, reads=50_000, spec=dict(
        log_uniform=(2000, 20000), err_range=(0.12, 0.12), sub_frac=0.5, candidates=4)),
}
DISTANCE = 0.1
# HBM bytes per launch of the dominant kernel under the auto engine (the queue
# launch for each), from FETCH_SIZE and WRITE_SIZE counter runs scaled by a 1 GiB
# copy calibration (profiles/verify_queue_kernel_config2.md,
# profiles/hbm_traffic_verify_queue_config3.md, profiles/hbm_traffic_verify_queue_config4.md).
MEASURED_TRAFFIC = {(2, "aligned"): 1.624e9, (3, "aligned"): 9.547e9, (4, "aligned"): 1.536e9}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def window_len(m):
    """n = (uint32)(m * (1 + 2*d)) in float32 (src-api/coreMyersGPU-CC35.cu:131)."""
    f = np.float32(1.0) + np.float32(2.0) * np.float32(DISTANCE)
    return np.floor(np.asarray(m).astype(np.float32) * f).astype(np.int64)


def algorithmic_bytes(qinfo, cands):
    """Text window (n/2 B packed) + candidate (16 B) + result (8 B) per
    candidate, plus Peq (80 B per 128 bases) and qinfo (8 B) per read."""
    m = qinfo[cands["query"], 1].astype(np.int64)
    n = window_len(m)
    per_cand = (n + 1) // 2 + 16 + 8
    per_read = 80 * ((qinfo[:, 1].astype(np.int64) + 127) // 128) + 8
    return int(per_cand.sum() + per_read.sum())


def threshold_of(m, e):
    """(uint32)((float)m * e), the float32 product of src/checksum-2.c:88."""
    return int(np.float32(m) * np.float32(e))


REF_PROGRAM = os.path.join(REPO, "oracle", "_ref", "myersCPU-GEM-padding")


def reference_cpu(ref, ref_size, b, idx, gpu_results, threads):
    """Run the reference's own CPU program -- src-integration/myers-reorderBuffer-padding.c,
    the API's aligned-window twin, compiled from its sources with the reference's
    flags into oracle/_ref by `make -C oracle ref` -- on the sampled candidates:
    it reads the .4bits.ref / .bin.qry layouts (myers-reorderBuffer-padding.c:211-272),
    times its compute (:486-496) and writes "<id> <score> <column>" lines (:404-430).
    Returns (seconds, None) after checking every result against the GPU's, or
    (None, reason) when the program cannot run this workload."""
    if not os.access(REF_PROGRAM, os.X_OK):
        return None, "oracle/_ref not built (needs the reference sources)"
    cands = b["cands"][idx]
    if int(cands["position"].max()) >= 1 << 32:
        return None, "positions beyond 32 bits (the program's candidates are u32)"
    if int(b["qinfo"][:, 1].max()) > 4096:
        return None, "reads over 4096 bases (the program skips them)"
    qinfo, peq = b["qinfo"], b["peq"]
    with tempfile.TemporaryDirectory(prefix="myers_ref_") as tmp:
        ref_path = os.path.join(tmp, "ref.4bits.ref")
        with open(ref_path, "wb") as f:
            f.write(struct.pack("<II", len(ref), int(ref_size)))
            np.ascontiguousarray(ref, dtype="<u4").tofile(f)
        qry = os.path.join(tmp, "sample.bin.qry")
        with open(qry, "wb") as f:
            f.write(struct.pack("<5I", int(qinfo[:, 1].astype(np.int64).sum()), len(peq), 0, len(qinfo), len(idx)))
            np.ascontiguousarray(peq, dtype="<u4").tofile(f)
            c8 = np.empty((len(idx), 2), dtype="<u4")
            c8[:, 0] = cands["query"]
            c8[:, 1] = cands["position"]
            c8.tofile(f)
            np.ascontiguousarray(qinfo, dtype="<u4").tofile(f)
        # The program's cleanup frees res->h_reorderResults, which initResults
        # never sets (myers-reorderBuffer-padding.c:278-296,370): a heap chunk
        # reused from a freed stdio buffer then hands free() file bytes, and the
        # process dies after writing its results.  glibc's MALLOC_PERTURB_=255
        # fills each new allocation with 0x00, so that free() gets NULL.
        run = subprocess.run([REF_PROGRAM, repr(DISTANCE), ref_path, qry], capture_output=True, text=True,
                             env=dict(os.environ, OMP_NUM_THREADS=str(threads), MALLOC_PERTURB_="255"),
                             timeout=900)
        if run.returncode != 0:
            return None, f"program exited with {run.returncode}: {run.stdout[-200:]} {run.stderr[-200:]}"
        secs = None
        for line in run.stdout.splitlines():
            if line.startswith("TIME COMPUTE MYERS"):
                secs = float(line.split(":")[1])
        with open(qry + ".gemres.pad.cpu") as f:
            tok = f.read().split()
    if secs is None or int(tok[0]) != len(idx):
        return None, "unexpected program output"
    rows = np.array(tok[1:], dtype=np.int64).reshape(-1, 3)       # id, score, column
    got = np.empty((len(idx), 2), dtype=np.uint32)
    got[rows[:, 0], 0] = rows[:, 2]
    got[rows[:, 0], 1] = rows[:, 1]
    bad = int((got != gpu_results[idx]).any(axis=1).sum())
    if bad:
        raise SystemExit(f"bench: GPU results differ from the reference program on {bad} of {len(idx)} candidates")
    return secs, None


def cpu_baseline(ref, ref_size, b, gpu_results, seconds, mode_exact=False):
    """Time the oracle (C restatement of src/myers-4bits*.c, OpenMP) on a
    seeded sample of the same candidates and check the GPU results on it;
    then, where it is built, the reference's own CPU program on the same
    sample (aligned windows only), which becomes the reported baseline."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    omode = O.MODE_EXACT if mode_exact else O.MODE_ALIGNED
    rng = np.random.default_rng(1234)
    cands = b["cands"]
    order = rng.permutation(len(cands))
    n = 256
    elapsed = 0.0
    checked = 0
    while True:
        idx = np.sort(order[:n])
        t0 = time.perf_counter()
        res = O.myers_batch(ref, ref_size, b["peq"], b["qinfo"], cands[idx], DISTANCE, omode)
        elapsed = time.perf_counter() - t0
        mismatches = int((res != gpu_results[idx]).any(axis=1).sum())
        checked = n
        if mismatches:
            raise SystemExit(f"bench: GPU results differ from the oracle on {mismatches} of {n} sampled candidates")
        if elapsed >= seconds * 0.5 or n >= len(cands):
            break
        n = min(len(cands), int(n * max(2.0, min(16.0, seconds / max(elapsed, 1e-3)))))
    m = b["qinfo"][cands["query"][idx], 1].astype(np.int64)
    cells = int((m * window_len(m)).sum())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # GEMTools' block BPM over 64-bit words on the same sample (a second CPU baseline)
    t0 = time.perf_counter()
    res64 = O.bpm64_batch(ref, ref_size, b["peq"], b["qinfo"], cands[idx], DISTANCE, omode)
    elapsed64 = time.perf_counter() - t0
    if not np.array_equal(res64, res):
        raise SystemExit("bench: the 64-bit BPM baseline differs from the oracle")
    port = {
        "value": checked / elapsed,
        "unit": "candidates/s",
        "gcups": cells / elapsed / 1e9,
        "cores": threads,
        "kind": "port",
        "sample": f"{checked} seeded-random candidates of the same batch ({elapsed:.1f} s, OpenMP static); "
                  f"GPU results bit-exact on all of them",
        "cpu_model": _cpu_model(),
        "algorithm": "u32 blocks with int8 carry (src/myers-4bits.c)",
    }
    bpm64 = {"value": checked / elapsed64, "unit": "candidates/s", "gcups": cells / elapsed64 / 1e9,
             "cores": threads, "kind": "port",
             "algorithm": "64-bit block BPM (GEMTools/src/gt_map_align_bpm_generic.c)"}
    secs, why = (None, "exact windows (the program is the aligned-window twin)") if mode_exact else \
        reference_cpu(ref, ref_size, b, idx, gpu_results, threads)
    if secs is None:
        return dict(port, bpm64=bpm64, reference={"skipped": why})
    return {
        "value": checked / secs,
        "unit": "candidates/s",
        "gcups": cells / secs / 1e9,
        "cores": threads,
        "kind": "reference",
        "sample": f"the same {checked} candidates; the program's own compute time ({secs:.1f} s, OpenMP); "
                  f"GPU results bit-exact on all of them",
        "cpu_model": _cpu_model(),
        "program": "oracle/_ref/myersCPU-GEM-padding: src-integration/myers-reorderBuffer-padding.c "
                   "built with -O3 -m64 -fopenmp",
        "port": port,
        "bpm64": bpm64,
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_steps(m, bufs, b, flags, steps, warmup):
    """Time `steps` passes of the pipeline over a device-resident batch."""
    dp, dq, dc, dr = bufs
    nq, nc = len(b["qinfo"]), len(b["cands"])

    def step():
        M.verify_device(m, dp, len(b["peq"]), dq, nq, dc, nc, DISTANCE, dr, flags=flags, h_qinfo=b["qinfo"])

    for _ in range(warmup):
        step()
    M.synchronize(m)
    M.lib().myersResetStats(m.buffer(0))
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    M.synchronize(m)     # waits for the last step; every step's events are folded in
    t1 = time.perf_counter()
    st = m.stats()
    if st.batchesTotal != steps:
        raise SystemExit(f"bench: {st.batchesTotal} timed batches for {steps} steps")
    return {"elapsed": t1 - t0, "kernel_ms": st.kernelMsTotal / st.batchesTotal,
            "verify_ms": st.verifyMsTotal / st.batchesTotal,
            "launches": st.kernelLaunchesTotal / st.batchesTotal, "binned": bool(st.binned),
            "cand_per_s": nc * steps / (t1 - t0)}


# ------------------------------------------------------------------ serving

def serving_pool(ref, ref_size, seed, n_requests):
    """Config-5 requests: 1..4096 reads each (log-uniform); 90% of reads from the
    mixed-length set (config 3), 10% long noisy reads (config 4)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0.0, np.log(4096.0), size=n_requests)).round().astype(int).clip(1, 4096)
    n_long = [int(rng.binomial(s, 0.1)) for s in sizes]
    short = M.synth_batch(M.make_spec(seed + 1, int(sum(sizes) - sum(n_long)) + 1, distance=DISTANCE,
                                      **CONFIGS[3]["spec"]), ref, ref_size)
    long = M.synth_batch(M.make_spec(seed + 2, max(1, sum(n_long)), distance=DISTANCE, **CONFIGS[4]["spec"]),
                         ref, ref_size)

    def index(b):
        per = np.bincount(b["cands"]["query"], minlength=len(b["qinfo"]))
        return np.concatenate([[0], np.cumsum(per)])

    cs, cl = index(short), index(long)
    reqs = []
    si = li = 0
    for s, nl in zip(sizes, n_long):
        parts = [(short, cs, si, si + s - nl), (long, cl, li, li + nl)]
        si += s - nl
        li += nl
        peqs, qis, cands, e_off, q_off = [], [], [], 0, 0
        for b, cix, q0, q1 in parts:
            if q1 <= q0:
                continue
            qi = b["qinfo"][q0:q1].copy()
            e0 = qi[0, 0]
            e1 = qi[-1, 0] + (qi[-1, 1] + 127) // 128
            peqs.append(b["peq"][e0:e1])
            qi[:, 0] = (qi[:, 0].astype(np.int64) - int(e0) + e_off).astype(np.uint32)
            qis.append(qi)
            c = b["cands"][cix[q0]:cix[q1]].copy()
            c["query"] = (c["query"].astype(np.int64) - q0 + q_off).astype(np.uint32)
            cands.append(c)
            e_off += e1 - e0
            q_off += q1 - q0
        reqs.append((np.concatenate(peqs), np.concatenate(qis), np.concatenate(cands)))
    return reqs


def concat_requests(reqs):
    """One batch holding every request (Peq offsets and query ids rebased)."""
    peqs, qis, cands, e, q = [], [], [], 0, 0
    for peq, qi, c in reqs:
        qi = qi.copy()
        qi[:, 0] += np.uint32(e)
        c = c.copy()
        c["query"] += np.uint32(q)
        peqs.append(peq)
        qis.append(qi)
        cands.append(c)
        e += len(peq)
        q += len(qi)
    return np.concatenate(peqs), np.concatenate(qis), np.concatenate(cands)


def measure_serving(m, ref, ref_size, loads, seconds, batch_cells=0, engine=0, seed=5, resident=True,
                    no_long_launch=False):
    pool = serving_pool(ref, ref_size, seed, 192)
    cells = [int((r[1][r[2]["query"], 1].astype(np.int64) * window_len(r[1][r[2]["query"], 1])).sum()) for r in pool]
    cands = [len(r[2]) for r in pool]
    alone = [m.verify(*r, DISTANCE) for r in pool[:24]]          # reference results for the mix check
    # bulk rate of the same mix: the whole pool as one batch (host copies included, as in serving)
    whole = concat_requests(pool)
    bulk = []
    for _ in range(3):
        t0 = time.perf_counter()
        whole_res = m.verify(*whole, DISTANCE)
        bulk.append(len(whole[2]) / (time.perf_counter() - t0))
    bulk_cps = max(bulk)
    offs = np.concatenate([[0], np.cumsum(cands)])
    for k in range(24):
        if not np.array_equal(whole_res[offs[k]:offs[k + 1]], alone[k]):
            raise SystemExit("bench: results of the pool as one batch differ from the requests verified alone")
    out = {"pool_requests": len(pool), "mean_reads_per_request": float(np.mean([len(r[1]) for r in pool])),
           "mean_candidates_per_request": float(np.mean(cands)),
           "server": "resident kernel" if resident else "continuous batching",
           "buffers": 1 if resident else m.num_buffers,
           "max_batch_cells": None if resident else (batch_cells or "library default (5e11)"),
           "engine": engine, "levels": []}
    flags = engine | (M.SERVE_RESIDENT if resident else 0) | (M.NO_LONG_LAUNCH if no_long_launch and not resident else 0)
    if no_long_launch and not resident:
        out["no_long_launch"] = True
    with M.Server(m, distance=DISTANCE, flags=flags, max_batch_cells=batch_cells) as sv:
        # saturated throughput: every pool request arrives at once (best of 3, as the bulk rate)
        walls = []
        for _ in range(3):
            made = [M.Server.request(*r) for r in pool]
            lat, wall = sv.run_open_loop([x[0] for x in made], np.zeros(len(made), dtype=np.uint64))
            walls.append(wall)
        wall = min(walls)
        sat_cps = sum(cands) / (wall / 1e3)
        sat_rps = len(pool) / (wall / 1e3)
        out["bulk_candidates_per_s"] = bulk_cps
        out["saturated_candidates_per_s"] = sat_cps
        out["saturated_of_bulk"] = sat_cps / bulk_cps
        out["saturated_gcups"] = sum(cells) / (wall / 1e3) / 1e9
        for (req, keep), a in zip(made[:24], alone):
            if not np.array_equal(keep[3], a):
                raise SystemExit("bench: served results differ from the same requests verified alone")
        rng = np.random.default_rng(seed + 1)
        for load in loads:
            # offered load as a fraction of the mix's bulk rate (the whole pool as one batch)
            rate = load * bulk_cps / float(np.mean(cands))
            n = max(50, int(rate * seconds))
            idx = rng.integers(0, len(pool), size=n)
            arrivals = (np.cumsum(rng.exponential(1.0 / rate, size=n)) * 1e9).astype(np.uint64)
            made = [M.Server.request(*pool[i]) for i in idx]
            lat, wall, qms = sv.run_open_loop([x[0] for x in made], arrivals, queue=True)
            tail = lat >= np.percentile(lat, 99)
            for (req, keep), i in zip(made, idx):
                if i < len(alone) and not np.array_equal(keep[3], alone[i]):
                    raise SystemExit("bench: served results differ from the same requests verified alone")
            nc = sum(cands[i] for i in idx)
            out["levels"].append({
                "load": load, "requests": n, "offered_requests_per_s": rate,
                "offered_of_bulk": rate * float(np.mean(cands)) / bulk_cps,
                "offered_of_saturated": rate / sat_rps,
                "candidates_per_s": nc / (wall / 1e3), "gcups": sum(cells[i] for i in idx) / (wall / 1e3) / 1e9,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "p999_ms": float(np.percentile(lat, 99.9)), "max_ms": float(lat.max()),
                "queue_p50_ms": float(np.percentile(qms, 50)), "queue_p99_ms": float(np.percentile(qms, 99)),
                "tail_queue_ms": float(qms[tail].mean()),
                "tail_reads": float(np.mean([len(pool[i][1]) for i in idx[tail]])),
                "max_reads_per_request": int(max(len(pool[i][1]) for i in idx)),
                "batches": int(sv.stats().batches)})
            log(f"serving load {load:.2f} ({out['levels'][-1]['offered_of_bulk']:.2f} of bulk): "
                f"{out['levels'][-1]['candidates_per_s']:.3e} cand/s, "
                f"p50 {out['levels'][-1]['p50_ms']:.2f} ms, p99 {out['levels'][-1]['p99_ms']:.2f} ms, "
                f"queue p99 {out['levels'][-1]['queue_p99_ms']:.2f} ms, slowest 1%: queue "
                f"{out['levels'][-1]['tail_queue_ms']:.2f} ms, {out['levels'][-1]['tail_reads']:.0f} reads")
    return out


def roofline(cells, kernel_ms, hbm_bytes, traffic, kernel):
    """VALU roofline of one launch: algorithmic ops (K_GFX950 per word-column
    of the nominal cells) over the verification launches' HIP-event time."""
    kernel_s = kernel_ms / 1e3
    achieved = cells * K_GFX950 / 32.0 / kernel_s
    hbm = hbm_bytes / kernel_s / 1e9
    return {
        "bound": "valu",
        "achieved": achieved / 1e12,
        "peak": PEAK_VALU_OPS / 1e12,
        "unit": "TOP/s",
        "frac": achieved / PEAK_VALU_OPS,
        "traffic": traffic,
        "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
        "kernel": kernel,
        "kernel_ms": kernel_ms,
        "ops_per_cell": K_GFX950 / 32.0,
        "peak_note": "256 CUs x 128 lane-ops/clk x 2.4 GHz: the dual-issue VALU rate",
        "op_mix_ceiling": {"peak": PEAK_OP_MIX_OPS / 1e12, "unit": "TOP/s", "frac": achieved / PEAK_OP_MIX_OPS,
                           "gcups": PEAK_OP_MIX_OPS / (K_GFX950 / 32.0) / 1e9,
                           "note": "7 double-rate + 3 single-rate ops per word: 6.5 single-issue slots"},
        "ratio_to_single_issue_peak": achieved / PEAK_SINGLE_ISSUE_OPS,
        "ratio_reference_ops_to_peak": cells * K_REF / 32.0 / kernel_s / PEAK_VALU_OPS,
        "hbm": {"achieved": hbm, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": hbm / PEAK_HBM_GBS,
                "algorithmic_bytes": hbm_bytes},
    }


def variant_names(m, b):
    info = m.variants()
    used = sorted({int(v) for v in m.bin_table()[np.unique((b["qinfo"][:, 1] + 31) // 32)]})
    return [f"L{info[v]['lanes']}xW{info[v]['words']}" for v in used]


def dominant_kernel(m, b, engine, run):
    """Name of the launch the roofline is taken over."""
    names = variant_names(m, b)
    if not run["binned"]:
        return f"verify_kernel<{names[0][1:].replace('xW', ', ')}, false> (one launch)"
    if engine in (M.ENGINE_PERSISTENT, M.ENGINE_AUTO) and run["launches"] == 1:
        return "verify_queue_kernel (one launch over every bin: " + ", ".join(names) + ")"
    return f"{run['launches']:.0f} launches ({', '.join(names)})"


def measure_config(m, ref, cfg_id, steps, warmup, engine, mode, graph, compare, cpu_seconds, threshold=-1.0):
    """Time one configuration on the device-resident batch; check a sample
    against the CPU; optionally compare the engines and time threshold mode."""
    cfg = CONFIGS[cfg_id]
    spec = M.make_spec(cfg["ref_seed"] + 1, cfg["reads"], distance=DISTANCE, **cfg["spec"])
    t0 = time.time()
    b = M.synth_batch(spec, ref, cfg["ref_size"])
    nc, nq = len(b["cands"]), len(b["qinfo"])
    log(f"{cfg['name']}: {nq} reads, {nc} candidates, {b['cells']:.3e} cells in {time.time() - t0:.1f}s")
    mode_flag = M.WINDOW_EXACT if mode == "exact" else M.WINDOW_ALIGNED
    gflag = M.USE_GRAPH if graph else 0
    flags = mode_flag | engine | gflag
    bufs = (M.DeviceArray(b["peq"]), M.DeviceArray(b["qinfo"]), M.DeviceArray(b["cands"]),
            M.DeviceArray.empty((nc, 2), np.uint32))
    try:
        run = run_steps(m, bufs, b, flags, steps, warmup)
        results = bufs[3].get()
        cells = b["cells"]
        value = nc * steps / run["elapsed"]
        out = {
            "workload": f"{cfg['name']}: seed-{cfg['ref_seed']} {cfg['ref_size'] / 1e9:.3g} Gbase reference, "
                        f"{nq} reads, {nc} candidates, window = len*1.2, {mode} windows",
            "candidates_per_s": value, "gcups": cells * steps / run["elapsed"] / 1e9,
            "ms_per_step": run["elapsed"] / steps * 1e3, "steps": steps, "warmup": warmup,
            "reads": nq, "candidates": nc, "cells": cells, "variants": variant_names(m, b),
            "launches_per_step": run["launches"], "binned": run["binned"],
            "pipeline_ms": run["kernel_ms"], "verify_ms": run["verify_ms"],
            "roofline": roofline(cells, run["verify_ms"], algorithmic_bytes(b["qinfo"], b["cands"]),
                                 MEASURED_TRAFFIC.get((cfg_id, mode)), dominant_kernel(m, b, engine, run)),
        }
        log(f"{cfg['name']}: {value:.4e} candidates/s, {out['gcups']:.0f} GCUPS, verify {run['verify_ms']:.3f} ms, "
            f"pipeline {run['kernel_ms']:.3f} ms, {run['launches']:.0f} launches, "
            f"frac {out['roofline']['frac']:.3f}")
        if compare:
            other = {}
            for name, eng in (("plain", M.ENGINE_PLAIN), ("binned", M.ENGINE_BINNED),
                              ("persistent", M.ENGINE_PERSISTENT)):
                r = run_steps(m, bufs, b, mode_flag | eng | gflag, max(3, steps // 2), 1)
                if not np.array_equal(bufs[3].get(), results):
                    raise SystemExit(f"bench: engine {name} differs from the timed run")
                other[name] = r
            # the plain kernel sized for the longest read, one launch over the whole batch
            table = m.bin_table()
            wmax = int((int(b["qinfo"][:, 1].max()) + 31) // 32)
            v_long = int(table[wmax])
            forced = table.copy()
            forced[1:wmax + 1] = v_long
            m.set_bin_table(forced)
            one = run_steps(m, bufs, b, mode_flag | M.ENGINE_PLAIN, max(3, steps // 2), 1)
            if not np.array_equal(bufs[3].get(), results):
                raise SystemExit("bench: the one-variant launch differs from the timed run")
            m.set_bin_table(table)
            vinfo = m.variants()[v_long]
            cps = {k: r["cand_per_s"] for k, r in other.items()}
            out["engine_comparison"] = {
                "plain_candidates_per_s": cps["plain"], "binned_candidates_per_s": cps["binned"],
                "persistent_candidates_per_s": cps["persistent"],
                "plain_verify_ms": other["plain"]["verify_ms"], "binned_verify_ms": other["binned"]["verify_ms"],
                "persistent_verify_ms": other["persistent"]["verify_ms"],
                "plain_launches": other["plain"]["launches"], "binned_launches": other["binned"]["launches"],
                "persistent_launches": other["persistent"]["launches"],
                "persistent_over_plain": cps["persistent"] / cps["plain"],
                "persistent_over_binned": cps["persistent"] / cps["binned"],
                "one_launch_variant": f"L{vinfo['lanes']}xW{vinfo['words']}",
                "one_launch_candidates_per_s": one["cand_per_s"],
                "persistent_over_one_launch": cps["persistent"] / one["cand_per_s"],
                "note": "plain = one launch per length bin, in turn, on one stream; binned = one resident grid "
                        "per bin on forked streams; persistent = one queue launch over every bin; one launch = "
                        "the kernel for the longest read over every candidate",
            }
            log(f"{cfg['name']} engines: " + ", ".join(f"{k} {v:.4e}" for k, v in cps.items()) +
                f", one launch {one['cand_per_s']:.4e}")
        if threshold >= 0:
            m.set_threshold(threshold)
            r = run_steps(m, bufs, b, flags, max(3, steps // 2), 1)
            m.set_threshold(-1.0)
            tres = bufs[3].get()
            # same results as the full run for candidates under the threshold, the no-match pair otherwise
            t = np.array([threshold_of(x, threshold) for x in range(int(b["qinfo"][:, 1].max()) + 1)],
                         dtype=np.int64)[b["qinfo"][b["cands"]["query"], 1]]
            under = (results[:, 1] != 0xFFFFFFFF) & (results[:, 1].astype(np.int64) <= t)
            skipped = results[:, 1] == 0xFFFFFFFF
            ok = (np.array_equal(tres[under], results[under]) and np.array_equal(tres[skipped], results[skipped])
                  and np.all(tres[~under & ~skipped] == 0xFFFFFFFF))
            if not ok:
                raise SystemExit("bench: threshold-mode results differ from the full run")
            out["threshold_mode"] = {"error_rate": threshold, "candidates_per_s": r["cand_per_s"],
                                     "verify_ms": r["verify_ms"], "under_threshold": int(under.sum()),
                                     "speedup_vs_full": r["cand_per_s"] / value}
            log(f"{cfg['name']} threshold mode: {out['threshold_mode']}")
        if cpu_seconds > 0:
            log(f"{cfg['name']}: CPU baseline / parity sample ...")
            out["cpu_baseline"] = cpu_baseline(ref, cfg["ref_size"], b, results, cpu_seconds, mode == "exact")
            log(f"{cfg['name']}: CPU {out['cpu_baseline']['value']:.4e} candidates/s "
                f"({out['cpu_baseline']['kind']})")
        return out
    finally:
        for a in bufs:
            a.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=sorted(CONFIGS), help="headline configuration")
    ap.add_argument("--legs", default="2,4", help="other configurations measured beside the headline ('' = none)")
    ap.add_argument("--engine", default="auto", choices=["auto", "plain", "persistent", "binned"])
    ap.add_argument("--mode", default="aligned", choices=["aligned", "exact"])
    ap.add_argument("--graph", action="store_true", help="replay the batch pipeline as a captured HIP graph")
    ap.add_argument("--no-compare", action="store_true", help="skip the headline's engine comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget per config (0 = skip)")
    ap.add_argument("--serve-loads", default="0.1,0.3,0.5,0.7,0.9",
                    help="serving load levels (fractions of the mix's bulk rate); '' = skip")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-buffers", type=int, default=3)
    ap.add_argument("--serve-batch-cells", type=float, default=0, help="serving batch cap on cells (0 = library default)")
    ap.add_argument("--serve-engine", default="auto", choices=["auto", "plain", "persistent", "binned"])
    ap.add_argument("--serve-no-long-launch", action="store_true",
                    help="serving batches keep long reads in the queue launch (no long-read launch)")
    ap.add_argument("--serve-mode", default="both", choices=["both", "resident", "batches"],
                    help="serving by continuous batching over the buffers, by one resident kernel, or both "
                         "(batching is the reported 'serving' line, the resident kernel 'serving_resident')")
    ap.add_argument("--variant", default=None, help="force a kernel variant for every read length, e.g. L2xW16")
    ap.add_argument("--threshold", type=float, default=0.2,
                    help="also time threshold mode at this error rate on every measured config (negative = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (candidates would shard by read "
              "across GPUs that each hold a copy of the reference). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    cfg = CONFIGS[args.config]
    legs = [int(x) for x in args.legs.split(",") if x.strip() and int(x) != args.config]
    legs = [c for c in legs if CONFIGS[c]["ref_seed"] == cfg["ref_seed"] and CONFIGS[c]["ref_size"] == cfg["ref_size"]]
    t0 = time.time()
    ref = M.synth_reference(cfg["ref_seed"], cfg["ref_size"])
    log(f"reference {cfg['ref_size']} bases in {time.time() - t0:.1f}s")

    engines = {"auto": M.ENGINE_AUTO, "plain": M.ENGINE_PLAIN, "persistent": M.ENGINE_PERSISTENT,
               "binned": M.ENGINE_BINNED}
    serve_loads = [float(x) for x in args.serve_loads.split(",") if x.strip()]
    num_buffers = args.serve_buffers if serve_loads else 1
    # buffers sized for the largest batch measured (config 3: 11M candidates)
    sizes = {c: CONFIGS[c]["reads"] * 12 for c in [args.config] + legs}
    max_cands = max(sizes.values())
    m = M.Myers(ref, cfg["ref_size"], num_buffers=num_buffers, avg_query=550, cands_per_query=11,
                buffer_bytes=max(512 << 20, int(max_cands * 130)))
    if args.variant:
        lanes, words = (int(x) for x in args.variant.lstrip("L").split("xW"))
        v = m.variant_index(lanes, words)
        table = m.bin_table()
        w = np.arange(len(table))
        table[(w >= 1) & (w <= lanes * words)] = v
        m.set_bin_table(table)

    head = measure_config(m, ref, args.config, args.steps, args.warmup, engines[args.engine], args.mode, args.graph,
                          not args.no_compare, args.cpu_seconds,
                          args.threshold)
    leg_out = {}
    for c in legs:
        steps = args.steps if c == 2 else max(3, args.steps // 4)
        leg_out[CONFIGS[c]["name"]] = measure_config(
            m, ref, c, steps, max(1, args.warmup // 2), engines[args.engine], args.mode, args.graph,
            c == 4 and not args.no_compare, args.cpu_seconds / 2 if c != 4 else args.cpu_seconds / 2,
            args.threshold)

    serving = serving_resident = None
    if serve_loads and args.serve_mode in ("both", "batches"):
        log("serving (continuous batching) ...")
        serving = measure_serving(m, ref, cfg["ref_size"], serve_loads, args.serve_seconds,
                                  int(args.serve_batch_cells), engines[args.serve_engine], resident=False,
                                  no_long_launch=args.serve_no_long_launch)
    if serve_loads and args.serve_mode in ("both", "resident"):
        log("serving (resident kernel) ...")
        serving_resident = measure_serving(m, ref, cfg["ref_size"], serve_loads, args.serve_seconds,
                                           int(args.serve_batch_cells), engines[args.serve_engine], resident=True)

    out = {
        "metric": "candidates/s & GCUPS on 1 MI355X (bit-exact); p99 latency under serving load",
        "value": head["candidates_per_s"],
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generator, see include/myers-tools.h)",
        "gcups": head["gcups"],
        "config": {
            "workload": head["workload"],
            "reads": head["reads"], "candidates": head["candidates"], "cells": head["cells"],
            "reference_bases": cfg["ref_size"], "engine": args.engine, "graph": args.graph,
            "variants": head["variants"], "launches_per_step": head["launches_per_step"], "binned": head["binned"],
            "pipeline_ms": head["pipeline_ms"], "verify_ms": head["verify_ms"],
        },
        "roofline": head["roofline"],
        "engine_comparison": head.get("engine_comparison"),
        "threshold_mode": {name: x["threshold_mode"] for name, x in
                           [(cfg["name"], head)] + list(leg_out.items()) if "threshold_mode" in x} or None,
        "cpu_baseline": head.get("cpu_baseline"),
        "legs": leg_out,
        "serving": serving,
        "serving_resident": serving_resident,
    }
    print(json.dumps(out), flush=True)
    m.close()


if __name__ == "__main__":
    main()
