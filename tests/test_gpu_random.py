"""Seeded random batches through every engine, both window modes and a range
of thresholds, each equal to the oracle (its full results filtered by the
threshold in threshold mode): mixed lengths over one-lane, multi-lane and
long-read variants, Poisson candidate counts, error rates up to 30%."""
import numpy as np
import pytest

import myers_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZE = 4_000_000
LENGTHS = [64, 150, 256, 300, 420, 520, 640, 777, 900, 1000, 1024, 1500, 3000]
ENGINES = [M.ENGINE_AUTO, M.ENGINE_PLAIN, M.ENGINE_PERSISTENT, M.ENGINE_BINNED]
RATES = [-1.0, 0.02, 0.08, 0.15, 0.2, 0.3]


def test_random_batches():
    rng = np.random.default_rng(20261203)
    ref = M.synth_reference(91, SIZE)
    with M.Myers(ref, SIZE, avg_query=800, cands_per_query=11, buffer_bytes=256 << 20) as m:
        for it in range(150):
            nq = int(rng.integers(20, 700))
            choices = sorted(set(int(x) for x in rng.choice(LENGTHS, size=int(rng.integers(1, 5)))))
            spec = M.make_spec(int(rng.integers(1 << 30)), nq, choices=choices,
                               err_range=(0.0, float(rng.uniform(0.02, 0.3))), sub_frac=0.7,
                               poisson=float(rng.uniform(2, 14)), cap=40, distance=0.1)
            b = M.synth_batch(spec, ref, SIZE)
            mode = int(rng.choice([M.WINDOW_ALIGNED, M.WINDOW_EXACT]))
            engine = int(rng.choice(ENGINES))
            e = float(rng.choice(RATES))
            if len(b["cands"]) == 0:
                continue
            print(f"[gpu-case] random batch {it}: reads={nq} lengths={choices} candidates={len(b['cands'])} "
                  f"mode={mode} engine={engine} e={e}", flush=True)
            m.set_threshold(e)
            got = m.verify(b["peq"], b["qinfo"], b["cands"], 0.1, flags=mode | engine)
            full = O.myers_batch(ref, SIZE, b["peq"], b["qinfo"], b["cands"], 0.1, mode)
            want = full if e < 0 else O.apply_threshold(b["qinfo"], b["cands"], full, e)
            np.testing.assert_array_equal(got, want, err_msg=f"random batch {it}")
        m.set_threshold(-1.0)
