"""CPU model of the multi-wave kernels' schedule (verify_wide in
myers_kernels.hip): lane d of a candidate computes column t - d at step t,
wave q runs 8 steps behind wave q-1 on top of that, a lane's hin is the top
Ph/Mh bit of the lane below from the previous step, and at a wave boundary it
is the top lane of the wave below from 9 steps earlier (passed through the
LDS ring).  The model must give the plain DP's (score, column); a wrong
hand-over delay must not."""
import numpy as np
import pytest

import oracle_lib as O

M32 = 0xFFFFFFFF


def run_schedule(qcodes, text, Q, W, delay=9):
    m = len(qcodes)
    L = 64 * Q
    pad = 32 * L * W - m
    eq = np.zeros((6, L, W), dtype=np.uint64)
    for i, c in enumerate(qcodes):
        r = pad + i
        eq[min(int(c), 4), (r // 32) // W, (r // 32) % W] |= 1 << (r % 32)
    tid = np.arange(L)
    lag = tid + 8 * (tid // 64)
    max_lag = L - 1 + 8 * (Q - 1)
    n = len(text)
    Pv = np.full((L, W), M32, dtype=np.uint64)
    Mv = np.zeros((L, W), dtype=np.uint64)
    ph_top = np.zeros(L, dtype=np.uint64)
    mh_top = np.zeros(L, dtype=np.uint64)
    top_bits = {}
    score = best = m
    best_col = 0
    text = np.asarray(text)
    for t in range(n + max_lag):
        col = t - lag
        codes = np.where(col < 0, 5, text[np.clip(col, 0, n - 1)])
        ph_in = np.zeros(L, dtype=np.uint64)
        mh_in = np.zeros(L, dtype=np.uint64)
        ph_in[1:], mh_in[1:] = ph_top[:-1], mh_top[:-1]
        for q in range(Q):
            p, mm = top_bits.get((q - 1, t - delay), (0, 0)) if q else (0, 0)
            ph_in[64 * q], mh_in[64 * q] = p << 31, mm << 31
        carry = np.zeros(L, dtype=np.uint64)
        ph_prev, mh_prev = ph_in, mh_in
        for k in range(W):
            e = eq[codes, tid, k]
            pv, mv = Pv[:, k], Mv[:, k]
            xv = e | mv
            eqh = (e | (mh_in >> 31)) if k == 0 else e
            s = (eqh & pv) + pv + carry
            carry, s = s >> 32, s & M32
            xh = (s ^ pv) | eqh
            ph = (mv | (~(xh | pv) & M32)) & M32
            mh = pv & xh
            php = ((ph << 1) & M32) | (ph_prev >> 31)
            mhp = ((mh << 1) & M32) | (mh_prev >> 31)
            ph_prev, mh_prev = ph, mh
            Pv[:, k] = (mhp | (~(xv | php) & M32)) & M32
            Mv[:, k] = php & xv
        ph_top, mh_top = ph_prev, mh_prev
        for q in range(Q):
            top_bits[(q, t)] = (int(ph_top[64 * q + 63] >> 31), int(mh_top[64 * q + 63] >> 31))
        score += int(ph_top[L - 1] >> 31) - int(mh_top[L - 1] >> 31)
        c = t - max_lag
        if 0 <= c < n and score < best:
            best, best_col = score, c
    return best, best_col


@pytest.mark.parametrize("Q,W,m", [(2, 1, 2500), (2, 1, 4096), (3, 1, 5000)])
def test_multiwave_schedule_equals_dp(Q, W, m):
    rng = np.random.default_rng(m)
    q = rng.integers(0, 4, m).astype(np.uint8)
    body = q.copy()
    flip = rng.random(m) < 0.05
    body[flip] = rng.integers(0, 4, int(flip.sum()))
    text = np.concatenate([rng.integers(0, 4, 150), body, rng.integers(0, 4, 150)]).astype(np.uint8)
    _, (score, col) = O.dp_trace(q, text)
    assert run_schedule(q, text, Q, W) == (score, col)
    assert run_schedule(q, text, Q, W, delay=8) != (score, col)
