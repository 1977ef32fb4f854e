// Candidate-verification kernels for MI355X (gfx950, wave64).
//
// Each candidate is verified by the Myers/Hyyro bit-vector recurrence over
// its reference window (reference kernel: src-api/coreMyersGPU-CC35.cu:110-240;
// CPU form: src/myers-4bits.c:111-135).  All work is integer VALU (no MFMA).
//
//   * A kernel variant <L, W> gives each candidate L lanes of a wave64 and
//     each lane W u32 words of the query bit-vector (32*L*W rows in all).
//     Inside a lane the W words form one full-width vector (add carries and
//     Ph/Mh shifts run through all of them).  Between the lanes of a group
//     the block form of the recurrence applies (src/myers-4bits.c:115-134):
//     a lane needs only the horizontal delta hin at its bottom boundary,
//     which is the top bit of the Ph/Mh of the lane below in the same
//     column.  The lanes run diagonally: lane b of a group computes column
//     t - b at step t, so its hin is the lower lane's output of step t - 1,
//     moved up one lane by a single DPP shift.  No carry crosses lanes, no
//     ballot, no branch.  A lane before its first column reads an all-zero
//     Peq row; with Eq = 0 and hin = 0 the initial state Pv = ~0, Mv = 0 is
//     a fixed point emitting no delta, so late-starting lanes need no reset.
//     L = 1 is the special case with hin = 0 and no lag.
//   * Query bit-vectors are top-justified: a query of m bases occupies the
//     top m rows and the rows below it never match (their Peq bits are 0).
//     Such rows stay at the same fixed point, so they emit no horizontal
//     delta and the query rows see the semi-global boundary D[0][j]=0.  The
//     score row m-1 is therefore bit 31 of the top lane's top word for every
//     m the variant covers.
//   * Peq masks are staged into LDS once per distinct read of a wave, as
//     [ceil(W/4) groups][read slot][rows][L lanes] x 128-bit (rows = 5 bases,
//     plus a zero row for L > 1), so one column's Eq words are ceil(W/4)
//     ds_read_b128 at an address chosen by the column's base code, and the
//     (slot, base) pairs of a wave fall on consecutive 16-byte bank groups.
//   * The packed reference (8 bases per u32) streams through a register
//     window per lane, two 8-column blocks ahead of use; exact-start windows
//     and the diagonal lag use a per-lane funnel shift.
//   * Per word and column the recurrence is 10 VALU ops: and, add-with-carry,
//     3 three-input logic ops, 2 funnel shifts, or, and, and.  and/or/bitop3
//     issue at twice the rate of the funnel shifts and the carry add on
//     gfx950, so variants are sized for at least two waves per SIMD.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "myers_kernels.h"

namespace {

struct CandRec {
    uint32_t query;
    uint32_t pad;
    uint64_t position;
};

constexpr uint32_t kSkipScore = 0xFFFFFFFFu;
constexpr uint32_t kNoMatch = 0xFFFFFFFFu;  // threshold mode: {column, score} of a candidate above it
constexpr uint32_t kZeroRow = 5u;          // LDS row of all-zero Peq words (L > 1)

// n = (uint32)(m * (1 + 2*d)) in float32 with round-to-nearest steps
// (src-api/coreMyersGPU-CC35.cu:131).
__device__ __forceinline__ uint32_t window_len(uint32_t m, float d) {
    float f = __fmul_rn((float)m, __fadd_rn(1.0f, __fmul_rn(2.0f, d)));
    return (uint32_t)f;
}

// Rows 5..7 of the Peq do not exist; the reference kernel's selectEq maps
// them to row 0 (src-api/coreMyersGPU-CC35.cu:82-95).  Applied to all eight
// nibbles of a text word at once: clear bit 3, then zero codes 5..7.
__device__ __forceinline__ uint32_t sanitize_text(uint32_t w) {
    uint32_t c = w & 0x77777777u;
    uint32_t hi = (c >> 2) & 0x11111111u;
    uint32_t lo = (c | (c >> 1)) & 0x11111111u;
    return c & ~((hi & lo) * 7u);
}

// Under-threshold bound k = (uint32)((float)m * e) (src/checksum-2.c:88).
__device__ __forceinline__ uint32_t threshold_of(uint32_t m, float e) {
    return (uint32_t)__fmul_rn((float)m, e);
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

template <int L, int W>
struct Shape {
    static constexpr int kGroups = (W + 3) / 4;        // 128-bit LDS groups per lane
    static constexpr int kTotal = L * W;                // words per candidate
    static constexpr int kSlots = 64 / L;               // candidates per wave
    static constexpr int kRows = L > 1 ? 6 : 5;         // A C G T N (+ zero row)
    // Peq is staged once per distinct query of the wave ("query slot"), not
    // once per candidate; up to kQSlots queries per pass.  LDS layout
    // [group][slot][row][lane] x 128 bit: the uint4s one column reads for
    // the (slot, base) pairs of a wave are consecutive in slot*rows + base,
    // so they spread evenly over the LDS banks.
    static constexpr int kQSlotsMax = L == 1 ? 16 : 8;
    static constexpr int kQSlots = kSlots < kQSlotsMax ? kSlots : kQSlotsMax;
    static constexpr int kRowU4 = L;                    // uint4 between rows of a slot
    static constexpr int kSlotU4 = kRows * L;           // uint4 between slots
    static constexpr int kGroupU4 = kQSlots * kSlotU4;  // uint4 between groups
    static constexpr int kLdsU4 = kGroups * kGroupU4;
};

// Loads and stores of a batch's arrays.  kHost: the arrays are host memory
// rewritten while a resident kernel runs (the serving arena), which the GPU
// may hold in its L2; system-scope accesses go around that cache, so a load
// never sees a line cached for an earlier request and a store reaches the
// host once it completes.  Otherwise plain (cached) accesses.
template <bool kHost>
__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
    if constexpr (kHost) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else return *p;
}
template <bool kHost>
__device__ __forceinline__ uint64_t ld64(const uint64_t *p) {
    if constexpr (kHost) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else return *p;
}
template <bool kHost>
__device__ __forceinline__ void st_result(uint2 *p, uint2 v) {
    if constexpr (kHost) {
        __hip_atomic_store(reinterpret_cast<uint64_t *>(p), ((uint64_t)v.y << 32) | v.x, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
        *p = v;
    }
}

// Stage the Peq words of the wave's query slots into LDS, top-justified,
// with all 64 lanes sharing the work.  Word gw (0 <= gw < L*W) of a slot
// holds virtual rows [32gw, 32gw+32); virtual row r is query row r - pad,
// pad = 32*L*W - m, so it is the funnel of query words lo and lo+1
// (lo = floor((32gw - pad)/32)) shifted right by (-pad) mod 32.  Query word
// j (rows 32j..32j+31) of base c is bitmap[c][j%4] of entry posEntry + j/4;
// words outside the query read 0.  Lane gl = gw / W of the group reads word
// k = gw % W at LDS uint4 (k/4)*kGroupU4 + slot*kSlotU4 + c*L + gl, lane k%4.
template <int L, int W, bool kHost = false, uint32_t kThreads = 64>
__device__ __forceinline__ void stage_peq(uint4 *lds, const uint2 *qtab, uint32_t nslots,
                                          const uint32_t *__restrict__ peq, uint32_t lane) {
    using S = Shape<L, W>;
    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
    const uint32_t total = nslots * 5u * (uint32_t)S::kTotal;
    for (uint32_t idx = lane; idx < total; idx += kThreads) {
        const uint32_t gw = idx % (uint32_t)S::kTotal;
        const uint32_t rest = idx / (uint32_t)S::kTotal;
        const uint32_t c = rest % 5u;
        const uint32_t slot = rest / 5u;
        const uint2 qi = qtab[slot];                   // {posEntry, m}
        const int32_t pad = 32 * S::kTotal - (int32_t)qi.y;
        const int32_t nq = (int32_t)((qi.y + 31u) >> 5);
        const int32_t b0 = 32 * (int32_t)gw - pad;
        const int32_t lo = b0 >> 5;                    // floor division
        const uint32_t sh = (uint32_t)b0 & 31u;
        const uint32_t *base = peq + (size_t)qi.x * 20u + 4u * c;
        uint32_t wlo = 0, whi = 0;
        if (lo >= 0 && lo < nq) wlo = ld32<kHost>(base + (uint32_t)(lo >> 2) * 20u + (uint32_t)(lo & 3));
        if (lo + 1 >= 0 && lo + 1 < nq)
            whi = ld32<kHost>(base + (uint32_t)((lo + 1) >> 2) * 20u + (uint32_t)((lo + 1) & 3));
        const uint32_t k = gw % (uint32_t)W, gl = gw / (uint32_t)W;
        const uint32_t u4 = (k >> 2) * (uint32_t)S::kGroupU4 + slot * (uint32_t)S::kSlotU4 + c * (uint32_t)L + gl;
        l32[u4 * 4u + (k & 3u)] = (uint32_t)((((uint64_t)whi << 32) | wlo) >> sh);
    }
}

// Truth tables of v_bitop3_b32(a, b, c): bit (4a+2b+c) of the immediate is
// f(a, b, c); with a = 0xF0, b = 0xCC, c = 0xAA the table is f applied
// bitwise to those constants.
constexpr uint32_t kXorOr = (0xF0u ^ 0xCCu) | 0xAAu;          // (a ^ b) | c
constexpr uint32_t kOrNor = (0xF0u | ~(0xCCu | 0xAAu)) & 0xFFu;  // a | ~(b | c)

// The lower lane's value of the previous step: DPP row_shr:1 (within rows
// of 16) or wave_shr:1, with 0 on the first lane of every group.
template <int L>
__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v, uint32_t not_first) {
    if constexpr (L > 16) {
        v = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);   // wave_shr:1
    } else {
        v = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);   // row_shr:1
    }
    if constexpr (L != 16 && L != 64) v &= not_first;   // rows and the wave already start at 0
    return v;
}

// One group of four words (g) of a one-lane column (L = 1), for the cut-off
// column below: the recurrence over words 4g..4g+3 with the Eq words e
// (src/myers-4bits.c:115-134 block form), in the phased order of column_grp.
// carry is the add's carry between words; ph_prev / mh_prev the Ph / Mh
// (before the shift) of the word below, whose top bit shifts into this
// group's first word.
template <int W>
__device__ __forceinline__ void group_update(int g, uint32_t (&Pv)[W], uint32_t (&Mv)[W], const uint4 &e,
                                             uint32_t &carry, uint32_t &ph_prev, uint32_t &mh_prev) {
    const uint32_t e4[4] = {e.x, e.y, e.z, e.w};
    const int k0 = 4 * g;
    const int n = (W - k0) < 4 ? (W - k0) : 4;
    uint32_t t[4], xv[4], sum[4], xh[4], ph[4], mh[4], php[4], mhp[4];
#pragma unroll
    for (int j = 0; j < 4 && j < n; j++) {
        t[j] = e4[j] & Pv[k0 + j];
        xv[j] = e4[j] | Mv[k0 + j];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 4 && j < n; j++) {
        uint32_t c_out;
        sum[j] = __builtin_addc(t[j], Pv[k0 + j], carry, &c_out);
        carry = c_out;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 4 && j < n; j++) xh[j] = __builtin_amdgcn_bitop3_b32(sum[j], Pv[k0 + j], e4[j], kXorOr);
#pragma unroll
    for (int j = 0; j < 4 && j < n; j++) {
        ph[j] = __builtin_amdgcn_bitop3_b32(Mv[k0 + j], xh[j], Pv[k0 + j], kOrNor);
        mh[j] = Pv[k0 + j] & xh[j];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 4 && j < n; j++) {
        const int k = k0 + j;
        php[j] = k == 0 ? (ph[j] << 1) : __builtin_amdgcn_alignbit(ph[j], ph_prev, 31u);
        mhp[j] = k == 0 ? (mh[j] << 1) : __builtin_amdgcn_alignbit(mh[j], mh_prev, 31u);
        ph_prev = ph[j];
        mh_prev = mh[j];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 4 && j < n; j++) {
        Pv[k0 + j] = __builtin_amdgcn_bitop3_b32(mhp[j], xv[j], php[j], kOrNor);
        Mv[k0 + j] = php[j] & xv[j];
    }
    __builtin_amdgcn_sched_barrier(0);
}

// One column for one lane's W words.  Eq words stream from LDS one group of
// four words ahead of use (group g+1 is read while group g computes; the
// caller reads the next column's group 0 at the start of the column), so 12
// Eq registers are live instead of 2W.  For L > 1, ph_in / mh_in carry the
// lower lane's hin in bit 31 (src/myers-4bits.c:115-134 block form).  The
// top word's Ph / Mh (before the shift) come back in ph_top / mh_top.
//
// kPhased: issue each group's word updates in phases (below).  With several
// waves per SIMD this raises the share of quad-cycles that issue two VALU
// instructions (18% -> 35% on 1000-base reads, 11% faster); a wave that runs
// a long chain alone on its SIMD (the long-read launch, multi-wave
// candidates) is 5-10% faster with the compiler's interleaved order, which
// overlaps the carry chain with the other words' operations.  The phased
// body is also group_update's; calling group_update here measured 2% slower
// (the compiler sinks Xv into the last phase), and pinning Xv in the first
// phase makes the cut-off column spill.
template <int W, bool kHin, uint32_t kGS, bool kPhased = true>
__device__ __forceinline__ void column_grp(uint32_t (&Pv)[W], uint32_t (&Mv)[W], const uint4 &e0,
                                           const uint4 *lds_lane, uint32_t cur_off, uint32_t ph_in,
                                           uint32_t mh_in, uint32_t &ph_top, uint32_t &mh_top) {
    constexpr int kG = (W + 3) / 4;     // kGS: uint4 between groups in LDS
    uint32_t carry = 0, ph_prev = ph_in, mh_prev = mh_in;
    uint4 cur = e0, nxt = e0;
#pragma unroll
    for (int g = 0; g < kG; g++) {
        {
            if (g + 1 < kG) nxt = lds_lane[cur_off + (uint32_t)(g + 1) * kGS];
            const uint32_t e4[4] = {cur.x, cur.y, cur.z, cur.w};
            if constexpr (kPhased) {
                const int k0 = 4 * g;
                const int n = (W - k0) < 4 ? (W - k0) : 4;
                uint32_t eqv[4], t[4], xv[4], sum[4], xh[4], ph[4], mh[4], php[4], mhp[4];
                // The group's words go through the recurrence in phases, each
                // phase over all of its words, so that the double-rate bit
                // operations of independent words issue back to back (a SIMD
                // issues two of them in one quad-cycle) and the single-rate ones
                // (add with carry, funnel shifts) come in runs:
                //   Eq & Pv, Xv | carry chain | Xh, Ph, Mh | shifts | Pv, Mv.
                // sched_barrier keeps the compiler from interleaving the phases.
    #pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (j >= n) break;
                    const int k = k0 + j;
                    eqv[j] = (kHin && k == 0) ? (e4[0] | (mh_in >> 31)) : e4[j];
                    t[j] = eqv[j] & Pv[k];
                    xv[j] = e4[j] | Mv[k];
                }
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (j >= n) break;
                    uint32_t c_out;
                    sum[j] = __builtin_addc(t[j], Pv[k0 + j], carry, &c_out);
                    carry = c_out;
                }
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int j = 0; j < 4 && j < n; j++) xh[j] = __builtin_amdgcn_bitop3_b32(sum[j], Pv[k0 + j], eqv[j], kXorOr);
    #pragma unroll
                for (int j = 0; j < 4 && j < n; j++) {
                    ph[j] = __builtin_amdgcn_bitop3_b32(Mv[k0 + j], xh[j], Pv[k0 + j], kOrNor);
                    mh[j] = Pv[k0 + j] & xh[j];
                }
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int j = 0; j < 4 && j < n; j++) {
                    const int k = k0 + j;
                    php[j] = (!kHin && k == 0) ? (ph[j] << 1) : __builtin_amdgcn_alignbit(ph[j], ph_prev, 31u);
                    mhp[j] = (!kHin && k == 0) ? (mh[j] << 1) : __builtin_amdgcn_alignbit(mh[j], mh_prev, 31u);
                    ph_prev = ph[j];
                    mh_prev = mh[j];
                }
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int j = 0; j < 4 && j < n; j++) {
                    Pv[k0 + j] = __builtin_amdgcn_bitop3_b32(mhp[j], xv[j], php[j], kOrNor);
                    Mv[k0 + j] = php[j] & xv[j];
                }
                __builtin_amdgcn_sched_barrier(0);
            } else {
#pragma unroll
                for (int k = 4 * g; k < 4 * g + 4 && k < W; k++) {
                    const uint32_t pv = Pv[k], mv = Mv[k];
                    const uint32_t xv = e4[k - 4 * g] | mv;
                    const uint32_t eq = (kHin && k == 0) ? (e4[0] | (mh_in >> 31)) : e4[k - 4 * g];
                    uint32_t c_out;
                    const uint32_t sum = __builtin_addc(eq & pv, pv, carry, &c_out);
                    carry = c_out;
                    const uint32_t xh = __builtin_amdgcn_bitop3_b32(sum, pv, eq, kXorOr);
                    const uint32_t ph = __builtin_amdgcn_bitop3_b32(mv, xh, pv, kOrNor);
                    const uint32_t mh = pv & xh;
                    const uint32_t php = (!kHin && k == 0) ? (ph << 1) : __builtin_amdgcn_alignbit(ph, ph_prev, 31u);
                    const uint32_t mhp = (!kHin && k == 0) ? (mh << 1) : __builtin_amdgcn_alignbit(mh, mh_prev, 31u);
                    ph_prev = ph;
                    mh_prev = mh;
                    Pv[k] = __builtin_amdgcn_bitop3_b32(mhp, xv, php, kOrNor);
                    Mv[k] = php & xv;
                }
            }
            cur = nxt;
        }
    }
    ph_top = ph_prev;
    mh_top = mh_prev;
}

// The same column (L = 1) with the Ukkonen cut-off of threshold mode (Myers
// 1999, blocks model; GEMTools gt_map_align_bpm.c:141-214): only groups
// 0..ga-1 are advanced, ga being the same in every lane of the wave.  Group
// G+1 runs inside group G's (scalar) branch, so a skipped group costs a
// compare and a branch, the carry stays in its scalar register, and no
// value needs merging between groups; ph_top / mh_top are those of group
// ga-1's top word.
template <int G, int W, uint32_t kGS>
__device__ __forceinline__ void column_cut(uint32_t (&Pv)[W], uint32_t (&Mv)[W], const uint4 &cur,
                                           const uint4 *lds_lane, uint32_t cur_off, uint32_t ga,
                                           uint32_t carry, uint32_t ph_prev, uint32_t mh_prev,
                                           uint32_t &ph_top, uint32_t &mh_top) {
    constexpr int kG = (W + 3) / 4;
    uint4 nxt = cur;
    if constexpr (G + 1 < kG) nxt = lds_lane[cur_off + (uint32_t)(G + 1) * kGS];
    group_update<W>(G, Pv, Mv, cur, carry, ph_prev, mh_prev);
    if constexpr (G + 1 < kG) {
        if ((uint32_t)(G + 1) < ga) {
            column_cut<G + 1, W, kGS>(Pv, Mv, nxt, lds_lane, cur_off, ga, carry, ph_prev, mh_prev, ph_top, mh_top);
            return;
        }
    }
    ph_top = ph_prev;
    mh_top = mh_prev;
}

template <int L, int W>
struct Workspace {
    uint4 lds[Shape<L, W>::kLdsU4];
    uint2 qtab[Shape<L, W>::kQSlots];
};

// LDS hand-over between the lanes of one wave when a workgroup holds several
// independent waves: a wave's LDS accesses execute in order, so it is enough
// to wait for them and keep the compiler from moving accesses across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Workgroups of kWaves independent waves (1: the workgroup is the wave).
template <int kWaves>
__device__ __forceinline__ void chunk_sync() {
    if constexpr (kWaves == 1) {
        __syncthreads();
    } else {
        wave_lds_sync();
    }
}

template <int kWaves>
__device__ __forceinline__ uint32_t wave_lane() {
    return kWaves == 1 ? threadIdx.x : (threadIdx.x & 63u);
}

// kThreads: the threads sharing the workspace (64, or 128 for a workgroup of
// two waves that stage together).
template <int L, int W, int kWaves = 1, uint32_t kThreads = 64>
__device__ __forceinline__ void zero_row_init(Workspace<L, W> &ws) {
    using S = Shape<L, W>;
    if constexpr (S::kRows > 5) {
        // the all-zero row read by lanes before their first column
        uint32_t *l32 = reinterpret_cast<uint32_t *>(ws.lds);
        constexpr uint32_t kPerGroup = (uint32_t)(S::kQSlots * L);    // (slot, lane) pairs
        const uint32_t t0 = kThreads == 64u ? wave_lane<kWaves>() : threadIdx.x;
        for (uint32_t i = t0; i < (uint32_t)S::kGroups * kPerGroup * 4u; i += kThreads) {
            const uint32_t u4 = i >> 2, g = u4 / kPerGroup, r = u4 % kPerGroup;
            const uint32_t slot = r / (uint32_t)L, gl = r % (uint32_t)L;
            l32[(g * (uint32_t)S::kGroupU4 + slot * (uint32_t)S::kSlotU4 + kZeroRow * (uint32_t)L + gl) * 4u +
                (i & 3u)] = 0u;
        }
    }
}

// Verify chunks chunk_lo, chunk_lo + stride, ... < chunk_hi of the slot range
// [start, start + count); a chunk is the kSlots candidates of one wave.
// kCut: threshold mode.  L = 1 (W >= 8): the Ukkonen cut-off; L > 1: a
// wave stops once its candidates cannot get under the threshold.  kHost: the
// batch's arrays are host memory of the resident server (ld32 / st_result).
// kShare = 2: a workgroup of two waves verifies chunk pairs (2c, 2c+1) with
// one workspace: the distinct reads of both waves' candidates share its
// query slots (one pass while they number at most kQSlots), so each wave
// needs half the LDS of a wave of its own.  xchg: three words of LDS for the
// hand-over between the waves (the last read of wave 0, each wave's count of
// reads).
template <int L, int W, bool kCut, int kWaves = 1, bool kHost = false, int kShare = 1>
__device__ __forceinline__ void verify_chunks(const myers_verify_args_t &a, uint32_t start, uint32_t count,
                                              uint32_t chunk_lo, uint32_t chunk_hi, uint32_t stride,
                                              Workspace<L, W> &ws, uint32_t *xchg = nullptr) {
    static_assert(kShare == 1 || (kShare == 2 && kWaves == 1 && !kHost), "shared workspace: one-wave chunks");
    using S = Shape<L, W>;
    constexpr bool kUkk = kCut && L == 1;       // the cut-off
    constexpr bool kStop = kCut && L > 1;       // the early stop
    uint4 *lds = ws.lds;
    uint2 *qtab = ws.qtab;
    const uint32_t lane = kShare == 1 ? wave_lane<kWaves>() : (threadIdx.x & 63u);
    const uint32_t wave = kShare == 1 ? 0u : (threadIdx.x >> 6);
    const uint32_t gl = lane % L;                      // lane within the candidate's group
    const uint32_t *__restrict__ ref = a.ref;
    const int32_t last_word = (int32_t)(a.ref_words - 1);
    const CandRec *__restrict__ cands = static_cast<const CandRec *>(a.cands);
    const uint2 *__restrict__ qinfo = reinterpret_cast<const uint2 *>(a.qinfo);
    uint2 *__restrict__ results = static_cast<uint2 *>(a.results);
    const uint32_t not_first = gl != 0u ? ~0u : 0u;
    const uint64_t lanes_upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);

    for (uint32_t chunk = chunk_lo; chunk < chunk_hi; chunk += stride) {
        const uint32_t slot = (chunk * (uint32_t)kShare + wave) * S::kSlots + lane / L;
        const bool valid = slot < count;
        uint32_t cand = 0, m = 0, pos_entry = 0, query = 0xFFFFFFFFu;
        uint64_t pos = 0;
        bool sane = false;
        if (valid) {
            const uint32_t s = start + slot;
            cand = a.order ? ld32<kHost>(a.order + s) : s;
            if constexpr (kHost) {
                pos = ld64<true>(&cands[cand].position);
                query = ld32<true>(&cands[cand].query);
            } else {
                const CandRec cr = cands[cand];
                pos = cr.position;
                query = cr.query;
            }
            if (query < a.num_queries) {
                const uint64_t qi = ld64<kHost>(reinterpret_cast<const uint64_t *>(qinfo + query));
                pos_entry = (uint32_t)qi;
                m = (uint32_t)(qi >> 32);
                // the query must fit this variant and its Peq entries the batch
                sane = m <= 32u * S::kTotal &&
                       (uint64_t)pos_entry + ((m + 127u) >> 7) <= (uint64_t)a.num_peq_entries;
            }
        }
        if (!sane) m = 0;
        const uint32_t n = window_len(m, a.distance);
        // skip rule: src-api/coreMyersGPU-CC35.cu:142
        const bool live = valid && sane && pos < a.ref_size && (a.ref_size - pos) > (uint64_t)n;
        uint64_t nib0 = 0;
        uint32_t ncols_live = 0;
        if (live) {
            if (a.mode == 0) {
                // aligned window: src-api/coreMyersGPU-CC35.cu:132,144
                nib0 = pos & ~7ull;
                ncols_live = 8u * ((n >> 3) + ((n & 7u) ? 2u : 1u));
            } else {
                nib0 = pos;
                ncols_live = n;
            }
        }

        // Query slots: consecutive candidates of one query share a slot (the
        // callers' batches and the binning keep a query's candidates adjacent).
        uint32_t query_below = (uint32_t)__shfl_up((int)query, L);
        if constexpr (kShare == 2) {
            // wave 1's first candidate continues wave 0's last read or heads a new one
            if (wave == 0u && lane == 63u) xchg[0] = query;
            __syncthreads();
            if (wave == 1u && lane < (uint32_t)L) query_below = xchg[0];
        }
        const bool first = kShare == 1 ? lane < (uint32_t)L : (wave == 0u && lane < (uint32_t)L);
        const bool head = (first || query_below != query) && gl == 0u;
        const uint64_t heads = __ballot(head);
        uint32_t slot_base = 0u, nqslots = (uint32_t)__builtin_popcountll(heads);
        if constexpr (kShare == 2) {
            if (lane == 0u) xchg[1u + wave] = nqslots;
            __syncthreads();
            const uint32_t h0 = xchg[1], h1 = xchg[2];
            slot_base = wave == 0u ? 0u : h0;
            nqslots = h0 + h1;
        }
        const uint32_t qslot = slot_base + (uint32_t)__builtin_popcountll(heads & lanes_upto) - 1u;

        // More distinct queries than LDS slots: one pass per kQSlots queries.
        for (uint32_t base_slot = 0; base_slot < nqslots; base_slot += S::kQSlots) {
            const bool in_pass = qslot >= base_slot && qslot < base_slot + S::kQSlots;
            const uint32_t my_slot = in_pass ? qslot - base_slot : 0u;
            chunk_sync<kWaves>();                        // previous readers of qtab / lds are done
            if (in_pass && head) qtab[my_slot] = make_uint2(pos_entry, m);
            chunk_sync<kWaves>();
            stage_peq<L, W, kHost, 64u * kShare>(lds, qtab, min(nqslots - base_slot, (uint32_t)S::kQSlots), a.peq,
                                                 kShare == 1 ? lane : threadIdx.x);
            chunk_sync<kWaves>();
            const uint4 *lds_lane = lds + my_slot * (uint32_t)S::kSlotU4 + gl;
            const uint32_t ncols = in_pass ? ncols_live : 0u;
            // threshold of the group's top lane (-1: no early stop for this lane)
            int32_t kthr = -1;
            if constexpr (kStop) {
                if (a.threshold >= 0.0f && live && in_pass && gl == (uint32_t)(L - 1))
                    kthr = (int32_t)threshold_of(m, a.threshold);
            }

            // Steps: the top lane of a group finishes column ncols-1 at step
            // ncols-1 + L-1.  Lane gl reads text from nibble nib0 - gl.
            const uint32_t max_cols = wave_max(ncols);
            const uint32_t nblocks = (max_cols + (uint32_t)(L - 1) + 7u) >> 3;
            const int64_t q = (int64_t)nib0 - (int64_t)gl;
            const int32_t w0 = (int32_t)(q >> 3);            // floor
            const uint32_t shift = (uint32_t)(q & 7) * 4u;

            uint32_t Pv[W], Mv[W];
#pragma unroll
            for (int k = 0; k < W; k++) {
                Pv[k] = ~0u;
                Mv[k] = 0u;
            }
            int32_t score = (int32_t)m, min_score = (int32_t)m;
            uint32_t min_col = 0;
            uint32_t ph_top = 0, mh_top = 0;
            // Cut-off state: rows with D <= k (k + pad in layout rows, where
            // the pad rows below the query hold D = row) must lie in the
            // first ga groups, 8 rows below its top.  Groups hold 128 rows,
            // the top one 32 (W - 4 (kG - 1)).  Start with every lane's rows
            // 0..kv+8 covered; D at the top row R of group ga-1 is then R.
            constexpr uint32_t kG = (uint32_t)((W + 3) / 4);
            constexpr int32_t kTopRows = 32 * (W - 4 * ((W + 3) / 4 - 1));
            int32_t kv = 0;
            uint32_t ga = kG;
            if constexpr (kUkk) {
                kv = (int32_t)threshold_of(m, a.threshold) + (32 * W - (int32_t)m);
                ga = min(wave_max(ncols > 0u ? ((uint32_t)kv + 135u) / 128u : 1u), kG);
                score = ga == kG ? 32 * W : 128 * (int32_t)ga;
            }

            // Text words stream through registers two blocks (16 columns) ahead
            // of their use; the Eq words of step t+1 are read from LDS while
            // step t computes.  Reads outside the stored reference are clamped
            // to its first / last (all-N guard) word; they only feed steps that
            // no lane keeps.
            auto text_word = [&](int32_t k) { return ref[min(max(w0 + k, 0), last_word)]; };
            auto text_block = [&](uint32_t lo, uint32_t hi, uint32_t b) {
                uint32_t t = sanitize_text(__builtin_amdgcn_alignbit(hi, lo, shift));
                if constexpr (L > 1) {
                    // steps before this lane's first column read the zero row
                    const int32_t early = (int32_t)gl - 8 * (int32_t)b;     // such steps in this block
                    if (early > 0) {
                        const uint32_t low = early >= 8 ? ~0u : ((1u << (4 * early)) - 1u);
                        t = (t & ~low) | (0x55555555u & low);
                    }
                }
                return t;
            };
            const uint32_t tw0 = text_word(0);
            const uint32_t tw1 = text_word(1);
            uint32_t tw2 = text_word(2);
            uint32_t tw3 = text_word(3);
            uint32_t text = text_block(tw0, tw1, 0);
            uint32_t text_next = text_block(tw1, tw2, 1);
            uint32_t cur_off = (text & 7u) * S::kRowU4;
            uint4 eq4 = lds_lane[cur_off];
            for (uint32_t b = 0; b < nblocks; b++) {
                const uint32_t tw4 = text_word((int32_t)b + 4);
                if constexpr (kUkk) {
                    // The extent changes only between blocks of 8 columns.  A row
                    // with D <= k in column j + t lies at most t + 1 rows above
                    // the last such row of column j - 1, so the extent must keep
                    // 8 rows of headroom over it: rows R-7..R with D <= k force
                    // D(R) <= k + 7 (one group is added), and a group can go
                    // when its rows and the 8 below it exceed k: D(R) > k + 135
                    // (for the shorter top group a stricter test than needed).
                    const bool on = 8u * b < ncols;
                    if (ga < kG && __ballot(on && score <= kv + 7) != 0ull) {
                        // new group: D rises by one per row above R (Pv = ~0, Mv = 0)
#pragma unroll
                        for (int g = 1; g < (int)kG; g++) {
                            const bool sel = (uint32_t)g == ga;
#pragma unroll
                            for (int k = 4 * g; k < 4 * g + 4 && k < W; k++) {
                                Pv[k] = sel ? ~0u : Pv[k];
                                Mv[k] = sel ? 0u : Mv[k];
                            }
                        }
                        score += ga + 1u == kG ? kTopRows : 128;
                        ga++;
                    } else if (ga > 1u && __ballot(on && score <= kv + 135) == 0ull) {
                        // drop group ga-1: S moves down by its vertical deltas
#pragma unroll
                        for (int g = 1; g < (int)kG; g++) {
                            if ((uint32_t)g + 1u == ga) {
                                int32_t dv = 0;
#pragma unroll
                                for (int k = 4 * g; k < 4 * g + 4 && k < W; k++)
                                    dv += __builtin_popcount(Pv[k]) - __builtin_popcount(Mv[k]);
                                score -= dv;
                            }
                        }
                        ga--;
                    }
                }
                const uint32_t ga_s = __builtin_amdgcn_readfirstlane(ga);
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const uint32_t cn = (i < 7) ? ((text >> (4 * (i + 1))) & 7u) : (text_next & 7u);
                    const uint32_t next_off = cn * S::kRowU4;
                    if constexpr (L == 1) {
                        const uint4 eq4n = lds_lane[next_off];        // next column, group 0
                        if constexpr (kUkk) {
                            column_cut<0, W, (uint32_t)S::kGroupU4>(Pv, Mv, eq4, lds_lane, cur_off, ga_s, 0u, 0u,
                                                                   0u, ph_top, mh_top);
                        } else {
                            column_grp<W, false, (uint32_t)S::kGroupU4>(Pv, Mv, eq4, lds_lane, cur_off, 0u, 0u,
                                                                       ph_top, mh_top);
                        }
                        eq4 = eq4n;
                        score += (int32_t)(ph_top >> 31) - (int32_t)(mh_top >> 31);
                        const uint32_t col = 8u * b + (uint32_t)i;
                        const bool on = col < ncols;
                        if constexpr (!kUkk) {
                            // first minimum of the score row (src-api/coreMyersGPU-CC35.cu:205-207,227-228)
                            const bool better = (score < min_score) && on;
                            min_col = better ? col : min_col;
                            min_score = better ? score : min_score;
                        } else {
                            if (ga == kG) {
                                // score is D at row m plus the pad rows below the query
                                const int32_t d = score - (32 * W - (int32_t)m);
                                const bool better = (d < min_score) && on;
                                min_col = better ? col : min_col;
                                min_score = better ? d : min_score;
                            }
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    } else {
                        const uint4 eq4n = lds_lane[next_off];        // next column, group 0
                        const uint32_t ph_in = from_lower_lane<L>(ph_top, not_first);
                        const uint32_t mh_in = from_lower_lane<L>(mh_top, not_first);
                        column_grp<W, true, (uint32_t)S::kGroupU4, kWaves == 1>(
                            Pv, Mv, eq4, lds_lane, cur_off, ph_in, mh_in, ph_top, mh_top);
                        eq4 = eq4n;
                        // score row and first minimum (src-api/coreMyersGPU-CC35.cu:205-207,227-228);
                        // only the group's top lane keeps them, at column t - (L-1)
                        score += (int32_t)(ph_top >> 31) - (int32_t)(mh_top >> 31);
                        const uint32_t col = 8u * b + (uint32_t)i - (uint32_t)(L - 1);
                        const bool better = (score < min_score) && (col < ncols);
                        min_col = better ? col : min_col;
                        min_score = better ? score : min_score;
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    cur_off = next_off;
                }
                text = text_next;
                text_next = text_block(tw2, tw3, b + 2);
                tw2 = tw3;
                tw3 = tw4;
                if constexpr (kStop) {
                    // Threshold mode: the score row moves by at most 1 per column, so a
                    // candidate whose minimum so far and whose score less the columns
                    // left both exceed the threshold cannot get under it.  The wave
                    // stops once each of its candidates has finished or is in that state.
                    if (a.threshold >= 0.0f) {
                        const int32_t col = (int32_t)(8u * b + 7u) - (L - 1);
                        const bool hopeless = col >= 0 && min_score > kthr &&
                                              score - ((int32_t)ncols - 1 - col) > kthr;
                        const bool running = kthr >= 0 && col < (int32_t)ncols - 1 && !hopeless;
                        if (__ballot(running) == 0ull) break;
                    }
                }
            }

            if (in_pass && valid && gl == (uint32_t)(L - 1)) {
                uint2 r = live ? make_uint2(min_col, (uint32_t)min_score) : make_uint2(0u, kSkipScore);
                // threshold mode: score <= (uint32)((float)m * e) (src/checksum-2.c:88,104-105)
                if (live && a.threshold >= 0.0f && (uint32_t)min_score > threshold_of(m, a.threshold))
                    r = make_uint2(kNoMatch, kNoMatch);
                st_result<kHost>(results + cand, r);
            }
        }
    }
}

// Take the next range of work items [lo, hi) for a persistent workgroup from
// the variant's queue: items [0, n) are split into one part per XCD, a
// workgroup starts on the part of the XCD it runs on (workgroups are
// dispatched round-robin over the 8 XCDs) and moves on to the other parts
// when its own is drained, so the atomics of one part stay within one XCD
// until the tail.  Each pull takes a guided share of what is left in the
// part, an eighth of an even split over the part's workgroups (at least
// min_take items), so long-running items still balance at the tail.  Thread 0 only; *part counts the parts
// this workgroup has found drained.  Returns false when all are.
__device__ __forceinline__ bool pull_work(uint32_t *heads, uint32_t n, uint32_t min_take, uint32_t &part,
                                          uint32_t &lo, uint32_t &hi) {
    const uint32_t xcd = blockIdx.x % (uint32_t)MYERS_QUEUE_PARTS;
    const uint32_t workers = max(1u, gridDim.x / (uint32_t)MYERS_QUEUE_PARTS);
    while (part < (uint32_t)MYERS_QUEUE_PARTS) {
        const uint32_t x = (xcd + part) % (uint32_t)MYERS_QUEUE_PARTS;
        const uint32_t p_lo = (uint32_t)(((uint64_t)n * x) / MYERS_QUEUE_PARTS);
        const uint32_t p_hi = (uint32_t)(((uint64_t)n * (x + 1u)) / MYERS_QUEUE_PARTS);
        uint32_t *head = heads + x * (uint32_t)MYERS_QUEUE_STRIDE;
        const uint32_t seen = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t left = (p_hi - p_lo) > seen ? (p_hi - p_lo) - seen : 0u;
        if (left == 0u) {                               // drained: no atomic needed
            part++;
            continue;
        }
        const uint32_t take = max(min_take, left / (8u * workers));
        const uint32_t got = atomicAdd(head, take);
        if (got < p_hi - p_lo) {
            lo = p_lo + got;
            hi = min(p_hi, lo + take);
            return true;
        }
        part++;
    }
    return false;
}

// Waves per SIMD the register allocator must leave room for (1 = no limit).
// The wide one-lane kernels fit 3 (<= 168 VGPRs) or 4 (<= 128); more
// resident waves keep the double-rate logic ops issuing.  The cut-off
// kernels for W >= 28 then spill a few per-chunk values, outside the column
// loop.
template <int L, int W, bool kCut, bool kPersistent>
constexpr int kWavesPerEU = L != 1 ? 1 : (W >= 28 ? 3 : (W >= 20 ? ((!kPersistent && !kCut) ? 4 : 3) : 1));

// Variants with a threshold-mode kernel: the cut-off (L = 1, W >= 8) or the
// early stop (multi-lane).
template <int L, int W>
constexpr bool kHasCut = (L == 1 && W >= 8) || (L > 1 && L <= 64);

// Issue priority of a variant's waves (s_setprio).  A wave of a long read
// runs one candidate's chain of dependent column steps, and the chain of the
// longest candidate sets the time of its batch; so the arbiter of a shared
// SIMD serves waves of longer reads first and short-read waves fill the other
// issue slots.  Where every wave of a SIMD is of one kind nothing changes.
template <int L>
constexpr int kWavePrio = L >= 64 ? 3 : (L >= 16 ? 2 : (L >= 4 ? 1 : 0));

template <int L>
__device__ __forceinline__ void set_wave_prio() {
    if constexpr (kWavePrio<L> > 0) __builtin_amdgcn_s_setprio(kWavePrio<L>);
}

// Plain launch: a fixed grid strides over the chunks of one slot range.
template <int L, int W, bool kCut>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<L, W, kCut, false>)))
void verify_kernel(myers_verify_args_t a) {
    using S = Shape<L, W>;
    __shared__ Workspace<L, W> ws;
    set_wave_prio<L>();
    uint32_t start = a.start, count = a.count;
    if (a.bin_ranges != nullptr) {
        start = a.bin_ranges[2 * a.variant];
        count = a.bin_ranges[2 * a.variant + 1];
    }
    const uint32_t nchunks = (count + S::kSlots - 1) / S::kSlots;
    zero_row_init<L, W>(ws);
    // Workgroups are dispatched round-robin over the 8 XCDs; renumber them so
    // that each XCD (own L2) takes a contiguous run of chunks, which keeps the
    // candidates of one query -- and their Peq -- in one L2.
    uint32_t first = blockIdx.x;
    if ((gridDim.x & 7u) == 0u) first = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    verify_chunks<L, W, kCut>(a, start, count, first, nchunks, gridDim.x, ws);
}

// Persistent launch: a resident grid pulls ranges of chunks (at least
// `chunk_factor`) from the variant's device queue (part heads reset by the
// binning) until the bin is drained; every wave reaches the exit test.
template <int L, int W, bool kCut>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<L, W, kCut, true>)))
void verify_persistent_kernel(myers_verify_args_t a) {
    using S = Shape<L, W>;
    __shared__ Workspace<L, W> ws;
    __shared__ uint32_t item_sh, item_hi_sh;
    const uint32_t start = a.bin_ranges[2 * a.variant];
    const uint32_t count = a.bin_ranges[2 * a.variant + 1];
    const uint32_t nchunks = (count + S::kSlots - 1) / S::kSlots;
    const uint32_t cf = a.chunk_factor ? a.chunk_factor : 1u;
    uint32_t *heads = a.queue_heads + a.variant * (uint32_t)(MYERS_QUEUE_PARTS * MYERS_QUEUE_STRIDE);
    uint32_t part = 0;
    if (blockIdx.x >= nchunks) return;      // more workgroups than chunks: the others cover them
    set_wave_prio<L>();
    zero_row_init<L, W>(ws);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t lo = 0, hi = 0;
            if (!pull_work(heads, nchunks, cf, part, lo, hi)) lo = hi = 0xFFFFFFFFu;
            item_sh = lo;
            item_hi_sh = hi;
        }
        __syncthreads();
        const uint32_t lo = __builtin_amdgcn_readfirstlane(item_sh);
        const uint32_t hi = __builtin_amdgcn_readfirstlane(item_hi_sh);
        if (lo == 0xFFFFFFFFu) break;
        verify_chunks<L, W, kCut>(a, start, count, lo, hi, 1u, ws);
    }
}

// ------------------------------------------------- multi-wave candidates
// A wave issues one instruction every few clocks, so the time of one long
// candidate grows with the steps times the words each lane carries: a
// 20,000-base read on one wave (64 lanes x 10 words) takes ~5-10 ms.  For
// latency, a candidate can span Q waves of one workgroup (64Q lanes, W words
// each).  Lanes run diagonally as in one wave (lane d computes column t - d);
// at a wave boundary the bottom lane's hin is the top lane of the wave below,
// passed through LDS: wave q runs kWaveLag steps behind wave q-1 on top of the
// per-lane lag, each wave's top lane packs its (Ph, Mh) top bits of a block
// of 8 steps into one word of an LDS ring, and one barrier per block orders
// writes and reads.  Every wave runs the same number of blocks.

constexpr uint32_t kWaveLag = 8;
constexpr uint32_t kRingWords = 4;

template <int Q, int W>
struct WideShape {
    static constexpr int kLanes = 64 * Q;               // lanes per candidate
    static constexpr int kGroups = (W + 3) / 4;
    static constexpr int kTotal = kLanes * W;           // words per candidate
    static constexpr int kRows = 6;                     // A C G T N + zero row
    static constexpr int kRowU4 = kLanes;
    static constexpr int kLdsU4 = kGroups * kRows * kRowU4;
    static constexpr uint32_t kMaxLag = (uint32_t)kLanes - 1u + kWaveLag * (uint32_t)(Q - 1);
};

template <int Q, int W>
struct WideWorkspace {
    uint4 lds[WideShape<Q, W>::kLdsU4];
    uint32_t ring[Q - 1][kRingWords];
};

// Peq of one read, top-justified, transposed for 64Q lanes (the layout of
// stage_peq with one slot).
template <int Q, int W>
__device__ __forceinline__ void stage_peq_wide(uint4 *lds, uint2 qi, const uint32_t *__restrict__ peq,
                                               uint32_t tid) {
    using S = WideShape<Q, W>;
    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
    const int32_t pad = 32 * S::kTotal - (int32_t)qi.y;
    const int32_t nq = (int32_t)((qi.y + 31u) >> 5);
    for (uint32_t idx = tid; idx < 5u * (uint32_t)S::kTotal; idx += (uint32_t)S::kLanes) {
        const uint32_t gw = idx % (uint32_t)S::kTotal;
        const uint32_t c = idx / (uint32_t)S::kTotal;
        const int32_t b0 = 32 * (int32_t)gw - pad;
        const int32_t lo = b0 >> 5;
        const uint32_t sh = (uint32_t)b0 & 31u;
        const uint32_t *base = peq + (size_t)qi.x * 20u + 4u * c;
        uint32_t wlo = 0, whi = 0;
        if (lo >= 0 && lo < nq) wlo = base[(uint32_t)(lo >> 2) * 20u + (uint32_t)(lo & 3)];
        if (lo + 1 >= 0 && lo + 1 < nq) whi = base[(uint32_t)((lo + 1) >> 2) * 20u + (uint32_t)((lo + 1) & 3)];
        const uint32_t k = gw % (uint32_t)W, gl = gw / (uint32_t)W;
        const uint32_t u4 = ((k >> 2) * (uint32_t)S::kRows + c) * (uint32_t)S::kRowU4 + gl;
        l32[u4 * 4u + (k & 3u)] = (uint32_t)((((uint64_t)whi << 32) | wlo) >> sh);
    }
}

// One candidate's columns on wave wq of its workgroup (Q > 1).  A wave's
// role is fixed for the whole loop, so each role is its own instantiation:
// kBelow -- lane 0's hin is the top lane of the wave below, read from the LDS
// ring; kAbove -- this wave's top lane feeds the wave above; kTop -- the top
// wave keeps the score row.  Ring word of block b: the top lane's Ph bit 31
// of step 8b+i at bit i, its Mh bit 31 at bit 8+i.  Returns {min_col,
// min_score}, meaningful on lane 63 of the top wave.
template <int Q, int W, bool kBelow, bool kAbove, bool kTop>
__device__ __forceinline__ uint2 wide_columns(WideWorkspace<Q, W> &ws, uint32_t wq, uint32_t tid,
                                              const uint32_t *__restrict__ ref, int32_t last_word,
                                              uint64_t nib0, uint32_t lag, uint32_t ncols, uint32_t m) {
    using S = WideShape<Q, W>;
    constexpr uint32_t kGS = (uint32_t)(S::kRows * S::kRowU4);
    const uint32_t lane = tid & 63u;
    const uint4 *lds_lane = ws.lds + tid;
    const uint32_t nblocks = (ncols + S::kMaxLag + 7u) >> 3;
    const int64_t qn = (int64_t)nib0 - (int64_t)lag;
    const int32_t w0 = (int32_t)(qn >> 3);
    const uint32_t shift = (uint32_t)(qn & 7) * 4u;

    uint32_t Pv[W], Mv[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
        Pv[k] = ~0u;
        Mv[k] = 0u;
    }
    // Score row and its first minimum.  Before its first column the top lane
    // keeps its initial state (every lane below it is before its own first
    // column too), so the score stays m and those steps never win; only the
    // last block can reach columns past the window and checks them.  Windows
    // of at most 65,536 columns keep the minimum as one key, (score << 16) |
    // column (score <= m <= 24,576): the smaller key is the smaller score, then
    // the earlier column; a column past the window is masked to all ones.
    int32_t score = (int32_t)m, min_score = (int32_t)m;
    uint32_t min_col = 0, min_key = m << 16;
    uint32_t ph_top = 0, mh_top = 0, acc_p = 0, acc_m = 0;
    auto text_word = [&](int32_t k) { return ref[min(max(w0 + k, 0), last_word)]; };
    auto text_block = [&](uint32_t lo, uint32_t hi, uint32_t b) {
        uint32_t t = sanitize_text(__builtin_amdgcn_alignbit(hi, lo, shift));
        // steps before this lane's first column read the zero row
        const int32_t early = (int32_t)lag - 8 * (int32_t)b;
        if (early > 0) {
            const uint32_t low = early >= 8 ? ~0u : ((1u << (4 * early)) - 1u);
            t = (t & ~low) | (0x55555555u & low);
        }
        return t;
    };
    const uint32_t tw0 = text_word(0);
    const uint32_t tw1 = text_word(1);
    uint32_t tw2 = text_word(2);
    uint32_t tw3 = text_word(3);
    uint32_t text = text_block(tw0, tw1, 0);
    uint32_t text_next = text_block(tw1, tw2, 1);
    uint32_t cur_off = (text & 7u) * (uint32_t)S::kRowU4;
    uint4 eq4 = lds_lane[cur_off];
    auto block = [&](uint32_t b, auto last, auto keyed) {
        constexpr bool kLast = decltype(last)::value;
        constexpr bool kKeyed = decltype(keyed)::value;
        const uint32_t tw4 = text_word((int32_t)b + 4);
        // ring words of the wave below: its top lane's steps 8b-9 .. 8b-2,
        // written in blocks b-2 and b-1 (the same value in every lane)
        uint32_t w1 = 0, w2 = 0;
        if constexpr (kBelow) {
            if (b >= 1u) w1 = ws.ring[wq - 1u][(b - 1u) % kRingWords];
            if (b >= 2u) w2 = ws.ring[wq - 1u][(b - 2u) % kRingWords];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t cn = (i < 7) ? ((text >> (4 * (i + 1))) & 7u) : (text_next & 7u);
            const uint32_t next_off = cn * (uint32_t)S::kRowU4;
            const uint4 eq4n = lds_lane[next_off];
            // hin from the lane below (DPP wave_shr:1); lane 0 keeps the fill:
            // 0 on the bottom wave, else the wave below's bit of step 8b+i-9
            // moved to bit 31 (only bit 31 of a hin is read)
            uint32_t fill_p = 0u, fill_m = 0u;
            if constexpr (kBelow) {
                fill_p = i == 0 ? (w2 << 24) : (w1 << (32 - i));
                fill_m = i == 0 ? (w2 << 16) : (w1 << (24 - i));
            }
            const uint32_t ph_in = (uint32_t)__builtin_amdgcn_update_dpp((int)fill_p, (int)ph_top, 0x138, 0xF, 0xF, false);
            const uint32_t mh_in = (uint32_t)__builtin_amdgcn_update_dpp((int)fill_m, (int)mh_top, 0x138, 0xF, 0xF, false);
            column_grp<W, true, kGS, false>(Pv, Mv, eq4, lds_lane, cur_off, ph_in, mh_in, ph_top, mh_top);
            eq4 = eq4n;
            if constexpr (kAbove) {
                acc_p = (acc_p >> 1) | (ph_top & 0x80000000u);
                acc_m = (acc_m >> 1) | (mh_top & 0x80000000u);
            }
            if constexpr (kTop) {
                // score row: the top lane of the top wave, at column t - kMaxLag
                score += (int32_t)(ph_top >> 31) - (int32_t)(mh_top >> 31);
                const uint32_t col = 8u * b + (uint32_t)i - S::kMaxLag;
                if constexpr (kKeyed) {
                    const uint32_t c = kLast ? (col < ncols ? col : ~0u) : col;
                    min_key = min(min_key, ((uint32_t)score << 16) | c);
                } else {
                    const bool better = kLast ? (score < min_score) && (col < ncols) : (score < min_score);
                    min_col = better ? col : min_col;
                    min_score = better ? score : min_score;
                }
            }
            cur_off = next_off;
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (kAbove) {
            if (lane == 63u) ws.ring[wq][b % kRingWords] = (acc_p >> 24) | ((acc_m >> 16) & 0xFF00u);
        }
        __syncthreads();
        text = text_next;
        text_next = text_block(tw2, tw3, b + 2);
        tw2 = tw3;
        tw3 = tw4;
    };
    auto run = [&](auto keyed) {
        for (uint32_t b = 0; b + 1u < nblocks; b++) block(b, std::false_type{}, keyed);
        if (nblocks) block(nblocks - 1u, std::true_type{}, keyed);
    };
    if (kTop && ncols <= 0x10000u) {
        run(std::true_type{});
        return make_uint2(min_key & 0xFFFFu, min_key >> 16);
    }
    run(std::false_type{});
    return make_uint2(min_col, (uint32_t)min_score);
}

// Candidates cand_lo, cand_lo + stride, ... < cand_hi of the slot range
// [start, start + count), one per workgroup at a time.
template <int Q, int W>
__device__ __forceinline__ void verify_wide(const myers_verify_args_t &a, uint32_t start, uint32_t count,
                                            uint32_t cand_lo, uint32_t cand_hi, uint32_t stride,
                                            WideWorkspace<Q, W> &ws, uint32_t &staged_query) {
    using S = WideShape<Q, W>;
    const uint32_t tid = threadIdx.x;
    const uint32_t wq = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave of the workgroup
    const uint32_t lag = tid + kWaveLag * wq;                      // this lane's column lag
    const uint32_t *__restrict__ ref = a.ref;
    const int32_t last_word = (int32_t)(a.ref_words - 1);
    const CandRec *__restrict__ cands = static_cast<const CandRec *>(a.cands);
    const uint2 *__restrict__ qinfo = reinterpret_cast<const uint2 *>(a.qinfo);
    uint2 *__restrict__ results = static_cast<uint2 *>(a.results);

    for (uint32_t ci = cand_lo; ci < cand_hi && ci < count; ci += stride) {
        // every lane reads the same candidate: all values below are uniform
        const uint32_t cand = a.order ? a.order[start + ci] : start + ci;
        const CandRec cr = cands[cand];
        uint32_t m = 0, pos_entry = 0;
        bool sane = false;
        if (cr.query < a.num_queries) {
            const uint2 qi = qinfo[cr.query];
            pos_entry = qi.x;
            m = qi.y;
            sane = m <= 32u * S::kTotal && (uint64_t)pos_entry + ((m + 127u) >> 7) <= (uint64_t)a.num_peq_entries;
        }
        if (!sane) m = 0;
        const uint32_t n = window_len(m, a.distance);
        const bool live = sane && cr.position < a.ref_size && (a.ref_size - cr.position) > (uint64_t)n;
        if (!live) {
            if (tid == 0u) results[cand] = make_uint2(0u, kSkipScore);
            continue;
        }
        uint64_t nib0;
        uint32_t ncols;
        if (a.mode == 0) {
            nib0 = cr.position & ~7ull;
            ncols = 8u * ((n >> 3) + ((n & 7u) ? 2u : 1u));
        } else {
            nib0 = cr.position;
            ncols = n;
        }
        // consecutive candidates of one read keep its Peq in LDS
        if (cr.query != staged_query) {
            __syncthreads();
            stage_peq_wide<Q, W>(ws.lds, make_uint2(pos_entry, m), a.peq, tid);
            __syncthreads();
            staged_query = cr.query;
        }
        uint2 r;
        if (wq == 0u) {
            r = wide_columns<Q, W, false, true, false>(ws, wq, tid, ref, last_word, nib0, lag, ncols, m);
        } else if (wq + 1u < (uint32_t)Q) {
            r = wide_columns<Q, W, true, true, false>(ws, wq, tid, ref, last_word, nib0, lag, ncols, m);
        } else {
            r = wide_columns<Q, W, true, false, true>(ws, wq, tid, ref, last_word, nib0, lag, ncols, m);
        }
        if (tid == (uint32_t)S::kLanes - 1u) {
            if (a.threshold >= 0.0f && r.y > threshold_of(m, a.threshold)) r = make_uint2(kNoMatch, kNoMatch);
            results[cand] = r;
        }
    }
}

template <int Q, int W>
__device__ __forceinline__ void zero_row_init_wide(WideWorkspace<Q, W> &ws) {
    using S = WideShape<Q, W>;
    uint32_t *l32 = reinterpret_cast<uint32_t *>(ws.lds);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(S::kGroups * S::kRowU4 * 4); i += (uint32_t)S::kLanes) {
        const uint32_t u4 = i >> 2, g = u4 / S::kRowU4, r = u4 % S::kRowU4;
        l32[((g * S::kRows + kZeroRow) * S::kRowU4 + r) * 4u + (i & 3u)] = 0u;
    }
    __syncthreads();
}

template <int Q, int W>
__global__ __launch_bounds__(64 * Q) void verify_wide_kernel(myers_verify_args_t a) {
    __shared__ WideWorkspace<Q, W> ws;
    uint32_t start = a.start, count = a.count;
    if (a.bin_ranges != nullptr) {
        start = a.bin_ranges[2 * a.variant];
        count = a.bin_ranges[2 * a.variant + 1];
    }
    set_wave_prio<64 * Q>();
    zero_row_init_wide<Q, W>(ws);
    uint32_t first = blockIdx.x;
    if ((gridDim.x & 7u) == 0u) first = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    uint32_t staged = 0xFFFFFFFFu;
    verify_wide<Q, W>(a, start, count, first, count, gridDim.x, ws, staged);
}

template <int Q, int W>
__global__ __launch_bounds__(64 * Q) void verify_wide_persistent_kernel(myers_verify_args_t a) {
    __shared__ WideWorkspace<Q, W> ws;
    __shared__ uint32_t item_sh, item_hi_sh;
    const uint32_t start = a.bin_ranges[2 * a.variant];
    const uint32_t count = a.bin_ranges[2 * a.variant + 1];
    const uint32_t cf = a.chunk_factor ? a.chunk_factor : 1u;
    uint32_t *heads = a.queue_heads + a.variant * (uint32_t)(MYERS_QUEUE_PARTS * MYERS_QUEUE_STRIDE);
    uint32_t part = 0;
    if (blockIdx.x >= count) return;        // more workgroups than candidates: the others cover them
    set_wave_prio<64 * Q>();
    zero_row_init_wide<Q, W>(ws);
    uint32_t staged = 0xFFFFFFFFu;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t lo = 0, hi = 0;
            if (!pull_work(heads, count, cf, part, lo, hi)) lo = hi = 0xFFFFFFFFu;
            item_sh = lo;
            item_hi_sh = hi;
        }
        __syncthreads();
        const uint32_t lo = __builtin_amdgcn_readfirstlane(item_sh);
        const uint32_t hi = __builtin_amdgcn_readfirstlane(item_hi_sh);
        if (lo == 0xFFFFFFFFu) break;
        verify_wide<Q, W>(a, start, count, lo, hi, 1u, ws, staged);
    }
}

// (L, W) pairs compiled into the library.  L = 1 covers queries up to
// 32*32 = 1024 bases in one lane; L > 1 covers long reads (up to 64*12*32 =
// 24576 bases).
#define MYERS_VARIANT_LIST(X)                                                                    \
    X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 5) X(1, 6) X(1, 7) X(1, 8) X(1, 9) X(1, 10) X(1, 11)     \
    X(1, 12) X(1, 13) X(1, 14) X(1, 16) X(1, 19) X(1, 20) X(1, 22) X(1, 24) X(1, 25) X(1, 28)    \
    X(1, 29) X(1, 32)                                                                            \
    X(2, 8) X(2, 12) X(2, 16) X(4, 8) X(4, 12) X(4, 16) X(8, 8) X(8, 12) X(8, 16)                \
    X(16, 8) X(16, 9) X(16, 10) X(16, 11) X(16, 12) X(16, 16) X(32, 7) X(32, 8) X(32, 9)          \
    X(32, 10) X(32, 11) X(32, 12) X(32, 16) X(64, 7) X(64, 8) X(64, 9) X(64, 10) X(64, 12)

// Multi-wave variants (lanes = 64 Q per candidate, one candidate per
// workgroup of Q waves): long reads with lower per-candidate latency.
#define MYERS_WIDE_LIST(X) X(2, 4) X(2, 6) X(4, 2) X(4, 3)

// ------------------------------------------------- long-read bins, one launch
// A bin of long reads holds few candidates, and each takes one chain of
// dependent column steps: its time is the chain, not the chip's throughput.
// Separate launches per bin stack such waves on shared SIMDs, where they
// slow each other.  One launch of MYERS_LONG_WAVES-wave workgroups (the
// waves of a workgroup sit on different SIMDs of one CU) runs all of them:
// each wave pulls the next chunk from one counter over the bins in plan
// order, longest reads first, so the longest chains start first and a SIMD
// holds one long chain at a time.
#define MYERS_LONG_CASE(L, W)                                                                    \
    case ((L) << 8) | (W):                                                                       \
        if constexpr ((L) >= 4 && (L) <= 64) {                                                   \
            auto &w = *reinterpret_cast<Workspace<(L), (W)> *>(ws);                              \
            if (prepared != j) {                                                                 \
                zero_row_init<(L), (W), MYERS_LONG_WAVES>(w);                                    \
                prepared = j;                                                                    \
            }                                                                                    \
            verify_chunks<(L), (W), false, MYERS_LONG_WAVES>(a, start, count, chunk, chunk + 1u, 1u, w); \
        }                                                                                        \
        break;

using LongWide = WideWorkspace<MYERS_LONG_WAVES, MYERS_WIDE_MAX_WORDS>;
static_assert(sizeof(WideWorkspace<MYERS_LONG_WAVES, 2>) == sizeof(LongWide) &&
                  WideShape<MYERS_LONG_WAVES, 2>::kGroups == WideShape<MYERS_LONG_WAVES, 3>::kGroups,
              "wide candidates of 2 and 3 words share one LDS layout");

__global__ __launch_bounds__(64 * MYERS_LONG_WAVES) void verify_long_kernel(myers_verify_args_t a,
                                                                           myers_long_plan_t p) {
    extern __shared__ uint4 long_lds[];
    __shared__ uint32_t wide_item_sh;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63u;
    set_wave_prio<64>();
    // Phase 1: the wide bins, one candidate per workgroup (longest reads).
    if (p.n_wide) {
        LongWide &wws = *reinterpret_cast<LongWide *>(long_lds);
        zero_row_init_wide<MYERS_LONG_WAVES, MYERS_WIDE_MAX_WORDS>(wws);
        uint32_t staged = 0xFFFFFFFFu;
        const uint32_t grid = gridDim.x;
        uint32_t witem = (blockIdx.x + grid - (p.rotate / (uint32_t)MYERS_LONG_WAVES) % grid) % grid;
        for (bool first = true;; first = false) {
            if (!first) {
                if (threadIdx.x == 0) wide_item_sh = atomicAdd(a.queue_heads + MYERS_WIDE_HEAD, 1u);
                __syncthreads();
                witem = grid + (uint32_t)__builtin_amdgcn_readfirstlane((int)wide_item_sh);
                __syncthreads();                 // every wave has read it before the next pull
            }
            uint32_t j = 0, base = 0, start = 0, count = 0;
            bool found = false;
            for (; j < p.n_wide; j++) {
                const uint32_t v = p.wide_variant[j];
                const uint32_t c = a.bin_ranges[2u * v + 1u];
                if (witem - base < c) {
                    start = a.bin_ranges[2u * v];
                    count = c;
                    found = true;
                    break;
                }
                base += c;
            }
            if (!found) break;                   // every wide bin drained
            const uint32_t slot = witem - base;
            if (p.wide_words[j] == 2u) {
                verify_wide<MYERS_LONG_WAVES, 2>(a, start, count, slot, slot + 1u, 1u,
                                                 *reinterpret_cast<WideWorkspace<MYERS_LONG_WAVES, 2> *>(long_lds),
                                                 staged);
            } else {
                verify_wide<MYERS_LONG_WAVES, 3>(a, start, count, slot, slot + 1u, 1u, wws, staged);
            }
        }
        __syncthreads();                         // phase 2 reuses the LDS
    }
    // Phase 2: the other bins, one chunk per wave.  The first round gives one
    // item to each wave free at the start -- past the workgroups that took a
    // wide candidate in phase 1's first round (from workgroup rotate / waves
    // on) -- so that the longest of these items start at once; later items
    // are pulled from the counter.
    uint4 *ws = long_lds + wave * p.ws_u4;
    uint32_t *counter = a.queue_heads + MYERS_LONG_HEAD;
    uint32_t prepared = 0xFFFFFFFFu;
    uint32_t wide_first = 0;
    for (uint32_t j = 0; j < p.n_wide; j++) wide_first += a.bin_ranges[2u * p.wide_variant[j] + 1u];
    wide_first = min(wide_first, gridDim.x);
    const uint32_t nw = gridDim.x * (uint32_t)MYERS_LONG_WAVES;
    const uint32_t free_waves = nw - wide_first * (uint32_t)MYERS_LONG_WAVES;
    const uint32_t slot = blockIdx.x * (uint32_t)MYERS_LONG_WAVES + wave;
    const uint32_t rel = (slot + nw - (p.rotate + wide_first * (uint32_t)MYERS_LONG_WAVES) % nw) % nw;
    uint32_t item = rel;
    for (bool first = rel < free_waves;; first = false) {
        if (!first) {
            uint32_t got = 0;
            if (lane == 0u) got = atomicAdd(counter, 1u);
            item = free_waves + (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)got, 0));
        }
        uint32_t j = 0, base = 0, start = 0, count = 0;
        bool found = false;
        for (; j < p.n; j++) {
            const uint32_t v = p.variant[j];
            const uint32_t c = a.bin_ranges[2u * v + 1u];
            const uint32_t slots = 64u / (p.key[j] >> 8);
            const uint32_t chunks = (c + slots - 1u) / slots;
            if (item - base < chunks) {
                start = a.bin_ranges[2u * v];
                count = c;
                found = true;
                break;
            }
            base += chunks;
        }
        if (!found) break;                       // every bin drained
        const uint32_t chunk = item - base;
        switch (p.key[j]) {
            MYERS_VARIANT_LIST(MYERS_LONG_CASE)
            default:
                break;
        }
    }
}

// ------------------------------------------------- one launch over all bins
// The queue launch: a resident grid of one-wave workgroups takes one chunk
// at a time from a single counter over the bins in plan order (the host
// orders them longest reads first, so the grid's tail is short chunks), and
// runs each chunk with its bin's variant.  No launch boundary separates the
// bins: a wave that finishes the last chunk of one bin takes the next bin's
// first.  The register budget of a kernel is that of its largest variant,
// so the one-lane and the multi-lane variants are two kernels (classes);
// the multi-lane class stops at 12 words per lane (W = 16 would leave two
// waves per SIMD for every bin of the launch).
template <int L, int W>
constexpr int kQueueClass = L == 1 ? MYERS_QUEUE_ONE_LANE : ((L <= 64 && W <= 12) ? MYERS_QUEUE_MULTI_LANE : -1);

// Waves per workgroup of a class's queue kernel.
template <int kClass>
constexpr int kQueueWaves = kClass == MYERS_QUEUE_MULTI_LANE ? 2 : 1;

#define MYERS_QUEUE_CASE(L, W)                                                                   \
    case ((L) << 8) | (W):                                                                       \
        if constexpr (kQueueClass<(L), (W)> == kClass) {                                         \
            auto &w = *reinterpret_cast<Workspace<(L), (W)> *>(queue_lds + kXchgU4);             \
            if (prepared != key) {                                                               \
                set_wave_prio<(L)>();                                                            \
                zero_row_init<(L), (W), 1, 64u * kShare>(w);                                     \
                prepared = key;                                                                  \
            }                                                                                    \
            verify_chunks<(L), (W), kCut && kHasCut<(L), (W)>, 1, false, kShare>(a, start, count, chunk, \
                                                                                 chunk + 1u, 1u, w, xchg); \
        }                                                                                        \
        break;

template <int kClass, bool kCut>
__global__ __launch_bounds__(64 * kQueueWaves<kClass>) __attribute__((amdgpu_waves_per_eu(3)))
void verify_queue_kernel(myers_verify_args_t a, myers_queue_plan_t p) {
    extern __shared__ uint4 queue_lds[];
    // Multi-lane class: workgroups of two waves that take chunk pairs and
    // share one workspace (verify_chunks, kShare = 2), after one uint4 of
    // hand-over words; the item counter then counts chunk pairs.
    constexpr int kShare = kQueueWaves<kClass>;
    constexpr uint32_t kXchgU4 = kShare == 2 ? 1u : 0u;
    uint32_t *xchg = reinterpret_cast<uint32_t *>(queue_lds);
    uint32_t *counter = a.queue_heads + MYERS_QUEUE_HEAD(kClass);
    uint32_t prepared = 0xFFFFFFFFu;
    for (;;) {
        uint32_t item;
        if constexpr (kShare == 1) {
            uint32_t got = 0;
            if (threadIdx.x == 0) got = atomicAdd(counter, 1u);
            item = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)got, 0));
        } else {
            if (threadIdx.x == 0) xchg[3] = atomicAdd(counter, 1u);
            __syncthreads();
            item = __builtin_amdgcn_readfirstlane(xchg[3]);
            __syncthreads();                     // read by all before the next claim
        }
        uint32_t j = 0, base = 0, start = 0, count = 0;
        bool found = false;
        for (; j < p.n; j++) {
            const uint32_t v = p.variant[j];
            const uint32_t c = a.bin_ranges[2u * v + 1u];
            const uint32_t slots = 64u / (p.key[j] >> 8);
            const uint32_t chunks = (c + slots * (uint32_t)kShare - 1u) / (slots * (uint32_t)kShare);
            if (item - base < chunks) {
                start = a.bin_ranges[2u * v];
                count = c;
                found = true;
                break;
            }
            base += chunks;
        }
        if (!found) break;                       // every bin drained: each wave exits here
        const uint32_t chunk = item - base;
        const uint32_t key = p.key[j];
        switch (key) {
            MYERS_VARIANT_LIST(MYERS_QUEUE_CASE)
            default:
                break;
        }
    }
}

// ------------------------------------------------- resident serving kernel
// One-wave workgroups that stay resident and verify the chunks a host thread
// publishes (the protocol is described with myers_serve_args_t).  Loads of
// host-written descriptors use system-scope atomics: vector loads that no
// cache keeps stale, which matters for a kernel that outlives many requests.
// Every atomic is relaxed: on a chip of several XCDs an agent- or
// system-scope acquire invalidates the XCD's L2 and a release writes it back,
// and idle waves polling with them would keep flushing the reference text of
// the working waves.  The ordering the protocol needs comes instead from
// (a) the host side: items and requests are written before the tail;
// (b) address dependencies: a wave reads an item only at an index it won,
//     and a request only at the slot the item names;
// (c) the request arena, results and tickets live in fine-grained host
//     memory, which the GPU does not cache, so a store is visible to the host
//     once it has completed; a wave waits for its result stores to complete
//     (s_waitcnt, the workgroup-scope release) before it counts its chunk,
//     so the wave that completes a request writes the ticket after every
//     result of the request has completed.

__device__ __forceinline__ uint32_t load_sys(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t load_sys64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// The same value is loaded by every lane: keep it in scalar registers, as a
// kernel argument would be, so it takes no vector registers in the column loops.
__device__ __forceinline__ uint32_t uniform_sys(const uint32_t *p) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)load_sys(p));
}
__device__ __forceinline__ uint64_t uniform_sys64(const void *p) {
    const uint64_t v = load_sys64(reinterpret_cast<const uint64_t *>(p));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// A key of the SIMD this wave runs on, from the hardware wave id (SIMD, CU,
// shader array and engine) and the XCD.  Two waves with the same key share a
// SIMD; a collision of two SIMDs only costs speed.
__device__ __forceinline__ uint32_t simd_key() {
    const uint32_t hw = (uint32_t)__builtin_amdgcn_s_getreg((31 << 11) | 4);     // HW_REG_HW_ID
    const uint32_t xcc = (uint32_t)__builtin_amdgcn_s_getreg((31 << 11) | 20);   // HW_REG_XCC_ID
    const uint32_t simd = (hw >> 4) & 3u, cu = (hw >> 8) & 15u, sh = (hw >> 12) & 1u, se = (hw >> 13) & 7u;
    return (simd | (cu << 2) | (sh << 6) | (se << 7) | ((xcc & 15u) << 10)) & (MYERS_SERVE_SIMD_KEYS - 1u);
}

// Try to take the next item of ring r: true and *item on success.
__device__ __forceinline__ bool serve_take(uint32_t *head, uint32_t *dtail, uint32_t &item) {
    for (;;) {
        const uint32_t h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t t = __hip_atomic_load(dtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int32_t)(t - h) <= 0) return false;
        uint32_t expect = h;
        if (__hip_atomic_compare_exchange_strong(head, &expect, h + 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            item = h;
            return true;
        }
    }
}

// Lane 0: the next item's index in its ring (`ring`), or false when the wave
// must exit.  A long item is taken only with the SIMD's count below
// long_per_simd; the caller lowers it when the item is done.  The wave
// leaves when the host's stop generation is seen, or after idle_ticks in
// which no item of either ring was claimed anywhere (the rings' heads did
// not move), measured on this wave's own real-time counter.  An idle wave
// backs off from ~3 us to ~220 us between polls, so that a grid of idle
// waves does not flood the memory system with atomics while a few waves run
// long chains; its count of polls without progress bounds the wait as well
// (under ~30 s), whatever the real-time counter does.
__device__ bool serve_claim(const myers_serve_args_t &s, uint32_t simd, uint32_t &item, uint32_t &ring) {
    uint32_t *dstop = s.dev + MYERS_SERVE_STOP;
    uint64_t *poll = reinterpret_cast<uint64_t *>(s.dev + MYERS_SERVE_POLL);
    uint32_t *busy = s.simd_busy + simd;
    uint32_t idle = 0, polls = 0, heads = 0xFFFFFFFFu;
    uint64_t t_idle = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        // a long item, if fewer than long_per_simd long items run on this SIMD
        uint32_t *hl = s.dev + MYERS_SERVE_HEAD, *tl = s.dev + MYERS_SERVE_DTAIL;
        uint32_t running = __hip_atomic_load(busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int32_t)(__hip_atomic_load(tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                      __hip_atomic_load(hl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) > 0 &&
            running < s.long_per_simd) {
            if (__hip_atomic_compare_exchange_strong(busy, &running, running + 1u, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                if (serve_take(hl, tl, item)) {
                    ring = MYERS_SERVE_LONG_RING;
                    return true;
                }
                __hip_atomic_fetch_sub(busy, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (serve_take(s.dev + MYERS_SERVE_HEAD + 64u, s.dev + MYERS_SERVE_DTAIL + 64u, item)) {
            ring = MYERS_SERVE_SHORT_RING;
            return true;
        }
        // nothing claimable here.  Progress anywhere (a head moved) restarts
        // the idle time; then refresh the mirrors of the host tails (one wave
        // per poll period).
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        uint32_t sum = 0;
        for (uint32_t r = 0; r < (uint32_t)MYERS_SERVE_RINGS; r++)
            sum += __hip_atomic_load(s.dev + MYERS_SERVE_HEAD + 64u * r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sum != heads) {
            heads = sum;
            t_idle = now;
            polls = 0;
        }
        uint64_t seen = __hip_atomic_load(poll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        polls++;
        if ((polls & 255u) == 0u ||
            (now - seen >= s.poll_ticks &&
             __hip_atomic_compare_exchange_strong(poll, &seen, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT))) {
            bool fresh = false;
            for (uint32_t r = 0; r < (uint32_t)MYERS_SERVE_RINGS; r++) {
                uint32_t *dtail = s.dev + MYERS_SERVE_DTAIL + 64u * r;
                const uint32_t ht = load_sys(s.host_tail + r * (uint32_t)MYERS_SERVE_TAIL_STRIDE);
                uint32_t old = __hip_atomic_load(dtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while ((int32_t)(ht - old) > 0) {
                    if (__hip_atomic_compare_exchange_strong(dtail, &old, ht, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        fresh = true;
                        break;
                    }
                }
            }
            if (load_sys(s.host_stop) == s.gen)
                __hip_atomic_store(dstop, s.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (fresh) {
                idle = 0;
                continue;
            }
        }
        if (__hip_atomic_load(dstop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == s.gen) return false;
        if (now - t_idle > s.idle_ticks || polls > (1u << 17)) return false;
        idle = idle < 64u ? idle + 1u : 64u;
        for (uint32_t k = 0; k < idle; k++) __builtin_amdgcn_s_sleep(127);
    }
}

template <int L, int W>
constexpr bool kServeVariant = kQueueClass<L, W> >= 0;
// A long item is a chain of column steps that often has its SIMD to itself:
// it runs with the interleaved word order (verify_chunks with the wave-local
// synchronisation of several-wave workgroups, which fits a one-wave
// workgroup as well), 5-10% faster alone than the phased order.
template <int L>
constexpr int kServeWaves = L >= MYERS_SERVE_LONG_LANES ? MYERS_LONG_WAVES : 1;

#define MYERS_SERVE_CASE(L, W)                                                                   \
    case ((L) << 8) | (W):                                                                       \
        if constexpr (kServeVariant<(L), (W)>) {                                                 \
            auto &w = *reinterpret_cast<Workspace<(L), (W)> *>(serve_lds);                       \
            if (prepared != key) {                                                               \
                set_wave_prio<(L)>();                                                            \
                zero_row_init<(L), (W)>(w);                                                      \
                prepared = key;                                                                  \
            }                                                                                    \
            verify_chunks<(L), (W), kCut && kHasCut<(L), (W)>, kServeWaves<(L)>, true>(a, first, count, 0u, 1u, 1u, w); \
        }                                                                                        \
        break;

template <bool kCut>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void verify_serve_kernel(myers_serve_args_t s) {
    extern __shared__ uint4 serve_lds[];
    uint32_t prepared = 0xFFFFFFFFu;
    const uint32_t simd = simd_key();
    for (;;) {
        uint32_t got = 0, got_ring = 2u;           // ring 2: exit
        if (threadIdx.x == 0 && !serve_claim(s, simd, got, got_ring)) got_ring = 2u;
        const uint32_t ring = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)got_ring, 0));
        if (ring > 1u) break;
        const uint32_t item = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)got, 0));
        const myers_serve_item_t *it = s.items + ring * (s.ring_mask + 1u) + (item & s.ring_mask);
        const uint32_t key = uniform_sys(&it->key);
        const uint32_t slot = uniform_sys(&it->req) & s.req_mask;
        const uint32_t first = uniform_sys(&it->first);
        const uint32_t count = min(uniform_sys(&it->count), 64u);
        const myers_serve_req_t *rq = s.reqs + slot;
        myers_verify_args_t a;
        a.ref = s.ref;
        a.ref_words = s.ref_words;
        a.ref_size = s.ref_size;
        a.peq = reinterpret_cast<const uint32_t *>(uniform_sys64(&rq->peq));
        a.qinfo = reinterpret_cast<const uint32_t *>(uniform_sys64(&rq->qinfo));
        a.cands = reinterpret_cast<const void *>(uniform_sys64(&rq->cands));
        a.results = reinterpret_cast<void *>(uniform_sys64(&rq->results));
        a.order = reinterpret_cast<const uint32_t *>(uniform_sys64(&rq->order));
        a.num_peq_entries = uniform_sys(&rq->num_peq_entries);
        a.num_queries = uniform_sys(&rq->num_queries);
        a.mode = uniform_sys(&rq->mode);
        a.distance = __uint_as_float(uniform_sys(reinterpret_cast<const uint32_t *>(&rq->distance)));
        a.threshold = __uint_as_float(uniform_sys(reinterpret_cast<const uint32_t *>(&rq->threshold)));
        a.bin_ranges = nullptr;
        a.start = first;
        a.count = count;
        a.variant = 0;
        a.chunk_factor = 1;
        a.queue_heads = nullptr;
        const uint32_t total = uniform_sys(&rq->num_items_cands);
        const uint64_t ticket = uniform_sys64(&rq->ticket);
        switch (key) {
            MYERS_VARIANT_LIST(MYERS_SERVE_CASE)
            default:
                break;
        }
        // this wave's result stores complete before its chunk is counted
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t before = __hip_atomic_fetch_add(s.done + slot, count, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
            if (before + count == total) {
                // the request's last item: reset its counter for the slot's next
                // request, then publish the ticket (every result has completed)
                __hip_atomic_store(s.done + slot, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_s_waitcnt(0);
                __hip_atomic_store(s.flags + slot, ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            // the SIMD can take another long item
            if (ring == MYERS_SERVE_LONG_RING)
                __hip_atomic_fetch_sub(s.simd_busy + simd, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------- binning

__global__ __launch_bounds__(64) void bin_reset_kernel(uint32_t *err) {
    if (threadIdx.x == 0) *err = 0;
}

constexpr uint32_t kTooLong = 0xFFFFFFFFu;
constexpr uint32_t kBadQuery = 0xFFFFFFFEu;

__device__ __forceinline__ uint32_t variant_of(const uint2 *qinfo, uint32_t num_queries, const CandRec *cands,
                                               uint32_t i, const uint8_t *bin_of_words, uint32_t max_words) {
    const uint32_t q = cands[i].query;
    if (q >= num_queries) return kBadQuery;
    const uint32_t w = (qinfo[q].y + 31u) >> 5;
    return w <= max_words ? (uint32_t)bin_of_words[w] : kTooLong;
}

// Binning is a stable partition of the candidates by variant with no global
// atomics: block b of the count and scatter kernels owns the contiguous span
// [b*span, (b+1)*span); bin_count writes each block's per-variant counts
// (variant-major), bin_scan (one block per variant) turns them into each
// block's first slot within the variant, bin_base lays the variants' ranges
// end to end, and bin_scatter writes every candidate's
// id at its slot in input order.  A read's candidates, adjacent in the
// input, stay adjacent in their bin.

__device__ __forceinline__ uint32_t bin_span(uint32_t n) {
    const uint32_t span = (n + gridDim.x - 1u) / gridDim.x;
    return (span + 255u) & ~255u;
}

__global__ __launch_bounds__(256) void bin_count_kernel(const uint2 *qinfo, uint32_t num_queries,
                                                        const CandRec *cands,
                                                        uint32_t n, const uint8_t *bin_of_words,
                                                        uint32_t max_words, uint32_t nv,
                                                        uint32_t *block_counts, uint32_t *err) {
    __shared__ uint32_t hist[MYERS_MAX_VARIANTS];
    for (uint32_t i = threadIdx.x; i < nv; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint32_t span = bin_span(n);
    const uint32_t lo = blockIdx.x * span, hi = min(n, lo + span);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint32_t v = variant_of(qinfo, num_queries, cands, i, bin_of_words, max_words);
        if (v < nv) {
            atomicAdd(&hist[v], 1u);
        } else {
            atomicOr(err, v == kBadQuery ? 2u : 1u);
        }
    }
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < nv; v += blockDim.x) block_counts[v * gridDim.x + blockIdx.x] = hist[v];
}

// One block of 256 threads per variant: an exclusive scan of the variant's
// per-block counts (16 consecutive blocks per thread, then a scan of the 256
// thread sums), in place, giving each count block's first slot relative to
// the variant's range; the variant's total goes to ranges[2v+1].
__global__ __launch_bounds__(256) void bin_scan_kernel(uint32_t *block_counts, uint32_t nblocks,
                                                       uint32_t *ranges) {
    __shared__ uint32_t part[256];
    const uint32_t t = threadIdx.x, v = blockIdx.x;
    const uint32_t per = (nblocks + 255u) / 256u;
    uint32_t *row = block_counts + v * nblocks;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t b = t * per + k;
        if (b < nblocks) sum += row[b];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 256u; o <<= 1) {       // inclusive scan (Hillis-Steele)
        const uint32_t x = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t b = t * per + k;
        if (b < nblocks) {
            const uint32_t c = row[b];
            row[b] = run;
            run += c;
        }
    }
    if (t == 255u) ranges[2 * v + 1] = part[255];
}

// Variant ranges from the totals, and the persistent queues' part heads.
__global__ __launch_bounds__(64) void bin_base_kernel(uint32_t nv, uint32_t *ranges, uint32_t *queue_heads) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t v = 0; v < nv; v++) {
            ranges[2 * v] = acc;
            acc += ranges[2 * v + 1];
        }
    }
    for (uint32_t i = threadIdx.x; i < nv * (uint32_t)MYERS_QUEUE_PARTS; i += blockDim.x)
        queue_heads[i * (uint32_t)MYERS_QUEUE_STRIDE] = 0;
    if (threadIdx.x == 0) {
        queue_heads[MYERS_LONG_HEAD] = 0;
        queue_heads[MYERS_WIDE_HEAD] = 0;
        for (uint32_t c = 0; c < (uint32_t)MYERS_QUEUE_CLASSES; c++) queue_heads[MYERS_QUEUE_HEAD(c)] = 0;
    }
}

// Write each candidate's id at its slot: per 256-candidate round, ranks within
// a wave come from ballots, the waves' counts are added in wave order in LDS,
// and the block's running first slot per variant advances by the round's
// counts.
__global__ __launch_bounds__(256) void bin_scatter_kernel(const uint2 *qinfo, uint32_t num_queries,
                                                          const CandRec *cands,
                                                          uint32_t n, const uint8_t *bin_of_words,
                                                          uint32_t max_words, uint32_t nv,
                                                          const uint32_t *block_first, const uint32_t *ranges,
                                                          uint32_t *order, uint2 *results) {
    __shared__ uint32_t next[MYERS_MAX_VARIANTS];
    __shared__ uint32_t wave_cnt[4][MYERS_MAX_VARIANTS];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t v = threadIdx.x; v < nv; v += blockDim.x)
        next[v] = ranges[2 * v] + block_first[v * gridDim.x + blockIdx.x];
    const uint32_t span = bin_span(n);
    const uint32_t lo = blockIdx.x * span, hi = min(n, lo + span);
    for (uint32_t r = lo; r < hi; r += 256u) {
        for (uint32_t i = threadIdx.x; i < 4u * nv; i += blockDim.x) wave_cnt[i / nv][i % nv] = 0;
        __syncthreads();
        const uint32_t i = r + threadIdx.x;
        const uint32_t v = i < hi ? variant_of(qinfo, num_queries, cands, i, bin_of_words, max_words) : kTooLong;
        if (i < hi && v >= nv) results[i] = make_uint2(0u, kSkipScore);
        uint32_t rank = 0;
        uint64_t pending = __ballot(v < nv);
        while (pending) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
            const uint32_t vv = (uint32_t)__shfl((int)v, (int)leader);
            const uint64_t mine = __ballot(v == vv);
            if (v == vv) rank = (uint32_t)__builtin_popcountll(mine & below);
            if (lane == leader) wave_cnt[wave][vv] = (uint32_t)__builtin_popcountll(mine);
            pending &= ~mine;
        }
        __syncthreads();
        if (v < nv) {
            uint32_t slot = next[v] + rank;
            for (uint32_t w = 0; w < wave; w++) slot += wave_cnt[w][v];
            order[slot] = i;
        }
        __syncthreads();
        for (uint32_t vv = threadIdx.x; vv < nv; vv += blockDim.x)
            next[vv] += wave_cnt[0][vv] + wave_cnt[1][vv] + wave_cnt[2][vv] + wave_cnt[3][vv];
        __syncthreads();
    }
}


// ------------------------------------------------------------- band probe
// Threshold mode (see myers_launch_probe).  probe_kernel: one thread per
// candidate, the read's first 32 bases as one bit-vector word (Peq word 0 of
// codes 0..4; the recurrence of src/myers-4bits.c:115-134), D at row 32
// tracked over the window's leading columns.  Each thread's five Eq words
// sit in its own LDS column and are read back by the text code of each
// step; the window's words are all requested before the first step (random
// windows: the loads' latency, not their bytes, bounds this kernel).
constexpr uint32_t kProbeLikely = 8u;      // D <= 8 somewhere on the first 32 bases: class 0
constexpr uint32_t kProbeSpan = 128u;      // slots reordered together: two one-lane chunks
constexpr uint32_t kProbeBlocks = 40u;     // at most 320 columns probed
__global__ __launch_bounds__(256) void probe_kernel(myers_verify_args_t a, uint32_t n,
                                                    const uint8_t *bin_of_words, uint32_t max_words,
                                                    uint64_t probe_mask) {
    __shared__ uint32_t eqs[5][256];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * 256u + tid;
    uint32_t cls = 0;
    uint32_t ncols = 0;
    uint64_t nib0 = 0;
    if (i < n) {
        const CandRec cr = static_cast<const CandRec *>(a.cands)[i];
        const uint2 *qinfo = reinterpret_cast<const uint2 *>(a.qinfo);
        if (cr.query < a.num_queries) {
            const uint2 qi = qinfo[cr.query];
            const uint32_t m = qi.y, w = (m + 31u) >> 5;
            const uint32_t v = w <= max_words ? (uint32_t)bin_of_words[w] : 64u;
            if (v < 64u && ((probe_mask >> v) & 1ull) && m >= 32u &&
                (uint64_t)qi.x + ((m + 127u) >> 7) <= (uint64_t)a.num_peq_entries) {
                const uint32_t nw = window_len(m, a.distance);
                // skip rule: src-api/coreMyersGPU-CC35.cu:142
                if (cr.position < a.ref_size && a.ref_size - cr.position > (uint64_t)nw) {
                    nib0 = a.mode == 0 ? (cr.position & ~7ull) : cr.position;
                    const uint32_t lead = (uint32_t)__fmul_rn(__fmul_rn(2.0f, a.distance), (float)m) + 72u;
                    ncols = min(min(nw, lead), 8u * kProbeBlocks);
                    for (uint32_t c = 0; c < 5u; c++) eqs[c][tid] = a.peq[(size_t)qi.x * 20u + 4u * c];
                }
            }
        }
    }
    const uint32_t wave_cols = wave_max(ncols);
    if (wave_cols > 0u) {
        const uint32_t *ref = a.ref;
        const int64_t last_word = (int64_t)a.ref_words - 1;
        const int64_t w0 = (int64_t)(nib0 >> 3);
        const uint32_t shift = (uint32_t)(nib0 & 7u) * 4u;
        uint32_t tw[kProbeBlocks + 1];
#pragma unroll
        for (uint32_t k = 0; k <= kProbeBlocks; k++) tw[k] = ref[min(w0 + (int64_t)k, last_word)];
        // keep the loads here, ahead of the steps (the compiler would sink
        // each one to its block, after the early exits)
#pragma unroll
        for (uint32_t k = 0; k <= kProbeBlocks; k++) asm volatile("" ::"v"(tw[k]));
        uint32_t pv = ~0u, mv = 0u;
        int32_t score = 32, best = 32;
#pragma unroll
        for (uint32_t b = 0; b < kProbeBlocks; b++) {
            if (8u * b >= wave_cols) break;
            const uint32_t t = sanitize_text(__builtin_amdgcn_alignbit(tw[b + 1], tw[b], shift));
            uint32_t eq[8];
#pragma unroll
            for (uint32_t j = 0; j < 8u; j++) eq[j] = eqs[(t >> (4u * j)) & 7u][tid];
#pragma unroll
            for (uint32_t j = 0; j < 8u; j++) {
                const uint32_t xv = eq[j] | mv;
                const uint32_t xh = __builtin_amdgcn_bitop3_b32((eq[j] & pv) + pv, pv, eq[j], kXorOr);
                const uint32_t ph = __builtin_amdgcn_bitop3_b32(mv, xh, pv, kOrNor);
                const uint32_t mh = pv & xh;
                score += (int32_t)(ph >> 31) - (int32_t)(mh >> 31);
                best = 8u * b + j < ncols ? min(best, score) : best;
                const uint32_t php = ph << 1, mhp = mh << 1;
                pv = __builtin_amdgcn_bitop3_b32(mhp, xv, php, kOrNor);
                mv = php & xv;
            }
        }
        if (ncols > 0u) cls = best <= (int32_t)kProbeLikely ? 0u : 1u;
    }
    if (i < n) static_cast<uint2 *>(a.results)[i].x = cls;
}

// Reorder the slots of each probed variant's range in spans of kProbeSpan
// (two chunks of 64 one-lane candidates, aligned to the range start): the
// first chunk takes the span's class-0 candidates and then its first class-1
// ones, the second chunk the rest, each in slot order.  A read's candidates,
// adjacent in the range, stay in runs, so the first chunk holds about as
// many distinct reads as class-0 candidates and its Peq still stages in one
// pass; the second chunk's waves then keep a shallow band.  One wave per
// span, spans in grid-stride order.
__global__ __launch_bounds__(64) void probe_order_kernel(const uint32_t *ranges, uint32_t nv, uint64_t probe_mask,
                                                         uint32_t *order, const uint2 *results) {
    const uint32_t lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t v = 0; v < nv && v < 64u; v++) {
        if (!((probe_mask >> v) & 1ull)) continue;
        const uint32_t start = ranges[2 * v], count = ranges[2 * v + 1];
        const uint32_t nspans = (count + kProbeSpan - 1u) / kProbeSpan;
        for (uint32_t sp = blockIdx.x; sp < nspans; sp += gridDim.x) {
            const uint32_t s0 = start + sp * kProbeSpan;
            const uint32_t here = min(kProbeSpan, count - sp * kProbeSpan);
            if (here <= 64u) continue;                  // one chunk: nothing to move
            const bool va = lane < here, vb = 64u + lane < here;
            const uint32_t ca = va ? order[s0 + lane] : 0u;
            const uint32_t cb = vb ? order[s0 + 64u + lane] : 0u;
            const uint32_t ka = va ? (results[ca].x & 1u) : 1u;
            const uint32_t kb = vb ? (results[cb].x & 1u) : 1u;
            const uint64_t z0a = __ballot(va && ka == 0u), z0b = __ballot(vb && kb == 0u);
            const uint32_t t = (uint32_t)(__builtin_popcountll(z0a) + __builtin_popcountll(z0b));
            if (t == 0u || t >= 64u) continue;          // wave-uniform
            const uint64_t o1a = __ballot(va && ka == 1u), o1b = __ballot(vb && kb == 1u);
            const uint32_t r1a = (uint32_t)__builtin_popcountll(o1a & below);
            const uint32_t r1b = (uint32_t)(__builtin_popcountll(o1a) + __builtin_popcountll(o1b & below));
            const uint32_t fill = 64u - t;
            const bool fa = va && (ka == 0u || r1a < fill);   // to the first chunk
            const bool fb = vb && (kb == 0u || r1b < fill);
            const uint64_t fma = __ballot(fa), fmb = __ballot(fb);
            const uint64_t sma = __ballot(va && !fa), smb = __ballot(vb && !fb);
            const uint32_t pa = fa ? (uint32_t)__builtin_popcountll(fma & below)
                                   : 64u + (uint32_t)__builtin_popcountll(sma & below);
            const uint32_t pb = fb ? (uint32_t)(__builtin_popcountll(fma) + __builtin_popcountll(fmb & below))
                                   : 64u + (uint32_t)(__builtin_popcountll(sma) + __builtin_popcountll(smb & below));
            if (va) order[s0 + pa] = ca;
            if (vb) order[s0 + pb] = cb;
        }
    }
}

}  // namespace

// ------------------------------------------------------------ variant table


namespace {
struct VariantEntry {
    uint32_t lanes, words;
    const void *fn;
    const void *persistent_fn;
    const void *cut_fn;              // threshold mode with the cut-off, or null
    const void *cut_persistent_fn;
    uint32_t block;                  // threads per workgroup
};
template <int L, int W>
constexpr const void *cut_kernel(bool persistent) {
    if constexpr (kHasCut<L, W>) {
        return persistent ? reinterpret_cast<const void *>(&verify_persistent_kernel<L, W, true>)
                          : reinterpret_cast<const void *>(&verify_kernel<L, W, true>);
    } else {
        return nullptr;
    }
}
#define MYERS_VARIANT_ENTRY(L, W)                                                                \
    {L, W, reinterpret_cast<const void *>(&verify_kernel<L, W, false>),                         \
     reinterpret_cast<const void *>(&verify_persistent_kernel<L, W, false>), cut_kernel<L, W>(false), \
     cut_kernel<L, W>(true), 64u},
#define MYERS_WIDE_ENTRY(Q, W)                                                                   \
    {64u * (Q), W, reinterpret_cast<const void *>(&verify_wide_kernel<Q, W>),                   \
     reinterpret_cast<const void *>(&verify_wide_persistent_kernel<Q, W>), nullptr, nullptr, 64u * (Q)},
const VariantEntry kVariants[] = {MYERS_VARIANT_LIST(MYERS_VARIANT_ENTRY) MYERS_WIDE_LIST(MYERS_WIDE_ENTRY)};
constexpr uint32_t kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
static_assert(kNumVariants < MYERS_MAX_VARIANTS, "variant table too large (the last queue row is the long-read counter)");
#define MYERS_COUNT_LONG(L, W) +((L) >= 4 && (L) <= 64 ? 1 : 0)
static_assert(0 MYERS_VARIANT_LIST(MYERS_COUNT_LONG) <= MYERS_LONG_MAX,
              "one long-read plan can hold a bin of every long-read variant");
#define MYERS_WS_U4(L, W) (uint32_t)((sizeof(Workspace<L, W>) + 15u) / 16u),
const uint32_t kWorkspaceU4[] = {MYERS_VARIANT_LIST(MYERS_WS_U4)};
}  // namespace

extern "C" uint32_t myers_num_variants(void) { return kNumVariants; }

extern "C" int myers_variant_desc(uint32_t v, myers_variant_desc_t *out) {
    if (v >= kNumVariants || out == nullptr) return -1;
    const uint32_t L = kVariants[v].lanes, W = kVariants[v].words;
    out->lanes = L;
    out->words = W;
    out->total_words = L * W;
    out->slots = L <= 64u ? 64u / L : 1u;
    out->block = kVariants[v].block;
    hipFuncAttributes attr;
    out->lds_bytes = hipFuncGetAttributes(&attr, kVariants[v].fn) == hipSuccess ? (uint32_t)attr.sharedSizeBytes : 0u;
    return 0;
}

extern "C" int myers_variant_occupancy(uint32_t v, int *wg_per_cu) {
    if (v >= kNumVariants || wg_per_cu == nullptr) return -1;
    int blocks = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kVariants[v].fn, (int)kVariants[v].block, 0);
    if (e != hipSuccess) return (int)e;
    *wg_per_cu = blocks;
    return 0;
}

extern "C" int myers_variant_occupancy_persistent(uint32_t v, int *wg_per_cu) {
    if (v >= kNumVariants || wg_per_cu == nullptr) return -1;
    int blocks = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kVariants[v].persistent_fn,
                                                                (int)kVariants[v].block, 0);
    if (e != hipSuccess) return (int)e;
    *wg_per_cu = blocks;
    return 0;
}

extern "C" int myers_launch_verify(uint32_t v, const myers_verify_args_t *args, uint32_t grid,
                                   int persistent, hipStream_t stream) {
    if (v >= kNumVariants || args == nullptr || grid == 0) return -1;
    if (persistent && (args->bin_ranges == nullptr || args->queue_heads == nullptr)) return -1;
    void *kargs[] = {const_cast<myers_verify_args_t *>(args)};
    const VariantEntry &ve = kVariants[v];
    const bool cut = args->threshold >= 0.0f && ve.cut_fn != nullptr;
    const void *fn = cut ? (persistent ? ve.cut_persistent_fn : ve.cut_fn) : (persistent ? ve.persistent_fn : ve.fn);
    hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(ve.block), kargs, 0, stream);
    return (int)e;
}

extern "C" int myers_variant_long_ok(uint32_t v, uint32_t *ws_u4) {
    constexpr uint32_t kOneWave = sizeof(kWorkspaceU4) / sizeof(kWorkspaceU4[0]);
    if (v >= kOneWave || kVariants[v].lanes < 4u || kVariants[v].lanes > 64u) return 0;
    if (ws_u4) *ws_u4 = kWorkspaceU4[v];
    return 1;
}

extern "C" uint32_t myers_long_wide_lds_bytes(void) { return (uint32_t)sizeof(LongWide); }

namespace {
#define MYERS_QUEUE_CLASS_OF(L, W) kQueueClass<L, W>,
const int kQueueClassOf[] = {MYERS_VARIANT_LIST(MYERS_QUEUE_CLASS_OF)};
const void *queue_kernel(uint32_t cls, bool cut) {
    if (cls == MYERS_QUEUE_ONE_LANE)
        return cut ? reinterpret_cast<const void *>(&verify_queue_kernel<MYERS_QUEUE_ONE_LANE, true>)
                   : reinterpret_cast<const void *>(&verify_queue_kernel<MYERS_QUEUE_ONE_LANE, false>);
    return cut ? reinterpret_cast<const void *>(&verify_queue_kernel<MYERS_QUEUE_MULTI_LANE, true>)
               : reinterpret_cast<const void *>(&verify_queue_kernel<MYERS_QUEUE_MULTI_LANE, false>);
}
}  // namespace

extern "C" int myers_variant_queue_class(uint32_t v, uint32_t *ws_u4) {
    constexpr uint32_t kOneWave = sizeof(kQueueClassOf) / sizeof(kQueueClassOf[0]);
    if (v >= kOneWave || kQueueClassOf[v] < 0) return -1;
    // the multi-lane class's workgroups keep one uint4 of hand-over words first
    if (ws_u4) *ws_u4 = kWorkspaceU4[v] + (kQueueClassOf[v] == MYERS_QUEUE_MULTI_LANE ? 1u : 0u);
    return kQueueClassOf[v];
}

extern "C" uint32_t myers_queue_waves(uint32_t cls) {
    return cls == (uint32_t)MYERS_QUEUE_MULTI_LANE ? (uint32_t)kQueueWaves<MYERS_QUEUE_MULTI_LANE>
                                                   : (uint32_t)kQueueWaves<MYERS_QUEUE_ONE_LANE>;
}

extern "C" int myers_queue_occupancy(uint32_t cls, int cut, uint32_t lds_bytes, int *wg_per_cu) {
    if (cls >= (uint32_t)MYERS_QUEUE_CLASSES || wg_per_cu == nullptr) return -1;
    int blocks = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, queue_kernel(cls, cut != 0),
                                                                64 * (int)myers_queue_waves(cls), (size_t)lds_bytes);
    if (e != hipSuccess) return (int)e;
    *wg_per_cu = blocks;
    return 0;
}

extern "C" int myers_launch_queue(const myers_verify_args_t *args, const myers_queue_plan_t *plan, uint32_t grid,
                                  hipStream_t stream) {
    if (args == nullptr || plan == nullptr || grid == 0 || plan->cls >= (uint32_t)MYERS_QUEUE_CLASSES ||
        plan->n == 0 || plan->n > (uint32_t)MYERS_QUEUE_MAX || args->bin_ranges == nullptr ||
        args->queue_heads == nullptr)
        return -1;
    uint32_t need = 0;
    for (uint32_t j = 0; j < plan->n; j++) {
        uint32_t u4 = 0;
        const uint32_t v = plan->variant[j];
        if (myers_variant_queue_class(v, &u4) != (int)plan->cls ||
            plan->key[j] != ((kVariants[v].lanes << 8) | kVariants[v].words))
            return -1;
        need = u4 > need ? u4 : need;
    }
    if (plan->ws_u4 < need || plan->ws_u4 * 16u > 64u * 1024u) return -1;
    void *kargs[] = {const_cast<myers_verify_args_t *>(args), const_cast<myers_queue_plan_t *>(plan)};
    return (int)hipLaunchKernel(queue_kernel(plan->cls, args->threshold >= 0.0f), dim3(grid),
                                dim3(64u * myers_queue_waves(plan->cls)), kargs, (size_t)plan->ws_u4 * 16u, stream);
}

extern "C" int myers_launch_long(const myers_verify_args_t *args, const myers_long_plan_t *plan, uint32_t grid,
                                 hipStream_t stream) {
    if (args == nullptr || plan == nullptr || grid == 0 || plan->n > (uint32_t)MYERS_LONG_MAX ||
        plan->n_wide > (uint32_t)MYERS_LONG_MAX || plan->n + plan->n_wide == 0 ||
        args->bin_ranges == nullptr || args->queue_heads == nullptr)
        return -1;
    uint32_t need = 0;
    for (uint32_t j = 0; j < plan->n_wide; j++) {
        const uint32_t v = plan->wide_variant[j], w = plan->wide_words[j];
        if (!myers_variant_long_ok(v, nullptr) || (w != 2u && w != 3u) ||
            kVariants[v].lanes * kVariants[v].words > 64u * MYERS_LONG_WAVES * w)
            return -1;
        need = (uint32_t)((sizeof(LongWide) + 16u * MYERS_LONG_WAVES - 1u) / (16u * MYERS_LONG_WAVES));
    }
    for (uint32_t j = 0; j < plan->n; j++) {
        uint32_t u4 = 0;
        if (!myers_variant_long_ok(plan->variant[j], &u4) ||
            plan->key[j] != ((kVariants[plan->variant[j]].lanes << 8) | kVariants[plan->variant[j]].words))
            return -1;
        need = u4 > need ? u4 : need;
    }
    if (plan->ws_u4 < need) return -1;
    const size_t lds = (size_t)plan->ws_u4 * 16u * MYERS_LONG_WAVES;
    if (lds > 64u * 1024u) {
        // above 64 KiB a launch needs the kernel's limit raised (gfx950: 160 KiB per workgroup)
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&verify_long_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    void *kargs[] = {const_cast<myers_verify_args_t *>(args), const_cast<myers_long_plan_t *>(plan)};
    return (int)hipLaunchKernel(reinterpret_cast<const void *>(&verify_long_kernel), dim3(grid),
                                dim3(64u * MYERS_LONG_WAVES), kargs, lds, stream);
}

// ------------------------------------------------- static device workspace
// The allocation-free entry point's workspace.  In host code the name
// `scratch` stands for the variable, not a device address: the address on
// each GPU comes from hipGetSymbolAddress, resolved once per device and kept.
static __device__ __align__(256) uint32_t scratch[MYERS_SCRATCH_WORDS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices];
}  // namespace

extern "C" int myers_scratch_init(int device, uint32_t **ptr) {
    if (device < 0 || device >= kMaxDevices || ptr == nullptr) return (int)hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device] == nullptr) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return (int)e;
        if (current != device) return (int)hipErrorInvalidDevice;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (p == nullptr) return (int)hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<uint32_t *>(p);
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}

extern "C" int myers_variant_serve_ok(uint32_t v, uint32_t *ws_u4) {
    if (myers_variant_queue_class(v, nullptr) < 0) return 0;
    if (ws_u4) *ws_u4 = kWorkspaceU4[v];                // one-wave workgroups: no hand-over words
    return 1;
}

extern "C" int myers_serve_occupancy(int cut, uint32_t lds_bytes, int *wg_per_cu) {
    if (wg_per_cu == nullptr) return -1;
    const void *fn = cut ? reinterpret_cast<const void *>(&verify_serve_kernel<true>)
                         : reinterpret_cast<const void *>(&verify_serve_kernel<false>);
    int blocks = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, 64, (size_t)lds_bytes);
    if (e != hipSuccess) return (int)e;
    *wg_per_cu = blocks;
    return 0;
}

extern "C" int myers_launch_serve(const myers_serve_args_t *args, int cut, uint32_t grid, uint32_t lds_bytes,
                                  hipStream_t stream) {
    if (args == nullptr || grid == 0 || lds_bytes > 64u * 1024u || args->dev == nullptr || args->done == nullptr ||
        args->simd_busy == nullptr ||
        ((args->ring_mask + 1u) & args->ring_mask) != 0u || ((args->req_mask + 1u) & args->req_mask) != 0u)
        return -1;
    const void *fn = cut ? reinterpret_cast<const void *>(&verify_serve_kernel<true>)
                         : reinterpret_cast<const void *>(&verify_serve_kernel<false>);
    void *kargs[] = {const_cast<myers_serve_args_t *>(args)};
    return (int)hipLaunchKernel(fn, dim3(grid), dim3(64), kargs, (size_t)lds_bytes, stream);
}

extern "C" uint64_t myers_probe_variants(void) {
    uint64_t mask = 0;
    for (uint32_t v = 0; v < kNumVariants && v < 64u; v++)
        if (kVariants[v].lanes == 1u && kVariants[v].cut_fn != nullptr) mask |= 1ull << v;
    return mask;
}

extern "C" int myers_launch_probe(const myers_verify_args_t *args, uint32_t num_cands, const uint8_t *bin_of_words,
                                  uint32_t max_words, uint64_t probe_mask, hipStream_t stream) {
    if (args == nullptr || args->results == nullptr || num_cands == 0u) return -1;
    hipLaunchKernelGGL(probe_kernel, dim3((num_cands + 255u) / 256u), dim3(256), 0, stream, *args, num_cands,
                       bin_of_words, max_words, probe_mask);
    return (int)hipGetLastError();
}

extern "C" int myers_launch_probe_order(const uint32_t *ranges, uint32_t num_variants, uint64_t probe_mask,
                                        uint32_t num_cands, uint32_t *order, const void *results,
                                        hipStream_t stream) {
    if (ranges == nullptr || order == nullptr || results == nullptr || num_variants > MYERS_MAX_VARIANTS) return -1;
    uint32_t grid = num_cands / kProbeSpan + 1u;
    if (grid > 4096u) grid = 4096u;
    hipLaunchKernelGGL(probe_order_kernel, dim3(grid), dim3(64), 0, stream, ranges, num_variants, probe_mask, order,
                       static_cast<const uint2 *>(results));
    return (int)hipGetLastError();
}

extern "C" int myers_launch_binning(const uint32_t *qinfo, uint32_t num_queries, const void *cands,
                                    uint32_t num_cands,
                                    const uint8_t *bin_of_words, uint32_t max_words,
                                    uint32_t num_variants, uint32_t *block_counts, uint32_t *ranges,
                                    uint32_t *order, uint32_t *err, void *results, uint32_t *queue_heads,
                                    hipStream_t stream) {
    if (num_variants == 0 || num_variants > MYERS_MAX_VARIANTS) return -1;
    const uint2 *qi = reinterpret_cast<const uint2 *>(qinfo);
    const CandRec *cr = static_cast<const CandRec *>(cands);
    uint32_t grid = (num_cands + 1023u) / 1024u;           // about 4 rounds of 256 per block
    if (grid > (uint32_t)MYERS_BIN_BLOCKS) grid = (uint32_t)MYERS_BIN_BLOCKS;
    if (grid == 0u) grid = 1u;
    hipLaunchKernelGGL(bin_reset_kernel, dim3(1), dim3(64), 0, stream, err);
    hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(256), 0, stream, qi, num_queries, cr, num_cands,
                       bin_of_words, max_words, num_variants, block_counts, err);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(num_variants), dim3(256), 0, stream, block_counts, grid, ranges);
    hipLaunchKernelGGL(bin_base_kernel, dim3(1), dim3(64), 0, stream, num_variants, ranges, queue_heads);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid), dim3(256), 0, stream, qi, num_queries, cr, num_cands,
                       bin_of_words, max_words, num_variants, block_counts, ranges, order,
                       static_cast<uint2 *>(results));
    return (int)hipGetLastError();
}
