/*
 * CPU oracle for the candidate-verification path -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, as the checker.  See myers_oracle.h for what each function restates
 * and how it is pinned.
 */
#include "myers_oracle.h"

#include <stdlib.h>
#include <string.h>

#define HIGH_BIT_32 0x80000000u
#define ALL_N_WORD  0x44444444u

uint32_t oracle_window_len(uint32_t m, float distance)
{
    /* src-api/coreMyersGPU-CC35.cu:131 and src/myers-4bits.c:163: the
     * uint32 query size times a float factor, truncated to uint32.
     * volatile keeps each step a float32 rounding. */
    volatile float factor = 1.0f + 2.0f * distance;
    volatile float prod = (float)m * factor;
    return (uint32_t)prod;
}

uint32_t oracle_threshold(uint32_t m, float e)
{
    /* src/checksum-2.c:88: `uint threshold = distance * querySize;` with a
     * float distance: the uint converts to float, the product is a float32
     * rounding, the assignment truncates. */
    volatile float prod = (float)m * e;
    return (uint32_t)prod;
}

void oracle_apply_threshold(const oracle_qinfo_t *qinfo, const oracle_cand_t *cands, uint64_t n, float e,
                            oracle_result_t *results)
{
    for (uint64_t i = 0; i < n; i++) {
        if (results[i].score == 0xFFFFFFFFu) continue;          /* skipped */
        const uint32_t m = qinfo[cands[i].query].size;
        if (results[i].score > oracle_threshold(m, e)) {
            results[i].column = 0xFFFFFFFFu;
            results[i].score = 0xFFFFFFFFu;
        }
    }
}

/* Nibble idx of the packed text (src/genReference-4bits.c:43-56 layout). */
static inline uint32_t text_code(const uint32_t *ref, uint64_t nw, uint64_t idx)
{
    uint64_t k = idx >> 3;
    uint32_t word = (k < nw) ? ref[k] : ALL_N_WORD;
    return (word >> ((uint32_t)(idx & 7u) * 4u)) & 0xFu;
}

/* The reference indexes the Peq row with (nibble & 7)
 * (src/myers-4bits.c:109, coreMyersGPU-CC35.cu:170); rows 5..7 do not
 * exist and the API kernel's selectEq falls back to row 0
 * (coreMyersGPU-CC35.cu:82-95). */
static inline uint32_t peq_row(uint32_t nibble)
{
    uint32_t c = nibble & 7u;
    return (c > 4u) ? 0u : c;
}

/*
 * One candidate with the block recurrence of src/myers-4bits.c:111-135
 * (u32 blocks, int8 horizontal carry between blocks), the window rules of
 * src/myers-4bits-padding.c:96-146 (aligned) or src/myers-4bits.c:100-147
 * (exact), and the skip rule of src/myers-4bits.c:92.
 */
static void verify_candidate(const uint32_t *ref, uint64_t nw, uint64_t size,
                             const oracle_peq_entry_t *qpeq, uint32_t m, uint64_t pos,
                             uint32_t n, int mode, uint32_t *Pv, uint32_t *Mv,
                             oracle_result_t *res)
{
    res->column = 0;
    res->score = 0xFFFFFFFFu;                          /* src/myers-4bits.c:254-255 */
    if (!((pos < size) && ((size - pos) > (uint64_t)n)))
        return;

    const uint32_t words = (m + 31u) / 32u;            /* src/myers-4bits.c:162 */
    const uint32_t final_mask = ((m % 32u) == 0u) ? HIGH_BIT_32 : (1u << ((m % 32u) - 1u));
    for (uint32_t w = 0; w < words; w++) {
        Pv[w] = 0xFFFFFFFFu;
        Mv[w] = 0u;
    }

    uint64_t start;
    uint32_t ncols;
    if (mode == ORACLE_MODE_ALIGNED) {
        /* numEntriesPerCandidate: src/myers-4bits-padding.c:163 */
        uint32_t entries = (n / 8u) + ((n % 8u) ? 2u : 1u);
        start = (pos / 8u) * 8u;
        ncols = 8u * entries;
    } else {
        start = pos;
        ncols = n;
    }

    int32_t score = (int32_t)m, min_score = (int32_t)m;
    uint32_t min_col = 0;
    for (uint32_t j = 0; j < ncols; j++) {
        const uint32_t c = peq_row(text_code(ref, nw, start + j));
        int8_t carry = 0, next;
        for (uint32_t w = 0; w < words; w++) {
            uint32_t pv = Pv[w], mv = Mv[w];
            uint32_t eq = qpeq[w / 4u].bitmap[c][w % 4u];
            uint32_t mask = (w + 1u == words) ? final_mask : HIGH_BIT_32;
            uint32_t xv = eq | mv;
            eq |= (uint32_t)((carry >> 1) & 1);
            uint32_t xh = (((eq & pv) + pv) ^ pv) | eq;
            uint32_t ph = mv | ~(xh | pv);
            uint32_t mh = pv & xh;
            next = (int8_t)(((ph & mask) != 0u) - ((mh & mask) != 0u));
            ph <<= 1;
            mh <<= 1;
            mh |= (uint32_t)((carry >> 1) & 1);
            ph |= (uint32_t)((carry + 1) >> 1);
            carry = next;
            Pv[w] = mh | ~(xv | ph);
            Mv[w] = ph & xv;
        }
        score += carry;
        if (score < min_score) {                        /* first minimum wins */
            min_score = score;
            min_col = j;
        }
    }
    res->column = min_col;
    res->score = (uint32_t)min_score;
}

/*
 * One candidate with GEMTools' block BPM over 64-bit words
 * (GEMTools/src/gt_map_align_bpm_generic.c:26-121): per block the
 * horizontal delta leaving the block's last query row (its level mask) is
 * the next block's hin, injected as Eq |= MHin, Ph |= PHin, Mh |= MHin, and
 * each block keeps the score of its last row.  Window and skip rules as
 * verify_candidate.  A second, independent formulation of the same DP, used
 * as a CPU baseline next to the reference's u32 code.
 */
static void verify_candidate_bpm64(const uint32_t *ref, uint64_t nw, uint64_t size,
                                   const oracle_peq_entry_t *qpeq, uint32_t m, uint64_t pos, uint32_t n,
                                   int mode, uint64_t *work, oracle_result_t *res)
{
    res->column = 0;
    res->score = 0xFFFFFFFFu;
    if (!((pos < size) && ((size - pos) > (uint64_t)n)))
        return;
    if (m == 0u) {                             /* empty query: D[0][j] = 0 */
        res->score = 0u;
        return;
    }
    const uint32_t blocks = (m + 63u) / 64u;
    uint64_t *peq = work;                      /* [5][blocks] */
    uint64_t *P = peq + 5u * blocks, *M = P + blocks, *mask = M + blocks;
    int64_t *score = (int64_t *)(mask + blocks);
    for (uint32_t c = 0; c < 5u; c++)
        for (uint32_t b = 0; b < blocks; b++) {
            const uint32_t w0 = 2u * b, w1 = 2u * b + 1u;
            const uint64_t lo = qpeq[w0 / 4u].bitmap[c][w0 % 4u];
            const uint64_t hi = (w1 < (m + 31u) / 32u) ? qpeq[w1 / 4u].bitmap[c][w1 % 4u] : 0u;
            peq[c * blocks + b] = lo | (hi << 32);
        }
    for (uint32_t b = 0; b < blocks; b++) {
        const uint32_t rows = (b + 1u == blocks) ? m - 64u * b : 64u;
        mask[b] = 1ull << (rows - 1u);
        P[b] = ~0ull;
        M[b] = 0u;
        score[b] = (b ? score[b - 1] : 0) + (int64_t)rows;
    }
    uint64_t start;
    uint32_t ncols;
    if (mode == ORACLE_MODE_ALIGNED) {
        start = (pos / 8u) * 8u;
        ncols = 8u * ((n / 8u) + ((n % 8u) ? 2u : 1u));
    } else {
        start = pos;
        ncols = n;
    }
    int64_t min_score = INT64_MAX;
    uint32_t min_col = 0;
    for (uint32_t j = 0; j < ncols; j++) {
        const uint32_t c = peq_row(text_code(ref, nw, start + j));
        uint64_t phin = 0, mhin = 0;
        for (uint32_t b = 0; b < blocks; b++) {
            const uint64_t pv = P[b], mv = M[b], eq = peq[c * blocks + b];
            const uint64_t xv = eq | mv;
            const uint64_t eqh = eq | mhin;
            const uint64_t xh = (((eqh & pv) + pv) ^ pv) | eqh;
            uint64_t ph = mv | ~(xh | pv);
            uint64_t mh = pv & xh;
            const uint64_t phout = (ph & mask[b]) != 0u, mhout = (mh & mask[b]) != 0u;
            ph = (ph << 1) | phin;
            mh = (mh << 1) | mhin;
            P[b] = mh | ~(xv | ph);
            M[b] = ph & xv;
            score[b] += (int64_t)phout - (int64_t)mhout;
            phin = phout;
            mhin = mhout;
        }
        if (score[blocks - 1] < min_score) {
            min_score = score[blocks - 1];
            min_col = j;
        }
    }
    /* the reference's convention when no column goes below m: (m, 0) */
    if (min_score >= (int64_t)m) {
        min_score = m;
        min_col = 0;
    }
    res->column = min_col;
    res->score = (uint32_t)min_score;
}

void oracle_bpm64_batch(const uint32_t *ref_words, uint64_t ref_num_words, uint64_t ref_size,
                        const oracle_peq_entry_t *peq, const oracle_qinfo_t *qinfo,
                        const oracle_cand_t *cands, uint32_t num_cands, float distance, int mode,
                        oracle_result_t *results)
{
    uint32_t max_m = 0;
    for (uint32_t i = 0; i < num_cands; i++)
        if (qinfo[cands[i].query].size > max_m) max_m = qinfo[cands[i].query].size;
    const uint32_t max_blocks = (max_m + 63u) / 64u + 1u;
    #pragma omp parallel
    {
        uint64_t *work = (uint64_t *)malloc(sizeof(uint64_t) * 9u * max_blocks);
        #pragma omp for schedule(static)
        for (uint32_t i = 0; i < num_cands; i++) {
            const oracle_qinfo_t qi = qinfo[cands[i].query];
            verify_candidate_bpm64(ref_words, ref_num_words, ref_size, peq + qi.pos_entry, qi.size,
                                   cands[i].position, oracle_window_len(qi.size, distance), mode, work,
                                   &results[i]);
        }
        free(work);
    }
}

void oracle_myers_batch(const uint32_t *ref_words, uint64_t ref_num_words, uint64_t ref_size,
                        const oracle_peq_entry_t *peq, const oracle_qinfo_t *qinfo,
                        const oracle_cand_t *cands, uint32_t num_cands, float distance,
                        int mode, uint32_t max_query_len, oracle_result_t *results)
{
    uint32_t max_m = 0;
    for (uint32_t i = 0; i < num_cands; i++) {
        uint32_t m = qinfo[cands[i].query].size;
        if (m > max_m) max_m = m;
    }
    const uint32_t max_words = (max_m + 31u) / 32u + 1u;

    #pragma omp parallel
    {
        uint32_t *Pv = (uint32_t *)malloc(sizeof(uint32_t) * max_words);
        uint32_t *Mv = (uint32_t *)malloc(sizeof(uint32_t) * max_words);
        /* src/myers-4bits.c:168 */
        #pragma omp for schedule(static)
        for (int64_t i = 0; i < (int64_t)num_cands; i++) {
            const oracle_qinfo_t qi = qinfo[cands[i].query];
            if (max_query_len != 0u && qi.size > max_query_len) {
                results[i].column = 0;
                results[i].score = 0xFFFFFFFFu;
                continue;
            }
            uint32_t n = oracle_window_len(qi.size, distance);
            verify_candidate(ref_words, ref_num_words, ref_size, peq + qi.pos_entry, qi.size,
                             cands[i].position, n, mode, Pv, Mv, &results[i]);
        }
        free(Pv);
        free(Mv);
    }
}

void oracle_pack_reference(const uint8_t *codes, uint64_t n, uint32_t *out)
{
    uint64_t nw = (n + 7u) / 8u;
    for (uint64_t k = 0; k < nw; k++) {
        uint32_t word = 0;
        for (uint32_t i = 0; i < 8u; i++) {
            uint64_t idx = 8u * k + i;
            uint32_t c = (idx < n) ? codes[idx] : 4u;
            if (c > 3u) c = 4u;                         /* base2number default */
            word |= c << (4u * i);
        }
        out[k] = word;
    }
}

void oracle_build_peq(const uint8_t *codes, uint32_t m, oracle_peq_entry_t *out)
{
    uint32_t entries = (m + 127u) / 128u;
    memset(out, 0, sizeof(oracle_peq_entry_t) * entries);
    for (uint32_t i = 0; i < m; i++) {
        uint32_t b = codes[i] > 3u ? 4u : codes[i];
        out[i / 128u].bitmap[b][(i % 128u) / 32u] |= 1u << (i % 32u);
    }
}

void oracle_myers_trace(const uint8_t *query, uint32_t m, const uint8_t *text, uint32_t n,
                        int32_t *scores, oracle_result_t *result)
{
    uint32_t entries = (m + 127u) / 128u;
    uint32_t words = (m + 31u) / 32u;
    oracle_peq_entry_t *peq = (oracle_peq_entry_t *)calloc(entries ? entries : 1u, sizeof(*peq));
    uint32_t *Pv = (uint32_t *)malloc(sizeof(uint32_t) * (words + 1u));
    uint32_t *Mv = (uint32_t *)malloc(sizeof(uint32_t) * (words + 1u));
    oracle_build_peq(query, m, peq);
    const uint32_t final_mask = ((m % 32u) == 0u) ? HIGH_BIT_32 : (1u << ((m % 32u) - 1u));
    for (uint32_t w = 0; w < words; w++) { Pv[w] = 0xFFFFFFFFu; Mv[w] = 0u; }
    int32_t score = (int32_t)m, min_score = (int32_t)m;
    uint32_t min_col = 0;
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t c = peq_row(text[j]);
        int8_t carry = 0, next;
        for (uint32_t w = 0; w < words; w++) {
            uint32_t pv = Pv[w], mv = Mv[w];
            uint32_t eq = peq[w / 4u].bitmap[c][w % 4u];
            uint32_t mask = (w + 1u == words) ? final_mask : HIGH_BIT_32;
            uint32_t xv = eq | mv;
            eq |= (uint32_t)((carry >> 1) & 1);
            uint32_t xh = (((eq & pv) + pv) ^ pv) | eq;
            uint32_t ph = mv | ~(xh | pv);
            uint32_t mh = pv & xh;
            next = (int8_t)(((ph & mask) != 0u) - ((mh & mask) != 0u));
            ph <<= 1;
            mh <<= 1;
            mh |= (uint32_t)((carry >> 1) & 1);
            ph |= (uint32_t)((carry + 1) >> 1);
            carry = next;
            Pv[w] = mh | ~(xv | ph);
            Mv[w] = ph & xv;
        }
        score += carry;
        if (scores) scores[j] = score;
        if (score < min_score) { min_score = score; min_col = j; }
    }
    result->column = min_col;
    result->score = (uint32_t)min_score;
    free(peq);
    free(Pv);
    free(Mv);
}

void oracle_dp_trace(const uint8_t *query, uint32_t m, const uint8_t *text, uint32_t n,
                     int32_t *scores, oracle_result_t *result)
{
    /* One DP column at a time: col[i] = D[i][j]. */
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (m + 1u));
    for (uint32_t i = 0; i <= m; i++) col[i] = (int32_t)i;   /* column "-1": D[i][0]=i */
    int32_t min_score = (int32_t)m;
    uint32_t min_col = 0;
    for (uint32_t j = 0; j < n; j++) {
        uint32_t t = peq_row(text[j]);
        int32_t diag = col[0];        /* D[0][j-1] */
        col[0] = 0;                   /* D[0][j] = 0 (semi-global) */
        for (uint32_t i = 1; i <= m; i++) {
            uint32_t q = query[i - 1u] > 3u ? 4u : query[i - 1u];
            int32_t sub = diag + (q == t ? 0 : 1);
            int32_t del = col[i] + 1;      /* D[i][j-1] + 1 */
            int32_t ins = col[i - 1u] + 1; /* D[i-1][j] + 1 */
            int32_t best = sub < del ? sub : del;
            best = best < ins ? best : ins;
            diag = col[i];
            col[i] = best;
        }
        if (scores) scores[j] = col[m];
        if (col[m] < min_score) { min_score = col[m]; min_col = j; }
    }
    result->column = min_col;
    result->score = (uint32_t)min_score;
    free(col);
}
