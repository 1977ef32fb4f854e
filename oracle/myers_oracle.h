/*
 * CPU oracle for the candidate-verification path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (never as the product path).
 *
 * It restates the reference CPU Myers in plain C:
 *   - aligned ("padding") windows, variable-length 128-base Peq entries:
 *       src-integration/myers-reorderBuffer-padding.c:97-208
 *       src/myers-4bits-padding.c:75-179
 *   - exact-start windows: src/myers-4bits.c:75-175
 *   - result/skip semantics of the API kernel: src-api/coreMyersGPU-CC35.cu:110-240
 * plus a brute-force semi-global Levenshtein DP used to pin the above.
 *
 * How each function is pinned (checked by tests/test_oracle*.py):
 *   oracle_myers_batch    golden vectors written by the compiled reference C
 *                         oracles (tests/golden/), the debug/ KAT, and
 *                         brute-force DP on random tiny inputs
 *   oracle_myers_trace    debug/SCORES-Myers-caos-code.txt per-column KAT, DP
 *   oracle_dp_trace       closed-form cases (empty query, identical strings,
 *                         all-mismatch) and the KAT
 *   oracle_pack_reference golden .4bits.ref written by src/genReference-4bits.c
 *   oracle_build_peq      golden .gem.qry written by src-api/genCandidate-reorder.c
 *   oracle_bpm64_batch    equal to oracle_myers_batch on the golden inputs and
 *                         on random batches (two formulations of one DP)
 *   oracle_threshold      thresholds read back from src/checksum-2.c's own
 *                         counts (tests/golden/threshold_kat.json)
 */
#ifndef MYERS_ORACLE_H
#define MYERS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Record layouts of src-api/myers-interface.h:10-27 (restated, own names). */
typedef struct { uint32_t bitmap[5][4]; } oracle_peq_entry_t;               /* 80 B */
typedef struct { uint32_t query; uint32_t pad; uint64_t position; } oracle_cand_t; /* 16 B */
typedef struct { uint32_t pos_entry; uint32_t size; } oracle_qinfo_t;       /* 8 B  */
typedef struct { uint32_t column; uint32_t score; } oracle_result_t;        /* 8 B  */

enum { ORACLE_MODE_ALIGNED = 0, ORACLE_MODE_EXACT = 1 };

/* The same verification with GEMTools' block BPM over 64-bit words
 * (GEMTools/src/gt_map_align_bpm_generic.c:26-121): a CPU baseline next to
 * the reference's u32 code.  No query length limit.  OpenMP over
 * candidates. */
void oracle_bpm64_batch(const uint32_t *ref_words, uint64_t ref_num_words, uint64_t ref_size,
                        const oracle_peq_entry_t *peq, const oracle_qinfo_t *qinfo,
                        const oracle_cand_t *cands, uint32_t num_cands, float distance, int mode,
                        oracle_result_t *results);

/* Under-threshold bound (uint32)((float)m * e), the float product of
 * src/checksum-2.c:88 (float distance times uint query size). */
uint32_t oracle_threshold(uint32_t m, float e);

/* Threshold mode: a verified candidate whose score exceeds
 * oracle_threshold(m, e) becomes {0xFFFFFFFF, 0xFFFFFFFF}; under-threshold
 * and skipped ({0, 0xFFFFFFFF}) candidates are left as they are
 * (src/checksum-2.c:104-105 counts score <= threshold as a match). */
void oracle_apply_threshold(const oracle_qinfo_t *qinfo, const oracle_cand_t *cands, uint64_t n, float e,
                            oracle_result_t *results);

/* Window length n = (uint32)(m * (1 + 2*d)) evaluated in float32
 * (src-api/coreMyersGPU-CC35.cu:131, src/myers-4bits.c:163). */
uint32_t oracle_window_len(uint32_t m, float distance);

/* Verify every candidate.  ref_words holds ref_num_words packed words
 * (8 bases per u32, base i at bits [4i,4i+3]); ref_size is the base count.
 * Words past ref_num_words read as all-N (0x44444444).
 * max_query_len: candidates whose query is longer get the skip sentinel
 * (4096 reproduces src-integration/myers-reorderBuffer-padding.c:200;
 * 0 means no limit).  Uses OpenMP (static schedule) over candidates. */
void oracle_myers_batch(const uint32_t *ref_words, uint64_t ref_num_words, uint64_t ref_size,
                        const oracle_peq_entry_t *peq, const oracle_qinfo_t *qinfo,
                        const oracle_cand_t *cands, uint32_t num_cands, float distance,
                        int mode, uint32_t max_query_len, oracle_result_t *results);

/* Per-column score trace (D[m][j] for j = 0..n-1) of one exact-start
 * alignment of query codes against text codes, computed by the block
 * Myers recurrence of src/myers-4bits.c:100-147.  Returns the final
 * (min score, first argmin column) through *result. */
void oracle_myers_trace(const uint8_t *query, uint32_t m, const uint8_t *text, uint32_t n,
                        int32_t *scores, oracle_result_t *result);

/* Brute-force semi-global Levenshtein DP (D[0][j]=0, D[i][0]=i, unit
 * costs); scores[j] = D[m][j].  Code 4 (N) matches only 4; codes 5..7 are
 * treated as 0 on the text side (the API kernel's selectEq). */
void oracle_dp_trace(const uint8_t *query, uint32_t m, const uint8_t *text, uint32_t n,
                     int32_t *scores, oracle_result_t *result);

/* 4-bit packing of base codes (src/genReference-4bits.c:21-72):
 * out[k] nibble i = code of base 8k+i; codes >3 are stored as 4 (N);
 * nibbles past n are 4.  out must hold ceil(n/8) words. */
void oracle_pack_reference(const uint8_t *codes, uint64_t n, uint32_t *out);

/* Peq entries of one query (src-api/genCandidate-reorder.c:258-318):
 * ceil(m/128) entries, bitmap[b][s] bit i set iff base 32s+i has code b
 * (codes >3 set the N mask, b=4); bits past m are 0. */
void oracle_build_peq(const uint8_t *codes, uint32_t m, oracle_peq_entry_t *out);

#ifdef __cplusplus
}
#endif
#endif
