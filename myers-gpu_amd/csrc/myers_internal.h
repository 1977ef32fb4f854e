/*
 * Internal entry points shared by the C host files of libmyers_amd.so
 * (not exported in the public headers).
 */
#ifndef MYERS_INTERNAL_H
#define MYERS_INTERNAL_H

#include <stdint.h>

/* 1 if every operation enqueued on the buffer's stream has finished, 0 if
 * work is pending, negative on a HIP error. */
int myers_buffer_ready(void *myersBuffer);

/* What the resident server needs from a library instance: the device, the
 * resident reference, the buffer's threshold, and for every read length in
 * words the variant the resident kernel runs it with (the bin table's
 * choice when that variant can run there, else the cheapest one that can),
 * given as the variant's key (lanes << 8 | words) and its pull rank (0 for
 * the variant with the most words per candidate). */
#define MYERS_SERVE_MAX_WORDS 768u       /* MYERS_MAX_QUERY_SIZE / 32 */
typedef struct {
    int             device;
    int             num_cus;
    const uint32_t *d_ref;
    uint64_t        ref_words;           /* readable words, guard included */
    uint64_t        ref_size;
    float           threshold;           /* < 0: off */
    uint32_t        max_words;           /* longest read the plan covers, in words */
    uint32_t        lds_bytes;           /* largest workspace of the plan's variants */
    uint32_t        num_ranks;
    uint32_t        rank_key[64];        /* key of each rank */
    uint32_t        rank_slots[64];      /* candidates per chunk of each rank */
    uint8_t         rank_of_words[MYERS_SERVE_MAX_WORDS + 1];
} myers_serve_plan_t;
int myers_serve_plan(void *myersBuffer, myers_serve_plan_t *out);

#endif
