/*
 * C-ABI between the C host library (myers_host.c) and the gfx950 kernels
 * (myers_kernels.hip).  Plain C types only.
 */
#ifndef MYERS_KERNELS_H
#define MYERS_KERNELS_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MYERS_MAX_VARIANTS = 64 };
/* Persistent work queues: each variant's chunks are split into one part per
 * XCD, each part with its own head counter on its own 128-byte line. */
enum { MYERS_QUEUE_PARTS = 8, MYERS_QUEUE_STRIDE = 32 };
#define MYERS_QUEUE_WORDS (MYERS_MAX_VARIANTS * MYERS_QUEUE_PARTS * MYERS_QUEUE_STRIDE)

/* Kernel variant: one query bit-vector of L*W u32 words split across L
 * lanes, W words per lane: L <= 64 lanes of one wave64, or L/64 waves of one
 * workgroup for L > 64. */
typedef struct {
    uint32_t lanes;          /* L: lanes per candidate (1..64, power of two) */
    uint32_t words;          /* W: u32 words per lane */
    uint32_t total_words;    /* L*W: longest query = 32*L*W bases */
    uint32_t slots;          /* candidates per workgroup: 64 / L, 1 for L > 64 */
    uint32_t lds_bytes;      /* LDS per workgroup */
    uint32_t block;          /* threads per workgroup: 64, or L for L > 64 (L/64 waves) */
} myers_variant_desc_t;

typedef struct {
    const uint32_t *ref;         /* packed reference + guard words (all-N) */
    uint64_t        ref_words;   /* words readable at ref (incl. guard) */
    uint64_t        ref_size;    /* reference length in bases */
    const uint32_t *peq;         /* qryEntry_t[] as u32[20] */
    const uint32_t *qinfo;       /* qryInfo_t[] as u32 pairs {posEntry, size} */
    uint32_t        num_peq_entries;
    uint32_t        num_queries;
    const void     *cands;       /* candInfo_t[] (16 B) */
    const uint32_t *order;       /* slot -> candidate, NULL = identity */
    const uint32_t *bin_ranges;  /* device u32 pairs {start, count} per variant, or NULL */
    uint32_t        start;       /* slot range used when bin_ranges == NULL */
    uint32_t        count;
    uint32_t        variant;     /* row of bin_ranges */
    uint32_t        mode;        /* 0 aligned window, 1 exact-start window */
    float           distance;
    float           threshold;   /* < 0: off; else results above (uint32)(m*threshold) get
                                    {0xFFFFFFFF, 0xFFFFFFFF}; L = 1 variants cut off rows and
                                    L = 4..64 variants stop hopeless waves */
    uint32_t        chunk_factor;  /* persistent launch: chunks per queue pull */
    void           *results;     /* resEntry_t[] in caller order */
    uint32_t       *queue_heads; /* persistent launch: MYERS_QUEUE_WORDS words of part heads */
} myers_verify_args_t;

/* Variant table compiled into the library. */
uint32_t myers_num_variants(void);
int      myers_variant_desc(uint32_t v, myers_variant_desc_t *out);
/* Workgroups of variant v resident per CU (occupancy), 0 on error. */
int      myers_variant_occupancy(uint32_t v, int *wg_per_cu);

/* Verify the slots of one variant.  grid = number of one-wave workgroups.
 * Plain launch: the grid strides over chunks of `slots` candidates.
 * Persistent launch (needs bin_ranges and queue_heads): each workgroup pulls
 * chunk_factor chunks at a time from the variant's queue head until the bin
 * is drained. */
int myers_launch_verify(uint32_t v, const myers_verify_args_t *args, uint32_t grid,
                        int persistent, hipStream_t stream);
/* Resident workgroups per CU of the persistent kernel of variant v. */
int myers_variant_occupancy_persistent(uint32_t v, int *wg_per_cu);

/* Long-read bins in one launch.  Bins of one-wave variants with 4..64 lanes
 * per candidate (myers_variant_long_ok) can run together in workgroups of
 * MYERS_LONG_WAVES waves: first the plan's wide bins, each candidate on a
 * whole workgroup, then the others (chunks of the plan's bins, in plan
 * order), each wave on its own.  The first round of each goes one item to a
 * workgroup / wave, from `rotate` on, so that successive launches start their
 * longest items on other CUs; later items are pulled from one counter each
 * (queue_heads + MYERS_WIDE_HEAD / MYERS_LONG_HEAD, reset by the binning).
 * Needs bin_ranges and queue_heads. */
enum { MYERS_LONG_MAX = 24, MYERS_LONG_WAVES = 4 };
#define MYERS_LONG_HEAD ((MYERS_MAX_VARIANTS - 1) * MYERS_QUEUE_PARTS * MYERS_QUEUE_STRIDE)
#define MYERS_WIDE_HEAD (MYERS_LONG_HEAD + MYERS_QUEUE_STRIDE)
/* LDS of one workgroup-wide candidate of the long-read launch (bytes), and
 * the most words per lane it takes: reads up to 64 * MYERS_LONG_WAVES * 32 *
 * MYERS_WIDE_MAX_WORDS bases. */
enum { MYERS_WIDE_MAX_WORDS = 3 };
uint32_t myers_long_wide_lds_bytes(void);
typedef struct {
    uint32_t n;                          /* one-wave bins, in pull order */
    uint32_t ws_u4;                      /* LDS per wave in 16-byte units (max over the bins) */
    uint32_t rotate;                     /* first round starts at wave `rotate` (a multiple of
                                            MYERS_LONG_WAVES): wide items on workgroups from
                                            rotate / MYERS_LONG_WAVES on, then one-wave items on
                                            the waves after them */
    uint32_t variant[MYERS_LONG_MAX];    /* variant ids (rows of bin_ranges) */
    uint32_t key[MYERS_LONG_MAX];        /* lanes << 8 | words of each */
    /* bins run first, one candidate per workgroup (MYERS_LONG_WAVES waves of
     * 64 lanes x wide_words[j] words); the first round starts at workgroup
     * rotate / MYERS_LONG_WAVES, later ones pull queue_heads + MYERS_WIDE_HEAD */
    uint32_t n_wide;
    uint32_t wide_variant[MYERS_LONG_MAX];
    uint32_t wide_words[MYERS_LONG_MAX]; /* 2 or 3 */
} myers_long_plan_t;
/* 1 if variant v can run in the long-read launch, and its LDS per wave. */
int myers_variant_long_ok(uint32_t v, uint32_t *ws_u4);
/* grid = workgroups of MYERS_LONG_WAVES waves. */
int myers_launch_long(const myers_verify_args_t *args, const myers_long_plan_t *plan, uint32_t grid,
                      hipStream_t stream);

/* One launch over all length bins ("queue launch").  A resident grid pulls
 * chunks from one device counter over the plan's bins in plan order (longest
 * reads first, so the tail is short chunks); each chunk runs the variant of
 * its bin through a switch over the compiled variants.  Variants fall into
 * two kernel classes with their own register budgets: MYERS_QUEUE_ONE_LANE
 * (L = 1; one-wave workgroups, one chunk per item) and MYERS_QUEUE_MULTI_LANE
 * (L = 2..64 with W <= 12; workgroups of myers_queue_waves() = 2 waves that
 * take a pair of chunks per item and share one staged Peq, so that each wave
 * needs half the LDS); other variants keep per-bin launches.  The class's
 * counter (queue_heads + MYERS_QUEUE_HEAD(cls)) is reset by the binning.
 * LDS per workgroup: ws_u4 x 16 bytes (the largest workspace of the plan's
 * variants, as myers_variant_queue_class reports it).  Needs bin_ranges and
 * queue_heads. */
enum { MYERS_QUEUE_ONE_LANE = 0, MYERS_QUEUE_MULTI_LANE = 1, MYERS_QUEUE_CLASSES = 2, MYERS_QUEUE_MAX = 48 };
#define MYERS_QUEUE_HEAD(cls) (MYERS_LONG_HEAD + (2u + (uint32_t)(cls)) * MYERS_QUEUE_STRIDE)
typedef struct {
    uint32_t cls;                        /* MYERS_QUEUE_ONE_LANE or MYERS_QUEUE_MULTI_LANE */
    uint32_t n;                          /* bins, in pull order */
    uint32_t ws_u4;                      /* LDS per workgroup in 16-byte units */
    uint32_t reserved;
    uint32_t variant[MYERS_QUEUE_MAX];   /* variant ids (rows of bin_ranges) */
    uint32_t key[MYERS_QUEUE_MAX];       /* lanes << 8 | words of each */
} myers_queue_plan_t;
/* Class of variant v in the queue launch (-1: none) and its workspace. */
int myers_variant_queue_class(uint32_t v, uint32_t *ws_u4);
/* Resident workgroups per CU of a class's queue kernel with `lds_bytes` of
 * dynamic LDS (cut: the threshold-mode kernel). */
int myers_queue_occupancy(uint32_t cls, int cut, uint32_t lds_bytes, int *wg_per_cu);
/* Waves per workgroup of a class's queue kernel. */
uint32_t myers_queue_waves(uint32_t cls);
/* grid = workgroups (normally CUs x occupancy). */
int myers_launch_queue(const myers_verify_args_t *args, const myers_queue_plan_t *plan, uint32_t grid,
                       hipStream_t stream);

/* Resident serving kernel.  A grid of one-wave workgroups stays on the
 * device and verifies work items published by a host thread: each item is a
 * chunk (up to 64 / L candidates of one variant) of one request.  Requests,
 * items, results and completion tickets live in coherent pinned host memory
 * that the kernel reads and writes directly; the claim counters live in
 * device memory (dev: MYERS_SERVE_DEV_WORDS words, zeroed once).
 *   rings:   two item rings, ring 0 for the chunks of long reads (one long
 *            chain of column steps each), ring 1 for the others.  A wave
 *            takes a long item only while no other long item runs on its
 *            SIMD (a per-SIMD flag in simd_busy, keyed by the hardware wave
 *            id: two chains on one SIMD would halve each other's speed), and
 *            a short item otherwise.
 *   publish: the host writes item i of ring r to items[r][i & ring_mask]
 *            and then the count of published items to host_tail[r *
 *            MYERS_SERVE_TAIL_STRIDE] (release).
 *   claim:   a wave takes item `head` of a ring by CAS while head < dtail,
 *            the device mirror of the ring's host tail; an idle wave
 *            refreshes the mirrors from host memory, at most once per
 *            poll_ticks over the whole grid.
 *   finish:  the wave adds its item's candidates to the request's device
 *            counter; the wave that completes the request writes the
 *            request's ticket to flags[slot] after its results are visible.
 *   exit:    every wave leaves when *host_stop equals the launch's `gen` and
 *            no item is left, or after idle_ticks (its own s_memrealtime
 *            ticks, 100 MHz) in which no ring's head moved; so every launch
 *            ends by itself.
 * Item and request counters are u32 and compared modulo 2^32. */
typedef struct {
    uint32_t key;          /* lanes << 8 | words of the chunk's variant */
    uint32_t req;          /* request slot (index into reqs, below req_mask + 1) */
    uint32_t first;        /* first slot of the chunk in the request's order array */
    uint32_t count;        /* candidates in the chunk */
} myers_serve_item_t;

typedef struct {
    const uint32_t *peq;       /* qryEntry_t[] as u32[20] */
    const uint32_t *qinfo;     /* qryInfo_t[] */
    const void     *cands;     /* candInfo_t[] */
    void           *results;   /* resEntry_t[], indexed like cands */
    const uint32_t *order;     /* slot -> candidate, grouped by variant */
    uint32_t        num_peq_entries;
    uint32_t        num_queries;
    uint32_t        num_items_cands;   /* candidates in the request's items */
    uint32_t        mode;      /* 0 aligned window, 1 exact-start window */
    float           distance;
    float           threshold; /* < 0: off */
    uint64_t        ticket;    /* written to flags[slot] on completion */
} myers_serve_req_t;

/* dev words: per ring r, head at MYERS_SERVE_HEAD + 64 r and device tail
 * at MYERS_SERVE_DTAIL + 64 r; then the poll time and the stop generation.
 * simd_busy: MYERS_SERVE_SIMD_KEYS words. */
enum { MYERS_SERVE_RINGS = 2, MYERS_SERVE_LONG_RING = 0, MYERS_SERVE_SHORT_RING = 1,
       MYERS_SERVE_HEAD = 0, MYERS_SERVE_DTAIL = 32, MYERS_SERVE_POLL = 128, MYERS_SERVE_STOP = 160,
       MYERS_SERVE_DEV_WORDS = 192, MYERS_SERVE_TAIL_STRIDE = 16,
       MYERS_SERVE_SIMD_KEYS = 16384, MYERS_SERVE_LONG_LANES = 16 };

typedef struct {
    const uint32_t *ref;
    uint64_t        ref_words;
    uint64_t        ref_size;
    const myers_serve_item_t *items;   /* host memory: ring r at items + r * (ring_mask + 1) */
    const myers_serve_req_t  *reqs;    /* request table in host memory */
    const uint32_t *host_tail;         /* host memory: items published, ring r at r * MYERS_SERVE_TAIL_STRIDE */
    const uint32_t *host_stop;         /* host memory: generation asked to stop */
    uint64_t       *flags;             /* host memory: per request slot, ticket when done */
    uint32_t       *done;              /* device memory: per request slot, candidates done (zeroed) */
    uint32_t       *dev;               /* device memory: claim counters (zeroed) */
    uint32_t       *simd_busy;         /* device memory: MYERS_SERVE_SIMD_KEYS flags (zeroed) */
    uint32_t        ring_mask;         /* items ring size - 1 (power of two) */
    uint32_t        req_mask;          /* request table size - 1 (power of two) */
    uint32_t        gen;               /* this launch's generation */
    uint32_t        idle_ticks;
    uint32_t        poll_ticks;
    uint32_t        long_per_simd;     /* long items that may run on one SIMD at once */
} myers_serve_args_t;

/* 1 if variant v can run in the resident kernel (the one-lane and the
 * multi-lane queue classes), and its LDS workspace in 16-byte units.  Its
 * chunks go to the long ring when it has MYERS_SERVE_LONG_LANES lanes or
 * more per candidate. */
int myers_variant_serve_ok(uint32_t v, uint32_t *ws_u4);
int myers_serve_occupancy(int cut, uint32_t lds_bytes, int *wg_per_cu);
/* grid = one-wave workgroups; lds_bytes >= the largest workspace of the
 * variants the host will publish. */
int myers_launch_serve(const myers_serve_args_t *args, int cut, uint32_t grid, uint32_t lds_bytes,
                       hipStream_t stream);

/* Device-side length binning: a stable partition of the candidates by
 * variant.  bin_of_words[w] (w = ceil(m/32), 0..max_words) gives the variant
 * of a query of w words.  block_counts holds MYERS_BIN_WORDS u32 of
 * workspace; ranges gets {start, count} pairs; order gets the candidate of
 * every slot (input order within each variant); the queue part heads are
 * reset.  A candidate whose query is longer than max_words words gets the
 * skip sentinel in results and sets err (u32) to 1; one naming a query >=
 * num_queries gets the sentinel and sets err to 2. */
enum { MYERS_BIN_BLOCKS = 4096 };
#define MYERS_BIN_WORDS (MYERS_BIN_BLOCKS * MYERS_MAX_VARIANTS)

/* Workspace of one batch's device pipeline, in u32 words: variant ranges
 * (2 per variant) and the binning error flag in the first
 * 8 * MYERS_MAX_VARIANTS words, then the persistent queues' part heads
 * (128-byte aligned), then the binning block counts.  The slot -> candidate
 * order (one word per candidate) is separate. */
#define MYERS_WORK_RANGES 0u
#define MYERS_WORK_ERR    (2u * MYERS_MAX_VARIANTS)
#define MYERS_WORK_HEADS  (8u * MYERS_MAX_VARIANTS)
#define MYERS_WORK_BLOCKS (MYERS_WORK_HEADS + (uint32_t)MYERS_QUEUE_WORDS)
#define MYERS_WORK_WORDS  (MYERS_WORK_BLOCKS + (uint32_t)MYERS_BIN_WORDS)

/* Static device workspace of the allocation-free entry point
 * (myersVerifyDeviceNoAlloc): a file-scope __device__ array in the kernel
 * library holding MYERS_WORK_WORDS workspace words followed by the order of
 * up to MYERS_NOALLOC_MAX_CANDIDATES candidates (9.1 MB per GPU, held until
 * the process ends).  myers_scratch_init resolves its device address on
 * `device`, which must be the current device, once per device (later calls
 * return the kept address); 0 on success, else a hipError_t. */
#define MYERS_NOALLOC_MAX_CANDIDATES (1u << 21)
#define MYERS_SCRATCH_WORDS (MYERS_WORK_WORDS + MYERS_NOALLOC_MAX_CANDIDATES)
int myers_scratch_init(int device, uint32_t **ptr);
int myers_launch_binning(const uint32_t *qinfo, uint32_t num_queries, const void *cands, uint32_t num_cands,
                         const uint8_t *bin_of_words, uint32_t max_words, uint32_t num_variants,
                         uint32_t *block_counts, uint32_t *ranges, uint32_t *order, uint32_t *err,
                         void *results, uint32_t *queue_heads, hipStream_t stream);

/* Threshold mode's band probe.  The cut-off advances a wave of one-lane
 * candidates as deep as its deepest lane, and a candidate whose read occurs
 * in its window keeps most of the column's rows within the threshold; in a
 * batch where such candidates are spread over every wave, every wave runs
 * deep.  myers_launch_probe (before the binning) classes every candidate
 * whose variant v has bit v set in probe_mask: the first 32 bases of the read
 * (one bit-vector word) run over the leading min(n, 2 d m + 72) columns of
 * its window, where a read placed by the filter begins, and results[i].x
 * gets 0 when some column has D <= 8 there, else 1; other candidates get 0.
 * The args' peq, qinfo, cands, ref, mode, distance and results are used.
 * myers_launch_probe_order (after the binning) reorders each probed
 * variant's range in spans of two 64-slot chunks so that the first chunk
 * takes the span's class-0 candidates (then class 1 in slot order), keeping
 * a read's candidates in runs.  Both only order the work: the results are
 * those of the batch without them. */
int myers_launch_probe(const myers_verify_args_t *args, uint32_t num_cands, const uint8_t *bin_of_words,
                       uint32_t max_words, uint64_t probe_mask, hipStream_t stream);
int myers_launch_probe_order(const uint32_t *ranges, uint32_t num_variants, uint64_t probe_mask,
                             uint32_t num_cands, uint32_t *order, const void *results, hipStream_t stream);
/* The variants worth probing: one-lane variants with a cut-off kernel. */
uint64_t myers_probe_variants(void);

#ifdef __cplusplus
}
#endif

#endif
