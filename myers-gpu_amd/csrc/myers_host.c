/*
 * Host side of the candidate-verification library: the buffer API of
 * src-api/myers-interface.h, re-designed for one MI355X.
 *
 * Plain C over the HIP runtime C API.  Kernels are launched through the
 * extern "C" shims of myers_kernels.hip (myers_kernels.h).
 *
 * Life of a batch (src-api/myersGPU.c:466-524 is the reference flow):
 *   send:    async H2D of Peq entries, qinfo and candidates on the buffer's
 *            stream -> kernel launches -> async D2H of the results.  The
 *            kernels write results in the caller's candidate order, so the
 *            reference's host reorder (myersGPU.c:351-422) and result
 *            scatter (myersGPU.c:500-513) do not exist here.
 *   launch:  the host scans qinfo (O(queries)) for the length bins present.
 *            One bin -> one launch over the candidates as they are.  More
 *            bins -> device-side binning (count, scan, base, scatter), then
 *            one launch per present bin in turn (plain engine), or one
 *            resident launch per bin on forked streams (persistent engine),
 *            where the long-read bins of a latency-bound batch share one
 *            launch (long_plan, myers_launch_long).
 *   receive: hipStreamSynchronize; results are in the pinned results buffer.
 *
 * All buffers of a library instance live on one device and share one copy
 * of the packed reference in HBM (the reference freed it once per buffer,
 * myersGPU.c:661).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "myers-interface.h"
#include "myers-amd.h"
#include "myers_internal.h"
#include "myers_kernels.h"

#define GUARD_WORDS      64u            /* all-N words after the reference */
#define ALL_N_WORD       0x44444444u
#define MAX_QUERY_WORDS  (MYERS_MAX_QUERY_SIZE / 32u)
#define STAGE_BYTES      (64u << 20)    /* pinned staging for the reference upload */
#define DEFAULT_BUFFER_CAP (2ull << 30)
#define EV_RING          32u            /* kernel-timing event pairs per buffer */
#define SIDE_STREAMS     4u             /* fork width of a persistent launch */

typedef struct myers_ctx {
    int       device;
    int       num_cus;
    uint32_t *d_ref;
    uint64_t  ref_words;         /* stored words */
    uint64_t  ref_words_alloc;   /* stored + guard */
    uint64_t  ref_size;          /* bases */
    uint32_t  num_variants;
    myers_variant_desc_t variants[MYERS_MAX_VARIANTS];
    int       occupancy[MYERS_MAX_VARIANTS];       /* plain kernel, workgroups per CU */
    int       occupancy_p[MYERS_MAX_VARIANTS];     /* persistent kernel */
    uint32_t  chunk_factor[MYERS_MAX_VARIANTS];    /* persistent: chunks per queue pull */
    uint32_t  long_seq;                            /* long-read launches so far (placement rotation) */
    int       queue_occ[MYERS_QUEUE_CLASSES][2];   /* queue kernel workgroups per CU, [class][cut], */
    uint32_t  queue_occ_lds[MYERS_QUEUE_CLASSES][2];  /* for this dynamic LDS size (0: not queried) */
    uint8_t   bin_of_words[MAX_QUERY_WORDS + 1];
    uint8_t  *d_bin_of_words;
    uint64_t  tune_gen;          /* bumped when the bin table or chunk factors change */
    uint32_t *d_scratch;         /* this device's static workspace (myersVerifyDeviceNoAlloc) */
    uint64_t  probe_mask;        /* threshold mode: variants whose candidates the band probe orders */
} myers_ctx_t;

typedef struct myers_buffer {
    myers_ctx_t *ctx;
    uint32_t     num_buffers;
    uint32_t     id_buffer;
    uint32_t     max_peq_entries;
    uint32_t     max_candidates;
    uint32_t     max_queries;
    hipStream_t  stream;
    qryEntry_t  *h_peq;
    qryEntry_t  *d_peq;
    qryInfo_t   *h_qinfo;
    qryInfo_t   *d_qinfo;
    candInfo_t  *h_cands;
    candInfo_t  *d_cands;
    resEntry_t  *h_results;
    resEntry_t  *d_results;
    uint32_t    *d_order;
    uint32_t    *d_work;          /* MYERS_WORK_WORDS: ranges | err | queue part heads | block counts */
    uint32_t    *h_err;           /* pinned copy of the binning error flag */
    hipEvent_t   ev_batch_start, ev_batch_end;
    /* HIP events around the verification launches of each batch, in a ring
     * so that batches enqueued back to back are all timed; folded into the
     * stats in order when the stream is synchronised. */
    hipEvent_t   ev_kernel[EV_RING][3];   /* start, end, start of the verification launches */
    uint32_t     ev_launches[EV_RING];
    uint8_t      ev_marked[EV_RING];      /* event 2 was recorded (eager batches after binning) */
    int          mark_slot;               /* slot whose event 2 enqueue_pipeline records, or -1 */
    uint32_t     ev_head, ev_pending;
    /* launches of one batch that run side by side: forked streams joined by events */
    hipStream_t  side[SIDE_STREAMS];
    hipEvent_t   ev_fork, ev_join[SIDE_STREAMS];
    /* steady-state graph of the last device-resident batch shape */
    hipGraphExec_t graph_exec;
    uint64_t     graph_key[12];
    int          graph_valid;
    int          graph_binned;
    uint32_t     graph_launches;
    int          exit_on_error;
    float        threshold;       /* < 0: off (myersSetThreshold) */
    int          last_error;
    int          pending;         /* a batch was enqueued and not yet received */
    int          binned;
    uint32_t     num_candidates;
    myersStats_t stats;
} myers_buffer_t;

static void load_default_tuning(myers_ctx_t *ctx);

/* ------------------------------------------------------------------ errors */

const char *myersGetErrorString(int error)
{
    switch (error) {
    case MYERS_SUCCESS:                return "Myers - Success";
    case MYERS_E_OPENING_FILE:         return "Myers - Error: opening file";
    case MYERS_E_READING_FILE:         return "Myers - Error: reading file";
    case MYERS_E_INSUFFICIENT_MEM_GPU: return "Myers - Error: there aren't enough GPU memory space";
    case MYERS_E_ALLOCATE_MEM:         return "Myers - Error: allocating data";
    case MYERS_E_INCOMPATIBLE_GPU:     return "Myers - Error: incompatible GPU";
    case MYERS_E_INVALID_ARGUMENT:     return "Myers - Error: invalid argument";
    case MYERS_E_CAPACITY:             return "Myers - Error: batch exceeds the buffer capacity";
    case MYERS_E_HIP:                  return "Myers - Error: HIP runtime call failed";
    case MYERS_E_QUERY_TOO_LONG:       return "Myers - Error: query longer than the supported maximum";
    case MYERS_E_NO_DEVICE:            return "Myers - Error: no HIP device";
    case MYERS_E_REFERENCE_TOO_LARGE:  return "Myers - Error: reference too large";
    default:                           return "Myers - Unknown error";
    }
}

static int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    if (e == hipSuccess) return MYERS_SUCCESS;
    fprintf(stderr, "Myers - HIP error '%s' in %s (%s:%d)\n", hipGetErrorString(e), what, file, line);
    return MYERS_E_HIP;
}
#define HIP_TRY(call) do { int _rc = hip_fail((call), #call, __FILE__, __LINE__); if (_rc) return _rc; } while (0)

/* The reference prints and exits on any error (src-api/myersGPU.c:24-47);
 * buffers created by initMyers keep that behaviour for the void calls. */
static void report(myers_buffer_t *b, int rc, const char *where)
{
    if (b) b->last_error = rc;
    if (rc != MYERS_SUCCESS && (b == NULL || b->exit_on_error)) {
        fprintf(stderr, "%s in %s\n", myersGetErrorString(rc), where);
        exit(EXIT_FAILURE);
    }
}

/* ---------------------------------------------------------------- getters */

qryEntry_t *getPEQBuffer(void *myersBuffer)        { return ((myers_buffer_t *)myersBuffer)->h_peq; }
candInfo_t *getCandidatesBuffer(void *myersBuffer) { return ((myers_buffer_t *)myersBuffer)->h_cands; }
qryInfo_t  *getPEQInfoBuffer(void *myersBuffer)    { return ((myers_buffer_t *)myersBuffer)->h_qinfo; }
resEntry_t *getResultsBuffer(void *myersBuffer)    { return ((myers_buffer_t *)myersBuffer)->h_results; }
uint32_t getMaxPEQEntries(void *myersBuffer)       { return ((myers_buffer_t *)myersBuffer)->max_peq_entries; }
uint32_t getMaxCandidates(void *myersBuffer)       { return ((myers_buffer_t *)myersBuffer)->max_candidates; }
uint32_t getMaxQueries(void *myersBuffer)          { return ((myers_buffer_t *)myersBuffer)->max_queries; }
uint32_t myersGetNumBuffers(void *myersBuffer)     { return ((myers_buffer_t *)myersBuffer)->num_buffers; }
uint64_t myersGetReferenceSize(void *myersBuffer)  { return ((myers_buffer_t *)myersBuffer)->ctx->ref_size; }
int myersGetLastError(void *myersBuffer)           { return myersBuffer ? ((myers_buffer_t *)myersBuffer)->last_error : MYERS_E_INVALID_ARGUMENT; }
void *myersGetStream(void *myersBuffer)            { return (void *)((myers_buffer_t *)myersBuffer)->stream; }

void myersDefaultConfig(myersConfig_t *cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    cfg->structSize = sizeof(*cfg);
    cfg->device = 0;
    cfg->exitOnError = 1;
    cfg->bufferBytes = 0;
}

/* ------------------------------------------------------------ bin table */

static uint32_t variant_cost(const myers_variant_desc_t *v)
{
    /* Estimated issue slots per candidate column: ~10 VALU ops per u32 word
     * plus ~12 per lane (text, score, minimum) and ~6 per lane for the
     * diagonal hand-over when L > 1.  Variants whose LDS leaves room for
     * fewer than two waves per SIMD (over 20 KB per one-wave workgroup)
     * cannot use the dual-issue rate of and/or/bitop3 and pay ~40%. */
    uint32_t cost = v->total_words * 10u + 12u * v->lanes + (v->lanes > 1u ? 6u * v->lanes : 0u);
    if (v->lds_bytes > 20u * 1024u) cost += (cost * 2u) / 5u;
    return cost;
}

/* Default length-bin table: for w query words pick the variant that covers
 * them with the least estimated VALU work per column. */
static void default_bin_table(myers_ctx_t *ctx)
{
    for (uint32_t w = 0; w <= MAX_QUERY_WORDS; w++) {
        uint32_t best = 0, best_cost = 0xFFFFFFFFu;
        for (uint32_t v = 0; v < ctx->num_variants; v++) {
            const myers_variant_desc_t *d = &ctx->variants[v];
            if (d->total_words < (w ? w : 1u)) continue;
            uint32_t cost = variant_cost(d);
            if (cost < best_cost) { best_cost = cost; best = v; }
        }
        ctx->bin_of_words[w] = (uint8_t)best;
    }
}

static uint32_t max_supported_words(const myers_ctx_t *ctx)
{
    uint32_t m = 0;
    for (uint32_t v = 0; v < ctx->num_variants; v++)
        if (ctx->variants[v].total_words > m) m = ctx->variants[v].total_words;
    return m < MAX_QUERY_WORDS ? m : MAX_QUERY_WORDS;
}

/* ------------------------------------------------------------- reference */

static int read_reference_file(const char *fn, uint32_t **words, uint64_t *num_words, uint64_t *size)
{
    FILE *fp = fopen(fn, "rb");
    if (fp == NULL) return MYERS_E_OPENING_FILE;
    uint32_t hdr[2];
    if (fread(hdr, sizeof(uint32_t), 2, fp) != 2) { fclose(fp); return MYERS_E_READING_FILE; }
    uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)hdr[0] + 1u));
    if (w == NULL) { fclose(fp); return MYERS_E_ALLOCATE_MEM; }
    if (fread(w, sizeof(uint32_t), hdr[0], fp) != hdr[0]) { free(w); fclose(fp); return MYERS_E_READING_FILE; }
    fclose(fp);
    *words = w;
    *num_words = hdr[0];
    *size = hdr[1];
    return MYERS_SUCCESS;
}

/* The kernels index reference words with signed 32-bit offsets (a lane's
 * first text word can lie before the window, and reads are clamped to
 * [0, last word]), so the stored words plus the guard must stay below 2^31:
 * 17 Gbases, four times the largest .4bits.ref header (u32 base count). */
static int check_reference_size(uint64_t num_words, uint64_t size)
{
    if (num_words + GUARD_WORDS > (uint64_t)INT32_MAX) return MYERS_E_REFERENCE_TOO_LARGE;
    if (size > num_words * 8u) return MYERS_E_INVALID_ARGUMENT;
    return MYERS_SUCCESS;
}

static int upload_reference(myers_ctx_t *ctx, const uint32_t *words, uint64_t num_words, uint64_t size)
{
    int rc0 = check_reference_size(num_words, size);
    if (rc0) return rc0;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const uint64_t bytes = (num_words + GUARD_WORDS) * sizeof(uint32_t);
    if (bytes > free_b) return MYERS_E_INSUFFICIENT_MEM_GPU;
    HIP_TRY(hipMalloc((void **)&ctx->d_ref, bytes));
    /* Stream the words through a pinned staging buffer. */
    void *stage = NULL;
    size_t stage_bytes = STAGE_BYTES;
    if (stage_bytes > num_words * sizeof(uint32_t)) stage_bytes = num_words * sizeof(uint32_t) + 4u;
    HIP_TRY(hipHostMalloc(&stage, stage_bytes, hipHostMallocDefault));
    uint64_t done = 0;
    int rc = MYERS_SUCCESS;
    while (done < num_words && rc == MYERS_SUCCESS) {
        uint64_t n = stage_bytes / sizeof(uint32_t);
        if (n > num_words - done) n = num_words - done;
        memcpy(stage, words + done, n * sizeof(uint32_t));
        rc = hip_fail(hipMemcpy(ctx->d_ref + done, stage, n * sizeof(uint32_t), hipMemcpyHostToDevice),
                      "hipMemcpy(reference)", __FILE__, __LINE__);
        done += n;
    }
    hipHostFree(stage);
    if (rc) return rc;
    HIP_TRY(hipMemsetD32((hipDeviceptr_t)(ctx->d_ref + num_words), (int)ALL_N_WORD, GUARD_WORDS));
    HIP_TRY(hipDeviceSynchronize());
    ctx->ref_words = num_words;
    ctx->ref_words_alloc = num_words + GUARD_WORDS;
    ctx->ref_size = size;
    return MYERS_SUCCESS;
}

/* ---------------------------------------------------------------- buffers */

static void free_buffer(myers_buffer_t *b)
{
    if (b == NULL) return;
    if (b->h_peq) hipHostFree(b->h_peq);
    if (b->h_qinfo) hipHostFree(b->h_qinfo);
    if (b->h_cands) hipHostFree(b->h_cands);
    if (b->h_results) hipHostFree(b->h_results);
    if (b->h_err) hipHostFree(b->h_err);
    if (b->d_peq) hipFree(b->d_peq);
    if (b->d_qinfo) hipFree(b->d_qinfo);
    if (b->d_cands) hipFree(b->d_cands);
    if (b->d_results) hipFree(b->d_results);
    if (b->d_order) hipFree(b->d_order);
    if (b->d_work) hipFree(b->d_work);
    if (b->ev_batch_start) hipEventDestroy(b->ev_batch_start);
    if (b->ev_batch_end) hipEventDestroy(b->ev_batch_end);
    for (uint32_t i = 0; i < EV_RING; i++) {
        for (uint32_t k = 0; k < 3; k++)
            if (b->ev_kernel[i][k]) hipEventDestroy(b->ev_kernel[i][k]);
    }
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    for (uint32_t i = 0; i < SIDE_STREAMS; i++) {
        if (b->side[i]) hipStreamDestroy(b->side[i]);
        if (b->ev_join[i]) hipEventDestroy(b->ev_join[i]);
    }
    if (b->ev_fork) hipEventDestroy(b->ev_fork);
    if (b->stream) hipStreamDestroy(b->stream);
    free(b);
}

static int init_buffer(myers_buffer_t *b)
{
    const size_t peq_b = (size_t)b->max_peq_entries * sizeof(qryEntry_t);
    const size_t qi_b = (size_t)b->max_queries * sizeof(qryInfo_t);
    const size_t ca_b = (size_t)b->max_candidates * sizeof(candInfo_t);
    const size_t re_b = (size_t)b->max_candidates * sizeof(resEntry_t);
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void **)&b->h_peq, peq_b, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&b->h_qinfo, qi_b, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&b->h_cands, ca_b, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&b->h_results, re_b, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&b->h_err, sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(hipMalloc((void **)&b->d_peq, peq_b));
    HIP_TRY(hipMalloc((void **)&b->d_qinfo, qi_b));
    HIP_TRY(hipMalloc((void **)&b->d_cands, ca_b));
    HIP_TRY(hipMalloc((void **)&b->d_results, re_b));
    HIP_TRY(hipMalloc((void **)&b->d_order, (size_t)b->max_candidates * sizeof(uint32_t)));
    HIP_TRY(hipMalloc((void **)&b->d_work, MYERS_WORK_WORDS * sizeof(uint32_t)));
    for (uint32_t i = 0; i < SIDE_STREAMS; i++) {
        HIP_TRY(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&b->ev_batch_start));
    HIP_TRY(hipEventCreate(&b->ev_batch_end));
    for (uint32_t i = 0; i < EV_RING; i++) {
        for (uint32_t k = 0; k < 3; k++) HIP_TRY(hipEventCreate(&b->ev_kernel[i][k]));
    }
    memset(b->h_results, 0, re_b);
    *b->h_err = 0;
    b->stats.structSize = sizeof(myersStats_t);
    b->mark_slot = -1;
    return MYERS_SUCCESS;
}

static int create_context(myers_ctx_t **out, int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MYERS_E_NO_DEVICE;
    if (device < 0 || device >= ndev) return MYERS_E_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(device));
    myers_ctx_t *ctx = (myers_ctx_t *)calloc(1, sizeof(myers_ctx_t));
    if (ctx == NULL) return MYERS_E_ALLOCATE_MEM;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { free(ctx); return MYERS_E_HIP; }
    ctx->num_cus = prop.multiProcessorCount;
    /* the allocation-free entry point's static workspace on this device */
    if (myers_scratch_init(device, &ctx->d_scratch) != 0) {
        free(ctx);
        return MYERS_E_HIP;
    }
    ctx->num_variants = myers_num_variants();
    ctx->probe_mask = myers_probe_variants();
    for (uint32_t v = 0; v < ctx->num_variants; v++) {
        myers_variant_desc(v, &ctx->variants[v]);
        int occ = 0;
        if (myers_variant_occupancy(v, &occ) != 0 || occ <= 0) occ = 1;
        ctx->occupancy[v] = occ;
        occ = 0;
        if (myers_variant_occupancy_persistent(v, &occ) != 0 || occ <= 0) occ = 1;
        ctx->occupancy_p[v] = occ;
        ctx->chunk_factor[v] = 1u;
    }
    default_bin_table(ctx);
    if (hipMalloc((void **)&ctx->d_bin_of_words, sizeof(ctx->bin_of_words)) != hipSuccess) {
        free(ctx);
        return MYERS_E_HIP;
    }
    if (hipMemcpy(ctx->d_bin_of_words, ctx->bin_of_words, sizeof(ctx->bin_of_words),
                  hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(ctx->d_bin_of_words);
        free(ctx);
        return MYERS_E_HIP;
    }
    load_default_tuning(ctx);
    *out = ctx;
    return MYERS_SUCCESS;
}

static void free_context(myers_ctx_t *ctx)
{
    if (ctx == NULL) return;
    if (ctx->d_ref) hipFree(ctx->d_ref);
    if (ctx->d_bin_of_words) hipFree(ctx->d_bin_of_words);
    free(ctx);
}

/* Capacities from the per-buffer byte budget, with the reference's formulas
 * (src-api/myersGPU.c:312-330). */
static int size_buffers(myers_ctx_t *ctx, uint32_t num_buffers, int32_t avg_query, int32_t cands_per_query,
                        uint64_t buffer_bytes, uint32_t *max_cands, uint32_t *max_queries, uint32_t *max_peq)
{
    if (avg_query <= 0) avg_query = 1024;
    if (cands_per_query <= 0) cands_per_query = 1;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    uint64_t budget = buffer_bytes;
    if (budget == 0) {
        budget = (uint64_t)((double)free_b * 0.95) / num_buffers;
        if (budget > DEFAULT_BUFFER_CAP) budget = DEFAULT_BUFFER_CAP;
    }
    if (budget * num_buffers > (uint64_t)free_b) return MYERS_E_INSUFFICIENT_MEM_GPU;
    const uint32_t entries = ((uint32_t)avg_query + 127u) / 128u;
    const double per_cand = ((double)entries * sizeof(qryEntry_t) + sizeof(qryInfo_t)) / cands_per_query
                            + sizeof(candInfo_t) + sizeof(resEntry_t) + sizeof(uint32_t);
    double mc = (double)budget / per_cand;
    if (mc > 4.0e9) mc = 4.0e9;
    *max_cands = (uint32_t)mc;
    *max_queries = *max_cands / (uint32_t)cands_per_query;
    double mp = (double)(*max_queries) * entries;
    if (mp > 4.0e9) mp = 4.0e9;
    *max_peq = (uint32_t)mp;
    (void)ctx;
    if (*max_cands < 1u || *max_queries < 1u) return MYERS_E_INSUFFICIENT_MEM_GPU;
    return MYERS_SUCCESS;
}

static int init_common(void ***out, uint32_t num_buffers, const uint32_t *words, uint64_t num_words,
                       uint64_t size, int32_t avg_query, int32_t cands_per_query, const myersConfig_t *cfg_in)
{
    if (out == NULL || num_buffers == 0 || words == NULL) return MYERS_E_INVALID_ARGUMENT;
    int rc = check_reference_size(num_words, size);      /* before any device work */
    if (rc) return rc;
    myersConfig_t cfg;
    myersDefaultConfig(&cfg);
    if (cfg_in) cfg = *cfg_in;
    *out = NULL;
    myers_ctx_t *ctx = NULL;
    rc = create_context(&ctx, cfg.device);
    if (rc) return rc;
    rc = upload_reference(ctx, words, num_words, size);
    if (rc) { free_context(ctx); return rc; }
    uint32_t mc = 0, mq = 0, mp = 0;
    rc = size_buffers(ctx, num_buffers, avg_query, cands_per_query, cfg.bufferBytes, &mc, &mq, &mp);
    if (rc) { free_context(ctx); return rc; }
    myers_buffer_t **bufs = (myers_buffer_t **)calloc(num_buffers, sizeof(myers_buffer_t *));
    if (bufs == NULL) { free_context(ctx); return MYERS_E_ALLOCATE_MEM; }
    for (uint32_t i = 0; i < num_buffers && rc == MYERS_SUCCESS; i++) {
        myers_buffer_t *b = (myers_buffer_t *)calloc(1, sizeof(myers_buffer_t));
        if (b == NULL) { rc = MYERS_E_ALLOCATE_MEM; break; }
        bufs[i] = b;
        b->ctx = ctx;
        b->num_buffers = num_buffers;
        b->id_buffer = i;
        b->max_candidates = mc;
        b->max_queries = mq;
        b->max_peq_entries = mp;
        b->exit_on_error = cfg.exitOnError;
        b->threshold = -1.0f;
        rc = init_buffer(b);
    }
    if (rc) {
        for (uint32_t i = 0; i < num_buffers; i++) free_buffer(bufs[i]);
        free(bufs);
        free_context(ctx);
        return rc;
    }
    *out = (void **)bufs;
    return MYERS_SUCCESS;
}

int myersInitFromMemory(void ***myersBuffer, uint32_t numBuffers, const uint32_t *refWords,
                        uint64_t numWords, uint64_t refSize, int32_t averageQuerySize,
                        int32_t candidatesPerQuery, const myersConfig_t *cfg)
{
    return init_common(myersBuffer, numBuffers, refWords, numWords, refSize, averageQuerySize,
                       candidatesPerQuery, cfg);
}

int myersInitEx(void ***myersBuffer, uint32_t numBuffers, const char *referenceFileName,
                int32_t averageQuerySize, int32_t candidatesPerQuery, const myersConfig_t *cfg)
{
    if (referenceFileName == NULL) return MYERS_E_INVALID_ARGUMENT;
    uint32_t *words = NULL;
    uint64_t nw = 0, size = 0;
    int rc = read_reference_file(referenceFileName, &words, &nw, &size);
    if (rc) return rc;
    rc = init_common(myersBuffer, numBuffers, words, nw, size, averageQuerySize, candidatesPerQuery, cfg);
    free(words);
    return rc;
}

int initMyers(void ***myersBuffer, uint32_t numBuffers, const char *referenceFileName,
              int32_t averageQuerySize, int32_t candidatesPerQuery)
{
    return myersInitEx(myersBuffer, numBuffers, referenceFileName, averageQuerySize,
                       candidatesPerQuery, NULL);
}

int myersEnd(void ***myersBuffer)
{
    if (myersBuffer == NULL || *myersBuffer == NULL) return MYERS_E_INVALID_ARGUMENT;
    myers_buffer_t **bufs = (myers_buffer_t **)(*myersBuffer);
    myers_ctx_t *ctx = bufs[0]->ctx;
    const uint32_t n = bufs[0]->num_buffers;
    hipSetDevice(ctx->device);
    int rc = hip_fail(hipDeviceSynchronize(), "hipDeviceSynchronize", __FILE__, __LINE__);
    for (uint32_t i = 0; i < n; i++) free_buffer(bufs[i]);
    free(bufs);
    free_context(ctx);   /* the shared reference is freed once */
    *myersBuffer = NULL;
    return rc;
}

void endMyers(void ***myersBuffer)
{
    int rc = myersEnd(myersBuffer);
    if (rc) {
        fprintf(stderr, "%s in endMyers\n", myersGetErrorString(rc));
        exit(EXIT_FAILURE);
    }
}

/* ----------------------------------------------------------------- launch */

typedef struct {
    uint64_t present;      /* bitmask of variants present */
    uint32_t num_present;
    uint32_t single;       /* the variant when num_present == 1 */
    int      too_long;
    int      sized;        /* reads[] / max_m[] below are filled in */
    uint32_t reads[MYERS_MAX_VARIANTS];   /* reads per variant */
    uint32_t max_m[MYERS_MAX_VARIANTS];   /* longest read per variant */
    int      counted;      /* cands[] below is filled in */
    uint32_t cands[MYERS_MAX_VARIANTS];   /* candidates per variant */
} bin_scan_t;

static void scan_qinfo(const myers_ctx_t *ctx, const qryInfo_t *qinfo, uint32_t nq, bin_scan_t *s)
{
    const uint32_t maxw = max_supported_words(ctx);
    memset(s, 0, sizeof(*s));
    s->sized = 1;
    for (uint32_t q = 0; q < nq; q++) {
        const uint32_t m = qinfo[q].size;
        uint32_t w = (m + 31u) >> 5;
        if (w > maxw) { s->too_long = 1; continue; }
        const uint32_t v = ctx->bin_of_words[w];
        s->present |= 1ull << v;
        s->reads[v]++;
        if (m > s->max_m[v]) s->max_m[v] = m;
    }
    for (uint32_t v = 0; v < ctx->num_variants; v++)
        if (s->present & (1ull << v)) { s->num_present++; s->single = v; }
}

static uint32_t grid_for(const myers_ctx_t *ctx, uint32_t v, uint32_t count, int persistent)
{
    const uint32_t slots = ctx->variants[v].slots;
    uint64_t chunks = ((uint64_t)count + slots - 1u) / slots;
    if (chunks == 0) chunks = 1;
    if (persistent) {
        uint64_t resident = (uint64_t)ctx->num_cus * (uint64_t)ctx->occupancy[v];
        /* keep the grid a multiple of 8 (one share per XCD) */
        resident = (resident + 7u) & ~7ull;
        if (chunks > resident) chunks = resident;
    }
    if (chunks > 0x7FFFFFFFull) chunks = 0x7FFFFFFFull;
    return (uint32_t)chunks;
}

/* Fold the oldest timed batch into the stats (waits for its end event). */
static int fold_oldest_timing(myers_buffer_t *b)
{
    const uint32_t slot = (b->ev_head + EV_RING - b->ev_pending) % EV_RING;
    int rc = hip_fail(hipEventSynchronize(b->ev_kernel[slot][1]), "hipEventSynchronize", __FILE__, __LINE__);
    if (rc) return rc;
    float ms = 0.0f, vms = 0.0f;
    if (hipEventElapsedTime(&ms, b->ev_kernel[slot][0], b->ev_kernel[slot][1]) == hipSuccess) {
        if (!b->ev_marked[slot] ||
            hipEventElapsedTime(&vms, b->ev_kernel[slot][2], b->ev_kernel[slot][1]) != hipSuccess)
            vms = ms;
        b->stats.kernelMs = ms;
        b->stats.kernelMsTotal += ms;
        b->stats.verifyMs = vms;
        b->stats.verifyMsTotal += vms;
        b->stats.kernelLaunchesTotal += b->ev_launches[slot];
        b->stats.batchesTotal++;
    }
    b->ev_pending--;
    return MYERS_SUCCESS;
}

static int timing_begin(myers_buffer_t *b, uint32_t *slot)
{
    if (b->ev_pending == EV_RING) {
        int rc = fold_oldest_timing(b);
        if (rc) return rc;
    }
    *slot = b->ev_head;
    b->ev_marked[*slot] = 0;
    return hip_fail(hipEventRecord(b->ev_kernel[*slot][0], b->stream), "hipEventRecord", __FILE__, __LINE__);
}

static int timing_end(myers_buffer_t *b, uint32_t slot, uint32_t launches)
{
    int rc = hip_fail(hipEventRecord(b->ev_kernel[slot][1], b->stream), "hipEventRecord", __FILE__, __LINE__);
    if (rc) return rc;
    b->ev_launches[slot] = launches;
    b->ev_head = (slot + 1u) % EV_RING;
    b->ev_pending++;
    return MYERS_SUCCESS;
}

typedef struct {
    const void *d_peq, *d_qinfo, *d_cands;
    void       *d_results;
    uint32_t    num_entries, num_queries, num_cands, flags;
    float       distance;
    float       threshold;
    uint32_t   *d_work;      /* MYERS_WORK_WORDS of pipeline workspace */
    uint32_t   *d_order;     /* slot -> candidate, num_cands words */
} batch_t;

/* Auto: mixed lengths go to the persistent engine.  A batch of one length bin
 * runs on the plain engine (no binning pass) unless it holds more than ~2.7
 * rounds of what the chip keeps resident (2048 candidates per CU): there the
 * queue launch's pulls leave a shorter tail than one workgroup per chunk
 * (1.1M candidates of 1,000 bases: 8.70 against 8.78 ms). */
static uint32_t resolve_engine(const myers_ctx_t *ctx, uint32_t flags, const bin_scan_t *scan, uint32_t num_cands)
{
    uint32_t engine = flags & MYERS_ENGINE_MASK;
    if (engine == MYERS_ENGINE_AUTO) {
        const int single = scan->num_present == 1 && !scan->too_long;
        const int large = (uint64_t)num_cands >= (uint64_t)ctx->num_cus * 2048u;
        engine = (single && !large) ? MYERS_ENGINE_PLAIN : MYERS_ENGINE_PERSISTENT;
    }
    return engine;
}

/* Variants present, most words first (longest candidates start first). */
static uint32_t present_by_cost(const myers_ctx_t *ctx, uint64_t present, uint32_t *out)
{
    uint32_t n = 0;
    for (uint32_t v = 0; v < ctx->num_variants; v++)
        if (present & (1ull << v)) out[n++] = v;
    for (uint32_t i = 1; i < n; i++) {
        uint32_t x = out[i], j = i;
        while (j > 0 && ctx->variants[out[j - 1]].total_words < ctx->variants[x].total_words) {
            out[j] = out[j - 1];
            j--;
        }
        out[j] = x;
    }
    return n;
}

/* Long-read bins of a persistent batch that go into one launch
 * (myers_launch_long): bins of 4..64 lanes per candidate, longest reads
 * first, when their chunks (counted from the host's candidates, or estimated
 * from the reads per bin and the batch's candidates per read) fit a few
 * rounds of one wave per SIMD.  Such bins are
 * bound by each candidate's chain of column steps; larger long-read batches
 * are throughput work and keep one launch per bin.
 *
 * A bin of reads over 256 words can instead run each candidate on a whole
 * 4-wave workgroup, which shortens its chain (measured: 20,000 bases in 2.9 ms
 * against 5.9 ms on one wave) for 4 SIMDs instead of one.  The plan widens
 * the bin with the longest estimated chain while that lowers the estimated
 * time of the launch, max(longest chain, SIMD-time of all items / SIMDs).
 * Per column step, from single-candidate runs: one wave of W words per lane
 * ~40 + 20 W ns, a workgroup of W words per lane ~63 + 18.5 W ns.  Returns the
 * bins taken (0: none). */
#define LONG_ROUNDS 4u
static double chain_ns(uint32_t max_m, uint32_t lanes, uint32_t words, int wide)
{
    const double steps = 1.2 * (double)max_m + (double)lanes;
    return steps * (wide ? 63.0 + 18.5 * (double)words : 40.0 + 20.0 * (double)words);
}

static uint32_t long_plan(const myers_ctx_t *ctx, const bin_scan_t *scan, const batch_t *bt,
                          const uint32_t *order, uint32_t n, myers_long_plan_t *lp)
{
    memset(lp, 0, sizeof(*lp));
    if (!scan->sized || bt->num_queries == 0) return 0;
    const double cpr = (double)bt->num_cands / (double)bt->num_queries;
    const double simds = (double)MYERS_LONG_WAVES * (double)ctx->num_cus;
    const uint32_t wide_lanes = 64u * MYERS_LONG_WAVES;
    uint32_t take[MYERS_LONG_MAX], nt = 0;
    double waves[MYERS_LONG_MAX], chain[MYERS_LONG_MAX], total_waves = 0.0, work = 0.0;
    int wide[MYERS_LONG_MAX] = {0};
    for (uint32_t i = 0; i < n && nt < MYERS_LONG_MAX; i++) {
        const uint32_t v = order[i];
        if (!myers_variant_long_ok(v, NULL)) continue;
        const myers_variant_desc_t *d = &ctx->variants[v];
        const double cands = scan->counted ? (double)scan->cands[v] : (double)scan->reads[v] * cpr;
        waves[nt] = cands / (double)d->slots;
        chain[nt] = cands > 0.0 ? chain_ns(scan->max_m[v], d->lanes, d->words, 0) : 0.0;
        total_waves += waves[nt];
        work += waves[nt] * chain[nt];
        take[nt++] = v;
    }
    if (nt == 0 || total_waves > (double)LONG_ROUNDS * simds) return 0;
    for (;;) {
        uint32_t top = 0;
        double longest = 0.0;
        for (uint32_t k = 0; k < nt; k++)
            if (chain[k] > longest) { longest = chain[k]; top = k; }
        const myers_variant_desc_t *d = &ctx->variants[take[top]];
        if (longest <= 0.0 || wide[top] || d->total_words <= wide_lanes ||
            d->total_words > wide_lanes * MYERS_WIDE_MAX_WORDS)
            break;
        const uint32_t ww = (d->total_words + wide_lanes - 1u) / wide_lanes;
        const double cands = waves[top] * (double)d->slots;
        const double c_wide = chain_ns(scan->max_m[take[top]], wide_lanes, ww, 1);
        const double work_wide = work - waves[top] * chain[top] + (double)MYERS_LONG_WAVES * cands * c_wide;
        double next = c_wide;
        for (uint32_t k = 0; k < nt; k++)
            if (k != top && chain[k] > next) next = chain[k];
        const double before = longest > work / simds ? longest : work / simds;
        const double after = next > work_wide / simds ? next : work_wide / simds;
        if (after >= before) break;
        wide[top] = 1;
        chain[top] = c_wide;
        work = work_wide;
    }
    for (uint32_t k = 0; k < nt; k++) {            /* wide bins keep the longest-first order */
        const uint32_t v = take[k];
        const uint32_t words = ctx->variants[v].total_words;
        if (wide[k]) {
            lp->wide_variant[lp->n_wide] = v;
            lp->wide_words[lp->n_wide] = (words + wide_lanes - 1u) / wide_lanes;
            lp->n_wide++;
            continue;
        }
        uint32_t u4 = 0;
        myers_variant_long_ok(v, &u4);
        lp->variant[lp->n] = v;
        lp->key[lp->n] = (ctx->variants[v].lanes << 8) | ctx->variants[v].words;
        lp->n++;
        if (u4 > lp->ws_u4) lp->ws_u4 = u4;
    }
    if (lp->n_wide) {
        const uint32_t u4 = (myers_long_wide_lds_bytes() + 16u * MYERS_LONG_WAVES - 1u) / (16u * MYERS_LONG_WAVES);
        if (u4 > lp->ws_u4) lp->ws_u4 = u4;
    }
    return nt;
}

/* Resident workgroups per CU of a queue kernel with `lds` bytes of
 * dynamic LDS (queried once per size). */
static uint32_t queue_grid(myers_ctx_t *ctx, uint32_t cls, int cut, uint32_t lds)
{
    if (ctx->queue_occ_lds[cls][cut] != lds) {
        int occ = 0;
        if (myers_queue_occupancy(cls, cut, lds, &occ) != 0 || occ <= 0) occ = 1;
        ctx->queue_occ[cls][cut] = occ;
        ctx->queue_occ_lds[cls][cut] = lds;
    }
    return (uint32_t)ctx->num_cus * (uint32_t)ctx->queue_occ[cls][cut];
}

/* Queue launches of a persistent batch: the bins (already in longest-first
 * order) of each kernel class go into one plan, in that order; bins of no
 * class (multi-lane W = 16, multi-wave variants) are left in `other`. */
static void queue_plans(const uint32_t *bins, uint32_t n, myers_queue_plan_t *qp, uint32_t *other,
                        uint32_t *nother)
{
    memset(qp, 0, sizeof(myers_queue_plan_t) * MYERS_QUEUE_CLASSES);
    *nother = 0;
    for (uint32_t c = 0; c < (uint32_t)MYERS_QUEUE_CLASSES; c++) qp[c].cls = c;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t u4 = 0;
        const int cls = myers_variant_queue_class(bins[i], &u4);
        if (cls < 0 || qp[cls].n >= (uint32_t)MYERS_QUEUE_MAX) {
            other[(*nother)++] = bins[i];
            continue;
        }
        myers_queue_plan_t *q = &qp[cls];
        q->variant[q->n] = bins[i];
        q->key[q->n] = 0;   /* filled by the caller from the variant table */
        q->n++;
        if (u4 > q->ws_u4) q->ws_u4 = u4;
    }
}

/* Enqueue the whole on-device pipeline of one batch on b->stream (no timing
 * events, so that it can be captured into a graph). */
static int enqueue_pipeline(myers_buffer_t *b, const batch_t *bt, const bin_scan_t *scan)
{
    myers_ctx_t *ctx = b->ctx;
    myers_verify_args_t a;
    memset(&a, 0, sizeof(a));
    a.ref = ctx->d_ref;
    a.ref_words = ctx->ref_words_alloc;
    a.ref_size = ctx->ref_size;
    a.peq = (const uint32_t *)bt->d_peq;
    a.qinfo = (const uint32_t *)bt->d_qinfo;
    a.num_peq_entries = bt->num_entries;
    a.num_queries = bt->num_queries;
    a.cands = bt->d_cands;
    a.mode = bt->flags & 1u;
    a.distance = bt->distance;
    a.threshold = bt->threshold;
    a.results = bt->d_results;

    const uint32_t engine = resolve_engine(ctx, bt->flags, scan, bt->num_cands);
    b->stats.launches = 0;
    b->stats.variantsUsed = 0;
    b->binned = 0;
    int rc;
    if (bt->num_cands == 0) return MYERS_SUCCESS;
    if (engine == MYERS_ENGINE_PLAIN && scan->num_present == 1 && !scan->too_long) {
        const uint32_t v = scan->single;
        a.start = 0;
        a.count = bt->num_cands;
        rc = myers_launch_verify(v, &a, grid_for(ctx, v, bt->num_cands, 0), 0, b->stream);
        if (rc) return hip_fail((hipError_t)rc, "verify launch", __FILE__, __LINE__);
        b->stats.launches = 1;
        b->stats.variantsUsed = 1;
        return MYERS_SUCCESS;
    }
    /* Threshold mode: the band probe classes the candidates of one-lane
     * cut-off variants (before the binning) and orders their bins (after). */
    const uint64_t probe = bt->threshold >= 0.0f ? ctx->probe_mask & scan->present : 0u;
    if (probe) {
        rc = myers_launch_probe(&a, bt->num_cands, ctx->d_bin_of_words, max_supported_words(ctx), probe, b->stream);
        if (rc) return hip_fail((hipError_t)rc, "probe launch", __FILE__, __LINE__);
    }
    /* device-side binning */
    uint32_t *ranges = bt->d_work + MYERS_WORK_RANGES;
    uint32_t *heads = bt->d_work + MYERS_WORK_HEADS;
    uint32_t *err = bt->d_work + MYERS_WORK_ERR;
    rc = myers_launch_binning((const uint32_t *)bt->d_qinfo, bt->num_queries, bt->d_cands, bt->num_cands,
                              ctx->d_bin_of_words, max_supported_words(ctx), ctx->num_variants,
                              bt->d_work + MYERS_WORK_BLOCKS, ranges, bt->d_order, err, bt->d_results, heads,
                              b->stream);
    if (rc) return hip_fail((hipError_t)rc, "binning launch", __FILE__, __LINE__);
    if (probe) {
        rc = myers_launch_probe_order(ranges, ctx->num_variants, probe, bt->num_cands, bt->d_order, bt->d_results,
                                      b->stream);
        if (rc) return hip_fail((hipError_t)rc, "probe order launch", __FILE__, __LINE__);
    }
    HIP_TRY(hipMemcpyAsync(b->h_err, err, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
    if (b->mark_slot >= 0) {
        /* eager batches: the verification launches are timed apart from the binning */
        HIP_TRY(hipEventRecord(b->ev_kernel[b->mark_slot][2], b->stream));
        b->ev_marked[b->mark_slot] = 1;
    }
    b->binned = 1;
    a.order = bt->d_order;
    a.bin_ranges = ranges;
    a.queue_heads = heads;
    uint32_t order[MYERS_MAX_VARIANTS];
    const uint32_t n = present_by_cost(ctx, scan->present, order);
    uint32_t launches = n;
    if (engine == MYERS_ENGINE_PLAIN) {
        /* one launch per length bin, in turn */
        for (uint32_t i = 0; i < n; i++) {
            a.variant = order[i];
            rc = myers_launch_verify(order[i], &a, grid_for(ctx, order[i], bt->num_cands, 1), 0, b->stream);
            if (rc) return hip_fail((hipError_t)rc, "verify launch", __FILE__, __LINE__);
        }
    } else {
        /* Persistent engines.  The long-read bins of a latency-bound batch
         * share one launch (long_plan).  MYERS_ENGINE_PERSISTENT puts the
         * other bins into one queue launch per kernel class (one launch for
         * a batch of one class); MYERS_ENGINE_BINNED gives every bin its own
         * resident grid.  Bins without a queue class keep their own grid.
         * When more than one launch results, they run on forked streams
         * (long-read launch first), so that each one's tail overlaps the
         * others. */
        myers_long_plan_t lp;
        uint32_t nlong = 0;
        if (bt->flags & MYERS_NO_LONG_LAUNCH) memset(&lp, 0, sizeof(lp));
        else nlong = long_plan(ctx, scan, bt, order, n, &lp);
        uint32_t rest[MYERS_MAX_VARIANTS], nrest = 0;
        for (uint32_t i = 0; i < n; i++) {
            int in_long = 0;
            for (uint32_t j = 0; j < lp.n; j++) in_long |= lp.variant[j] == order[i];
            for (uint32_t j = 0; j < lp.n_wide; j++) in_long |= lp.wide_variant[j] == order[i];
            if (!in_long) rest[nrest++] = order[i];
        }
        myers_queue_plan_t qp[MYERS_QUEUE_CLASSES];
        uint32_t other[MYERS_MAX_VARIANTS], nother = 0, nqueue = 0;
        if (engine == MYERS_ENGINE_PERSISTENT) {
            queue_plans(rest, nrest, qp, other, &nother);
            for (uint32_t c = 0; c < (uint32_t)MYERS_QUEUE_CLASSES; c++) {
                for (uint32_t j = 0; j < qp[c].n; j++)
                    qp[c].key[j] = (ctx->variants[qp[c].variant[j]].lanes << 8) | ctx->variants[qp[c].variant[j]].words;
                nqueue += qp[c].n ? 1u : 0u;
            }
        } else {
            memset(qp, 0, sizeof(qp));
            memcpy(other, rest, nrest * sizeof(uint32_t));
            nother = nrest;
        }
        launches = (nlong ? 1u : 0u) + nqueue + nother;
        const int fork = launches > 1u;
        const uint32_t width = !fork ? 0u : (launches < SIDE_STREAMS ? launches : SIDE_STREAMS);
        uint32_t next = 0;    /* launches placed so far: launch k goes on side stream k % width */
#define NEXT_STREAM() (fork ? b->side[(next++) % width] : b->stream)
        if (fork) {
            HIP_TRY(hipEventRecord(b->ev_fork, b->stream));
            for (uint32_t j = 0; j < width; j++) HIP_TRY(hipStreamWaitEvent(b->side[j], b->ev_fork, 0));
        }
        if (nlong) {
            /* successive launches start their first items a golden-ratio step
             * of workgroups apart (overlapping batches use other CUs) */
            const uint32_t seq = __atomic_fetch_add(&ctx->long_seq, 1u, __ATOMIC_RELAXED);
            const uint32_t wg = (uint32_t)(((uint64_t)seq * 2654435769ull & 0xFFFFFFFFull) * ctx->num_cus >> 32);
            lp.rotate = wg * MYERS_LONG_WAVES;
            rc = myers_launch_long(&a, &lp, ctx->num_cus, NEXT_STREAM());
            if (rc) return hip_fail((hipError_t)rc, "long-read launch", __FILE__, __LINE__);
        }
        for (uint32_t c = 0; c < (uint32_t)MYERS_QUEUE_CLASSES; c++) {
            if (qp[c].n == 0) continue;
            const int cut = bt->threshold >= 0.0f;
            rc = myers_launch_queue(&a, &qp[c], queue_grid(ctx, c, cut, qp[c].ws_u4 * 16u), NEXT_STREAM());
            if (rc) return hip_fail((hipError_t)rc, "queue launch", __FILE__, __LINE__);
        }
        for (uint32_t i = 0; i < nother; i++) {
            const uint32_t v = other[i];
            a.variant = v;
            a.chunk_factor = ctx->chunk_factor[v];
            uint64_t grid = (uint64_t)ctx->num_cus * (uint64_t)ctx->occupancy_p[v];
            rc = myers_launch_verify(v, &a, (uint32_t)grid, 1, NEXT_STREAM());
            if (rc) return hip_fail((hipError_t)rc, "persistent launch", __FILE__, __LINE__);
        }
#undef NEXT_STREAM
        for (uint32_t j = 0; j < width; j++) {
            HIP_TRY(hipEventRecord(b->ev_join[j], b->side[j]));
            HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_join[j], 0));
        }
    }
    b->stats.launches = launches;
    b->stats.variantsUsed = n;
    return MYERS_SUCCESS;
}

/* Candidates per variant, for batches with long-read bins (the long-read
 * plan sizes its rounds with them; reads carry very different candidate
 * counts across lengths).  Short-read batches skip this pass. */
static void scan_cands(const myers_ctx_t *ctx, const qryInfo_t *qinfo, uint32_t nq, const candInfo_t *cands,
                       uint32_t nc, bin_scan_t *s)
{
    uint64_t long_mask = 0;
    for (uint32_t v = 0; v < ctx->num_variants; v++)
        if (myers_variant_long_ok(v, NULL)) long_mask |= 1ull << v;
    if (!(s->present & long_mask)) return;
    const uint32_t maxw = max_supported_words(ctx);
    for (uint32_t i = 0; i < nc; i++) {
        const uint32_t q = cands[i].query;
        if (q >= nq) continue;
        const uint32_t w = (qinfo[q].size + 31u) >> 5;
        if (w <= maxw) s->cands[ctx->bin_of_words[w]]++;
    }
    s->counted = 1;
}

static void make_scan(myers_ctx_t *ctx, const qryInfo_t *h_qinfo, uint32_t num_queries, const candInfo_t *h_cands,
                      uint32_t num_cands, bin_scan_t *scan)
{
    if (h_qinfo) {
        scan_qinfo(ctx, h_qinfo, num_queries, scan);
        if (h_cands) scan_cands(ctx, h_qinfo, num_queries, h_cands, num_cands, scan);
    } else {
        memset(scan, 0, sizeof(*scan));
        scan->present = ctx->num_variants >= 64 ? ~0ull : ((1ull << ctx->num_variants) - 1u);
        scan->num_present = ctx->num_variants;
        scan->too_long = 1;   /* unknown: let the device check */
    }
}

/* Enqueue the verification of a batch on b->stream, timed with HIP events;
 * with MYERS_USE_GRAPH the pipeline is captured once per batch shape and
 * replayed as a graph. */
static int enqueue_verify(myers_buffer_t *b, const batch_t *bt, const qryInfo_t *h_qinfo, const candInfo_t *h_cands)
{
    bin_scan_t scan;
    make_scan(b->ctx, h_qinfo, bt->num_queries, h_cands, bt->num_cands, &scan);
    uint32_t slot = 0;
    int rc;
    if (!(bt->flags & MYERS_USE_GRAPH) || bt->num_cands == 0) {
        if ((rc = timing_begin(b, &slot))) return rc;
        b->mark_slot = (int)slot;
        rc = enqueue_pipeline(b, bt, &scan);
        b->mark_slot = -1;
        if (rc) return rc;
        return timing_end(b, slot, b->stats.launches);
    }
    uint32_t dist_bits, thr_bits;
    memcpy(&dist_bits, &bt->distance, sizeof(dist_bits));
    memcpy(&thr_bits, &bt->threshold, sizeof(thr_bits));
    uint32_t order[MYERS_MAX_VARIANTS];
    myers_long_plan_t lp;
    const uint64_t plan_key = (bt->flags & MYERS_NO_LONG_LAUNCH) ? 0u :
                              ((uint64_t)long_plan(b->ctx, &scan, bt, order,
                                                   present_by_cost(b->ctx, scan.present, order), &lp) |
                               ((uint64_t)lp.n_wide << 32));
    const uint64_t key[12] = {(uint64_t)(uintptr_t)bt->d_peq, (uint64_t)(uintptr_t)bt->d_qinfo,
                              (uint64_t)(uintptr_t)bt->d_cands, (uint64_t)(uintptr_t)bt->d_results,
                              ((uint64_t)bt->num_entries << 32) | bt->num_queries,
                              ((uint64_t)bt->num_cands << 32) | bt->flags,
                              (uint64_t)dist_bits, scan.present,
                              (uint64_t)scan.too_long, b->ctx->tune_gen, (uint64_t)thr_bits, plan_key};
    if (!b->graph_valid || memcmp(key, b->graph_key, sizeof(key)) != 0) {
        if (b->graph_exec) {
            hipGraphExecDestroy(b->graph_exec);
            b->graph_exec = NULL;
        }
        b->graph_valid = 0;
        hipGraph_t graph = NULL;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_pipeline(b, bt, &scan);
        hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        HIP_TRY(e);
        e = hipGraphInstantiate(&b->graph_exec, graph, NULL, NULL, 0);
        hipGraphDestroy(graph);
        HIP_TRY(e);
        memcpy(b->graph_key, key, sizeof(key));
        b->graph_valid = 1;
        b->graph_binned = b->binned;
        b->graph_launches = b->stats.launches;
        b->stats.graphCaptures++;
    }
    if ((rc = timing_begin(b, &slot))) return rc;
    HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
    b->binned = b->graph_binned;
    b->stats.launches = b->graph_launches;
    b->stats.variantsUsed = b->graph_launches;
    return timing_end(b, slot, b->graph_launches);
}

static uint64_t nominal_cells(const qryInfo_t *qinfo, const candInfo_t *cands, uint32_t nc, float d)
{
    uint64_t cells = 0;
    for (uint32_t i = 0; i < nc; i++) {
        uint32_t m = qinfo[cands[i].query].size;
        volatile float f = (float)m * (1.0f + 2.0f * d);
        cells += (uint64_t)m * (uint64_t)(uint32_t)f;
    }
    return cells;
}

int myersSendEx(void *myersBuffer, float distance, uint32_t numPEQEntries, uint32_t numQueries,
                uint32_t numCandidates, uint32_t flags)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL) return MYERS_E_INVALID_ARGUMENT;
    if (flags & ~(1u | MYERS_ENGINE_MASK | MYERS_USE_GRAPH | MYERS_NO_LONG_LAUNCH))
        return (b->last_error = MYERS_E_INVALID_ARGUMENT);
    if (numPEQEntries > b->max_peq_entries || numQueries > b->max_queries || numCandidates > b->max_candidates)
        return (b->last_error = MYERS_E_CAPACITY);
    int rc = hip_fail(hipSetDevice(b->ctx->device), "hipSetDevice", __FILE__, __LINE__);
    if (rc) return (b->last_error = rc);
    /* candidates must name a query of this batch */
    for (uint32_t i = 0; i < numCandidates; i++)
        if (b->h_cands[i].query >= numQueries) return (b->last_error = MYERS_E_INVALID_ARGUMENT);
    *b->h_err = 0;
    b->num_candidates = numCandidates;
    hipStream_t s = b->stream;
#define SEND_TRY(call) do { rc = hip_fail((call), #call, __FILE__, __LINE__); if (rc) return (b->last_error = rc); } while (0)
    SEND_TRY(hipEventRecord(b->ev_batch_start, s));
    if (numPEQEntries)
        SEND_TRY(hipMemcpyAsync(b->d_peq, b->h_peq, (size_t)numPEQEntries * sizeof(qryEntry_t), hipMemcpyHostToDevice, s));
    if (numQueries)
        SEND_TRY(hipMemcpyAsync(b->d_qinfo, b->h_qinfo, (size_t)numQueries * sizeof(qryInfo_t), hipMemcpyHostToDevice, s));
    if (numCandidates)
        SEND_TRY(hipMemcpyAsync(b->d_cands, b->h_cands, (size_t)numCandidates * sizeof(candInfo_t), hipMemcpyHostToDevice, s));
    const batch_t bt = {b->d_peq, b->d_qinfo, b->d_cands, b->d_results, numPEQEntries, numQueries,
                        numCandidates, flags, distance, b->threshold, b->d_work, b->d_order};
    rc = enqueue_verify(b, &bt, b->h_qinfo, b->h_cands);
    if (rc) return (b->last_error = rc);
    if (numCandidates)
        SEND_TRY(hipMemcpyAsync(b->h_results, b->d_results, (size_t)numCandidates * sizeof(resEntry_t), hipMemcpyDeviceToHost, s));
    SEND_TRY(hipEventRecord(b->ev_batch_end, s));
#undef SEND_TRY
    b->stats.cells = nominal_cells(b->h_qinfo, b->h_cands, numCandidates, distance);
    b->stats.candidates = numCandidates;
    b->pending = 1;
    b->last_error = MYERS_SUCCESS;
    return MYERS_SUCCESS;
}

static int finish_batch(myers_buffer_t *b, int with_batch_events)
{
    int rc = hip_fail(hipStreamSynchronize(b->stream), "hipStreamSynchronize", __FILE__, __LINE__);
    if (rc) return rc;
    while (b->ev_pending) {
        rc = fold_oldest_timing(b);
        if (rc) return rc;
    }
    if (!b->pending) return MYERS_SUCCESS;
    b->pending = 0;
    float ms = 0.0f;
    if (with_batch_events && hipEventElapsedTime(&ms, b->ev_batch_start, b->ev_batch_end) == hipSuccess)
        b->stats.batchMs = ms;
    b->stats.binned = (uint32_t)b->binned;
    if (b->binned && (*b->h_err & 1u)) return MYERS_E_QUERY_TOO_LONG;
    if (b->binned && (*b->h_err & 2u)) return MYERS_E_INVALID_ARGUMENT;
    return MYERS_SUCCESS;
}

int myers_serve_plan(void *myersBuffer, myers_serve_plan_t *out)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || out == NULL) return MYERS_E_INVALID_ARGUMENT;
    const myers_ctx_t *ctx = b->ctx;
    memset(out, 0, sizeof(*out));
    out->device = ctx->device;
    out->num_cus = ctx->num_cus;
    out->d_ref = ctx->d_ref;
    out->ref_words = ctx->ref_words_alloc;
    out->ref_size = ctx->ref_size;
    out->threshold = b->threshold;
    uint32_t var_of[MYERS_SERVE_MAX_WORDS + 1];
    uint64_t used = 0;
    const uint32_t maxw = max_supported_words(ctx);
    for (uint32_t w = 0; w <= maxw && w <= MYERS_SERVE_MAX_WORDS; w++) {
        uint32_t v = ctx->bin_of_words[w];
        if (!myers_variant_serve_ok(v, NULL)) {
            uint32_t best = 0xFFFFFFFFu, best_cost = 0xFFFFFFFFu;
            for (uint32_t u = 0; u < ctx->num_variants; u++) {
                if (!myers_variant_serve_ok(u, NULL) || ctx->variants[u].total_words < (w ? w : 1u)) continue;
                const uint32_t c = variant_cost(&ctx->variants[u]);
                if (c < best_cost) { best_cost = c; best = u; }
            }
            if (best == 0xFFFFFFFFu) break;      /* longer reads: none can run there */
            v = best;
        }
        var_of[w] = v;
        used |= 1ull << v;
        out->max_words = w;
    }
    /* ranks: the used variants, most words per candidate first */
    uint32_t order[MYERS_MAX_VARIANTS];
    const uint32_t n = present_by_cost(ctx, used, order);
    uint32_t rank_of_var[MYERS_MAX_VARIANTS];
    for (uint32_t r = 0; r < n; r++) {
        const myers_variant_desc_t *d = &ctx->variants[order[r]];
        uint32_t u4 = 0;
        myers_variant_serve_ok(order[r], &u4);
        rank_of_var[order[r]] = r;
        out->rank_key[r] = (d->lanes << 8) | d->words;
        out->rank_slots[r] = d->slots;
        if (u4 * 16u > out->lds_bytes) out->lds_bytes = u4 * 16u;
    }
    out->num_ranks = n;
    for (uint32_t w = 0; w <= out->max_words; w++) out->rank_of_words[w] = (uint8_t)rank_of_var[var_of[w]];
    return MYERS_SUCCESS;
}

int myers_buffer_ready(void *myersBuffer)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return -(int)e;
}

int myersReceiveEx(void *myersBuffer)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL) return MYERS_E_INVALID_ARGUMENT;
    int rc = hip_fail(hipSetDevice(b->ctx->device), "hipSetDevice", __FILE__, __LINE__);
    if (rc == MYERS_SUCCESS) rc = finish_batch(b, 1);
    b->last_error = rc;
    return rc;
}

void sendMyersBuffer(void *myersBuffer, float distance, uint32_t numPEQEntries, uint32_t numQueries,
                     uint32_t numCandidates)
{
    int rc = myersSendEx(myersBuffer, distance, numPEQEntries, numQueries, numCandidates, 0u);
    report((myers_buffer_t *)myersBuffer, rc, "sendMyersBuffer");
}

void receiveMyersBuffer(void *myersBuffer)
{
    int rc = myersReceiveEx(myersBuffer);
    report((myers_buffer_t *)myersBuffer, rc, "receiveMyersBuffer");
}

/* ------------------------------------------------------- device batches */

static int verify_device(myers_buffer_t *b, const void *d_peq, uint32_t numPEQEntries, const void *d_qinfo,
                         uint32_t numQueries, const void *d_candidates, uint32_t numCandidates,
                         float distance, uint32_t flags, void *d_results, const qryInfo_t *h_qinfo,
                         uint32_t *d_work, uint32_t *d_order)
{
    int rc = hip_fail(hipSetDevice(b->ctx->device), "hipSetDevice", __FILE__, __LINE__);
    if (rc) return (b->last_error = rc);
    *b->h_err = 0;
    const batch_t bt = {d_peq, d_qinfo, d_candidates, d_results, numPEQEntries, numQueries, numCandidates,
                        flags, distance, b->threshold, d_work, d_order};
    rc = enqueue_verify(b, &bt, h_qinfo, NULL);
    b->stats.cells = 0;
    b->stats.candidates = numCandidates;
    b->pending = (rc == MYERS_SUCCESS);
    b->last_error = rc;
    return rc;
}

int myersVerifyDevice(void *myersBuffer, const void *d_peq, uint32_t numPEQEntries, const void *d_qinfo,
                      uint32_t numQueries, const void *d_candidates, uint32_t numCandidates,
                      float distance, uint32_t flags, void *d_results, const qryInfo_t *h_qinfo)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL) return MYERS_E_INVALID_ARGUMENT;
    if ((flags & ~(1u | MYERS_ENGINE_MASK | MYERS_USE_GRAPH | MYERS_NO_LONG_LAUNCH)) ||
        (numCandidates && (!d_peq || !d_qinfo || !d_candidates || !d_results)))
        return (b->last_error = MYERS_E_INVALID_ARGUMENT);
    if (numCandidates > b->max_candidates) return (b->last_error = MYERS_E_CAPACITY);
    return verify_device(b, d_peq, numPEQEntries, d_qinfo, numQueries, d_candidates, numCandidates, distance,
                         flags, d_results, h_qinfo, b->d_work, b->d_order);
}

/* myersVerifyDevice with the pipeline workspace in the device's static
 * array instead of the buffer's: nothing is allocated.  Every call on one
 * GPU shares that array (the binning kernels re-initialise the parts that
 * must start at zero), so calls that overlap on one GPU race. */
int myersVerifyDeviceNoAlloc(void *myersBuffer, const void *d_peq, uint32_t numPEQEntries, const void *d_qinfo,
                             uint32_t numQueries, const void *d_candidates, uint32_t numCandidates,
                             float distance, uint32_t flags, void *d_results, const qryInfo_t *h_qinfo)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL) return MYERS_E_INVALID_ARGUMENT;
    /* graph instantiation may allocate: not on this path */
    if ((flags & ~(1u | MYERS_ENGINE_MASK | MYERS_NO_LONG_LAUNCH)) ||
        (numCandidates && (!d_peq || !d_qinfo || !d_candidates || !d_results)))
        return (b->last_error = MYERS_E_INVALID_ARGUMENT);
    if (numCandidates > MYERS_NOALLOC_MAX_CANDIDATES) {
        fprintf(stderr, "Myers - myersVerifyDeviceNoAlloc: %u candidates exceed the static workspace "
                "maximum of %u\n", numCandidates, (unsigned)MYERS_NOALLOC_MAX_CANDIDATES);
        return (b->last_error = MYERS_E_CAPACITY);
    }
    uint32_t *scratch = b->ctx->d_scratch;
    return verify_device(b, d_peq, numPEQEntries, d_qinfo, numQueries, d_candidates, numCandidates, distance,
                         flags, d_results, h_qinfo, scratch, scratch + MYERS_WORK_WORDS);
}

int myersSetThreshold(void *myersBuffer, float errorRate)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || errorRate != errorRate) return MYERS_E_INVALID_ARGUMENT;
    b->threshold = errorRate < 0.0f ? -1.0f : errorRate;
    return MYERS_SUCCESS;
}

float myersGetThreshold(void *myersBuffer)
{
    const myers_buffer_t *b = (const myers_buffer_t *)myersBuffer;
    return b == NULL ? -1.0f : b->threshold;
}

int myersSynchronize(void *myersBuffer)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL) return MYERS_E_INVALID_ARGUMENT;
    int rc = hip_fail(hipSetDevice(b->ctx->device), "hipSetDevice", __FILE__, __LINE__);
    if (rc == MYERS_SUCCESS) rc = finish_batch(b, 0);
    b->last_error = rc;
    return rc;
}

int myersGetStats(void *myersBuffer, myersStats_t *out)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || out == NULL) return MYERS_E_INVALID_ARGUMENT;
    /* callers built against an older, shorter struct get its fields only */
    const size_t n = out->structSize && out->structSize < sizeof(*out) ? out->structSize : sizeof(*out);
    memcpy(out, &b->stats, n);
    out->structSize = (uint32_t)n;
    return MYERS_SUCCESS;
}

void myersResetStats(void *myersBuffer)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL) return;
    b->stats.kernelMsTotal = 0.0;
    b->stats.verifyMsTotal = 0.0;
    b->stats.kernelLaunchesTotal = 0;
    b->stats.batchesTotal = 0;
}

/* -------------------------------------------------------------- helpers */

int myersDeviceMalloc(void **ptr, size_t bytes)
{
    if (ptr == NULL) return MYERS_E_INVALID_ARGUMENT;
    return hip_fail(hipMalloc(ptr, bytes ? bytes : 1u), "hipMalloc", __FILE__, __LINE__);
}

int myersDeviceFree(void *ptr)
{
    return hip_fail(hipFree(ptr), "hipFree", __FILE__, __LINE__);
}

int myersMemcpyHtoD(void *dst, const void *src, size_t bytes)
{
    return hip_fail(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(HtoD)", __FILE__, __LINE__);
}

int myersMemcpyDtoH(void *dst, const void *src, size_t bytes)
{
    return hip_fail(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy(DtoH)", __FILE__, __LINE__);
}

uint32_t myersNumVariants(void) { return myers_num_variants(); }

int myersGetVariantInfo(void *myersBuffer, uint32_t variant, myersVariantInfo_t *out)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || out == NULL || variant >= b->ctx->num_variants) return MYERS_E_INVALID_ARGUMENT;
    const myers_variant_desc_t *d = &b->ctx->variants[variant];
    out->lanes = d->lanes;
    out->words = d->words;
    out->maxQuerySize = 32u * d->total_words;
    out->candidatesPerWave = d->slots;
    out->ldsBytes = d->lds_bytes;
    out->workgroupsPerCU = b->ctx->occupancy[variant];
    return MYERS_SUCCESS;
}

int myersGetBinTable(void *myersBuffer, uint8_t *bin, uint32_t maxWords)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || bin == NULL) return MYERS_E_INVALID_ARGUMENT;
    for (uint32_t w = 0; w <= maxWords && w <= MAX_QUERY_WORDS; w++) bin[w] = b->ctx->bin_of_words[w];
    return MYERS_SUCCESS;
}

int myersSetBinTable(void *myersBuffer, const uint8_t *bin, uint32_t maxWords)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || bin == NULL || maxWords > MAX_QUERY_WORDS) return MYERS_E_INVALID_ARGUMENT;
    myers_ctx_t *ctx = b->ctx;
    for (uint32_t w = 0; w <= maxWords; w++) {
        if (bin[w] >= ctx->num_variants) return MYERS_E_INVALID_ARGUMENT;
        if (ctx->variants[bin[w]].total_words < (w ? w : 1u)) return MYERS_E_INVALID_ARGUMENT;
    }
    int rc = hip_fail(hipSetDevice(ctx->device), "hipSetDevice", __FILE__, __LINE__);
    if (rc) return rc;
    /* the table is shared by every buffer: wait for in-flight work first */
    rc = hip_fail(hipDeviceSynchronize(), "hipDeviceSynchronize", __FILE__, __LINE__);
    if (rc) return rc;
    for (uint32_t w = 0; w <= maxWords; w++) ctx->bin_of_words[w] = bin[w];
    ctx->tune_gen++;
    return hip_fail(hipMemcpy(ctx->d_bin_of_words, ctx->bin_of_words, sizeof(ctx->bin_of_words),
                              hipMemcpyHostToDevice), "hipMemcpy(bin table)", __FILE__, __LINE__);
}

int myersSetChunkFactor(void *myersBuffer, uint32_t variant, uint32_t chunks)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || variant >= b->ctx->num_variants || chunks < 1u || chunks > 64u) return MYERS_E_INVALID_ARGUMENT;
    b->ctx->chunk_factor[variant] = chunks;
    b->ctx->tune_gen++;   /* cached graphs baked in the old factor */
    return MYERS_SUCCESS;
}

int myersGetChunkFactor(void *myersBuffer, uint32_t variant, uint32_t *chunks)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || chunks == NULL || variant >= b->ctx->num_variants) return MYERS_E_INVALID_ARGUMENT;
    *chunks = b->ctx->chunk_factor[variant];
    return MYERS_SUCCESS;
}

static int find_variant(const myers_ctx_t *ctx, uint32_t lanes, uint32_t words)
{
    for (uint32_t v = 0; v < ctx->num_variants; v++)
        if (ctx->variants[v].lanes == lanes && ctx->variants[v].words == words) return (int)v;
    return -1;
}

/* Apply a tuning table to a context (not yet shared with running work). */
static int load_tuning_ctx(myers_ctx_t *ctx, const char *path)
{
    FILE *fp = fopen(path, "r");
    if (fp == NULL) return MYERS_E_OPENING_FILE;
    uint8_t bin[MAX_QUERY_WORDS + 1];
    memcpy(bin, ctx->bin_of_words, sizeof(bin));
    uint32_t cf[MYERS_MAX_VARIANTS];
    memcpy(cf, ctx->chunk_factor, sizeof(cf));
    char line[256];
    uint32_t lo = 0;
    int rc = MYERS_SUCCESS;
    while (fgets(line, sizeof(line), fp)) {
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        unsigned mw, lanes, words, chunks;
        int got = sscanf(line, "%u %u %u %u", &mw, &lanes, &words, &chunks);
        if (got <= 0) continue;
        if (got != 4 || mw < lo || mw > MAX_QUERY_WORDS || chunks < 1u || chunks > 64u) { rc = MYERS_E_READING_FILE; break; }
        int v = find_variant(ctx, lanes, words);
        if (v < 0 || ctx->variants[v].total_words < mw) { rc = MYERS_E_READING_FILE; break; }
        for (uint32_t w = lo; w <= mw; w++) bin[w] = (uint8_t)v;
        cf[v] = chunks;
        lo = mw + 1u;
    }
    fclose(fp);
    if (rc) return rc;
    memcpy(ctx->bin_of_words, bin, sizeof(bin));
    memcpy(ctx->chunk_factor, cf, sizeof(cf));
    ctx->tune_gen++;
    return hip_fail(hipMemcpy(ctx->d_bin_of_words, ctx->bin_of_words, sizeof(ctx->bin_of_words),
                              hipMemcpyHostToDevice), "hipMemcpy(bin table)", __FILE__, __LINE__);
}

int myersLoadTuning(void *myersBuffer, const char *path)
{
    myers_buffer_t *b = (myers_buffer_t *)myersBuffer;
    if (b == NULL || path == NULL) return MYERS_E_INVALID_ARGUMENT;
    int rc = hip_fail(hipSetDevice(b->ctx->device), "hipSetDevice", __FILE__, __LINE__);
    if (rc) return rc;
    rc = hip_fail(hipDeviceSynchronize(), "hipDeviceSynchronize", __FILE__, __LINE__);
    if (rc) return rc;
    return load_tuning_ctx(b->ctx, path);
}

/* $MYERS_TUNING_FILE, else <library dir>/../tuning/gfx950.tsv; silently
 * keeps the built-in table when neither exists. */
static void load_default_tuning(myers_ctx_t *ctx)
{
    const char *env = getenv("MYERS_TUNING_FILE");
    if (env && *env) {
        if (load_tuning_ctx(ctx, env) != MYERS_SUCCESS)
            fprintf(stderr, "Myers - warning: could not apply tuning file %s\n", env);
        return;
    }
    Dl_info info;
    if (dladdr((void *)&load_default_tuning, &info) == 0 || info.dli_fname == NULL) return;
    char path[4096];
    size_t n = strlen(info.dli_fname);
    if (n + 32 >= sizeof(path)) return;
    memcpy(path, info.dli_fname, n + 1);
    char *slash = strrchr(path, '/');
    if (slash == NULL) return;
    strcpy(slash, "/../tuning/gfx950.tsv");
    FILE *fp = fopen(path, "r");
    if (fp == NULL) return;
    fclose(fp);
    if (load_tuning_ctx(ctx, path) != MYERS_SUCCESS)
        fprintf(stderr, "Myers - warning: could not apply tuning file %s\n", path);
}
