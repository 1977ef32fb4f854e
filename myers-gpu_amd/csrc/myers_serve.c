/*
 * Continuous-batching server over the candidate-verification buffers.
 *
 * The reference has no serving mode: its driver fills one buffer per host
 * thread and blocks on it (src-api/test.c:184-193).  Here one server thread
 * keeps up to N batches in flight, one per buffer (each buffer has its own
 * HIP stream, so batches overlap on the device):
 *
 *   submit   -> FIFO of pending requests (input arrays referenced, not copied)
 *   admit    -> while a slot is free: pop requests in FIFO order while the
 *               batch still fits the slot's capacity and the device bytes of
 *               admitted, unfinished requests stay within the HBM budget;
 *               copy their Peq / qinfo / candidates into the slot's pinned
 *               arrays, rebasing posEntry and query ids; myersSendEx
 *   complete -> poll each busy slot's stream; on completion copy every
 *               request's results out of the slot and wake its waiters
 *
 * A request's candidates go through exactly the kernels a lone request
 * would, so its results do not depend on what it was batched with.
 *
 * Resident mode (MYERS_SERVE_RESIDENT) replaces the batches by one kernel
 * that stays on the device (verify_serve_kernel): the server thread copies
 * each admitted request into a pinned arena, bins its candidates by read
 * length on the host (longest reads first) into chunk items, appends them to
 * a work queue in pinned memory and publishes the new tail; the kernel's
 * waves pull items in queue order, and the wave that finishes a request's
 * last item writes the request's ticket, which the server thread polls.
 * Admission: a free request slot, room in the item ring and the arena, and
 * the HBM budget.  The same per-chunk verification runs as in a batch, so a
 * request's results are again independent of what else is in flight.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime_api.h>

#include "myers-interface.h"
#include "myers-amd.h"
#include "myers_internal.h"
#include "myers_kernels.h"

enum { REQ_QUEUED = 0, REQ_ADMITTED = 1, REQ_DONE = 2 };

typedef struct req {
    myersRequest_t r;
    uint64_t       ticket;
    uint64_t       t_submit_ns;
    uint64_t       t_admit_ns;
    uint64_t       t_done_ns;
    uint64_t       bytes;
    uint64_t       cells;
    int            state;
    int            error;
    struct req    *next;
    /* resident mode */
    uint64_t       arena_start, arena_end;    /* monotone arena offsets held */
    uint64_t       admit_idx;                 /* admission order */
    uint32_t       first_item[MYERS_SERVE_RINGS], nitems[MYERS_SERVE_RINGS];
    uint32_t       items;                     /* in both rings */
    resEntry_t    *res;                       /* results in the arena */
    int            too_long;
    int            in_ring;                   /* held in the resident in-flight ring */
    int            waiters;                   /* threads inside a wait on it */
    int            released;                  /* waited on: the caller is done with the ticket */
} req_t;

/* Resident mode state (MYERS_SERVE_RESIDENT). */
#define RES_RING_ITEMS  (1u << 20)       /* work items per ring (16 MB) */
#define RES_REQ_SLOTS   4096u            /* requests in flight at most */
#define RES_IDLE_TICKS  5000000u         /* 50 ms of the 100 MHz real-time counter */
#define RES_POLL_TICKS  200u             /* 2 us between reads of the host tail */
#define RES_DEFAULT_ARENA (1ull << 30)
#define RES_CTL_TAIL    0u               /* ring r's tail at RES_CTL_TAIL + r * MYERS_SERVE_TAIL_STRIDE */
#define RES_CTL_STOP    32u

typedef struct {
    myers_serve_plan_t  plan;
    int                 cut;
    uint32_t            grid;
    hipStream_t         stream;
    hipEvent_t          ev_done;          /* recorded after each launch */
    int                 running;          /* launched and its end not yet seen */
    uint32_t            gen;              /* generation of the current launch */
    myers_serve_item_t *items;            /* pinned, coherent */
    myers_serve_req_t  *reqs;
    uint64_t           *flags;
    uint32_t           *ctl;              /* ring tails (published items), stop generation */
    uint8_t            *arena;
    void               *d_items, *d_reqs, *d_flags, *d_ctl, *d_arena;   /* device addresses of the above */
    uint64_t            arena_bytes, arena_head, arena_tail;
    uint32_t           *d_dev, *d_done;   /* device memory: claim counters, per-slot counts */
    uint32_t           *d_simd;           /* device memory: per-SIMD long-item flags */
    uint32_t            published[MYERS_SERVE_RINGS];
    uint32_t            long_per_simd;
    uint32_t            active;           /* admitted requests with items, not yet done */
    uint64_t            admitted, oldest;
    req_t             **inflight;         /* RES_REQ_SLOTS ring by admission index */
} resident_t;

typedef struct {
    void     *buf;
    int       busy;
    uint32_t  nreq, cap_req;
    req_t   **reqs;
    uint32_t *cand_off;
    uint64_t  bytes;
    uint32_t  max_entries, max_queries, max_cands;
} slot_t;

typedef struct {
    pthread_t          thread;
    pthread_mutex_t    mu;
    pthread_cond_t     cv_work;
    pthread_cond_t     cv_done;
    myersServeConfig_t cfg;
    slot_t            *slots;
    uint32_t           nslots;
    req_t             *head, *tail;
    uint64_t           queued_cells;      /* cells of the requests in the queue */
    /* Requests by ticket: table[t - table_base] for table_base <= t <
     * next_ticket, NULL once freed.  A request is freed when its waiter has
     * released it and the server thread no longer refers to it; the table
     * drops its freed prefix when it fills, so it spans the tickets between
     * the oldest unreleased request and the newest. */
    req_t            **table;
    uint64_t           table_cap, table_base, next_ticket;
    uint64_t           inflight_bytes;
    int                stopping;
    int                thread_started;
    myersServeStats_t  stats;
    resident_t        *res;               /* MYERS_SERVE_RESIDENT */
} server_t;

static uint64_t now_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

/* The request of a ticket, NULL if it was never issued or is already
 * released.  Caller holds the lock. */
static req_t *req_of(server_t *sv, uint64_t ticket)
{
    if (ticket < sv->table_base || ticket >= sv->next_ticket) return NULL;
    req_t *q = sv->table[ticket - sv->table_base];
    return (q != NULL && !q->released) ? q : NULL;
}

/* Free a request once a wait released it, no other wait is still inside,
 * and the server thread holds no reference (it is done and out of the
 * resident ring).  Caller holds the lock. */
static void req_try_free(server_t *sv, req_t *q)
{
    if (!q->released || q->waiters != 0 || q->state != REQ_DONE || q->in_ring) return;
    sv->table[q->ticket - sv->table_base] = NULL;
    sv->stats.liveRequests--;
    free(q);
}

/* About 5 ms of verification at 100 TCUPS.  Batches carry enough long
 * candidates (one chain per wave each) to keep the chip busy with three in
 * flight: on the serving mix at 0.7 of the bulk rate, p99 19 ms against 28
 * ms with 2.5e11-cell batches (3 buffers, same box, two runs each). */
/* Backlog drain: more than DRAIN_BACKLOG capped batches queued -> batches of
 * up to DRAIN_FACTOR times the cap. */
#define DRAIN_BACKLOG 3u
#define DRAIN_FACTOR  4u
#define DEFAULT_BATCH_CELLS 500000000000ull

/* Sum of m * n over the request's candidates (nominal windows). */
static uint64_t request_cells(const myersRequest_t *r, float distance)
{
    volatile float factor = 1.0f + 2.0f * distance;
    uint64_t cells = 0;
    for (uint32_t i = 0; i < r->numCandidates; i++) {
        const uint32_t m = r->qinfo[r->cands[i].query].size;
        volatile float n = (float)m * factor;
        cells += (uint64_t)m * (uint64_t)(uint32_t)n;
    }
    return cells;
}

/* Device bytes a request occupies while admitted: Peq, qinfo, candidates,
 * results and the binning slot of each candidate. */
static uint64_t request_bytes(const myersRequest_t *r)
{
    return (uint64_t)r->numPEQEntries * sizeof(qryEntry_t) + (uint64_t)r->numQueries * sizeof(qryInfo_t) +
           (uint64_t)r->numCandidates * (sizeof(candInfo_t) + sizeof(resEntry_t) + sizeof(uint32_t));
}

void myersServeDefaultConfig(myersServeConfig_t *cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    cfg->structSize = sizeof(*cfg);
    cfg->distance = 0.1f;
    cfg->flags = 0;
    cfg->hbmBudgetBytes = 200000000000ull;
    cfg->pollMicros = 20;
    cfg->maxBatchCells = DEFAULT_BATCH_CELLS;
}

/* Copy finished batches' results out; caller holds the lock. */
static int complete_ready(server_t *sv)
{
    int progressed = 0;
    for (uint32_t i = 0; i < sv->nslots; i++) {
        slot_t *s = &sv->slots[i];
        if (!s->busy) continue;
        int ready = myers_buffer_ready(s->buf);
        if (ready == 0) continue;
        const uint64_t t0 = now_ns();
        int rc = myersReceiveEx(s->buf);
        const resEntry_t *res = getResultsBuffer(s->buf);
        const uint64_t t = now_ns();
        myersStats_t st;
        if (myersGetStats(s->buf, &st) == MYERS_SUCCESS) sv->stats.busyMs += st.kernelMs;
        for (uint32_t k = 0; k < s->nreq; k++) {
            req_t *q = s->reqs[k];
            if (rc == MYERS_SUCCESS || rc == MYERS_E_QUERY_TOO_LONG)
                memcpy(q->r.results, res + s->cand_off[k], (size_t)q->r.numCandidates * sizeof(resEntry_t));
            q->error = rc;
            q->t_done_ns = t;
            q->state = REQ_DONE;
            sv->inflight_bytes -= q->bytes;
            sv->stats.requests++;
            sv->stats.candidates += q->r.numCandidates;
        }
        s->busy = 0;
        s->nreq = 0;
        progressed = 1;
        sv->stats.completeMs += (double)(now_ns() - t0) / 1e6;
    }
    if (progressed) pthread_cond_broadcast(&sv->cv_done);
    return progressed;
}

/* Form and launch one batch in a free slot; caller holds the lock. */
static int launch_batch(server_t *sv, slot_t *s)
{
    const uint64_t t_start = now_ns();
    qryEntry_t *peq = getPEQBuffer(s->buf);
    qryInfo_t *qinfo = getPEQInfoBuffer(s->buf);
    candInfo_t *cands = getCandidatesBuffer(s->buf);
    uint32_t ne = 0, nq = 0, nc = 0;
    uint64_t cells = 0;
    s->nreq = 0;
    s->bytes = 0;
    const uint32_t cap_c = (sv->cfg.maxBatchCandidates && sv->cfg.maxBatchCandidates < s->max_cands)
                               ? sv->cfg.maxBatchCandidates : s->max_cands;
    /* A backlog of more than DRAIN_BACKLOG capped batches is drained in
     * batches of up to DRAIN_FACTOR times the cap: large batches run their
     * long reads at the throughput of one queue launch per kernel class
     * (what a batch of the whole backlog does), where capped batches take
     * the long-read launch built for latency. */
    uint64_t cap_cells = sv->cfg.maxBatchCells;
    int drain = 0;
    if (cap_cells <= UINT64_MAX / (DRAIN_FACTOR > DRAIN_BACKLOG ? DRAIN_FACTOR : DRAIN_BACKLOG) &&
        sv->queued_cells > DRAIN_BACKLOG * cap_cells) {
        cap_cells *= DRAIN_FACTOR;
        drain = 1;
    }
    while (sv->head) {
        req_t *q = sv->head;
        const myersRequest_t *r = &q->r;
        if (r->numPEQEntries > s->max_entries || r->numQueries > s->max_queries || r->numCandidates > cap_c) {
            /* can never fit a slot: fail it */
            sv->head = q->next;
            if (!sv->head) sv->tail = NULL;
            sv->queued_cells -= q->cells;
            q->error = MYERS_E_CAPACITY;
            q->t_done_ns = now_ns();
            q->t_admit_ns = q->t_done_ns;
            q->state = REQ_DONE;
            pthread_cond_broadcast(&sv->cv_done);
            continue;
        }
        if (ne + r->numPEQEntries > s->max_entries || nq + r->numQueries > s->max_queries ||
            nc + r->numCandidates > cap_c)
            break;
        if (s->nreq > 0 && cells + q->cells > cap_cells) break;
        if (sv->inflight_bytes + q->bytes > sv->cfg.hbmBudgetBytes && sv->inflight_bytes > 0) break;
        if (s->nreq == s->cap_req) {
            uint32_t cap = s->cap_req ? 2u * s->cap_req : 64u;
            req_t **nr = (req_t **)realloc(s->reqs, cap * sizeof(req_t *));
            uint32_t *no = (uint32_t *)realloc(s->cand_off, cap * sizeof(uint32_t));
            if (!nr || !no) {
                if (nr) s->reqs = nr;
                if (no) s->cand_off = no;
                break;
            }
            s->reqs = nr;
            s->cand_off = no;
            s->cap_req = cap;
        }
        sv->head = q->next;
        if (!sv->head) sv->tail = NULL;
        sv->queued_cells -= q->cells;
        memcpy(peq + ne, r->peq, (size_t)r->numPEQEntries * sizeof(qryEntry_t));
        for (uint32_t i = 0; i < r->numQueries; i++) {
            qinfo[nq + i].posEntry = r->qinfo[i].posEntry + ne;
            qinfo[nq + i].size = r->qinfo[i].size;
        }
        for (uint32_t i = 0; i < r->numCandidates; i++) {
            cands[nc + i].query = r->cands[i].query + nq;
            cands[nc + i].position = r->cands[i].position;
        }
        s->reqs[s->nreq] = q;
        s->cand_off[s->nreq] = nc;
        s->nreq++;
        ne += r->numPEQEntries;
        nq += r->numQueries;
        nc += r->numCandidates;
        cells += q->cells;
        s->bytes += q->bytes;
        sv->inflight_bytes += q->bytes;
        q->t_admit_ns = t_start;
        q->state = REQ_ADMITTED;
    }
    if (s->nreq == 0) return 0;
    int rc = myersSendEx(s->buf, sv->cfg.distance, ne, nq, nc, sv->cfg.flags & ~MYERS_USE_GRAPH);
    if (rc != MYERS_SUCCESS) {
        const uint64_t t = now_ns();
        for (uint32_t k = 0; k < s->nreq; k++) {
            s->reqs[k]->error = rc;
            s->reqs[k]->t_done_ns = t;
            s->reqs[k]->state = REQ_DONE;
            sv->inflight_bytes -= s->reqs[k]->bytes;
        }
        s->nreq = 0;
        pthread_cond_broadcast(&sv->cv_done);
        return 0;
    }
    s->busy = 1;
    sv->stats.batches++;
    if (drain && s->nreq > 1 && cells > sv->cfg.maxBatchCells) sv->stats.drainBatches++;
    sv->stats.admitMs += (double)(now_ns() - t_start) / 1e6;
    if (s->nreq > sv->stats.maxBatchRequests) sv->stats.maxBatchRequests = s->nreq;
    return 1;
}

/* ------------------------------------------------------- resident mode */

static void finish_request(server_t *sv, req_t *q, int error, uint64_t t)
{
    q->error = error;
    q->t_done_ns = t;
    q->state = REQ_DONE;
    sv->stats.requests++;
    sv->stats.candidates += q->r.numCandidates;
}

static void res_free(resident_t *rs)
{
    if (rs == NULL) return;
    if (rs->running && rs->ev_done) hipEventSynchronize(rs->ev_done);
    if (rs->ev_done) hipEventDestroy(rs->ev_done);
    if (rs->stream) hipStreamDestroy(rs->stream);
    if (rs->items) hipHostFree(rs->items);
    if (rs->reqs) hipHostFree(rs->reqs);
    if (rs->flags) hipHostFree(rs->flags);
    if (rs->ctl) hipHostFree(rs->ctl);
    if (rs->arena) hipHostFree(rs->arena);
    if (rs->d_dev) hipFree(rs->d_dev);
    if (rs->d_done) hipFree(rs->d_done);
    if (rs->d_simd) hipFree(rs->d_simd);
    free(rs->inflight);
    free(rs);
}

#define RES_TRY(call) do { if ((call) != hipSuccess) { res_free(rs); return MYERS_E_HIP; } } while (0)

/* Pinned, coherent host memory the kernel reads and writes in place. */
static hipError_t res_host_alloc(void **host, void **dev, size_t bytes)
{
    hipError_t e = hipHostMalloc(host, bytes, hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) return e;
    memset(*host, 0, bytes);
    return hipHostGetDevicePointer(dev, *host, 0);
}

static int res_init(server_t *sv, void *buf)
{
    resident_t *rs = (resident_t *)calloc(1, sizeof(resident_t));
    if (rs == NULL) return MYERS_E_ALLOCATE_MEM;
    int rc = myers_serve_plan(buf, &rs->plan);
    if (rc) { free(rs); return rc; }
    if (rs->plan.num_ranks == 0) { free(rs); return MYERS_E_INVALID_ARGUMENT; }
    rs->inflight = (req_t **)calloc(RES_REQ_SLOTS, sizeof(req_t *));
    if (rs->inflight == NULL) { free(rs); return MYERS_E_ALLOCATE_MEM; }
    rs->cut = rs->plan.threshold >= 0.0f;
    /* One long item per SIMD: the grid holds two waves per SIMD (registers),
     * and a second long chain would take the wave that serves short items,
     * which then starve under load (measured: the served rate halves at 0.7
     * of the bulk rate). */
    rs->long_per_simd = 1u;
    rs->arena_bytes = sv->cfg.arenaBytes ? sv->cfg.arenaBytes : RES_DEFAULT_ARENA;
    RES_TRY(hipSetDevice(rs->plan.device));
    RES_TRY(hipStreamCreateWithFlags(&rs->stream, hipStreamNonBlocking));
    RES_TRY(hipEventCreateWithFlags(&rs->ev_done, hipEventDisableTiming));
    RES_TRY(res_host_alloc((void **)&rs->items, &rs->d_items,
                           sizeof(myers_serve_item_t) * RES_RING_ITEMS * MYERS_SERVE_RINGS));
    RES_TRY(res_host_alloc((void **)&rs->reqs, &rs->d_reqs, sizeof(myers_serve_req_t) * RES_REQ_SLOTS));
    RES_TRY(res_host_alloc((void **)&rs->flags, &rs->d_flags, sizeof(uint64_t) * RES_REQ_SLOTS));
    RES_TRY(res_host_alloc((void **)&rs->ctl, &rs->d_ctl, 256));
    RES_TRY(res_host_alloc((void **)&rs->arena, &rs->d_arena, rs->arena_bytes));
    RES_TRY(hipMalloc((void **)&rs->d_dev, sizeof(uint32_t) * MYERS_SERVE_DEV_WORDS));
    RES_TRY(hipMalloc((void **)&rs->d_done, sizeof(uint32_t) * RES_REQ_SLOTS));
    RES_TRY(hipMalloc((void **)&rs->d_simd, sizeof(uint32_t) * MYERS_SERVE_SIMD_KEYS));
    RES_TRY(hipMemset(rs->d_simd, 0, sizeof(uint32_t) * MYERS_SERVE_SIMD_KEYS));
    RES_TRY(hipMemset(rs->d_dev, 0, sizeof(uint32_t) * MYERS_SERVE_DEV_WORDS));
    RES_TRY(hipMemset(rs->d_done, 0, sizeof(uint32_t) * RES_REQ_SLOTS));
    RES_TRY(hipDeviceSynchronize());
    int occ = 0;
    if (myers_serve_occupancy(rs->cut, rs->plan.lds_bytes, &occ) != 0 || occ <= 0) occ = 1;
    rs->grid = (uint32_t)rs->plan.num_cus * (uint32_t)occ;
    sv->stats.residentWorkgroups = rs->grid;
    sv->stats.residentLdsBytes = rs->plan.lds_bytes;
    sv->res = rs;
    return MYERS_SUCCESS;
}

#define RES_ALIGN(x) (((x) + 255u) & ~(uint64_t)255u)

/* Admit the queue's head request if it fits now: 1 admitted (or failed for
 * good), 0 it must wait for completions.  Caller holds the lock. */
static int res_admit(server_t *sv, req_t *q)
{
    resident_t *rs = sv->res;
    const myers_serve_plan_t *pl = &rs->plan;
    const myersRequest_t *r = &q->r;
    const uint32_t nr = pl->num_ranks;
    uint32_t counts[64] = {0}, pos[64];
    uint32_t too_long = 0, nitems = 0, nring[MYERS_SERVE_RINGS] = {0, 0};
    for (uint32_t i = 0; i < r->numCandidates; i++) {
        const uint32_t w = (r->qinfo[r->cands[i].query].size + 31u) >> 5;
        if (w > pl->max_words) too_long++;
        else counts[pl->rank_of_words[w]]++;
    }
    for (uint32_t k = 0; k < nr; k++) {
        const uint32_t n = (counts[k] + pl->rank_slots[k] - 1u) / pl->rank_slots[k];
        nring[(pl->rank_key[k] >> 8) >= (uint32_t)MYERS_SERVE_LONG_LANES ? 0 : 1] += n;
        nitems += n;
    }
    const uint64_t o_peq = 0;
    const uint64_t o_qi = RES_ALIGN(o_peq + (uint64_t)r->numPEQEntries * sizeof(qryEntry_t));
    const uint64_t o_ca = RES_ALIGN(o_qi + (uint64_t)r->numQueries * sizeof(qryInfo_t));
    const uint64_t o_or = RES_ALIGN(o_ca + (uint64_t)r->numCandidates * sizeof(candInfo_t));
    const uint64_t o_re = RES_ALIGN(o_or + (uint64_t)r->numCandidates * sizeof(uint32_t));
    const uint64_t len = RES_ALIGN(o_re + (uint64_t)r->numCandidates * sizeof(resEntry_t));
    const uint64_t t = now_ns();
    if (len > rs->arena_bytes || nring[0] > RES_RING_ITEMS || nring[1] > RES_RING_ITEMS) {
        q->t_admit_ns = t;
        finish_request(sv, q, MYERS_E_CAPACITY, t);
        return 1;
    }
    const int busy = rs->oldest != rs->admitted;
    if (busy) {
        if (rs->admitted - rs->oldest >= RES_REQ_SLOTS) return 0;
        const req_t *old = rs->inflight[rs->oldest % RES_REQ_SLOTS];
        for (uint32_t g = 0; g < (uint32_t)MYERS_SERVE_RINGS; g++)
            if ((uint64_t)(uint32_t)(rs->published[g] - old->first_item[g]) + nring[g] > RES_RING_ITEMS) return 0;
        if (sv->inflight_bytes + q->bytes > sv->cfg.hbmBudgetBytes) return 0;
    }
    /* arena: contiguous bytes, skipping the end of the ring when they do not fit there */
    uint64_t start = rs->arena_head;
    const uint64_t at = start % rs->arena_bytes;
    if (at + len > rs->arena_bytes) start += rs->arena_bytes - at;
    if (start + len - rs->arena_tail > rs->arena_bytes) return 0;
    rs->arena_head = start + len;
    const uint64_t base = start % rs->arena_bytes;
    uint8_t *h = rs->arena + base;
    const uintptr_t d = (uintptr_t)rs->d_arena + base;
    memcpy(h + o_peq, r->peq, (size_t)r->numPEQEntries * sizeof(qryEntry_t));
    memcpy(h + o_qi, r->qinfo, (size_t)r->numQueries * sizeof(qryInfo_t));
    memcpy(h + o_ca, r->cands, (size_t)r->numCandidates * sizeof(candInfo_t));
    uint32_t *order = (uint32_t *)(h + o_or);
    resEntry_t *res = (resEntry_t *)(h + o_re);
    /* stable counting sort by rank (longest reads first); too-long reads get the sentinel */
    uint32_t acc = 0;
    for (uint32_t k = 0; k < nr; k++) { pos[k] = acc; acc += counts[k]; }
    for (uint32_t i = 0; i < r->numCandidates; i++) {
        const uint32_t w = (r->qinfo[r->cands[i].query].size + 31u) >> 5;
        if (w > pl->max_words) {
            res[i].column = 0;
            res[i].score = 0xFFFFFFFFu;
            continue;
        }
        order[pos[pl->rank_of_words[w]]++] = i;
    }
    const uint32_t slot = (uint32_t)(rs->admitted % RES_REQ_SLOTS);
    myers_serve_req_t *rq = &rs->reqs[slot];
    rq->peq = (const uint32_t *)(d + o_peq);
    rq->qinfo = (const uint32_t *)(d + o_qi);
    rq->cands = (const void *)(d + o_ca);
    rq->results = (void *)(d + o_re);
    rq->order = (const uint32_t *)(d + o_or);
    rq->num_peq_entries = r->numPEQEntries;
    rq->num_queries = r->numQueries;
    rq->num_items_cands = r->numCandidates - too_long;
    rq->mode = sv->cfg.flags & 1u;
    rq->distance = sv->cfg.distance;
    rq->threshold = pl->threshold;
    rq->ticket = q->ticket + 1u;       /* flags start at 0 */
    uint32_t item[MYERS_SERVE_RINGS] = {rs->published[0], rs->published[1]};
    acc = 0;
    for (uint32_t k = 0; k < nr; k++) {
        const uint32_t g = (pl->rank_key[k] >> 8) >= (uint32_t)MYERS_SERVE_LONG_LANES ? 0u : 1u;
        for (uint32_t c = 0; c < counts[k]; c += pl->rank_slots[k]) {
            myers_serve_item_t *it = &rs->items[(size_t)g * RES_RING_ITEMS + (item[g] & (RES_RING_ITEMS - 1u))];
            it->key = pl->rank_key[k];
            it->req = slot;
            it->first = acc + c;
            it->count = counts[k] - c < pl->rank_slots[k] ? counts[k] - c : pl->rank_slots[k];
            item[g]++;
        }
        acc += counts[k];
    }
    q->arena_start = start;
    q->arena_end = start + len;
    q->admit_idx = rs->admitted;
    for (uint32_t g = 0; g < (uint32_t)MYERS_SERVE_RINGS; g++) {
        q->first_item[g] = rs->published[g];
        q->nitems[g] = nring[g];
    }
    q->items = nitems;
    q->res = res;
    q->too_long = too_long != 0;
    q->t_admit_ns = t;
    q->state = REQ_ADMITTED;
    q->in_ring = 1;
    rs->inflight[slot] = q;
    rs->admitted++;
    sv->inflight_bytes += q->bytes;
    if (nitems == 0) return 1;                  /* completed by the next poll */
    rs->active++;
    sv->stats.residentItems += nitems;
    /* publish: the items and the request are written before the tails */
    for (uint32_t g = 0; g < (uint32_t)MYERS_SERVE_RINGS; g++) {
        if (item[g] == rs->published[g]) continue;
        rs->published[g] = item[g];
        __atomic_store_n(&rs->ctl[RES_CTL_TAIL + g * MYERS_SERVE_TAIL_STRIDE], item[g], __ATOMIC_RELEASE);
    }
    return 1;
}

/* Hand back finished requests and release their arena and slots, oldest
 * first.  Caller holds the lock. */
static int res_complete(server_t *sv)
{
    resident_t *rs = sv->res;
    int progressed = 0;
    const uint64_t t0 = now_ns();
    for (uint64_t i = rs->oldest; i != rs->admitted; i++) {
        req_t *q = rs->inflight[i % RES_REQ_SLOTS];
        if (q == NULL || q->state != REQ_ADMITTED) continue;
        const uint32_t slot = (uint32_t)(i % RES_REQ_SLOTS);
        if (q->items && __atomic_load_n(&rs->flags[slot], __ATOMIC_ACQUIRE) != q->ticket + 1u) continue;
        if (q->items) rs->active--;
        memcpy(q->r.results, q->res, (size_t)q->r.numCandidates * sizeof(resEntry_t));
        sv->inflight_bytes -= q->bytes;
        finish_request(sv, q, q->too_long ? MYERS_E_QUERY_TOO_LONG : MYERS_SUCCESS, now_ns());
        progressed = 1;
    }
    while (rs->oldest != rs->admitted) {
        req_t *q = rs->inflight[rs->oldest % RES_REQ_SLOTS];
        if (q != NULL && q->state != REQ_DONE) break;
        rs->inflight[rs->oldest % RES_REQ_SLOTS] = NULL;
        rs->oldest++;
        rs->arena_tail = rs->oldest == rs->admitted ? rs->arena_head
                                                    : rs->inflight[rs->oldest % RES_REQ_SLOTS]->arena_start;
        if (q != NULL) {
            q->in_ring = 0;
            req_try_free(sv, q);
        }
    }
    if (progressed) {
        sv->stats.completeMs += (double)(now_ns() - t0) / 1e6;
        pthread_cond_broadcast(&sv->cv_done);
    }
    return progressed;
}

/* Launch the kernel when requests wait on it and no launch is running. */
static int res_ensure_running(server_t *sv)
{
    resident_t *rs = sv->res;
    if (rs->active == 0) return MYERS_SUCCESS;
    if (rs->running) {
        hipError_t e = hipEventQuery(rs->ev_done);
        if (e == hipErrorNotReady) return MYERS_SUCCESS;
        if (e != hipSuccess) return MYERS_E_HIP;
        rs->running = 0;                        /* it went idle and left: start another */
    }
    rs->gen = rs->gen + 1u ? rs->gen + 1u : 1u;
    myers_serve_args_t a;
    memset(&a, 0, sizeof(a));
    a.ref = rs->plan.d_ref;
    a.ref_words = rs->plan.ref_words;
    a.ref_size = rs->plan.ref_size;
    a.items = (const myers_serve_item_t *)rs->d_items;
    a.reqs = (const myers_serve_req_t *)rs->d_reqs;
    a.host_tail = (const uint32_t *)rs->d_ctl + RES_CTL_TAIL;
    a.host_stop = (const uint32_t *)rs->d_ctl + RES_CTL_STOP;
    a.flags = (uint64_t *)rs->d_flags;
    a.done = rs->d_done;
    a.dev = rs->d_dev;
    a.simd_busy = rs->d_simd;
    a.ring_mask = RES_RING_ITEMS - 1u;
    a.req_mask = RES_REQ_SLOTS - 1u;
    a.gen = rs->gen;
    a.idle_ticks = RES_IDLE_TICKS;
    a.long_per_simd = rs->long_per_simd;
    a.poll_ticks = RES_POLL_TICKS;
    if (myers_launch_serve(&a, rs->cut, rs->grid, rs->plan.lds_bytes, rs->stream) != 0) return MYERS_E_HIP;
    if (hipEventRecord(rs->ev_done, rs->stream) != hipSuccess) return MYERS_E_HIP;
    rs->running = 1;
    sv->stats.batches++;
    return MYERS_SUCCESS;
}

/* A failed launch fails every request in flight. */
static void res_fail_inflight(server_t *sv, int rc)
{
    resident_t *rs = sv->res;
    const uint64_t t = now_ns();
    for (uint64_t i = rs->oldest; i != rs->admitted; i++) {
        req_t *q = rs->inflight[i % RES_REQ_SLOTS];
        if (q == NULL || q->state != REQ_ADMITTED) continue;
        sv->inflight_bytes -= q->bytes;
        finish_request(sv, q, rc, t);
    }
    rs->active = 0;
    pthread_cond_broadcast(&sv->cv_done);
}

static void *resident_main(void *arg)
{
    server_t *sv = (server_t *)arg;
    resident_t *rs = sv->res;
    hipSetDevice(rs->plan.device);
    pthread_mutex_lock(&sv->mu);
    for (;;) {
        res_complete(sv);
        const uint64_t t0 = now_ns();
        int admitted = 0;
        while (sv->head) {
            req_t *q = sv->head;
            if (!res_admit(sv, q)) break;
            sv->head = q->next;
            if (!sv->head) sv->tail = NULL;
            sv->queued_cells -= q->cells;
            admitted = 1;
        }
        if (admitted) {
            sv->stats.admitMs += (double)(now_ns() - t0) / 1e6;
            pthread_cond_broadcast(&sv->cv_done);     /* requests failed at admission */
        }
        int rc = res_ensure_running(sv);
        if (rc != MYERS_SUCCESS) res_fail_inflight(sv, rc);
        if (rs->oldest == rs->admitted) {
            if (sv->head) continue;
            if (sv->stopping) break;
            pthread_cond_wait(&sv->cv_work, &sv->mu);
            continue;
        }
        pthread_mutex_unlock(&sv->mu);
        struct timespec ts = {0, (long)sv->cfg.pollMicros * 1000L};
        nanosleep(&ts, NULL);
        pthread_mutex_lock(&sv->mu);
    }
    /* every request is done: ask the running launch to leave, and wait for it */
    __atomic_store_n(&rs->ctl[RES_CTL_STOP], rs->gen, __ATOMIC_RELEASE);
    pthread_mutex_unlock(&sv->mu);
    if (rs->running) {
        hipEventSynchronize(rs->ev_done);
        rs->running = 0;
    }
    return NULL;
}

static void *server_main(void *arg)
{
    server_t *sv = (server_t *)arg;
    pthread_mutex_lock(&sv->mu);
    for (;;) {
        complete_ready(sv);
        int launched = 0;
        for (uint32_t i = 0; i < sv->nslots && sv->head; i++)
            if (!sv->slots[i].busy) launched |= launch_batch(sv, &sv->slots[i]);
        int busy = 0;
        for (uint32_t i = 0; i < sv->nslots; i++) busy |= sv->slots[i].busy;
        if (launched) continue;
        if (!busy) {
            if (sv->head) continue;           /* a freed slot can take them now */
            if (sv->stopping) break;
            pthread_cond_wait(&sv->cv_work, &sv->mu);
            continue;
        }
        pthread_mutex_unlock(&sv->mu);
        struct timespec ts = {0, (long)sv->cfg.pollMicros * 1000L};
        nanosleep(&ts, NULL);
        pthread_mutex_lock(&sv->mu);
    }
    pthread_mutex_unlock(&sv->mu);
    return NULL;
}

int myersServeStart(void **server, void **myersBuffers, uint32_t numBuffers, const myersServeConfig_t *cfg)
{
    if (server == NULL || myersBuffers == NULL || numBuffers == 0) return MYERS_E_INVALID_ARGUMENT;
    server_t *sv = (server_t *)calloc(1, sizeof(server_t));
    if (sv == NULL) return MYERS_E_ALLOCATE_MEM;
    myersServeDefaultConfig(&sv->cfg);
    if (cfg) {
        /* callers built against an older, shorter config keep the defaults of the new fields */
        const size_t n = cfg->structSize && cfg->structSize < sizeof(*cfg) ? cfg->structSize : sizeof(*cfg);
        memcpy(&sv->cfg, cfg, n);
        sv->cfg.structSize = sizeof(sv->cfg);
    }
    if (sv->cfg.maxBatchCells == 0) sv->cfg.maxBatchCells = DEFAULT_BATCH_CELLS;
    if (sv->cfg.hbmBudgetBytes == 0) sv->cfg.hbmBudgetBytes = 200000000000ull;
    if (sv->cfg.pollMicros == 0) sv->cfg.pollMicros = 20;
    sv->stats.structSize = sizeof(sv->stats);
    sv->slots = (slot_t *)calloc(numBuffers, sizeof(slot_t));
    if (sv->slots == NULL) {
        free(sv);
        return MYERS_E_ALLOCATE_MEM;
    }
    sv->nslots = numBuffers;
    if (sv->cfg.flags & MYERS_SERVE_RESIDENT) {
        int rc = res_init(sv, myersBuffers[0]);
        if (rc != MYERS_SUCCESS) {
            free(sv->slots);
            free(sv);
            return rc;
        }
    }
    for (uint32_t i = 0; i < numBuffers; i++) {
        sv->slots[i].buf = myersBuffers[i];
        sv->slots[i].max_entries = getMaxPEQEntries(myersBuffers[i]);
        sv->slots[i].max_queries = getMaxQueries(myersBuffers[i]);
        sv->slots[i].max_cands = getMaxCandidates(myersBuffers[i]);
    }
    pthread_mutex_init(&sv->mu, NULL);
    pthread_cond_init(&sv->cv_work, NULL);
    pthread_cond_init(&sv->cv_done, NULL);
    if (pthread_create(&sv->thread, NULL, sv->res ? resident_main : server_main, sv) != 0) {
        res_free(sv->res);
        free(sv->slots);
        free(sv);
        return MYERS_E_ALLOCATE_MEM;
    }
    sv->thread_started = 1;
    *server = sv;
    return MYERS_SUCCESS;
}

int myersServeSubmit(void *server, const myersRequest_t *req, uint64_t *ticket)
{
    server_t *sv = (server_t *)server;
    if (sv == NULL || req == NULL || ticket == NULL) return MYERS_E_INVALID_ARGUMENT;
    if ((req->numPEQEntries && !req->peq) || (req->numQueries && !req->qinfo) ||
        (req->numCandidates && (!req->cands || !req->results)))
        return MYERS_E_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < req->numCandidates; i++)
        if (req->cands[i].query >= req->numQueries) return MYERS_E_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < req->numQueries; i++)
        if ((uint64_t)req->qinfo[i].posEntry + (req->qinfo[i].size + 127u) / 128u > req->numPEQEntries)
            return MYERS_E_INVALID_ARGUMENT;
    req_t *q = (req_t *)calloc(1, sizeof(req_t));
    if (q == NULL) return MYERS_E_ALLOCATE_MEM;
    q->t_submit_ns = now_ns();
    q->r = *req;
    q->bytes = request_bytes(req);
    q->cells = request_cells(req, sv->cfg.distance);
    pthread_mutex_lock(&sv->mu);
    if (sv->stopping) {
        pthread_mutex_unlock(&sv->mu);
        free(q);
        return MYERS_E_INVALID_ARGUMENT;
    }
    if (sv->next_ticket - sv->table_base == sv->table_cap) {
        /* drop the freed prefix; grow only when the live span fills the table */
        uint64_t k = 0;
        while (k < sv->table_cap && sv->table[k] == NULL) k++;
        if (k > 0) {
            memmove(sv->table, sv->table + k, (size_t)(sv->table_cap - k) * sizeof(req_t *));
            sv->table_base += k;
        }
        if (sv->next_ticket - sv->table_base == sv->table_cap) {
            uint64_t cap = sv->table_cap ? 2u * sv->table_cap : 1024u;
            req_t **t = (req_t **)realloc(sv->table, cap * sizeof(req_t *));
            if (t == NULL) {
                pthread_mutex_unlock(&sv->mu);
                free(q);
                return MYERS_E_ALLOCATE_MEM;
            }
            sv->table = t;
            sv->table_cap = cap;
        }
    }
    q->ticket = sv->next_ticket++;
    sv->table[q->ticket - sv->table_base] = q;
    sv->stats.liveRequests++;
    if (sv->tail) sv->tail->next = q;
    else sv->head = q;
    sv->tail = q;
    sv->queued_cells += q->cells;
    *ticket = q->ticket;
    pthread_cond_signal(&sv->cv_work);
    pthread_mutex_unlock(&sv->mu);
    return MYERS_SUCCESS;
}

/* Wait for a ticket and release it (a ticket is waited on once).  Caller
 * holds the lock. */
static int wait_locked(server_t *sv, uint64_t ticket, uint64_t *t_submit, uint64_t *t_admit, uint64_t *t_done)
{
    req_t *q = req_of(sv, ticket);
    if (q == NULL) return MYERS_E_INVALID_ARGUMENT;
    q->waiters++;
    while (q->state != REQ_DONE) pthread_cond_wait(&sv->cv_done, &sv->mu);
    q->waiters--;
    if (t_submit) *t_submit = q->t_submit_ns;
    if (t_admit) *t_admit = q->t_admit_ns;
    if (t_done) *t_done = q->t_done_ns;
    const int error = q->error;
    q->released = 1;                      /* a later wait finds the ticket gone */
    req_try_free(sv, q);
    return error;
}

int myersServeWait(void *server, uint64_t ticket, double *latencyMs)
{
    server_t *sv = (server_t *)server;
    if (sv == NULL) return MYERS_E_INVALID_ARGUMENT;
    uint64_t ts = 0, td = 0;
    pthread_mutex_lock(&sv->mu);
    int rc = wait_locked(sv, ticket, &ts, NULL, &td);
    pthread_mutex_unlock(&sv->mu);
    if (latencyMs) *latencyMs = (double)(td - ts) / 1e6;
    return rc;
}

int myersServeGetStats(void *server, myersServeStats_t *out)
{
    server_t *sv = (server_t *)server;
    if (sv == NULL || out == NULL) return MYERS_E_INVALID_ARGUMENT;
    /* callers built against an older, shorter struct get its fields only */
    const size_t n = out->structSize && out->structSize < sizeof(*out) ? out->structSize : sizeof(*out);
    pthread_mutex_lock(&sv->mu);
    memcpy(out, &sv->stats, n);
    pthread_mutex_unlock(&sv->mu);
    out->structSize = (uint32_t)n;
    return MYERS_SUCCESS;
}

int myersServeStop(void *server)
{
    server_t *sv = (server_t *)server;
    if (sv == NULL) return MYERS_E_INVALID_ARGUMENT;
    pthread_mutex_lock(&sv->mu);
    sv->stopping = 1;
    pthread_cond_signal(&sv->cv_work);
    pthread_mutex_unlock(&sv->mu);
    if (sv->thread_started) pthread_join(sv->thread, NULL);
    res_free(sv->res);
    for (uint64_t i = sv->table_base; i < sv->next_ticket; i++) free(sv->table[i - sv->table_base]);
    free(sv->table);
    for (uint32_t i = 0; i < sv->nslots; i++) {
        free(sv->slots[i].reqs);
        free(sv->slots[i].cand_off);
    }
    free(sv->slots);
    pthread_mutex_destroy(&sv->mu);
    pthread_cond_destroy(&sv->cv_work);
    pthread_cond_destroy(&sv->cv_done);
    free(sv);
    return MYERS_SUCCESS;
}

typedef struct {
    server_t             *sv;
    uint32_t              n;
    const myersRequest_t *reqs;
    const uint64_t       *arrival;
    uint64_t              t0;
    uint64_t             *tickets;
    int                   rc;
} openloop_t;

static void *submitter_main(void *arg)
{
    openloop_t *ol = (openloop_t *)arg;
    for (uint32_t i = 0; i < ol->n; i++) {
        const uint64_t due = ol->t0 + ol->arrival[i];
        struct timespec ts = {(time_t)(due / 1000000000ull), (long)(due % 1000000000ull)};
        while (clock_nanosleep(CLOCK_MONOTONIC, TIMER_ABSTIME, &ts, NULL) == EINTR) {
        }
        int rc = myersServeSubmit(ol->sv, &ol->reqs[i], &ol->tickets[i]);
        if (rc != MYERS_SUCCESS) {
            ol->rc = rc;
            ol->tickets[i] = UINT64_MAX;
        }
    }
    return NULL;
}

int myersServeRunOpenLoop(void *server, uint32_t numRequests, const myersRequest_t *reqs,
                          const uint64_t *arrivalNs, double *latencyMs, double *wallMs)
{
    return myersServeRunOpenLoopEx(server, numRequests, reqs, arrivalNs, latencyMs, NULL, wallMs);
}

int myersServeRunOpenLoopEx(void *server, uint32_t numRequests, const myersRequest_t *reqs,
                            const uint64_t *arrivalNs, double *latencyMs, double *queueMs, double *wallMs)
{
    server_t *sv = (server_t *)server;
    if (sv == NULL || (numRequests && (!reqs || !arrivalNs || !latencyMs))) return MYERS_E_INVALID_ARGUMENT;
    for (uint32_t i = 1; i < numRequests; i++)
        if (arrivalNs[i] < arrivalNs[i - 1]) return MYERS_E_INVALID_ARGUMENT;
    openloop_t ol;
    memset(&ol, 0, sizeof(ol));
    ol.sv = sv;
    ol.n = numRequests;
    ol.reqs = reqs;
    ol.arrival = arrivalNs;
    ol.tickets = (uint64_t *)malloc(sizeof(uint64_t) * (numRequests ? numRequests : 1u));
    if (ol.tickets == NULL) return MYERS_E_ALLOCATE_MEM;
    for (uint32_t i = 0; i < numRequests; i++) ol.tickets[i] = UINT64_MAX;
    ol.t0 = now_ns() + 2000000ull;   /* first arrival 2 ms from now */
    pthread_t th;
    if (pthread_create(&th, NULL, submitter_main, &ol) != 0) {
        free(ol.tickets);
        return MYERS_E_ALLOCATE_MEM;
    }
    pthread_join(th, NULL);
    int rc = ol.rc;
    uint64_t last_done = ol.t0;
    pthread_mutex_lock(&sv->mu);
    for (uint32_t i = 0; i < numRequests; i++) {
        if (ol.tickets[i] == UINT64_MAX) {
            latencyMs[i] = -1.0;
            if (queueMs) queueMs[i] = -1.0;
            continue;
        }
        uint64_t ta = 0, td = 0;
        int e = wait_locked(sv, ol.tickets[i], NULL, &ta, &td);
        if (e != MYERS_SUCCESS && rc == MYERS_SUCCESS) rc = e;
        latencyMs[i] = (double)(td - (ol.t0 + arrivalNs[i])) / 1e6;
        if (queueMs) queueMs[i] = (double)(int64_t)(ta - (ol.t0 + arrivalNs[i])) / 1e6;
        if (td > last_done) last_done = td;
    }
    pthread_mutex_unlock(&sv->mu);
    if (wallMs) *wallMs = (double)(last_done - ol.t0) / 1e6;
    free(ol.tickets);
    return rc;
}
