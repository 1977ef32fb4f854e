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
 */
#define _GNU_SOURCE
#include <errno.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "myers-interface.h"
#include "myers-amd.h"
#include "myers_internal.h"

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
} req_t;

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
    req_t            **table;
    uint64_t           table_cap, next_ticket;
    uint64_t           inflight_bytes;
    int                stopping;
    int                thread_started;
    myersServeStats_t  stats;
} server_t;

static uint64_t now_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

/* Device bytes a request occupies while admitted: Peq, qinfo, candidates,
 * results and the binning slot of each candidate. */
/* About 2.5 ms of verification at 100 TCUPS: short enough that several
 * batches are in flight and a request does not wait behind a long one. */
#define DEFAULT_BATCH_CELLS 250000000000ull

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
    while (sv->head) {
        req_t *q = sv->head;
        const myersRequest_t *r = &q->r;
        if (r->numPEQEntries > s->max_entries || r->numQueries > s->max_queries || r->numCandidates > cap_c) {
            /* can never fit a slot: fail it */
            sv->head = q->next;
            if (!sv->head) sv->tail = NULL;
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
        if (s->nreq > 0 && cells + q->cells > sv->cfg.maxBatchCells) break;
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
    sv->stats.admitMs += (double)(now_ns() - t_start) / 1e6;
    if (s->nreq > sv->stats.maxBatchRequests) sv->stats.maxBatchRequests = s->nreq;
    return 1;
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
    for (uint32_t i = 0; i < numBuffers; i++) {
        sv->slots[i].buf = myersBuffers[i];
        sv->slots[i].max_entries = getMaxPEQEntries(myersBuffers[i]);
        sv->slots[i].max_queries = getMaxQueries(myersBuffers[i]);
        sv->slots[i].max_cands = getMaxCandidates(myersBuffers[i]);
    }
    pthread_mutex_init(&sv->mu, NULL);
    pthread_cond_init(&sv->cv_work, NULL);
    pthread_cond_init(&sv->cv_done, NULL);
    if (pthread_create(&sv->thread, NULL, server_main, sv) != 0) {
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
    if (sv->next_ticket == sv->table_cap) {
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
    q->ticket = sv->next_ticket++;
    sv->table[q->ticket] = q;
    if (sv->tail) sv->tail->next = q;
    else sv->head = q;
    sv->tail = q;
    *ticket = q->ticket;
    pthread_cond_signal(&sv->cv_work);
    pthread_mutex_unlock(&sv->mu);
    return MYERS_SUCCESS;
}

static int wait_locked(server_t *sv, uint64_t ticket, uint64_t *t_submit, uint64_t *t_admit, uint64_t *t_done)
{
    if (ticket >= sv->next_ticket) return MYERS_E_INVALID_ARGUMENT;
    req_t *q = sv->table[ticket];
    while (q->state != REQ_DONE) pthread_cond_wait(&sv->cv_done, &sv->mu);
    if (t_submit) *t_submit = q->t_submit_ns;
    if (t_admit) *t_admit = q->t_admit_ns;
    if (t_done) *t_done = q->t_done_ns;
    return q->error;
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
    for (uint64_t i = 0; i < sv->next_ticket; i++) free(sv->table[i]);
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
