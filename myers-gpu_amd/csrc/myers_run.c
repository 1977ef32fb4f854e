/*
 * myers-run: batch verification of query files through the buffer API, the
 * counterpart of the reference's harness src-api/test.c:30-207.
 *
 *   myers-run [-x] [-o OUT] [-t ERR] [-b MIB] [-n CANDS] DISTANCE REF.4bits.ref QRY [QRY ...]
 *
 * Each QRY (.gem.qry, or a GEM .prof converted on load) gets its own buffer;
 * buffers run concurrently, one OpenMP thread each, as test.c:184 does.  The
 * results go to OUT (default "<first QRY>.gem-int.pad.gpu", test.c:80) as
 * "<count>\n" then "<id> <score> <column>\n" per candidate, ids running over
 * the files in order (test.c:85-94).
 *
 * Differences from test.c: a file larger than the buffer's candidate capacity
 * is sent in several batches instead of being cut to the capacity
 * (test.c:114-117); -x selects exact-start windows (src/myers-4bits.c)
 * instead of the API's aligned windows; -t ERR also reports how many
 * candidates score <= (uint32)((float)m * ERR) (src/checksum-2.c:88-106);
 * -b sets the device bytes per buffer used to size capacities; -n caps the
 * candidates per batch.
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "myers-amd.h"
#include "myers-interface.h"
#include "myers-tools.h"

typedef struct {
    const char   *path;
    myersGemQry_t q;
    resEntry_t   *results;
    int           error;
    double        kernel_ms;
    uint32_t      batches;
} job_t;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int ends_with(const char *s, const char *suffix)
{
    const size_t a = strlen(s), b = strlen(suffix);
    return a >= b && strcmp(s + a - b, suffix) == 0;
}

static void usage(void)
{
    fprintf(stderr,
            "usage: myers-run [-x] [-o OUT] [-t ERR] [-b MIB] [-n CANDS] DISTANCE REF.4bits.ref QRY [QRY ...]\n"
            "  QRY   .gem.qry batch, or a GEM .prof file (converted on load)\n"
            "  -x    exact-start windows (default: windows aligned to 8 bases, as the API)\n"
            "  -o    results file (default <first QRY>.gem-int.pad.gpu)\n"
            "  -t    also count candidates with score <= (uint32)(m * ERR)\n"
            "  -b    device MiB per buffer used to size capacities (default: library's)\n"
            "  -n    at most CANDS candidates per batch (default: the buffer's capacity)\n");
    exit(2);
}

/* Send the whole file through one buffer, in as many batches as the candidate
 * capacity needs.  Peq and qinfo must fit in one batch. */
static int run_job(void *buf, job_t *j, float distance, uint32_t flags, uint32_t max_batch)
{
    const myersGemQry_t *q = &j->q;
    if (q->numEntries > getMaxPEQEntries(buf) || q->numQueries > getMaxQueries(buf)) {
        fprintf(stderr, "%s: %u Peq entries / %u queries exceed the buffer capacity (%u / %u); "
                "raise -b\n", j->path, q->numEntries, q->numQueries, getMaxPEQEntries(buf), getMaxQueries(buf));
        return MYERS_E_CAPACITY;
    }
    uint32_t cap = getMaxCandidates(buf);
    if (max_batch != 0 && max_batch < cap) cap = max_batch;
    if (cap == 0) return MYERS_E_CAPACITY;
    memcpy(getPEQBuffer(buf), q->peq, sizeof(qryEntry_t) * q->numEntries);
    memcpy(getPEQInfoBuffer(buf), q->qinfo, sizeof(qryInfo_t) * q->numQueries);
    for (uint32_t c0 = 0; c0 < q->numCandidates || (c0 == 0 && q->numCandidates == 0); c0 += cap) {
        const uint32_t n = q->numCandidates - c0 < cap ? q->numCandidates - c0 : cap;
        memcpy(getCandidatesBuffer(buf), q->cands + c0, sizeof(candInfo_t) * n);
        int rc = myersSendEx(buf, distance, q->numEntries, q->numQueries, n, flags);
        if (rc == 0) rc = myersReceiveEx(buf);
        if (rc != 0) return rc;
        memcpy(j->results + c0, getResultsBuffer(buf), sizeof(resEntry_t) * n);
        myersStats_t st;
        st.structSize = sizeof(st);
        if (myersGetStats(buf, &st) == 0) j->kernel_ms += st.kernelMs;
        j->batches++;
        if (q->numCandidates == 0) break;
    }
    return 0;
}

int main(int argc, char **argv)
{
    uint32_t flags = MYERS_WINDOW_ALIGNED;
    const char *out_path = NULL;
    double err = -1.0, mib = 0.0;
    uint32_t max_batch = 0;
    int opt;
    while ((opt = getopt(argc, argv, "xo:t:b:n:h")) != -1) {
        switch (opt) {
        case 'x': flags = MYERS_WINDOW_EXACT; break;
        case 'o': out_path = optarg; break;
        case 't': err = atof(optarg); break;
        case 'b': mib = atof(optarg); break;
        case 'n': max_batch = (uint32_t)strtoul(optarg, NULL, 10); break;
        default: usage();
        }
    }
    if (argc - optind < 3) usage();
    const float distance = (float)atof(argv[optind]);
    const char *ref_path = argv[optind + 1];
    const int num = argc - optind - 2;
    job_t *jobs = (job_t *)calloc((size_t)num, sizeof(job_t));
    if (jobs == NULL) return 1;

    printf("numBuffers: %d \n", num);
    uint64_t bases = 0, queries = 0, cands = 0;
    for (int i = 0; i < num; i++) {
        job_t *j = &jobs[i];
        j->path = argv[optind + 2 + i];
        printf("qryFile: %s \n", j->path);
        const int rc = ends_with(j->path, ".prof") ? myersParseProf(j->path, &j->q)
                                                   : myersReadGemQry(j->path, &j->q);
        if (rc != 0) {
            fprintf(stderr, "cannot read queries from %s\n", j->path);
            return 1;
        }
        j->results = (resEntry_t *)calloc((size_t)j->q.numCandidates + 1u, sizeof(resEntry_t));
        if (j->results == NULL) return 1;
        bases += j->q.totalSizeQueries;
        queries += j->q.numQueries;
        cands += j->q.numCandidates;
    }

    /* shape the capacities after the data (test.c:188 passes fixed 1000 / 10) */
    const int32_t avg_query = queries ? (int32_t)((bases + queries - 1) / queries) : 1000;
    const int32_t cpq = queries ? (int32_t)((cands + queries - 1) / queries) : 10;
    myersConfig_t cfg;
    myersDefaultConfig(&cfg);
    cfg.bufferBytes = (uint64_t)(mib * 1048576.0);
    void **buffers = NULL;
    int rc = myersInitEx(&buffers, (uint32_t)num, ref_path, avg_query > 0 ? avg_query : 1000,
                         cpq > 0 ? cpq : 10, &cfg);
    if (rc != 0) {
        fprintf(stderr, "initMyers: %s\n", myersGetErrorString(rc));
        return 1;
    }

    const double t0 = now_s();
    #pragma omp parallel for schedule(static, 1)
    for (int i = 0; i < num; i++) jobs[i].error = run_job(buffers[i], &jobs[i], distance, flags, max_batch);
    const double t1 = now_s();
    myersEnd(&buffers);
    printf("TOTAL TIME: \t %f \n", t1 - t0);

    int failed = 0;
    for (int i = 0; i < num; i++) {
        if (jobs[i].error != 0) {
            fprintf(stderr, "%s: %s\n", jobs[i].path, myersGetErrorString(jobs[i].error));
            failed = 1;
        }
    }
    if (failed) return 1;

    char *default_out = NULL;
    if (out_path == NULL) {
        if (asprintf(&default_out, "%s.gem-int.pad.gpu", jobs[0].path) < 0) return 1;
        out_path = default_out;
    }
    FILE *fp = fopen(out_path, "w");
    if (fp == NULL) {
        fprintf(stderr, "cannot write %s\n", out_path);
        return 1;
    }
    fprintf(fp, "%llu\n", (unsigned long long)cands);
    uint64_t id = 0, under = 0;
    double kernel_ms = 0.0;
    uint64_t cells = 0;
    for (int i = 0; i < num; i++) {
        const myersGemQry_t *q = &jobs[i].q;
        kernel_ms += jobs[i].kernel_ms;
        for (uint32_t c = 0; c < q->numCandidates; c++, id++) {
            const resEntry_t r = jobs[i].results[c];
            fprintf(fp, "%llu %u %u\n", (unsigned long long)id, r.score, r.column);
            const uint32_t m = q->qinfo[q->cands[c].query].size;
            volatile float n_f = (float)m * (1.0f + 2.0f * distance);
            cells += (uint64_t)m * (uint64_t)(uint32_t)n_f;
            if (err >= 0.0) {
                volatile float t_f = (float)m * (float)err;
                if (r.score <= (uint32_t)t_f) under++;
            }
        }
    }
    fclose(fp);
    printf("candidates: %llu  kernel: %.3f ms  %.1f GCUPS  results: %s\n", (unsigned long long)cands,
           kernel_ms, kernel_ms > 0 ? (double)cells / (kernel_ms * 1e-3) / 1e9 : 0.0, out_path);
    if (err >= 0.0) printf("under threshold (e=%g): %llu of %llu\n", err, (unsigned long long)under,
                           (unsigned long long)cands);
    for (int i = 0; i < num; i++) {
        myersFreeGemQry(&jobs[i].q);
        free(jobs[i].results);
    }
    free(jobs);
    free(default_out);
    return 0;
}
