/*
 * Data preparation for the candidate-verification library: seeded synthetic
 * references and read batches, Peq construction, 4-bit packing and the
 * .4bits.ref format.  See include/myers-tools.h for the generator's
 * definition.  OpenMP; every random draw comes from counter-based streams,
 * so the output does not depend on the number of threads.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "myers-tools.h"

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t hash3(uint64_t seed, uint64_t tag, uint64_t i)
{
    return mix64(seed ^ mix64(tag * 0x100000001B3ull + i));
}

typedef struct { uint64_t s; } rng_t;

static inline uint64_t rng_next(rng_t *r)
{
    r->s += 0x9E3779B97F4A7C15ull;
    uint64_t z = r->s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double rng_unif(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static inline uint64_t rng_below(rng_t *r, uint64_t n)
{
    return (uint64_t)(((unsigned __int128)rng_next(r) * n) >> 64);
}

/* P(k) = p (1-p)^(k-1), k >= 1 */
static inline uint32_t rng_geometric(rng_t *r, double p)
{
    double u = 1.0 - rng_unif(r);                 /* (0, 1] */
    double k = floor(log(u) / log1p(-p));
    return 1u + (uint32_t)(k > 1e6 ? 1e6 : k);
}

static inline uint32_t rng_poisson(rng_t *r, double mean)
{
    const double limit = exp(-mean);
    double prod = rng_unif(r);
    uint32_t k = 0;
    while (prod > limit && k < 100000u) {
        prod *= rng_unif(r);
        k++;
    }
    return k;
}

static inline uint32_t window_len(uint32_t m, float d)
{
    volatile float factor = 1.0f + 2.0f * d;
    volatile float prod = (float)m * factor;
    return (uint32_t)prod;
}

static inline uint32_t ref_code(const uint32_t *words, uint64_t i)
{
    return (words[i >> 3] >> ((uint32_t)(i & 7u) * 4u)) & 0xFu;
}

/* ------------------------------------------------------------- reference */

void myersSynthReference(uint64_t seed, uint64_t size, double nRunProb, uint32_t *words)
{
    const uint64_t nwords = (size + 7u) / 8u;
    const int64_t nchunks = (int64_t)((size + 63u) / 64u);
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nchunks; c++) {
        const uint64_t b0 = hash3(seed, 'B', 2u * (uint64_t)c);
        const uint64_t b1 = hash3(seed, 'B', 2u * (uint64_t)c + 1u);
        const uint64_t hr = hash3(seed, 'R', (uint64_t)c);
        uint32_t run_lo = 64, run_hi = 64;
        if ((double)(hr >> 11) * 0x1.0p-53 < nRunProb) {
            rng_t r = { hash3(seed, 'L', (uint64_t)c) };
            run_lo = (uint32_t)rng_below(&r, 64);
            uint32_t len = rng_geometric(&r, 0.1);
            run_hi = run_lo + len > 64u ? 64u : run_lo + len;
        }
        for (uint32_t w = 0; w < 8u; w++) {
            const uint64_t k = 8u * (uint64_t)c + w;
            if (k >= nwords) break;
            uint32_t word = 0;
            for (uint32_t i = 0; i < 8u; i++) {
                const uint32_t j = 8u * w + i;
                const uint64_t g = 64u * (uint64_t)c + j;
                uint32_t code;
                if (g >= size || (j >= run_lo && j < run_hi)) code = 4u;
                else code = (uint32_t)(((j < 32u ? b0 >> (2u * j) : b1 >> (2u * (j - 32u)))) & 3u);
                word |= code << (4u * i);
            }
            words[k] = word;
        }
    }
}

/* ------------------------------------------------------------------ Peq */

void myersBuildPeq(const uint8_t *codes, uint32_t m, qryEntry_t *out)
{
    const uint32_t entries = (m + 127u) / 128u;
    memset(out, 0, sizeof(qryEntry_t) * entries);
    for (uint32_t i = 0; i < m; i++) {
        const uint32_t b = codes[i] > 3u ? 4u : codes[i];
        out[i >> 7].bitmap[b][(i >> 5) & 3u] |= 1u << (i & 31u);
    }
}

void myersPackCodes(const uint8_t *codes, uint64_t n, uint32_t *out)
{
    const int64_t nw = (int64_t)((n + 7u) / 8u);
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nw; k++) {
        uint32_t word = 0;
        for (uint32_t i = 0; i < 8u; i++) {
            const uint64_t idx = 8u * (uint64_t)k + i;
            uint32_t c = idx < n ? codes[idx] : 4u;
            if (c > 3u) c = 4u;
            word |= c << (4u * i);
        }
        out[k] = word;
    }
}

static inline uint32_t fasta_code(unsigned char ch)
{
    switch (ch) {
    case 'A': case 'a': return 0u;
    case 'C': case 'c': return 1u;
    case 'G': case 'g': return 2u;
    case 'T': case 't': return 3u;
    default: return 4u;
    }
}

uint64_t myersPackFasta(const char *text, uint64_t textLen, uint32_t *out, uint64_t maxWords)
{
    uint64_t n = 0, i = 0;
    uint32_t word = 0;
    int header = 0, line_start = 1;
    for (i = 0; i < textLen; i++) {
        const unsigned char ch = (unsigned char)text[i];
        if (line_start && ch == '>') header = 1;
        if (ch == '\n' || ch == '\r') {
            header = 0;
            line_start = 1;
            continue;
        }
        line_start = 0;
        if (header) continue;
        word |= fasta_code(ch) << (4u * (uint32_t)(n & 7u));
        n++;
        if ((n & 7u) == 0u) {
            if (n / 8u <= maxWords) out[n / 8u - 1u] = word;
            word = 0;
        }
    }
    if (n & 7u) {
        for (uint64_t j = n & 7u; j < 8u; j++) word |= 4u << (4u * (uint32_t)j);
        if (n / 8u < maxWords) out[n / 8u] = word;
    }
    return n;
}

int myersWriteReferenceFile(const char *path, const uint32_t *words, uint32_t numEntries, uint32_t size)
{
    FILE *fp = fopen(path, "wb");
    if (fp == NULL) return -1;
    int ok = fwrite(&numEntries, sizeof(uint32_t), 1, fp) == 1 &&
             fwrite(&size, sizeof(uint32_t), 1, fp) == 1 &&
             fwrite(words, sizeof(uint32_t), numEntries, fp) == numEntries;
    return (fclose(fp) == 0 && ok) ? 0 : -1;
}

/* ---------------------------------------------------------------- reads */

typedef struct {
    uint32_t m;
    uint32_t ncand;
    double   sub, ins, del;
    rng_t    rng;       /* stream state after the plan draws */
} read_plan_t;

static void plan_read(const myersSynthSpec_t *s, uint64_t q, read_plan_t *p)
{
    rng_t r = { hash3(s->seed, 'Q', q) };
    uint32_t m;
    switch (s->lengthMode) {
    case MYERS_SYNTH_LEN_CHOICE:
        m = s->choices[rng_below(&r, s->numChoices)];
        break;
    case MYERS_SYNTH_LEN_LOGUNIFORM: {
        const double lo = log((double)s->minLength), hi = log((double)s->maxLength);
        m = (uint32_t)floor(exp(lo + (hi - lo) * rng_unif(&r)) + 0.5);
        if (m < s->minLength) m = s->minLength;
        if (m > s->maxLength) m = s->maxLength;
        break;
    }
    default:
        m = s->length;
    }
    uint32_t nc;
    if (s->candMode == MYERS_SYNTH_CAND_POISSON) {
        nc = rng_poisson(&r, s->candMean);
        if (nc < 1u) nc = 1u;
        if (nc > s->candCap) nc = s->candCap;
    } else {
        nc = s->candidates ? s->candidates : 1u;
    }
    if (s->errMax > 0.0) {
        const double e = s->errMin + (s->errMax - s->errMin) * rng_unif(&r);
        p->sub = e * s->subFrac;
        p->ins = p->del = e * (1.0 - s->subFrac) * 0.5;
    } else {
        p->sub = s->subRate;
        p->ins = s->insRate;
        p->del = s->delRate;
    }
    p->m = m;
    p->ncand = nc;
    p->rng = r;
}

/* Sample read codes from the reference at `start` with the error model. */
static void sample_read(const uint32_t *ref, uint64_t size, uint64_t start, const read_plan_t *p, double geom_p,
                        rng_t *r, uint8_t *out)
{
    uint64_t pos = start;
    uint32_t k = 0;
    const double t_sub = p->sub, t_ins = p->sub + p->ins, t_del = p->sub + p->ins + p->del;
    while (k < p->m) {
        const double u = rng_unif(r);
        const uint32_t b = pos < size ? ref_code(ref, pos) : 4u;
        if (u < t_sub) {
            uint32_t nb = (uint32_t)rng_below(r, 3);
            out[k++] = (uint8_t)(b > 3u ? nb : (b + 1u + nb) & 3u);
            pos++;
        } else if (u < t_ins) {
            uint32_t len = rng_geometric(r, geom_p);
            for (uint32_t i = 0; i < len && k < p->m; i++) out[k++] = (uint8_t)rng_below(r, 4);
        } else if (u < t_del) {
            pos += rng_geometric(r, geom_p);
        } else {
            out[k++] = (uint8_t)b;
            pos++;
        }
    }
}

void myersSynthFreeBatch(myersSynthBatch_t *b)
{
    if (b == NULL) return;
    free(b->peq);
    free(b->qinfo);
    free(b->cands);
    free(b->codes);
    free(b->codeOffset);
    free(b->trueStart);
    memset(b, 0, sizeof(*b));
}

int myersSynthBatch(const myersSynthSpec_t *s, const uint32_t *ref, uint64_t size, myersSynthBatch_t *out)
{
    if (s == NULL || ref == NULL || out == NULL || s->numQueries == 0) return -1;
    if (s->lengthMode == MYERS_SYNTH_LEN_CHOICE && (s->numChoices == 0 || s->numChoices > 16)) return -1;
    memset(out, 0, sizeof(*out));
    const uint64_t nq = s->numQueries;
    read_plan_t *plan = (read_plan_t *)malloc(sizeof(read_plan_t) * nq);
    uint64_t *entry_off = (uint64_t *)malloc(sizeof(uint64_t) * (nq + 1u));
    uint64_t *cand_off = (uint64_t *)malloc(sizeof(uint64_t) * (nq + 1u));
    uint64_t *code_off = (uint64_t *)malloc(sizeof(uint64_t) * (nq + 1u));
    if (!plan || !entry_off || !cand_off || !code_off) goto fail;

    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)nq; q++) plan_read(s, (uint64_t)q, &plan[q]);

    uint32_t max_m = 0;
    entry_off[0] = cand_off[0] = code_off[0] = 0;
    for (uint64_t q = 0; q < nq; q++) {
        const uint32_t m = plan[q].m;
        if (m > max_m) max_m = m;
        if ((uint64_t)window_len(m, s->distance) + 2u * (uint64_t)m + 2u >= size) goto fail;
        entry_off[q + 1] = entry_off[q] + (m + 127u) / 128u;
        cand_off[q + 1] = cand_off[q] + plan[q].ncand;
        code_off[q + 1] = code_off[q] + m;
    }
    if (entry_off[nq] > 0xFFFFFFFFull) goto fail;

    out->numQueries = nq;
    out->numEntries = entry_off[nq];
    out->numCandidates = cand_off[nq];
    out->totalBases = code_off[nq];
    out->peq = (qryEntry_t *)malloc(sizeof(qryEntry_t) * (out->numEntries ? out->numEntries : 1u));
    out->qinfo = (qryInfo_t *)malloc(sizeof(qryInfo_t) * nq);
    out->cands = (candInfo_t *)malloc(sizeof(candInfo_t) * (out->numCandidates ? out->numCandidates : 1u));
    out->trueStart = (uint32_t *)malloc(sizeof(uint32_t) * nq);
    if (!out->peq || !out->qinfo || !out->cands || !out->trueStart) goto fail_out;
    if (s->keepCodes) {
        out->codes = (uint8_t *)malloc(out->totalBases ? out->totalBases : 1u);
        out->codeOffset = (uint64_t *)malloc(sizeof(uint64_t) * (nq + 1u));
        if (!out->codes || !out->codeOffset) goto fail_out;
        memcpy(out->codeOffset, code_off, sizeof(uint64_t) * (nq + 1u));
    }

    uint64_t cells = 0;
    const double geom_p = s->indelGeomP > 0.0 ? s->indelGeomP : 0.7;
    #pragma omp parallel reduction(+ : cells)
    {
        uint8_t *tmp = (uint8_t *)malloc(max_m ? max_m : 1u);
        #pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < (int64_t)nq; q++) {
            read_plan_t *p = &plan[q];
            rng_t r = p->rng;
            const uint32_t m = p->m;
            const uint32_t n = window_len(m, s->distance);
            const uint64_t start = rng_below(&r, size - 2u * (uint64_t)m - 1u);
            uint8_t *codes = s->keepCodes ? out->codes + code_off[q] : tmp;
            sample_read(ref, size, start, p, geom_p, &r, codes);
            myersBuildPeq(codes, m, out->peq + entry_off[q]);
            out->qinfo[q].posEntry = (uint32_t)entry_off[q];
            out->qinfo[q].size = m;
            out->trueStart[q] = (uint32_t)(start > 0xFFFFFFFFull ? 0xFFFFFFFFull : start);
            candInfo_t *c = out->cands + cand_off[q];
            const uint64_t margin = (uint64_t)(m / 10u);
            c[0].query = (uint32_t)q;
            c[0].position = start >= margin ? start - margin : 0u;
            for (uint32_t k = 1; k < p->ncand; k++) {
                c[k].query = (uint32_t)q;
                c[k].position = rng_below(&r, size - n - 1u);
            }
            cells += (uint64_t)p->ncand * (uint64_t)m * (uint64_t)n;
        }
        free(tmp);
    }
    out->cells = cells;
    free(plan);
    free(entry_off);
    free(cand_off);
    free(code_off);
    return 0;

fail_out:
    myersSynthFreeBatch(out);
fail:
    free(plan);
    free(entry_off);
    free(cand_off);
    free(code_off);
    return -1;
}

/* ------------------------------------------------------------- formats */

int myersReadReferenceFile(const char *path, uint32_t **words, uint32_t *numEntries, uint32_t *size)
{
    FILE *fp = fopen(path, "rb");
    if (fp == NULL) return -1;
    uint32_t hdr[2];
    int ok = fread(hdr, sizeof(uint32_t), 2, fp) == 2;
    uint32_t *w = ok ? (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)hdr[0] + 1u)) : NULL;
    ok = ok && w && fread(w, sizeof(uint32_t), hdr[0], fp) == hdr[0];
    fclose(fp);
    if (!ok) {
        free(w);
        return -1;
    }
    *words = w;
    *numEntries = hdr[0];
    *size = hdr[1];
    return 0;
}

void myersFreeGemQry(myersGemQry_t *q)
{
    if (q == NULL) return;
    free(q->peq);
    free(q->cands);
    free(q->qinfo);
    free(q->gemScores);
    free(q->bases);
    memset(q, 0, sizeof(*q));
}

int myersReadGemQry(const char *path, myersGemQry_t *out)
{
    memset(out, 0, sizeof(*out));
    FILE *fp = fopen(path, "rb");
    if (fp == NULL) return -1;
    uint32_t hdr[5];
    int ok = fread(hdr, sizeof(uint32_t), 5, fp) == 5;
    if (ok) {
        out->totalSizeQueries = hdr[0];
        out->numEntries = hdr[1];
        out->numQueries = hdr[3];
        out->numCandidates = hdr[4];
        out->peq = (qryEntry_t *)malloc(sizeof(qryEntry_t) * ((size_t)hdr[1] + 1u));
        out->cands = (candInfo_t *)malloc(sizeof(candInfo_t) * ((size_t)hdr[4] + 1u));
        out->qinfo = (qryInfo_t *)malloc(sizeof(qryInfo_t) * ((size_t)hdr[3] + 1u));
        out->gemScores = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)hdr[4] + 1u));
        out->bases = (char *)malloc((size_t)hdr[0] + 1u);
        ok = out->peq && out->cands && out->qinfo && out->gemScores && out->bases &&
             fread(out->peq, sizeof(qryEntry_t), hdr[1], fp) == hdr[1] &&
             fread(out->cands, sizeof(candInfo_t), hdr[4], fp) == hdr[4] &&
             fread(out->qinfo, sizeof(qryInfo_t), hdr[3], fp) == hdr[3];
        /* scores and bases are optional trailers (src-api/test.c stops after qinfo) */
        if (ok) {
            size_t got = fread(out->gemScores, sizeof(uint32_t), hdr[4], fp);
            if (got != hdr[4]) memset(out->gemScores, 0, sizeof(uint32_t) * hdr[4]);
            got = fread(out->bases, 1, hdr[0], fp);
            if (got != hdr[0]) memset(out->bases, 0, hdr[0]);
        }
    }
    fclose(fp);
    if (!ok) {
        myersFreeGemQry(out);
        return -1;
    }
    return 0;
}

int myersWriteGemQry(const char *path, const myersGemQry_t *in)
{
    FILE *fp = fopen(path, "wb");
    if (fp == NULL) return -1;
    const uint32_t hdr[5] = {in->totalSizeQueries, in->numEntries, 0u, in->numQueries, in->numCandidates};
    int ok = fwrite(hdr, sizeof(uint32_t), 5, fp) == 5 &&
             fwrite(in->peq, sizeof(qryEntry_t), in->numEntries, fp) == in->numEntries &&
             fwrite(in->cands, sizeof(candInfo_t), in->numCandidates, fp) == in->numCandidates &&
             fwrite(in->qinfo, sizeof(qryInfo_t), in->numQueries, fp) == in->numQueries &&
             (in->gemScores == NULL ||
              fwrite(in->gemScores, sizeof(uint32_t), in->numCandidates, fp) == in->numCandidates) &&
             (in->bases == NULL || fwrite(in->bases, 1, in->totalSizeQueries, fp) == in->totalSizeQueries);
    return (fclose(fp) == 0 && ok) ? 0 : -1;
}

static int grow(void **p, size_t *cap, size_t need, size_t elem)
{
    if (need <= *cap) return 0;
    size_t c = *cap ? *cap : 1024u;
    while (c < need) c *= 2u;
    void *n = realloc(*p, c * elem);
    if (n == NULL) return -1;
    *p = n;
    *cap = c;
    return 0;
}

int myersParseProf(const char *path, myersGemQry_t *out)
{
    memset(out, 0, sizeof(*out));
    FILE *fp = fopen(path, "rb");
    if (fp == NULL) return -1;
    size_t cap_e = 0, cap_c = 0, cap_s = 0, cap_q = 0, cap_b = 0, ne = 0, nc = 0, nq = 0, nb = 0;
    char *line = NULL;
    size_t line_cap = 0;
    ssize_t len;
    int ok = 1;
    while (ok && (len = getline(&line, &line_cap, fp)) > 0) {
        while (len > 0 && (line[len - 1] == '\n' || line[len - 1] == '\r')) line[--len] = 0;
        if (len == 0) continue;
        /* the read is the first whitespace-free token, as sscanf("%s") takes it */
        const size_t m = strcspn(line, " \t");
        char *tab = line[m] ? line + m : NULL;
        const uint32_t entries = (uint32_t)((m + 127u) / 128u);
        if (grow((void **)&out->peq, &cap_e, ne + entries, sizeof(qryEntry_t)) ||
            grow((void **)&out->qinfo, &cap_q, nq + 1u, sizeof(qryInfo_t)) ||
            grow((void **)&out->bases, &cap_b, nb + m, 1u)) {
            ok = 0;
            break;
        }
        uint8_t *codes = (uint8_t *)malloc(m ? m : 1u);
        if (codes == NULL) {
            ok = 0;
            break;
        }
        for (size_t i = 0; i < m; i++) {
            codes[i] = (uint8_t)fasta_code((unsigned char)line[i]);
            out->bases[nb + i] = line[i];
        }
        myersBuildPeq(codes, (uint32_t)m, out->peq + ne);
        free(codes);
        out->qinfo[nq].posEntry = (uint32_t)ne;
        out->qinfo[nq].size = (uint32_t)m;
        /* candidates: whitespace-separated "area:strand:position:score" */
        char *p = tab;
        while (p && *p) {
            while (*p == '\t' || *p == ' ') p++;
            if (!*p) break;
            char *end = p;
            while (*end && *end != '\t' && *end != ' ') end++;
            char save = *end;
            *end = 0;
            char *c1 = strchr(p, ':');
            char *c2 = c1 ? strchr(c1 + 1, ':') : NULL;
            char *c3 = c2 ? strchr(c2 + 1, ':') : NULL;
            if (c3 == NULL) {
                ok = 0;
                break;
            }
            if (grow((void **)&out->cands, &cap_c, nc + 1u, sizeof(candInfo_t)) ||
                grow((void **)&out->gemScores, &cap_s, nc + 1u, sizeof(uint32_t))) {
                ok = 0;
                break;
            }
            out->cands[nc].query = (uint32_t)nq;
            out->cands[nc].position = (uint64_t)strtoull(c2 + 1, NULL, 10);
            out->gemScores[nc] = (uint32_t)strtoul(c3 + 1, NULL, 10);
            nc++;
            *end = save;
            p = end;
        }
        ne += entries;
        nq++;
        nb += m;
    }
    free(line);
    fclose(fp);
    if (!ok) {
        myersFreeGemQry(out);
        return -1;
    }
    out->numEntries = (uint32_t)ne;
    out->numQueries = (uint32_t)nq;
    out->numCandidates = (uint32_t)nc;
    out->totalSizeQueries = (uint32_t)nb;
    return 0;
}
