/*
 * A caller written only against myers-interface.h, the way the reference's
 * own driver src-api/test.c:100-207 uses the API: load a packed reference and
 * a query batch, fill the buffer through the getters, send, receive, read
 * the results.  Built with plain -std=c99 (the reference's inline getters
 * could not be linked this way).
 *
 * usage: api_caller <distance> <ref.4bits.ref> <batch.bin> <results.txt>
 *   batch.bin = u32 numEntries, numQueries, numCandidates;
 *               qryEntry_t[numEntries]; qryInfo_t[numQueries]; candInfo_t[numCandidates]
 *   results.txt = "<count>\n" then "<id> <score> <column>\n" (src-api/test.c:85-94)
 */
#include <stdio.h>
#include <stdlib.h>

#include "myers-interface.h"

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s distance ref batch out\n", argv[0]);
        return 2;
    }
    float distance = (float)atof(argv[1]);
    FILE *fp = fopen(argv[3], "rb");
    if (!fp) return 3;
    uint32_t hdr[3];
    if (fread(hdr, sizeof(uint32_t), 3, fp) != 3) return 4;
    qryEntry_t *peq = malloc(sizeof(qryEntry_t) * (hdr[0] + 1));
    qryInfo_t *qinfo = malloc(sizeof(qryInfo_t) * (hdr[1] + 1));
    candInfo_t *cands = malloc(sizeof(candInfo_t) * (hdr[2] + 1));
    if (fread(peq, sizeof(qryEntry_t), hdr[0], fp) != hdr[0]) return 5;
    if (fread(qinfo, sizeof(qryInfo_t), hdr[1], fp) != hdr[1]) return 6;
    if (fread(cands, sizeof(candInfo_t), hdr[2], fp) != hdr[2]) return 7;
    fclose(fp);

    void **buffer = NULL;
    int rc = initMyers(&buffer, 2, argv[2], 1000, 10);
    if (rc != 0) {
        fprintf(stderr, "initMyers failed: %d\n", rc);
        return 8;
    }
    /* Two buffers: the first half of the candidates in buffer 0, the rest in
     * buffer 1, both in flight at once. */
    uint32_t half = hdr[2] / 2;
    uint32_t counts[2] = {half, hdr[2] - half};
    uint32_t firsts[2] = {0, half};
    for (int b = 0; b < 2; b++) {
        if (hdr[0] > getMaxPEQEntries(buffer[b]) || hdr[1] > getMaxQueries(buffer[b]) ||
            counts[b] > getMaxCandidates(buffer[b]))
            return 9;
        qryEntry_t *bp = getPEQBuffer(buffer[b]);
        qryInfo_t *bq = getPEQInfoBuffer(buffer[b]);
        candInfo_t *bc = getCandidatesBuffer(buffer[b]);
        for (uint32_t i = 0; i < hdr[0]; i++) bp[i] = peq[i];
        for (uint32_t i = 0; i < hdr[1]; i++) bq[i] = qinfo[i];
        for (uint32_t i = 0; i < counts[b]; i++) bc[i] = cands[firsts[b] + i];
        sendMyersBuffer(buffer[b], distance, hdr[0], hdr[1], counts[b]);
    }
    FILE *out = fopen(argv[4], "w");
    if (!out) return 10;
    fprintf(out, "%u\n", hdr[2]);
    for (int b = 0; b < 2; b++) {
        receiveMyersBuffer(buffer[b]);
        resEntry_t *r = getResultsBuffer(buffer[b]);
        for (uint32_t i = 0; i < counts[b]; i++)
            fprintf(out, "%u %u %u\n", firsts[b] + i, r[i].score, r[i].column);
    }
    fclose(out);
    endMyers(&buffer);
    if (buffer != NULL) return 11;
    free(peq);
    free(qinfo);
    free(cands);
    return 0;
}
