/*
 * myers-prep: the reference's input preparation programs in one tool.
 *
 *   myers-prep fasta2ref IN.fasta [OUT]   FASTA -> .4bits.ref
 *       (src/genReference-4bits.c:21-140; default OUT "IN.<size>.4bits.ref")
 *   myers-prep prof2qry IN.prof [OUT]     GEM .prof -> .gem.qry
 *       (src-api/genCandidate-reorder.c:114-390; default OUT
 *        "IN.<queries>.<candidates/queries>.0.gem.qry")
 *
 * Differences: the last partial reference word is padded with N instead of
 * bytes read past the FASTA buffer (src/genReference-4bits.c:67-68); .prof
 * lines are not limited to 20,000 characters (genCandidate-reorder.c:6), and
 * each read's first Peq entry index is the running sum of the previous reads'
 * entries (genCandidate-reorder.c:227 adds to an uninitialised value).
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "myers-tools.h"

static int fasta2ref(const char *in, const char *out)
{
    FILE *fp = fopen(in, "rb");
    if (fp == NULL) {
        fprintf(stderr, "cannot open %s\n", in);
        return 1;
    }
    fseek(fp, 0, SEEK_END);
    const long len = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    char *text = (char *)malloc((size_t)(len > 0 ? len : 1));
    if (text == NULL || (len > 0 && fread(text, 1, (size_t)len, fp) != (size_t)len)) {
        fprintf(stderr, "cannot read %s\n", in);
        fclose(fp);
        free(text);
        return 1;
    }
    fclose(fp);
    const uint64_t max_words = ((uint64_t)len + 7u) / 8u + 1u;
    uint32_t *words = (uint32_t *)malloc(sizeof(uint32_t) * max_words);
    if (words == NULL) {
        free(text);
        return 1;
    }
    const uint64_t size = myersPackFasta(text, (uint64_t)len, words, max_words);
    free(text);
    if (size > UINT32_MAX) {
        fprintf(stderr, "%s: %llu bases do not fit the 32-bit size of the .4bits.ref header\n", in,
                (unsigned long long)size);
        free(words);
        return 1;
    }
    char *name = NULL;
    if (out == NULL) {
        if (asprintf(&name, "%s.%llu.4bits.ref", in, (unsigned long long)size) < 0) return 1;
        out = name;
    }
    const uint32_t nwords = (uint32_t)((size + 7u) / 8u);
    const int rc = myersWriteReferenceFile(out, words, nwords, (uint32_t)size);
    if (rc == 0) printf("%s: %llu bases, %u words\n", out, (unsigned long long)size, nwords);
    else fprintf(stderr, "cannot write %s\n", out);
    free(words);
    free(name);
    return rc == 0 ? 0 : 1;
}

static int prof2qry(const char *in, const char *out)
{
    myersGemQry_t q;
    if (myersParseProf(in, &q) != 0) {
        fprintf(stderr, "cannot parse %s\n", in);
        return 1;
    }
    char *name = NULL;
    if (out == NULL) {
        const unsigned ratio = q.numQueries ? q.numCandidates / q.numQueries : 0u;
        if (asprintf(&name, "%s.%u.%u.0.gem.qry", in, q.numQueries, ratio) < 0) return 1;
        out = name;
    }
    const int rc = myersWriteGemQry(out, &q);
    if (rc == 0) printf("%s: %u queries, %u candidates, %u Peq entries\n", out, q.numQueries,
                        q.numCandidates, q.numEntries);
    else fprintf(stderr, "cannot write %s\n", out);
    myersFreeGemQry(&q);
    free(name);
    return rc == 0 ? 0 : 1;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && strcmp(argv[1], "fasta2ref") == 0) return fasta2ref(argv[2], argc > 3 ? argv[3] : NULL);
    if (argc >= 3 && strcmp(argv[1], "prof2qry") == 0) return prof2qry(argv[2], argc > 3 ? argv[3] : NULL);
    fprintf(stderr, "usage: myers-prep fasta2ref IN.fasta [OUT.4bits.ref]\n"
                    "       myers-prep prof2qry IN.prof [OUT.gem.qry]\n");
    return 2;
}
