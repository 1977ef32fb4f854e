/*
 * myers-tools.h -- data preparation for the candidate-verification library
 * (libmyers_tools.so): seeded synthetic references and read batches in the
 * library's input layouts, and the reference's on-disk formats.
 *
 * Reference counterparts: the FASTA -> 4-bit packer src/genReference-4bits.c
 * and the GEM .prof -> Peq converter src-api/genCandidate-reorder.c.  The
 * synthetic generator has no counterpart (the reference's inputs were GEM
 * runs on human data, README.md:10-44).
 *
 * Generator definition (counter-based, so results do not depend on the
 * thread count):
 *   mix(x) = SplitMix64 finaliser; h(seed, tag, i) = mix(seed ^ mix(tag*K + i)).
 *   Reference base i: chunk c = i/64 holds one N run with probability
 *   nRunProb (start uniform in the chunk, length 1 + Geometric(0.1), cut at
 *   the chunk end); every other base is 2 bits of h(seed, 'B', i/32).
 *   Read q uses its own SplitMix64 stream seeded with h(seed, 'Q', q).
 */
#ifndef MYERS_TOOLS_H
#define MYERS_TOOLS_H

#include <stdint.h>

#include "myers-interface.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Packed words of a synthetic reference: ceil(size/8) words, nibbles past
 * size are N (4).  OpenMP over 64-base chunks. */
void myersSynthReference(uint64_t seed, uint64_t size, double nRunProb, uint32_t *words);

enum { MYERS_SYNTH_LEN_FIXED = 0, MYERS_SYNTH_LEN_CHOICE = 1, MYERS_SYNTH_LEN_LOGUNIFORM = 2 };
enum { MYERS_SYNTH_CAND_FIXED = 0, MYERS_SYNTH_CAND_POISSON = 1 };

typedef struct {
    uint64_t seed;
    uint32_t numQueries;
    uint32_t lengthMode;        /* MYERS_SYNTH_LEN_* */
    uint32_t length;            /* FIXED */
    uint32_t numChoices;        /* CHOICE: uniform over choices[0..numChoices) */
    uint32_t choices[16];
    uint32_t minLength;         /* LOGUNIFORM on [minLength, maxLength] */
    uint32_t maxLength;
    double   subRate;           /* per-base error probabilities when errMax == 0 */
    double   insRate;
    double   delRate;
    double   errMin;            /* if errMax > 0: per-read error rate e ~ U[errMin, errMax], */
    double   errMax;            /*   sub = e*subFrac, ins = del = e*(1-subFrac)/2 */
    double   subFrac;
    double   indelGeomP;        /* indel length ~ 1 + Geometric(p) - 1, i.e. P(len=k) = p(1-p)^(k-1) */
    uint32_t candMode;          /* MYERS_SYNTH_CAND_* */
    uint32_t candidates;        /* FIXED: candidates per read (1 true + rest random) */
    double   candMean;          /* POISSON mean (the true locus included) */
    uint32_t candCap;           /* POISSON cap */
    float    distance;          /* window n = (uint32)(m*(1+2d)); true locus = start - floor(0.1*m) */
    uint32_t keepCodes;         /* 1: also return the read bases (codes 0..4) */
} myersSynthSpec_t;

typedef struct {
    qryEntry_t *peq;            /* Peq entries (80 B per 128 bases) */
    qryInfo_t  *qinfo;          /* {posEntry, size} per read */
    candInfo_t *cands;          /* candidates, read-major, the true locus first */
    uint8_t    *codes;          /* read bases, concatenated (keepCodes only) */
    uint64_t   *codeOffset;     /* start of read q in codes (keepCodes only) */
    uint32_t   *trueStart;      /* reference position the read was sampled from */
    uint64_t    numEntries;
    uint64_t    numQueries;
    uint64_t    numCandidates;
    uint64_t    totalBases;
    uint64_t    cells;          /* sum over candidates of m * n */
} myersSynthBatch_t;

/* Generate a read batch from a packed reference.  Returns 0, or -1 on bad
 * arguments / allocation failure.  Free with myersSynthFreeBatch. */
int  myersSynthBatch(const myersSynthSpec_t *spec, const uint32_t *refWords, uint64_t refSize,
                     myersSynthBatch_t *out);
void myersSynthFreeBatch(myersSynthBatch_t *batch);

/* Peq entries of one read (codes 0..3 = A C G T, anything else = N):
 * ceil(m/128) entries (src-api/genCandidate-reorder.c:258-318 layout). */
void myersBuildPeq(const uint8_t *codes, uint32_t m, qryEntry_t *out);

/* 4-bit packing of base codes (src/genReference-4bits.c:21-72 layout):
 * out[k] nibble i = code of base 8k+i; codes > 3 and nibbles past n are N. */
void myersPackCodes(const uint8_t *codes, uint64_t n, uint32_t *out);

/* FASTA text -> packed words, as src/genReference-4bits.c:21-119 does:
 * header lines ('>') are skipped, A/C/G/T in either case map to 0..3 and
 * every other character to N.  Returns the number of bases, and writes at
 * most maxWords words. */
uint64_t myersPackFasta(const char *text, uint64_t textLen, uint32_t *out, uint64_t maxWords);

/* .4bits.ref file: {u32 numEntries; u32 size; u32 words[numEntries]}
 * (src/genReference-4bits.c:121-140).  Returns 0 or -1. */
int myersWriteReferenceFile(const char *path, const uint32_t *words, uint32_t numEntries, uint32_t size);

/* Read a .4bits.ref file: *words is malloc'ed (caller frees); returns 0 or -1. */
int myersReadReferenceFile(const char *path, uint32_t **words, uint32_t *numEntries, uint32_t *size);

/* A .gem.qry batch (src-api/genCandidate-reorder.c:353-390, read by
 * src-api/test.c:30-66): u32 totalSizeQueries, totalQueriesEntries,
 * sizeQueries, numQueries, numCandidates; qryEntry_t[E]; candInfo_t[C]
 * (16 B); qryInfo_t[Q]; u32 gemScore[C]; char bases[totalSizeQueries]. */
typedef struct {
    uint32_t    totalSizeQueries;
    uint32_t    numEntries;
    uint32_t    numQueries;
    uint32_t    numCandidates;
    qryEntry_t *peq;
    candInfo_t *cands;
    qryInfo_t  *qinfo;
    uint32_t   *gemScores;
    char       *bases;
} myersGemQry_t;

int  myersReadGemQry(const char *path, myersGemQry_t *out);        /* 0 or -1 */
int  myersWriteGemQry(const char *path, const myersGemQry_t *in);  /* 0 or -1 */
void myersFreeGemQry(myersGemQry_t *q);

/* Convert a GEM .prof file (one read per line: "<bases>\t<area>:<strand>:
 * <position>:<score>\t..." as parsed by src-api/genCandidate-reorder.c:114-235)
 * into a .gem.qry batch in memory: Peq entries per read (A/C/G/T in either
 * case, every other letter N), candidates in file order. 0 or -1. */
int  myersParseProf(const char *path, myersGemQry_t *out);

#ifdef __cplusplus
}
#endif

#endif /* MYERS_TOOLS_H */
