/*
 * myers-interface.h -- drop-in C API of the candidate-verification library
 * (libmyers_amd.so, MI355X / gfx950).
 *
 * Source-compatible with the reference header src-api/myers-interface.h:1-46:
 * same macros, record layouts and function names, so callers written against
 * the reference (src-api/test.c, the GEM mapper) compile and link unchanged.
 * Deliberate differences, each fixing a reference defect:
 *   - the seven getters are ordinary external functions (the reference
 *     declares them `inline`, src-api/myers-interface.h:31-39, which under
 *     C99 emits no symbol, so callers failed to link);
 *   - initMyers returns 0 on success or a myersError_t code (the reference
 *     declares int but defines void, src-api/myersGPU.c:284).
 * Extensions (error returns, window modes, device-resident batches, timing)
 * are in myers-amd.h.
 *
 * No HIP, CUDA or C++ types appear here: plain C99 with <stdint.h>.
 */
#ifndef MYERS_INTERFACE_H
#define MYERS_INTERFACE_H

#include <stdint.h>

#define     NUM_BITS            4
#define     NUM_BASES           5
#define     NUM_BASES_ENTRY     128
#define     SIZE_GPU_HW_WORD    32
#define     NUM_SUB_ENTRIES     (NUM_BASES_ENTRY / SIZE_GPU_HW_WORD)

#ifdef __cplusplus
extern "C" {
#endif

/* Peq entry: 128 query bases.  bitmap[b][s] bit i is set iff query base
 * 128*e + 32*s + i has base code b (0=A 1=C 2=G 3=T 4=other/N).  80 bytes.
 * (src-api/myers-interface.h:10-12) */
typedef struct {
    uint32_t bitmap[NUM_BASES][NUM_SUB_ENTRIES];
} qryEntry_t;

/* Result of one candidate: column of the first minimum and the minimum
 * edit distance.  Skipped candidates hold {0, 0xFFFFFFFF}.  8 bytes.
 * (src-api/myers-interface.h:14-17) */
typedef struct {
    uint32_t column;
    uint32_t score;
} resEntry_t;

/* Candidate: query index and reference position (in bases).  16 bytes.
 * (src-api/myers-interface.h:19-22) */
typedef struct {
    uint32_t query;
    uint64_t position;
} candInfo_t;

/* Query: index of its first Peq entry and its length in bases.  8 bytes.
 * (src-api/myers-interface.h:24-27) */
typedef struct {
    uint32_t posEntry;
    uint32_t size;
} qryInfo_t;

/* Buffer accessors (src-api/myers-interface.h:31-39; src-api/myersGPU.c:55-88).
 * The pointers are pinned host memory owned by the library; the caller fills
 * the PEQ, PEQ-info and candidate buffers before sendMyersBuffer and reads
 * the results buffer after receiveMyersBuffer. */
qryEntry_t *getPEQBuffer        (void *myersBuffer);
candInfo_t *getCandidatesBuffer (void *myersBuffer);
qryInfo_t  *getPEQInfoBuffer    (void *myersBuffer);
resEntry_t *getResultsBuffer    (void *myersBuffer);

uint32_t getMaxPEQEntries (void *myersBuffer);
uint32_t getMaxCandidates (void *myersBuffer);
uint32_t getMaxQueries    (void *myersBuffer);

/* Load the 4-bit packed reference file ({u32 numEntries; u32 size;
 * u32 words[numEntries]}) into HBM once and create numBuffers independent
 * buffers, each with its own HIP stream (src-api/myersGPU.c:284-344).
 * averageQuerySize / candidatesPerQuery size the buffers (<= 0 selects
 * 1024 / 1).  Returns 0 or a myersError_t code. */
int  initMyers          (void ***myersBuffer, uint32_t numBuffers, const char *referenceFileName,
                         int32_t averageQuerySize, int32_t candidatesPerQuery);

/* Enqueue one batch: copy the first numPEQEntries / numQueries /
 * numCandidates records of the buffer's host arrays to the GPU, verify every
 * candidate with window n = (uint32)(size * (1 + 2*distance)) and copy the
 * results back.  Asynchronous (src-api/myersGPU.c:466-494). */
void sendMyersBuffer    (void *myersBuffer, float distance, uint32_t numPEQEntries,
                         uint32_t numQueries, uint32_t numCandidates);

/* Wait for the buffer's batch; afterwards getResultsBuffer()[i] holds the
 * result of candidate i in the caller's order (src-api/myersGPU.c:515-524). */
void receiveMyersBuffer (void *myersBuffer);

/* Release every buffer and the device reference; sets *myersBuffer to NULL
 * (src-api/myersGPU.c:651-676). */
void endMyers           (void ***myersBuffer);

#ifdef __cplusplus
}
#endif

#endif /* MYERS_INTERFACE_H */
