/*
 * myers-amd.h -- extensions of the candidate-verification library
 * (libmyers_amd.so, MI355X / gfx950) beyond the reference API of
 * myers-interface.h.  Plain C99; no HIP or C++ types in any signature
 * (device pointers and streams are passed as void *).
 *
 * Reference counterparts are cited per entry point; entry points marked
 * "extension" have none in the reference.
 */
#ifndef MYERS_AMD_H
#define MYERS_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "myers-interface.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  0..5 keep the values of the reference's myersError_t
 * (src-api/myers-common.h:13-24); the rest are extensions. */
typedef enum {
    MYERS_SUCCESS                = 0,
    MYERS_E_OPENING_FILE         = 1,
    MYERS_E_READING_FILE         = 2,
    MYERS_E_INSUFFICIENT_MEM_GPU = 3,
    MYERS_E_ALLOCATE_MEM         = 4,
    MYERS_E_INCOMPATIBLE_GPU     = 5,
    MYERS_E_INVALID_ARGUMENT     = 16,  /* NULL handle, zero buffers, bad flag */
    MYERS_E_CAPACITY             = 17,  /* batch larger than the buffer's capacity */
    MYERS_E_HIP                  = 18,  /* a HIP runtime call failed */
    MYERS_E_QUERY_TOO_LONG       = 19,  /* a query exceeds the longest supported length;
                                           its candidates got the skip sentinel */
    MYERS_E_NO_DEVICE            = 20,  /* no HIP device visible */
    MYERS_E_REFERENCE_TOO_LARGE  = 21   /* stored words + 64 guard words exceed 2^31 - 1
                                           (about 17 Gbases): the kernels index words
                                           with signed 32-bit offsets */
} myersError_t;

/* Window semantics (flags of myersSendEx / myersVerifyDevice). */
#define MYERS_WINDOW_ALIGNED   0u  /* API kernel: window starts at the 8-base word holding the
                                      position, 8*(n/8 + (n%8 ? 2 : 1)) columns, columns counted
                                      from that word (src-api/coreMyersGPU-CC35.cu:132,144,164) */
#define MYERS_WINDOW_EXACT     1u  /* exact start: n columns from the position
                                      (src/myers-4bits.c:100-147) */
/* Engine selection (extension).  AUTO: the plain engine for a batch of one
 * length bin of up to 2048 candidates per CU, the persistent engine
 * otherwise. */
#define MYERS_ENGINE_AUTO       0u
#define MYERS_ENGINE_PLAIN      (1u << 4)  /* one launch per length bin */
#define MYERS_ENGINE_PERSISTENT (2u << 4)  /* one persistent launch over a length-binned device queue
                                              (per kernel class: one-lane / multi-lane variants) */
#define MYERS_ENGINE_BINNED     (3u << 4)  /* one resident grid per length bin, bins on forked streams */
#define MYERS_ENGINE_MASK       (3u << 4)
/* Capture the batch's on-device pipeline (binning + kernels) into a HIP
 * graph once per batch shape and replay it (extension). */
#define MYERS_USE_GRAPH         (1u << 6)
/* Persistent engine: do not give the long-read bins of a small batch their
 * own latency-ordered launch; every bin goes into the queue launches
 * (extension; throughput scheduling only). */
#define MYERS_NO_LONG_LAUNCH    (1u << 7)

/* Longest query supported, in bases (extension: the reference stops at 4096,
 * src-api/myersGPU.c:366,406-410). */
#define MYERS_MAX_QUERY_SIZE   24576u

typedef struct {
    uint32_t structSize;        /* sizeof(myersConfig_t) */
    int32_t  device;            /* HIP device index (default 0) */
    int32_t  exitOnError;       /* 1: print and exit on errors in the void API calls,
                                   as the reference does (src-api/myersGPU.c:24-47) */
    uint32_t reserved;
    uint64_t bufferBytes;       /* device bytes per buffer used to size capacities;
                                   0 = min(95% of free HBM / numBuffers, 2 GiB) */
} myersConfig_t;

/* Timing and counters of a buffer's last batch (extension). */
typedef struct {
    uint32_t structSize;
    uint32_t launches;          /* verification kernel launches in the last batch */
    uint32_t variantsUsed;      /* distinct kernel variants in the last batch */
    uint32_t binned;            /* 1 if the batch went through device-side binning */
    double   kernelMs;          /* HIP-event time of the verification kernels, last timed batch */
    double   batchMs;           /* HIP-event time from first copy to last copy */
    uint64_t cells;             /* sum of m * n over verified candidates (nominal window) */
    uint64_t candidates;
    /* accumulated over the buffer's life (or since myersResetStats), folded
     * in when the stream is synchronised (receive / myersSynchronize) */
    double   kernelMsTotal;         /* sum over batches of their kernel time */
    uint64_t kernelLaunchesTotal;   /* verification launches */
    uint64_t batchesTotal;          /* timed batches */
    uint64_t graphCaptures;         /* graphs captured (MYERS_USE_GRAPH) */
    /* the verification launches alone (from the end of the device-side
     * binning; graph replays: the whole pipeline, as kernelMs) */
    double   verifyMs;              /* last timed batch */
    double   verifyMsTotal;         /* accumulated as kernelMsTotal */
} myersStats_t;

void myersDefaultConfig(myersConfig_t *cfg);

/* initMyers with a configuration (extension of src-api/myersGPU.c:284-344). */
int myersInitEx(void ***myersBuffer, uint32_t numBuffers, const char *referenceFileName,
                int32_t averageQuerySize, int32_t candidatesPerQuery, const myersConfig_t *cfg);

/* As myersInitEx, with the packed reference taken from host memory:
 * numWords u32 words of 8 bases each (base i of a word in bits [4i, 4i+3]),
 * refSize bases (extension; 64-bit size lifts the 4.29 Gbase header limit
 * of src/genReference-4bits.c:134-135). */
int myersInitFromMemory(void ***myersBuffer, uint32_t numBuffers, const uint32_t *refWords,
                        uint64_t numWords, uint64_t refSize, int32_t averageQuerySize,
                        int32_t candidatesPerQuery, const myersConfig_t *cfg);

/* sendMyersBuffer returning an error code, with window/engine flags
 * (src-api/myersGPU.c:466-494). */
int myersSendEx(void *myersBuffer, float distance, uint32_t numPEQEntries, uint32_t numQueries,
                uint32_t numCandidates, uint32_t flags);

/* receiveMyersBuffer returning an error code (src-api/myersGPU.c:515-524). */
int myersReceiveEx(void *myersBuffer);

/* Threshold mode (extension; semantics of src/checksum-2.c:88,104-105 and the
 * cut-off search of GEMTools/src/gt_map_align_bpm.c:141-214).  With
 * errorRate >= 0, every later batch on this buffer reports a candidate whose
 * minimum score is <= (uint32)((float)m * errorRate) exactly as without the
 * threshold, and every other verified candidate as
 * {column 0xFFFFFFFF, score 0xFFFFFFFF}; skipped candidates keep
 * {0, 0xFFFFFFFF}.  Kernels for queries of up to 1024 bases then skip the
 * rows whose distance already exceeds the threshold (Ukkonen cut-off), and
 * the multi-lane kernels of per-bin launches stop a wave once none of its
 * candidates can get under the threshold.  In binned batches a probe of each
 * such query's first 32 bases orders those candidates so that the ones
 * likely under the threshold share waves (the order of the work only; the
 * results stay in the caller's order).  errorRate < 0 switches the mode off
 * (the default). */
#define MYERS_NO_MATCH 0xFFFFFFFFu
int   myersSetThreshold(void *myersBuffer, float errorRate);
float myersGetThreshold(void *myersBuffer);

/* endMyers returning an error code (src-api/myersGPU.c:651-676). */
int myersEnd(void ***myersBuffer);

const char *myersGetErrorString(int error);     /* src-api/myersGPU.c:31-40 */
int         myersGetLastError(void *myersBuffer);  /* extension */
int         myersGetStats(void *myersBuffer, myersStats_t *out);  /* extension */
void        myersResetStats(void *myersBuffer);                   /* extension */
uint32_t    myersGetNumBuffers(void *myersBuffer);                /* extension */
uint64_t    myersGetReferenceSize(void *myersBuffer);             /* extension */

/* ---- Device-resident batches (extension) ------------------------------
 * For callers whose queries and candidates already sit in HBM: no host
 * copies; the batch runs on the buffer's stream and uses its binning
 * workspace (so it needs numCandidates <= getMaxCandidates(buffer)).
 * h_qinfo (optional, host copy of the qinfo array) lets the library pick
 * the kernel variants, and plan the long-read launch, without reading device
 * memory; NULL launches every variant over the device-side bins.  Candidates naming a query >=
 * numQueries, or a query whose entries run past numPEQEntries, get the skip
 * sentinel.  Returns after enqueueing. */
int myersVerifyDevice(void *myersBuffer, const void *d_peq, uint32_t numPEQEntries, const void *d_qinfo,
                      uint32_t numQueries, const void *d_candidates, uint32_t numCandidates,
                      float distance, uint32_t flags, void *d_results, const qryInfo_t *h_qinfo);
/* myersVerifyDevice that allocates nothing (extension, no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * binning workspace is a static device array of the library, one per GPU and
 * held until the process ends, instead of the buffer's workspace.  At most
 * MYERS_NOALLOC_MAX_CANDIDATES candidates (more: MYERS_E_CAPACITY with a
 * message); MYERS_USE_GRAPH is refused (MYERS_E_INVALID_ARGUMENT).  Every
 * call on one GPU shares that array, so calls that overlap on one GPU --
 * from different buffers or threads -- race: serialise them. */
#define MYERS_NOALLOC_MAX_CANDIDATES (1u << 21)
int myersVerifyDeviceNoAlloc(void *myersBuffer, const void *d_peq, uint32_t numPEQEntries, const void *d_qinfo,
                             uint32_t numQueries, const void *d_candidates, uint32_t numCandidates,
                             float distance, uint32_t flags, void *d_results, const qryInfo_t *h_qinfo);
int myersSynchronize(void *myersBuffer);          /* wait for the buffer's stream */
void *myersGetStream(void *myersBuffer);          /* the buffer's hipStream_t */

/* Device memory helpers so that hosts without a HIP toolchain (ctypes,
 * JNI, cgo) can stage device-resident batches. */
int myersDeviceMalloc(void **ptr, size_t bytes);
int myersDeviceFree(void *ptr);
int myersMemcpyHtoD(void *dst, const void *src, size_t bytes);
int myersMemcpyDtoH(void *dst, const void *src, size_t bytes);

/* Kernel variants compiled into the library and the length-bin table
 * (extension; used by the autotuner).  bin[w] is the variant used for
 * queries of w = ceil(m/32) words, 0 <= w <= maxWords. */
typedef struct {
    uint32_t lanes;          /* lanes of a wave64 per candidate */
    uint32_t words;          /* u32 words per lane */
    uint32_t maxQuerySize;   /* 32 * lanes * words */
    uint32_t candidatesPerWave;
    uint32_t ldsBytes;
    int32_t  workgroupsPerCU;
} myersVariantInfo_t;

uint32_t myersNumVariants(void);
int      myersGetVariantInfo(void *myersBuffer, uint32_t variant, myersVariantInfo_t *out);
int      myersGetBinTable(void *myersBuffer, uint8_t *bin, uint32_t maxWords);
int      myersSetBinTable(void *myersBuffer, const uint8_t *bin, uint32_t maxWords);
/* Chunks (of candidatesPerWave candidates) a persistent workgroup takes per
 * pull from its bin's queue (1..64). */
int      myersSetChunkFactor(void *myersBuffer, uint32_t variant, uint32_t chunks);
int      myersGetChunkFactor(void *myersBuffer, uint32_t variant, uint32_t *chunks);
/* Load an autotuner table: text lines "maxWords lanes words chunkFactor"
 * (maxWords ascending; '#' starts a comment).  Queries of w words use the
 * first line with maxWords >= w; longer ones keep the built-in choice.
 * initMyers loads $MYERS_TUNING_FILE, else tuning/gfx950.tsv next to the
 * library's directory, if present. */
int      myersLoadTuning(void *myersBuffer, const char *path);

/* ---- Serving: continuous batching (extension) -------------------------
 * A server thread owns the buffers handed to it (each is one batch slot with
 * its own stream).  Requests are admitted FIFO into a forming batch while the
 * batch fits a slot and the device bytes of admitted, unfinished requests
 * stay within hbmBudgetBytes; a batch is launched as soon as a slot is free,
 * so requests arriving while others run join the next batch and batches run
 * concurrently on the device.  Results are copied to each request's output
 * when its batch completes; every candidate's result is the same as when its
 * request is verified alone. */
/* Server flag: verify with one resident kernel instead of batches.  Each
 * admitted request is binned by read length on the host into chunks that go
 * straight into a device-visible work queue, its longest reads first; a
 * kernel that stays on the device pulls chunks of every admitted request in
 * arrival order and reports each request as soon as its own chunks are done,
 * so a request waits neither for a free buffer nor for other requests.
 * Requests' inputs, results and the queue live in pinned host memory that the
 * kernel reads and writes directly (arenaBytes); the kernel leaves the device
 * after ~50 ms without work and is launched again when work arrives. */
#define MYERS_SERVE_RESIDENT   (1u << 8)

typedef struct {
    uint32_t structSize;
    uint32_t reserved;
    float    distance;          /* window n = (uint32)(m*(1+2*distance)) for every request */
    uint32_t flags;             /* MYERS_WINDOW_* | MYERS_ENGINE_* | MYERS_SERVE_RESIDENT */
    uint64_t hbmBudgetBytes;    /* admission budget (0 = 200e9 bytes) */
    uint32_t pollMicros;        /* completion poll period while batches run (0 = 20) */
    uint32_t maxBatchCandidates;/* cap per batch (0 = slot capacity) */
    uint64_t maxBatchCells;     /* cap per batch on the sum of m * n over its candidates, so
                                   that batches stay short and several run at once
                                   (0 = 5e11; UINT64_MAX = no cap); a request larger than
                                   the cap runs as a batch of its own.  While more than three
                                   capped batches' worth of requests wait, a batch may take
                                   up to four times the cap (draining a backlog at the
                                   throughput of large batches) */
    uint64_t arenaBytes;        /* MYERS_SERVE_RESIDENT: pinned host memory for admitted
                                   requests' inputs and results (0 = 1 GiB) */
} myersServeConfig_t;

typedef struct {
    const qryEntry_t *peq;
    const qryInfo_t  *qinfo;    /* posEntry relative to this request's peq */
    const candInfo_t *cands;    /* query relative to this request's qinfo */
    resEntry_t       *results;  /* numCandidates results, written on completion */
    uint32_t          numPEQEntries;
    uint32_t          numQueries;
    uint32_t          numCandidates;
    uint32_t          reserved;
} myersRequest_t;

typedef struct {
    uint32_t structSize;
    uint32_t reserved;
    uint64_t requests;          /* completed */
    uint64_t batches;           /* launched */
    uint64_t candidates;        /* verified */
    uint64_t maxBatchRequests;
    double   busyMs;            /* sum of batch kernel time */
    double   admitMs;           /* server-thread time forming and sending batches */
    double   completeMs;        /* server-thread time handing results to requests */
    /* MYERS_SERVE_RESIDENT: the resident kernel's shape (batches counts its launches) */
    uint32_t residentWorkgroups;   /* one-wave workgroups per launch */
    uint32_t residentLdsBytes;     /* LDS per workgroup */
    uint64_t residentItems;        /* work items published */
    uint64_t liveRequests;         /* submitted and not yet released by a wait */
    uint64_t drainBatches;         /* batches of several requests above maxBatchCells (backlog drain) */
} myersServeStats_t;

void myersServeDefaultConfig(myersServeConfig_t *cfg);
/* Start a server over numBuffers buffers of one initMyers call (the server
 * uses them exclusively until myersServeStop). */
int  myersServeStart(void **server, void **myersBuffers, uint32_t numBuffers, const myersServeConfig_t *cfg);
/* Queue a request; the input arrays must stay valid until the request is
 * admitted (myersServeWait returns), results until it completes. */
int  myersServeSubmit(void *server, const myersRequest_t *req, uint64_t *ticket);
/* Block until the request completes; latencyMs gets submit-to-completion
 * time.  The wait releases the ticket: the server frees the request, and a
 * later wait on the same ticket returns MYERS_E_INVALID_ARGUMENT (waits that
 * are already blocked on it all return normally). */
int  myersServeWait(void *server, uint64_t ticket, double *latencyMs);
int  myersServeGetStats(void *server, myersServeStats_t *out);
/* Drain every queued request, stop the server thread, release the buffers. */
int  myersServeStop(void *server);
/* Open-loop driver: submit request i at arrivalNs[i] (ns after the call
 * starts, ascending) from a submitter thread, whatever the server's progress;
 * wait for all; latencyMs[i] = completion - arrival. */
int  myersServeRunOpenLoop(void *server, uint32_t numRequests, const myersRequest_t *reqs,
                           const uint64_t *arrivalNs, double *latencyMs, double *wallMs);
/* As myersServeRunOpenLoop, and when queueMs is not NULL, queueMs[i] = time
 * from arrival to the start of forming the batch that took request i (the
 * rest of latencyMs[i] is that batch's copy, launch and verification). */
int  myersServeRunOpenLoopEx(void *server, uint32_t numRequests, const myersRequest_t *reqs,
                             const uint64_t *arrivalNs, double *latencyMs, double *queueMs, double *wallMs);

#ifdef __cplusplus
}
#endif

#endif /* MYERS_AMD_H */
