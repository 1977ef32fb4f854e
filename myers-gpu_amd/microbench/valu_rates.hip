// Issue rates of gfx950 VALU instructions, the building blocks of the
// bit-vector recurrence (32-bit bit operations, the add with carry, the
// funnel shift) and the 64-bit forms that could replace some of them
// (v_lshlrev_b64, v_lshl_add_u64, v_lshl_or_b32, v_add3_u32, ...).
//
// Each wave runs eight independent chains of one instruction, 32
// instructions per loop iteration, written in inline asm so that the
// instruction is exactly the one named.  The register operands are all
// written before they are read again eight instructions later, so no
// hazard wait is needed inside a chain.  The kernels touch memory only to
// store one word per lane (to keep the chains alive) and one record per
// wave; both buffers are sized for the largest grid.
//
// Output: one line per (instruction, waves per SIMD): wave64 instructions
// issued per SIMD per nanosecond and per shader clock (the clock from the
// waves' own s_memtime / s_memrealtime deltas).
//
//   make -C myers-gpu_amd microbench && myers-gpu_amd/bin/valu-rates
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x,             \
                    hipGetErrorString(e_));                                       \
            exit(1);                                                              \
        }                                                                         \
    } while (0)

enum Op {
    OP_AND, OP_OR, OP_XOR, OP_BITOP3, OP_ADD, OP_ADD3, OP_LSHL, OP_LSHL_OR, OP_LSHL_ADD,
    OP_ALIGNBIT, OP_BFI, OP_PERM, OP_CNDMASK, OP_ADD_CO, OP_ADDC, OP_DPP_MOV,
    OP_LSHL_B64, OP_LSHR_B64, OP_LSHL_ADD_U64, OP_MOV_B64,
    OP_MIX_RECURRENCE, OP_MIX_PAIR64,
    OP_COUNT
};

static const char *kOpName[OP_COUNT] = {
    "v_and_b32", "v_or_b32", "v_xor_b32", "v_bitop3_b32", "v_add_u32", "v_add3_u32",
    "v_lshlrev_b32", "v_lshl_or_b32", "v_lshl_add_u32", "v_alignbit_b32", "v_bfi_b32",
    "v_perm_b32", "v_cndmask_b32(sgpr)", "v_add_co_u32(sgpr)", "v_addc_co_u32(sgpr)",
    "v_mov_b32_dpp", "v_lshlrev_b64", "v_lshrrev_b64", "v_lshl_add_u64", "v_mov_b64",
    "mix:recurrence(10/word)", "mix:pair64(shift via lshl_add_u64)",
};

// Instructions per unrolled repetition (four repetitions per loop iteration).
static int ops_per_rep(int op) {
    return (op == OP_MIX_RECURRENCE || op == OP_MIX_PAIR64) ? 4 * 10 : 8;
}

#define R8(M) M(0) M(1) M(2) M(3) M(4) M(5) M(6) M(7)

template <int kOp>
__global__ void __launch_bounds__(256) rate_kernel(uint32_t *sink, uint64_t *rec, int iters, uint32_t seed) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a0 = seed ^ tid, a1 = a0 * 3u, a2 = a0 * 5u, a3 = a0 * 7u;
    uint32_t a4 = a0 * 11u, a5 = a0 * 13u, a6 = a0 * 17u, a7 = a0 * 19u;
    uint32_t b = seed * 0x9e3779b9u + tid, c = b ^ 0x5bd1e995u;
    uint64_t u0 = ((uint64_t)a1 << 32) | a0, u1 = ((uint64_t)a3 << 32) | a2;
    uint64_t u2 = ((uint64_t)a5 << 32) | a4, u3 = ((uint64_t)a7 << 32) | a6;
    uint64_t u4 = u0 ^ 1u, u5 = u1 ^ 3u, u6 = u2 ^ 5u, u7 = u3 ^ 7u;
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0;   // carries / masks
    const uint64_t bc = ((uint64_t)c << 32) | b;
    const uint64_t smask = ((uint64_t)(seed * 0x2545f491u) << 32) | (seed * 0x9e3779b9u);   // uniform
    __syncthreads();
    const uint64_t t0 = __builtin_amdgcn_s_memtime();
    const uint64_t r0 = __builtin_amdgcn_s_memrealtime();
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int rep = 0; rep < 4; rep++) {
            if constexpr (kOp == OP_AND) {
#define M(k) asm volatile("v_and_b32 %0, %0, %1" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_OR) {
#define M(k) asm volatile("v_or_b32 %0, %0, %1" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_XOR) {
#define M(k) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_BITOP3) {
#define M(k) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xbe" : "+v"(a##k) : "v"(b), "v"(c));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_ADD) {
#define M(k) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_ADD3) {
#define M(k) asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(a##k) : "v"(b), "v"(c));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_LSHL) {
#define M(k) asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(a##k));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_LSHL_OR) {
#define M(k) asm volatile("v_lshl_or_b32 %0, %0, 1, %1" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_LSHL_ADD) {
#define M(k) asm volatile("v_lshl_add_u32 %0, %0, 1, %1" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_ALIGNBIT) {
#define M(k) asm volatile("v_alignbit_b32 %0, %0, %1, 31" : "+v"(a##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_BFI) {
#define M(k) asm volatile("v_bfi_b32 %0, %0, %1, %2" : "+v"(a##k) : "v"(b), "v"(c));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_PERM) {
#define M(k) asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(a##k) : "v"(b), "v"(c));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_CNDMASK) {
#define M(k) asm volatile("v_cndmask_b32 %0, %0, %1, %2" : "+v"(a##k) : "v"(b), "s"(smask));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_ADD_CO) {
#define M(k) asm volatile("v_add_co_u32 %0, %1, %0, %2" : "+v"(a##k), "+s"(s##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_ADDC) {
#define M(k) asm volatile("v_addc_co_u32 %0, %1, %0, %2, %1" : "+v"(a##k), "+s"(s##k) : "v"(b));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_DPP_MOV) {
#define M(k) asm volatile("v_mov_b32_dpp %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf" : "+v"(a##k) : "v"(a##k));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_LSHL_B64) {
#define M(k) asm volatile("v_lshlrev_b64 %0, 1, %0" : "+v"(u##k));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_LSHR_B64) {
#define M(k) asm volatile("v_lshrrev_b64 %0, 1, %0" : "+v"(u##k));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_LSHL_ADD_U64) {
#define M(k) asm volatile("v_lshl_add_u64 %0, %0, 1, %1" : "+v"(u##k) : "v"(bc));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_MOV_B64) {
#define M(k) asm volatile("v_mov_b64 %0, %1" : "=v"(u##k) : "v"(u##k));
                R8(M)
#undef M
            } else if constexpr (kOp == OP_MIX_RECURRENCE) {
                // One word of the recurrence per chain (4 chains, 10 instructions
                // each): and, or, addc, bitop3 x3, and, alignbit x2, and; the
                // four chains' instructions come phase by phase.
#define P1(k, h) asm volatile("v_and_b32 %0, %0, %2\n\tv_or_b32 %1, %1, %3" : "+v"(a##k), "+v"(a##h) : "v"(b), "v"(c));
                P1(0, 4) P1(1, 5) P1(2, 6) P1(3, 7)
#undef P1
#define P2(k) asm volatile("v_addc_co_u32 %0, %1, %0, %2, %1" : "+v"(a##k), "+s"(s##k) : "v"(b));
                P2(0) P2(1) P2(2) P2(3)
#undef P2
#define P3(k) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xbe\n\tv_bitop3_b32 %0, %0, %1, %2 bitop3:0xf1\n\tv_and_b32 %0, %0, %1" : "+v"(a##k) : "v"(b), "v"(c));
                P3(0) P3(1) P3(2) P3(3)
#undef P3
#define P4(k) asm volatile("v_alignbit_b32 %0, %0, %1, 31\n\tv_alignbit_b32 %0, %0, %2, 31" : "+v"(a##k) : "v"(b), "v"(c));
                P4(0) P4(1) P4(2) P4(3)
#undef P4
#define P5(k) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xf1\n\tv_and_b32 %0, %0, %1" : "+v"(a##k) : "v"(b), "v"(c));
                P5(0) P5(1) P5(2) P5(3)
#undef P5
            } else if constexpr (kOp == OP_MIX_PAIR64) {
                // The same word count with the two funnel shifts of each word
                // pair replaced by one v_lshl_add_u64 per pair and vector
                // (2 per pair of words) plus two v_lshrrev_b32 carry bits.
#define P1(k, h) asm volatile("v_and_b32 %0, %0, %2\n\tv_or_b32 %1, %1, %3" : "+v"(a##k), "+v"(a##h) : "v"(b), "v"(c));
                P1(0, 4) P1(1, 5) P1(2, 6) P1(3, 7)
#undef P1
#define P2(k) asm volatile("v_addc_co_u32 %0, %1, %0, %2, %1" : "+v"(a##k), "+s"(s##k) : "v"(b));
                P2(0) P2(1) P2(2) P2(3)
#undef P2
#define P3(k) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xbe\n\tv_bitop3_b32 %0, %0, %1, %2 bitop3:0xf1\n\tv_and_b32 %0, %0, %1" : "+v"(a##k) : "v"(b), "v"(c));
                P3(0) P3(1) P3(2) P3(3)
#undef P3
                asm volatile("v_lshl_add_u64 %0, %0, 1, %2\n\tv_lshl_add_u64 %1, %1, 1, %2" : "+v"(u0), "+v"(u1) : "v"(bc));
                asm volatile("v_lshl_add_u64 %0, %0, 1, %2\n\tv_lshl_add_u64 %1, %1, 1, %2" : "+v"(u2), "+v"(u3) : "v"(bc));
                asm volatile("v_lshrrev_b32 %0, 31, %0\n\tv_lshrrev_b32 %1, 31, %1" : "+v"(a4), "+v"(a5));
                asm volatile("v_lshrrev_b32 %0, 31, %0\n\tv_lshrrev_b32 %1, 31, %1" : "+v"(a6), "+v"(a7));
#define P5(k) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xf1\n\tv_and_b32 %0, %0, %1" : "+v"(a##k) : "v"(b), "v"(c));
                P5(0) P5(1) P5(2) P5(3)
#undef P5
            }
        }
    }
    const uint64_t t1 = __builtin_amdgcn_s_memtime();
    const uint64_t r1 = __builtin_amdgcn_s_memrealtime();
    uint64_t acc = u0 ^ u1 ^ u2 ^ u3 ^ u4 ^ u5 ^ u6 ^ u7 ^ s0 ^ s1 ^ s2 ^ s3 ^ s4 ^ s5 ^ s6 ^ s7;
    sink[tid] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7 ^ b ^ c ^ (uint32_t)acc ^ (uint32_t)(acc >> 32);
    if ((threadIdx.x & 63u) == 0) {
        const uint32_t wave = tid >> 6;
        rec[4 * wave + 0] = t0;
        rec[4 * wave + 1] = t1;
        rec[4 * wave + 2] = r0;
        rec[4 * wave + 3] = r1;
    }
}

typedef void (*kernel_t)(uint32_t *, uint64_t *, int, uint32_t);

template <int... kOps>
static std::vector<kernel_t> table(std::integer_sequence<int, kOps...>) {
    return {rate_kernel<kOps>...};
}

int main(int argc, char **argv) {
    int iters = argc > 1 ? atoi(argv[1]) : 4000;
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const int simds = cus * 4;
    const int kMaxWavesPerSimd = 8;
    const size_t max_threads = (size_t)simds * kMaxWavesPerSimd * 64;
    uint32_t *d_sink;
    uint64_t *d_rec;
    CHECK(hipMalloc(&d_sink, max_threads * sizeof(uint32_t)));
    CHECK(hipMalloc(&d_rec, max_threads / 64 * 4 * sizeof(uint64_t)));
    std::vector<uint64_t> rec(max_threads / 64 * 4);
    auto kernels = table(std::make_integer_sequence<int, OP_COUNT>{});
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    printf("# %s, %d CUs, %d SIMDs; %d loop iterations per wave\n", prop.name, cus, simds, iters);
    printf("# %-36s %6s %10s %12s %12s %10s\n", "instruction", "w/SIMD", "ms", "inst/ns/SIMD", "inst/clk/SIMD", "clk GHz");
    const int waves_list[] = {1, 2, 3, 4, 8};
    for (int op = 0; op < OP_COUNT; op++) {
        for (int w : waves_list) {
            // 256-thread workgroups (one wave per SIMD of a CU); w of them per CU
            const int blocks = cus * w;
            const size_t threads = (size_t)blocks * 256;
            if (threads > max_threads) { fprintf(stderr, "grid too large\n"); return 1; }
            hipLaunchKernelGGL(kernels[op], dim3(blocks), dim3(256), 0, 0, d_sink, d_rec, 16, 1u);   // warm-up
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(kernels[op], dim3(blocks), dim3(256), 0, 0, d_sink, d_rec, iters, 7u);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms = 0;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            CHECK(hipMemcpy(rec.data(), d_rec, threads / 64 * 4 * sizeof(uint64_t), hipMemcpyDeviceToHost));
            double clk = 0, real = 0;
            for (size_t wv = 0; wv < threads / 64; wv++) {
                clk += (double)(rec[4 * wv + 1] - rec[4 * wv + 0]);
                real += (double)(rec[4 * wv + 3] - rec[4 * wv + 2]);
            }
            const double ghz = real > 0 ? clk / real * 0.1 : 0.0;     // s_memrealtime ticks at 100 MHz
            const double inst = (double)threads / 64 * iters * 4.0 * ops_per_rep(op);
            const double per_ns = inst / simds / (ms * 1e6);
            printf("%-38s %6d %10.3f %12.3f %12.3f %10.3f\n", kOpName[op], w, ms, per_ns,
                   ghz > 0 ? per_ns / ghz : 0.0, ghz);
        }
    }
    CHECK(hipFree(d_sink));
    CHECK(hipFree(d_rec));
    return 0;
}
