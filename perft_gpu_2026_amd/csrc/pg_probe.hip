// pg_probe.hip -- calibration kernel for the VALU roofline of the move-
// counting kernels (no reference counterpart).
//
// The fused leaf is integer VALU work: no contraction, a few bytes of HBM per
// entry.  Its ceiling is the rate at which the SIMDs issue VALU instructions,
// which the published figures do not give for integer bit operations (gfx950
// also issues two VALU instructions of different waves in one quad-cycle for
// some forms).  This kernel measures it: every wave runs a loop whose body is
// exactly PROBE_BODY VALU instructions written in inline asm, as 8 independent
// dependency chains, so the count per launch is known exactly and latency is
// hidden.  `kind` picks the instruction:
//   0 v_xor_b32        1 v_bitop3_b32     2 v_lshlrev_b64    3 v_bcnt_u32_b32
//   4 v_lshl_add_u64   5 the leaf's mix: per 8 instructions four 32-bit
//     bitwise forms (v_bitop3_b32, v_and_or_b32, v_and_b32, v_or_b32), one
//     v_lshlrev_b64, one v_cndmask_b32, one v_add_u32 and one v_bcnt_u32_b32
//   6 v_lshlrev_b32    7 v_alignbit_b32   8 v_cndmask_b32    9 v_add_u32
//   10 v_and_or_b32    11 v_sub_co_u32 (carry out to vcc)      12 v_and_b32
//   13 v_or_b32        14 v_mov_b32       15 v_bfrev_b32       16 v_or3_b32
//   17 v_ffbl_b32      18 v_add3_u32      19 v_cmp_eq_u32 (to vcc)
//   20 v_xor_b32 in its 64-bit (VOP3) encoding
// (the leaf's static code is ~48 % 32-bit bitwise, ~9 % 64-bit shifts, ~5 %
// selects, ~5 % adds and ~3 % popcounts, the rest moves and compares).
#include "pg_internal.h"

namespace pg {

constexpr int PROBE_BS = 256;
constexpr int PROBE_BODY = 32;  // VALU instructions per loop iteration

// written only if a chain ends on a value it cannot reach, so the loop stays
__device__ u32 g_probeSink;

#define P8(op) op(a0) op(a1) op(a2) op(a3) op(a4) op(a5) op(a6) op(a7)

template <int KIND>
__global__ __launch_bounds__(PROBE_BS) void k_valu_probe(u64 iters)
{
    u32 a0 = threadIdx.x, a1 = a0 * 3, a2 = a0 ^ 5, a3 = a0 + 7, a4 = a0 * 11, a5 = a0 ^ 13, a6 = a0 + 17,
        a7 = a0 * 19;
    u32 b = blockIdx.x | 1, c = a0 ^ 0x5A5A5A5Au;
    u64 s0 = a0, s1 = a1, s2 = a2, s3 = a3, s4 = a4, s5 = a5, s6 = a6, s7 = a7;
    const u64 m = 0x5555555555555555ull;
    for (u64 i = 0; i < iters; i++) {
#pragma unroll
        for (int r = 0; r < PROBE_BODY / 8; r++) {
            if constexpr (KIND == 0) {
#define OP(x) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x) : "v"(b));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 1) {
#define OP(x) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(x) : "v"(b), "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 2) {
#define OP(x) asm volatile("v_lshlrev_b64 %0, 3, %0" : "+v"(x));
                OP(s0) OP(s1) OP(s2) OP(s3) OP(s4) OP(s5) OP(s6) OP(s7)
#undef OP
            } else if constexpr (KIND == 3) {
#define OP(x) asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(x) : "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 4) {
#define OP(x) asm volatile("v_lshl_add_u64 %0, %0, 1, %1" : "+v"(x) : "v"(m));
                OP(s0) OP(s1) OP(s2) OP(s3) OP(s4) OP(s5) OP(s6) OP(s7)
#undef OP
            } else if constexpr (KIND == 6) {
#define OP(x) asm volatile("v_lshlrev_b32 %0, 3, %0" : "+v"(x));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 7) {
#define OP(x) asm volatile("v_alignbit_b32 %0, %0, %1, 24" : "+v"(x) : "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 8) {
#define OP(x) asm volatile("v_cndmask_b32 %0, %0, %1, %2" : "+v"(x) : "v"(c), "s"(m));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 9) {
#define OP(x) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(b));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 10) {
#define OP(x) asm volatile("v_and_or_b32 %0, %0, %1, %2" : "+v"(x) : "v"(b), "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 11) {
#define OP(x) asm volatile("v_sub_co_u32 %0, vcc, %0, %1" : "+v"(x) : "v"(b) : "vcc");
                P8(OP)
#undef OP
            } else if constexpr (KIND == 12) {
#define OP(x) asm volatile("v_and_b32 %0, %0, %1" : "+v"(x) : "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 13) {
#define OP(x) asm volatile("v_or_b32 %0, %0, %1" : "+v"(x) : "v"(b));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 14) {
#define OP(x) asm volatile("v_mov_b32 %0, %1" : "=v"(x) : "v"(x));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 15) {
#define OP(x) asm volatile("v_bfrev_b32 %0, %0" : "+v"(x));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 16) {
#define OP(x) asm volatile("v_or3_b32 %0, %0, %1, %2" : "+v"(x) : "v"(b), "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 17) {
#define OP(x) asm volatile("v_ffbl_b32 %0, %0" : "+v"(x));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 18) {
#define OP(x) asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(x) : "v"(b), "v"(c));
                P8(OP)
#undef OP
            } else if constexpr (KIND == 19) {
#define OP(x) asm volatile("v_cmp_eq_u32 vcc, %0, %1" : : "v"(x), "v"(b) : "vcc");
                P8(OP)
#undef OP
            } else if constexpr (KIND == 20) {
#define OP(x) asm volatile("v_xor_b32_e64 %0, %0, %1" : "+v"(x) : "v"(b));
                P8(OP)
#undef OP
            } else {
                asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(a0) : "v"(b), "v"(c));
                asm volatile("v_and_or_b32 %0, %0, %1, %2" : "+v"(a1) : "v"(b), "v"(c));
                asm volatile("v_and_b32 %0, %0, %1" : "+v"(a2) : "v"(c));
                asm volatile("v_or_b32 %0, %0, %1" : "+v"(a3) : "v"(b));
                asm volatile("v_lshlrev_b64 %0, 3, %0" : "+v"(s4));
                asm volatile("v_cndmask_b32 %0, %0, %1, %2" : "+v"(a5) : "v"(c), "s"(m));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(a6) : "v"(b));
                asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(a7) : "v"(c));
            }
        }
    }
    u32 r = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7 ^ (u32)(s0 ^ s1 ^ s2 ^ s3 ^ s4 ^ s5 ^ s6 ^ s7);
    if (r == 0x9E3779B9u && iters == 0x7FFFFFFFull) g_probeSink = r;
}
#undef P8

constexpr int PROBE_KINDS = 21;
template <int K>
static void launchProbe(int kind, unsigned blocks, hipStream_t st, u64 iters)
{
    if (kind == K) hipLaunchKernelGGL(k_valu_probe<K>, dim3(blocks), dim3(PROBE_BS), 0, st, iters);
    if constexpr (K + 1 < PROBE_KINDS) launchProbe<K + 1>(kind, blocks, st, iters);
}

}  // namespace pg

using namespace pg;

// Runs the probe on wavesPerSimd waves per SIMD of every CU (blocks of four
// waves, one per SIMD) for `iters` loop iterations; *ms = the launch's time
// (HIP events on the library stream), *waveInsts = the VALU instructions it
// issued (blocks x 4 waves x iters x 32).
extern "C" int pg_valu_probe(int kind, int wavesPerSimd, uint64_t iters, double *ms, uint64_t *waveInsts)
{
    Lib &L = lib();
    if (!L.init) return PG_ERR_NOT_INIT;
    if (kind < 0 || kind >= PROBE_KINDS || wavesPerSimd < 1 || wavesPerSimd > 8 || !iters || !ms || !waveInsts) {
        setError("pg_valu_probe: bad argument");
        return PG_ERR_INVALID;
    }
    const unsigned blocks = (unsigned)(L.cuCount ? L.cuCount : 256) * (unsigned)wavesPerSimd;
    hipEvent_t a, b;
    PG_HIP(hipEventCreate(&a));
    PG_HIP(hipEventCreate(&b));
    PG_HIP(hipEventRecord(a, L.stream));
    launchProbe<0>(kind, blocks, L.stream, iters);
    PG_HIP(hipGetLastError());
    PG_HIP(hipEventRecord(b, L.stream));
    PG_HIP(hipEventSynchronize(b));
    float t = 0;
    PG_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t;
    *waveInsts = (uint64_t)blocks * (PROBE_BS / 64) * iters * PROBE_BODY;
    return PG_OK;
}
