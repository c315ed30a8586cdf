// pg_launcher.h -- the library's internal host API (launcher, GPU BFS, -cpu
// path, batch) that the C ABI in include/perft_gpu.h and the reference-named
// C++ layer in include/perft_launcher.h are built on.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/perft_gpu.h"
#include "pg_chess.h"

namespace pg {

// 128-bit accumulator for host sums above the launch depth (uint128.h:7-109)
struct u128 {
    uint64_t lo = 0, hi = 0;
    u128() = default;
    explicit u128(uint64_t v) : lo(v), hi(0) {}
    u128 &operator+=(const u128 &o)
    {
        uint64_t old = lo;
        lo += o.lo;
        hi += o.hi + (lo < old ? 1 : 0);
        return *this;
    }
    double toDouble() const { return (double)hi * 18446744073709551616.0 + (double)lo; }
    // decimal text by repeated division by 10^9 on 32-bit limbs
    void toDecimal(char *buf, int n) const;
};

int initLib(int device, uint64_t arenaBytes);
int shutdownLib();
int gpuBfs(const pg_position *pos, int depth, bool useTT, void *buf, uint64_t bytes, uint64_t *count);
int perftLaunch(const pg_position *pos, int depth, int launchDepth, bool useTT, u128 *out);
int perftCpu(const pg_position *pos, int depth, bool useTT, uint64_t *out);
int estimateLD(const pg_position *pos, float *bf, int *ld);
int getEffectiveLD();
void setEffectiveLD(int ld);
uint64_t getOomFallbackCount();
size_t lastBfsPeak();
// largest per-call arena peak since resetPeakSinceReset (printTTStats)
size_t peakSinceReset();
void resetPeakSinceReset();
// arena bytes a GPU call may use (the arena, or the test limit if set)
size_t usableArena();
extern int g_leafPlies;
int perftBatch(int n, const pg_position *pos, const uint8_t *depths, const pg_batch_options *opt, uint64_t *counts);
// a batched engine over a caller's arena partition, stream and pinned buffer
// (a serving slot's), for merging queued requests into one multi-root pass
struct BatchState;
BatchState *batchSlotCreate(void *arena, size_t bytes, hipStream_t st, u64 *pinned);
void batchSlotDestroy(BatchState *B);
int batchSlotRun(BatchState *B, int n, const pg_position *pos, const uint8_t *depths, bool useTT, uint64_t *counts);

}  // namespace pg
