// pg_device.h -- wave64 building blocks shared by the gfx950 kernels.
#pragma once
#include "pg_internal.h"

namespace pg {

constexpr int BS = 256;          // threads per block: 4 wave64
constexpr int EXPAND_TILE = 2048;  // children per expand block (8 per thread)
constexpr int EXPAND_WCAP = 2048;  // parents staged in LDS per expand block
constexpr int GEN_LCAP = 16384;    // moves staged in LDS per genmoves block (32 KB)

// ------------------------------------------------------------ wave64 helpers
__device__ __forceinline__ u32 waveSum(u32 v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ u64 waveSum64(u64 v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ u32 waveInclScan(u32 v)
{
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        u32 t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}
__device__ __forceinline__ u64 waveInclScan64(u64 v)
{
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        u64 t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// block-wide sum (BS threads), result valid in thread 0
__device__ __forceinline__ u64 blockSum64(u64 v)
{
    __shared__ u64 part[BS / 64];
    v = waveSum64(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    u64 t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BS / 64; w++) t += part[w];
    return t;
}

// first p in [0, n) with incl[p] > c (n if none)
__device__ __forceinline__ u64 upperBound(const u32 *__restrict__ incl, u64 n, u64 c)
{
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if ((u64)incl[mid] <= c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Intra-wave LDS hand-off: order this wave's LDS writes before its later reads
// by other lanes (no block barrier: the waves of a block run independently).
__device__ __forceinline__ void waveLdsSync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

}  // namespace pg
