// pg_device.h -- wave64 building blocks shared by the gfx950 kernels.
#pragma once
#include "pg_internal.h"

namespace pg {

constexpr int BS = 256;          // threads per block: 4 wave64
constexpr int EXPAND_TILE = 2048;  // children per expand block (8 per thread)
// moves staged in LDS per genmoves block (20 KB: eight blocks, 32 waves, per
// CU; a 32 KB tile left five blocks and took the 15-million-board level of
// the config-2 tree 11 % longer; 16 KB sent too many blocks to direct writes)
constexpr int GEN_LCAP = 10240;

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

// Transposition-table counters: each lane tallies its probes, hits and stores
// in registers, and a wave adds its totals to ctr[0..2] once, when it is
// done.  Atomics per chunk on one line (in the batched engine, the line that
// also holds the table pointer every wave reads) serialised the GPU on it and
// halved the speed of the TT leaf.
struct TTTally {
    u32 probes = 0, hits = 0, stores = 0;
    __device__ void add(bool p, bool h, bool s) { probes += p, hits += h, stores += s; }
};
__device__ __forceinline__ void ttFlush(u64 *ctr, const TTTally &t)
{
    u64 p = waveSum64(t.probes), h = waveSum64(t.hits), s = waveSum64(t.stores);
    if ((threadIdx.x & 63) == 0) {
        if (p) atomicAdd((unsigned long long *)&ctr[0], (unsigned long long)p);
        if (h) atomicAdd((unsigned long long *)&ctr[1], (unsigned long long)h);
        if (s) atomicAdd((unsigned long long *)&ctr[2], (unsigned long long)s);
    }
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

// Exclusive scan of `tiles` tile sums into tileOffs[0..tiles], tileOffs[tiles]
// = total, by one block of TOP_BS threads.  The sums are taken in rows of
// 4 * TOP_BS: thread t reads sums 4t..4t+3 of a row with one 16-byte load, so
// each wave load covers 1 KB of consecutive sums; a wave scan of the thread
// totals and the waves' totals in LDS give every sum its offset.  ROWS
// rows share one barrier and have their loads in flight together.
// ACC is the type the waves' totals of a row are added in: u32 for tile sums
// (a row of 4,096 at most 256 * 218 each), u64 for group sums (64 tiles each).
constexpr int TOP_BS = 1024;
constexpr int TOP_ROWS = 8;
// Returns the total in every thread.
template <typename ACC = u32, int ROWS = TOP_ROWS>
__device__ __forceinline__ u64 scanTileSums(const u32 *__restrict__ tileSums, u64 tiles, u64 *__restrict__ tileOffs)
{
    __shared__ u32 wtot[ROWS][TOP_BS / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    constexpr u64 ROW = 4ull * TOP_BS, PASS = ROW * ROWS;
    const bool aligned = ((uintptr_t)tileSums & 15) == 0;
    u64 carry = 0;
    for (u64 c0 = 0; c0 < tiles; c0 += PASS) {
        const u64 left = tiles - c0;
        const int rows = left >= PASS ? ROWS : (int)((left + ROW - 1) / ROW);  // uniform
        u32 x[ROWS][4], ex[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const u64 i0 = c0 + r * ROW + 4u * tid;
            x[r][0] = x[r][1] = x[r][2] = x[r][3] = 0;
            if (r < rows) {
                if (aligned && i0 + 4 <= tiles) {
                    uint4 q = *(const uint4 *)(tileSums + i0);
                    x[r][0] = q.x, x[r][1] = q.y, x[r][2] = q.z, x[r][3] = q.w;
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (i0 + k < tiles) x[r][k] = tileSums[i0 + k];
                }
            }
        }
        // a thread's four sums and a wave's 256 fit u32 for tile sums (at
        // most 256 * 218) and group sums (at most 64 tiles of those)
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            if (r < rows) {
                u32 t = x[r][0] + x[r][1] + x[r][2] + x[r][3];
                u32 incl = waveInclScan(t);
                ex[r] = incl - t;
                if (lane == 63) wtot[r][w] = incl;
            }
        }
        __syncthreads();
        u64 rowBase = carry;
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            if (r < rows) {
                ACC before = 0, all = 0;
#pragma unroll
                for (int k = 0; k < TOP_BS / 64; k++) {
                    ACC t = wtot[r][k];
                    all += t;
                    before += k < w ? t : 0u;
                }
                const u64 i0 = c0 + r * ROW + 4u * tid;
                u64 o = rowBase + before + ex[r];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (i0 + k < tiles) tileOffs[i0 + k] = o;
                    o += x[r][k];
                }
                rowBase += all;
            }
        }
        carry = rowBase;
        __syncthreads();  // wtot is rewritten by the next pass
    }
    if (tid == 0) tileOffs[tiles] = carry;
    return carry;
}

// Publishes a value to a fine-grained pinned host slot that a host thread
// polls while the stream runs on: one system-scope store, which orders no
// other write (what later kernels read is ordered by the stream).
__device__ __forceinline__ void publishToHost(u64 *slot, u64 v)
{
    __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------- grouped scan (plain BFS)
// The tile sums are scanned in groups of SCAN_GROUP tiles, in three passes:
// scanGroups (all CUs) writes every tile's offset inside its group and the
// group sums; one block scans the group sums (scanTileSums over them); each
// apply wave adds its group's offset and its in-group offset.  A single-block
// pass over the tile sums themselves was store-bound on one CU (26 us for the
// 59,155 tiles of a 15-million-entry level with per-tile 64-bit offsets,
// 12.7 us with per-group offsets and a wave sum in the apply).
constexpr int SCAN_GROUP = 64;
__host__ __device__ constexpr u64 scanGroupCount(u64 tiles) { return (tiles + SCAN_GROUP - 1) / SCAN_GROUP; }

// One wave per four groups: lane l holds tiles 4l..4l+3 of the wave's 256
// (one 16-byte load when whole and aligned), so a group is one 16-lane row;
// a row scan gives the in-group offsets (at most 64 * 256 * 218: u32).
__device__ __forceinline__ void scanGroups(const u32 *__restrict__ tileSums, u64 tiles, u32 *__restrict__ tileIn,
                                           u32 *__restrict__ groupSums, u64 quad)
{
    const int lane = threadIdx.x & 63;
    const u64 t0 = quad * 4 * SCAN_GROUP;
    if (t0 >= tiles) return;
    const u64 t = t0 + 4ull * lane;
    u32 x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    if (t + 4 <= tiles && ((uintptr_t)tileSums & 15) == 0) {
        uint4 q = *(const uint4 *)(tileSums + t);
        x0 = q.x, x1 = q.y, x2 = q.z, x3 = q.w;
    } else {
        if (t < tiles) x0 = tileSums[t];
        if (t + 1 < tiles) x1 = tileSums[t + 1];
        if (t + 2 < tiles) x2 = tileSums[t + 2];
        if (t + 3 < tiles) x3 = tileSums[t + 3];
    }
    const u32 sum = x0 + x1 + x2 + x3;
    u32 inc = sum;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        const u32 v = __shfl_up(inc, o, 16);
        if ((lane & 15) >= o) inc += v;
    }
    const u32 e0 = inc - sum, e1 = e0 + x0, e2 = e1 + x1, e3 = e2 + x2;
    if (t + 4 <= tiles && ((uintptr_t)tileIn & 15) == 0) {
        *(uint4 *)(tileIn + t) = make_uint4(e0, e1, e2, e3);
    } else {
        if (t < tiles) tileIn[t] = e0;
        if (t + 1 < tiles) tileIn[t + 1] = e1;
        if (t + 2 < tiles) tileIn[t + 2] = e2;
        if (t + 3 < tiles) tileIn[t + 3] = e3;
    }
    const u64 g = quad * 4 + (lane >> 4);
    if ((lane & 15) == 15 && g * SCAN_GROUP < tiles) groupSums[g] = inc;
}

// The expand partitions from the scan itself: parent p is the first parent of
// expand tile t iff excl(p) <= t * EXPAND_TILE < incl(p) (a parent has fewer
// than EXPAND_TILE children, so it starts at most one tile), which is
// upper_bound(incl, t * EXPAND_TILE); the entry n - 1 also writes the end
// entry parts[ceil(C / EXPAND_TILE)] = n - 1.  Writes stay below `cap`.
__device__ __forceinline__ void markTileStart(u32 *__restrict__ parts, u64 cap, u32 excl, u32 incl, u64 p, u64 n)
{
    const u64 t = ((u64)excl + EXPAND_TILE - 1) / EXPAND_TILE;
    if (t * EXPAND_TILE < incl && t < cap) parts[t] = (u32)p;
    if (p == n - 1) {
        const u64 te = ((u64)incl + EXPAND_TILE - 1) / EXPAND_TILE;
        if (te < cap) parts[te] = (u32)p;
    }
}

// scanApplyTile with the grouped offsets: the tile's offset is its group's
// offset plus its offset inside the group.  With `parts` non-null it also
// writes the expand partitions (markTileStart).
__device__ __forceinline__ void scanApplyTileGrouped(u32 *__restrict__ d, u64 n, const u32 *__restrict__ tileIn,
                                                     const u64 *__restrict__ groupOffs, u64 tile,
                                                     u32 *__restrict__ parts, u64 partsCap)
{
    const int lane = threadIdx.x & 63;
    if (tile * BS >= n) return;
    const u32 tileOff = (u32)groupOffs[tile / SCAN_GROUP] + tileIn[tile];
    const u64 i0 = tile * BS + (u64)lane * 4;
    const bool vec = i0 + 4 <= n && ((uintptr_t)d & 15) == 0;
    u32 x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    if (vec) {
        uint4 q = *(const uint4 *)(d + i0);
        x0 = q.x, x1 = q.y, x2 = q.z, x3 = q.w;
    } else {
        if (i0 < n) x0 = d[i0];
        if (i0 + 1 < n) x1 = d[i0 + 1];
        if (i0 + 2 < n) x2 = d[i0 + 2];
        if (i0 + 3 < n) x3 = d[i0 + 3];
    }
    u32 s1 = x0 + x1, s2 = s1 + x2, s3 = s2 + x3;
    u32 base = tileOff + waveInclScan(s3) - s3;
    if (vec) {
        *(uint4 *)(d + i0) = make_uint4(base + x0, base + s1, base + s2, base + s3);
    } else {
        if (i0 < n) d[i0] = base + x0;
        if (i0 + 1 < n) d[i0 + 1] = base + s1;
        if (i0 + 2 < n) d[i0 + 2] = base + s2;
        if (i0 + 3 < n) d[i0 + 3] = base + s3;
    }
    if (parts) {
        // entries past n have no children (x = 0) and mark nothing
        markTileStart(parts, partsCap, base, base + x0, i0, n);
        markTileStart(parts, partsCap, base + x0, base + s1, i0 + 1, n);
        markTileStart(parts, partsCap, base + s1, base + s2, i0 + 2, n);
        markTileStart(parts, partsCap, base + s2, base + s3, i0 + 3, n);
    }
}

// One wave turns the 256 counts of `tile` into inclusive counts: lane l holds
// entries 4l..4l+3 of the tile (a 16-byte access when whole and aligned),
// scans them locally, and adds the wave's exclusive lane prefix and the
// tile's offset.  Waves whose tile starts past n do nothing.
__device__ __forceinline__ void scanApplyTile(u32 *__restrict__ d, u64 n, const u64 *__restrict__ tileOffs, u64 tile)
{
    const int lane = threadIdx.x & 63;
    if (tile * BS >= n) return;
    const u64 i0 = tile * BS + (u64)lane * 4;
    const bool vec = i0 + 4 <= n && ((uintptr_t)d & 15) == 0;
    u32 x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    if (vec) {
        uint4 q = *(const uint4 *)(d + i0);
        x0 = q.x, x1 = q.y, x2 = q.z, x3 = q.w;
    } else {
        if (i0 < n) x0 = d[i0];
        if (i0 + 1 < n) x1 = d[i0 + 1];
        if (i0 + 2 < n) x2 = d[i0 + 2];
        if (i0 + 3 < n) x3 = d[i0 + 3];
    }
    u32 s1 = x0 + x1, s2 = s1 + x2, s3 = s2 + x3;
    u32 base = (u32)tileOffs[tile] + waveInclScan(s3) - s3;
    if (vec) {
        *(uint4 *)(d + i0) = make_uint4(base + x0, base + s1, base + s2, base + s3);
    } else {
        if (i0 < n) d[i0] = base + x0;
        if (i0 + 1 < n) d[i0 + 1] = base + s1;
        if (i0 + 2 < n) d[i0 + 2] = base + s2;
        if (i0 + 3 < n) d[i0 + 3] = base + s3;
    }
}

// ------------------------------------------------------- expand (one tile)
// Parents of the children c0 .. c0+EXPAND_TILE-1, given the parents p0..p1
// whose child ranges meet the tile (incl = inclusive child counts).  Each
// parent with children writes its index + 1 at the LDS position of its first
// child (the parent holding the tile's first child at position 0); an
// inclusive max-scan then gives every child its parent.  Thread t returns the
// parents of children c0 + 8t .. c0 + 8t + 7 in o[].  `tile` must hold zeros
// on entry; every thread's reads of it are done when this returns.
constexpr int EXPAND_PER = EXPAND_TILE / BS;
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void expandTileParents(const u32 *__restrict__ incl, u64 p0, u64 p1, u64 c0,
                                                  u32 *__restrict__ tile, u32 *__restrict__ wmax,
                                                  u32 (&o)[EXPAND_PER])
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (u64 p = p0 + tid; p <= p1; p += BS) {
        u64 ex = p ? incl[p - 1] : 0, in = incl[p];
        if (in > ex) {
            u64 st = ex > c0 ? ex - c0 : 0;
            if (st < EXPAND_TILE) tile[st] = (u32)p + 1;
        }
    }
    __syncthreads();
    u32 v[EXPAND_PER];
    u32 m = 0;
#pragma unroll
    for (int k = 0; k < EXPAND_PER; k++) {
        u32 x = tile[tid * EXPAND_PER + k];
        m = x > m ? x : m;
        v[k] = m;
    }
    // running max across the wave, then across the block's waves
    u32 wm = m;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        u32 t = __shfl_up(wm, d);
        if (lane >= d) wm = t > wm ? t : wm;
    }
    if (lane == 63) wmax[w] = wm;
    u32 prev = __shfl_up(wm, 1);
    if (lane == 0) prev = 0;
    __syncthreads();
    for (int k = 0; k < w; k++) prev = wmax[k] > prev ? wmax[k] : prev;
#pragma unroll
    for (int k = 0; k < EXPAND_PER; k++) o[k] = (v[k] > prev ? v[k] : prev) - 1;
}

// Writes one tile's values (thread t holding entries 8t .. 8t+7 in o[]) to
// dst[c0 .. min(c0 + EXPAND_TILE, c)).  The values go through LDS so that each
// store instruction of a wave covers 1 KB of consecutive entries, written with
// non-temporal 16-byte stores: the output is far larger than the caches and is
// read back by the next level's kernels, not by this one.  No thread may
// still read `tile` on entry; the caller places a barrier before reusing it.
__device__ __forceinline__ void storeTileStreaming(u32 *__restrict__ tile, const u32 (&o)[EXPAND_PER],
                                                   u32 *__restrict__ dst, u64 c0, u64 c)
{
    const int tid = threadIdx.x;
    u32x4 *t4 = (u32x4 *)tile;
#pragma unroll
    for (int k = 0; k < EXPAND_PER / 4; k++)
        t4[tid * (EXPAND_PER / 4) + k] = u32x4{o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]};
    __syncthreads();
    if (c0 + EXPAND_TILE <= c && ((uintptr_t)(dst + c0) & 15) == 0) {
        u32x4 *d4 = (u32x4 *)(dst + c0);
#pragma unroll
        for (int k = 0; k < EXPAND_PER / 4; k++) __builtin_nontemporal_store(t4[k * BS + tid], d4 + k * BS + tid);
    } else {
        for (int k = tid; k < EXPAND_TILE; k += BS)
            if (c0 + k < c) dst[c0 + k] = tile[k];
    }
}

// Moves of one genmoves block from its LDS tile to out[gs, gs + total).  The
// tile holds global index g at LDS index g - gs + (gs & 7), so that every
// 8-move global chunk starts at a multiple of 8 in LDS: whole chunks go out
// as one 16-byte LDS read and one 16-byte store per lane; the partial chunks
// at either end move element by element (their other entries belong to the
// neighbouring blocks).  out must be 16-byte aligned, buf too.
__device__ __forceinline__ void copyMovesOut(const u16 *__restrict__ buf, u16 *__restrict__ out, u64 gs, u64 total)
{
    const u64 ge = gs + total, c0 = gs >> 3, c1 = (ge + 7) >> 3;
    for (u64 c = c0 + threadIdx.x; c < c1; c += BS) {
        const u32 l = (u32)((c - c0) << 3);
        const u64 g = c << 3;
        if (g >= gs && g + 8 <= ge) {
            *(uint4 *)(out + g) = *(const uint4 *)(buf + l);
        } else {
            for (int j = 0; j < 8; j++)
                if (g + j >= gs && g + j < ge) out[g + j] = buf[l + j];
        }
    }
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
