// pg_kernels.hip -- gfx950 kernels of the host-driven BFS and the plain
// per-position driver (perft_gpu_host_bfs, perft_kernels.cu:620-990).
//
// One BFS level (entries = (parent index, move) pairs):
//   k_makecount   child board = make(parent board, move); count its moves;
//                 per-256-entry tile sums for the scan            (VALU bound)
//   k_scan_groups in-group tile offsets and sums of 64-tile groups
//   k_scan_top    exclusive scan of the group sums (one block)
//   k_scan_small  both in one block, for levels of at most 64 groups
//   k_scan_apply  wave64 tile scan + group offset + in-group offset
//                 -> inclusive counts                            (HBM)
//   k_partitions  one upper_bound per 2048-child tile
//   k_expand      parent index of every child: first-child scatter + max-scan
//                 in LDS, coalesced non-temporal 16-byte stores    (HBM)
//   k_genmoves    moves of every board, staged in LDS, written coalesced
//   k_expand_genmoves  expand and genmoves in one launch on levels of at
//                 most FUSE_MAX_CHILDREN children
//   k_level_small all of the above in one block, for levels of at most
//                 SMALL_LEVEL entries (the first levels of every call)
//   k_level_root  the same for the first level, from the root and its moves
//                 passed as kernel arguments
// and after the last level a leaf kernel sums perft(1) or perft(2) of every
// entry (k_leaf; the 2-ply form is the reference's fused leaf, here with the
// child moves consumed as they are generated -- no per-thread move list).
// The host learns each level's size, and at the end the count, from values
// the scan (and k_publish_u64) store into a pinned host slot that it polls.
//
// Indices are u32 (a level holds < 2^32 entries; larger levels report
// PG_ERR_ARENA_OOM and the caller splits one ply down), sizes are 64-bit.
#include <string.h>

#include <memory>
#include <type_traits>
#include <unordered_map>

#include "pg_internal.h"
#include "pg_device.h"
#include "pg_tt.h"

namespace pg {

// ------------------------------------------------------------- level kernels
template <int C>
__global__ __launch_bounds__(BS) void k_makecount(const QuadBitBoard *__restrict__ pb, const u8 *__restrict__ ps,
                                                  const u32 *__restrict__ pidx, const u16 *__restrict__ mv, u64 n,
                                                  QuadBitBoard *__restrict__ ob, u8 *__restrict__ os,
                                                  u32 *__restrict__ cnt, u32 *__restrict__ tileSums)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    u32 c = 0;
    if (i < n) {
        u32 p = pidx[i];
        QuadBitBoard q = pb[p];
        u8 st = ps[p];
        makeMove<C>(q, st, mv[i]);
        ob[i] = q;
        os[i] = st;
        c = countChildren<1 - C>(q, st);
        cnt[i] = c;
    }
    u64 s = blockSum64(c);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = (u32)s;
}

// per-tile sums when the caller has none (pg_scan_u32 on arbitrary data)
__global__ __launch_bounds__(BS) void k_tile_sums(const u32 *__restrict__ d, u64 n, u32 *__restrict__ tileSums)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    u64 s = blockSum64(i < n ? d[i] : 0);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = (u32)s;
}

// in-group tile offsets and group sums, one wave per four groups
__global__ __launch_bounds__(BS) void k_scan_groups(const u32 *__restrict__ tileSums, u64 tiles,
                                                    u32 *__restrict__ tileIn, u32 *__restrict__ groupSums)
{
    scanGroups(tileSums, tiles, tileIn, groupSums, (u64)blockIdx.x * (BS / 64) + (threadIdx.x >> 6));
}

// Levels of at most 64 groups (4,096 tiles, 1,048,576 entries): the group
// pass and the scan of the group sums in one block, one launch fewer.  With
// `hostTotal` non-null the total is also published to that pinned host slot.
__global__ __launch_bounds__(TOP_BS) void k_scan_small(const u32 *__restrict__ tileSums, u64 tiles,
                                                       u32 *__restrict__ tileIn, u64 *__restrict__ groupOffs,
                                                       u64 *hostTotal)
{
    __shared__ u32 gs[SCAN_GROUP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 G = scanGroupCount(tiles);
    if (threadIdx.x < SCAN_GROUP) gs[threadIdx.x] = 0;
    __syncthreads();
    scanGroups(tileSums, tiles, tileIn, gs, (u64)w);  // 16 waves x 4 groups
    __syncthreads();
    if (w == 0) {
        const u32 v = gs[lane], inc = waveInclScan(v);  // at most 64 groups of 64 * 256 * 218
        if ((u64)lane < G) groupOffs[lane] = inc - v;
        if (lane == 63) {
            groupOffs[G] = inc;
            if (hostTotal) publishToHost(hostTotal, inc);
        }
    }
}

// exclusive scan of the G group sums into groupOffs[0..G] (one block), the
// total also to `hostTotal` when non-null
__global__ __launch_bounds__(TOP_BS) void k_scan_top(const u32 *__restrict__ groupSums, u64 groups,
                                                     u64 *__restrict__ groupOffs, u64 *hostTotal)
{
    const u64 total = scanTileSums<u64, 2>(groupSums, groups, groupOffs);
    if (hostTotal && threadIdx.x == 0) publishToHost(hostTotal, total);
}

// Inclusive counts from the group and in-group offsets: one wave per
// 256-entry tile, four consecutive entries per lane (one 16-byte load and
// store), a lane-local scan of the four and one wave scan of the lane totals;
// no block barrier.  With `parts` non-null it also writes the partitions of
// the next expand (what k_partitions computes by binary search).
__global__ __launch_bounds__(BS) void k_scan_apply(u32 *__restrict__ d, u64 n, const u32 *__restrict__ tileIn,
                                                   const u64 *__restrict__ groupOffs, u32 *__restrict__ parts,
                                                   u64 partsCap)
{
    scanApplyTileGrouped(d, n, tileIn, groupOffs, (u64)blockIdx.x * (BS / 64) + (threadIdx.x >> 6), parts,
                         partsCap);
}

// parts[t] = parent of child t*EXPAND_TILE, for t in [0, tiles]
__global__ __launch_bounds__(BS) void k_partitions(const u32 *__restrict__ incl, u64 n, u64 c, u64 tiles,
                                                   u32 *__restrict__ parts)
{
    u64 t = (u64)blockIdx.x * BS + threadIdx.x;
    if (t > tiles) return;
    u64 child = t * EXPAND_TILE;
    if (child > c) child = c;
    u64 p = upperBound(incl, n, child);
    parts[t] = (u32)(p < n ? p : n - 1);
}

// Parent index of every child of one 2048-child tile (first-child scatter and
// max-scan in LDS, then streaming 16-byte stores).  Per child: one LDS clear,
// one read, a max, a 4-byte LDS round trip and 4 bytes of a coalesced store --
// no per-child search.
__device__ __forceinline__ void expandBlock(const u32 *__restrict__ incl, u64 c, const u32 *__restrict__ parts,
                                            u32 *__restrict__ parentIdx, u64 blk, u32 *tile, u32 *wmax)
{
    const u64 c0 = blk * EXPAND_TILE;
    for (int k = threadIdx.x; k < EXPAND_TILE; k += BS) tile[k] = 0;
    __syncthreads();
    u32 o[EXPAND_PER];
    expandTileParents(incl, parts[blk], parts[blk + 1], c0, tile, wmax, o);
    storeTileStreaming(tile, o, parentIdx, c0, c);
}
__global__ __launch_bounds__(BS) void k_expand(const u32 *__restrict__ incl, u64 n, u64 c,
                                               const u32 *__restrict__ parts, u32 *__restrict__ parentIdx)
{
    __shared__ __attribute__((aligned(16))) u32 tile[EXPAND_TILE];
    __shared__ u32 wmax[BS / 64];
    expandBlock(incl, c, parts, parentIdx, blockIdx.x, tile, wmax);
}

// A whole small level in one block of SMALL_LEVEL threads, one entry each:
// make + exact count, a block scan of the counts, the total published to the
// pinned host slot, then every child written as (parent index, move) at its
// offset.  The child arrays are sized for n * MAX_MOVES before the launch, so
// nothing waits for the total; this replaces five launches (make+count, two
// scan passes, expand, genmoves) where each would run a few microseconds.
constexpr u32 SMALL_LEVEL = 1024;
// the block's part after the make: child i (valid when i < n) is written,
// counted, scanned, published and expanded
template <int C>  // side to move in the parent boards
__device__ __forceinline__ void levelSmallRest(const QuadBitBoard &q, u8 st, bool valid, QuadBitBoard *__restrict__ ob,
                                               u8 *__restrict__ os, u32 *__restrict__ nidx, u16 *__restrict__ nmv,
                                               u64 *hostTotal)
{
    __shared__ u32 wsum[SMALL_LEVEL / 64];
    const u32 i = threadIdx.x, lane = i & 63, w = i >> 6;
    u32 cnt = 0;
    if (valid) {
        ob[i] = q;
        os[i] = st;
        cnt = countChildren<1 - C>(q, st);
    }
    const u32 incl = waveInclScan(cnt);  // at most 1024 * 218 children
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    u32 before = 0, total = 0;
#pragma unroll
    for (u32 k = 0; k < SMALL_LEVEL / 64; k++) {
        const u32 t = wsum[k];
        before += k < w ? t : 0u;
        total += t;
    }
    if (i == 0) publishToHost(hostTotal, total);
    if (cnt) {
        const u32 off = before + incl - cnt;
        generateMoves<1 - C>(q, st, [&](u32 k, u16 m) {
            if (k < cnt) nmv[off + k] = m, nidx[off + k] = i;
        });
    }
}

template <int C>
__global__ __launch_bounds__(SMALL_LEVEL) void k_level_small(const QuadBitBoard *__restrict__ pb,
                                                             const u8 *__restrict__ ps, const u32 *__restrict__ pidx,
                                                             const u16 *__restrict__ mv, u32 n,
                                                             QuadBitBoard *__restrict__ ob, u8 *__restrict__ os,
                                                             u32 *__restrict__ nidx, u16 *__restrict__ nmv,
                                                             u64 *hostTotal)
{
    const u32 i = threadIdx.x;
    QuadBitBoard q{};
    u8 st = 0;
    if (i < n) {
        const u32 p = pidx[i];
        q = pb[p];
        st = ps[p];
        makeMove<C>(q, st, mv[i]);
    }
    levelSmallRest<C>(q, st, i < n, ob, os, nidx, nmv, hostTotal);
}

// The first level of a call, from the root and its moves passed as kernel
// arguments (no host-to-device copy before the first kernel); it also clears
// the call's two device counters (the leaf's sum and chunk queue).
struct RootArgs {
    QuadBitBoard board;
    u32 n;
    u8 state;
    u16 moves[MAX_MOVES];
};
template <int C>
__global__ __launch_bounds__(SMALL_LEVEL) void k_level_root(const RootArgs r, QuadBitBoard *__restrict__ ob,
                                                            u8 *__restrict__ os, u32 *__restrict__ nidx,
                                                            u16 *__restrict__ nmv, u64 *hostTotal,
                                                            u64 *__restrict__ counters)
{
    const u32 i = threadIdx.x;
    if (i < 2) counters[i] = 0;
    QuadBitBoard q = r.board;
    u8 st = r.state;
    if (i < r.n) makeMove<C>(q, st, r.moves[i]);
    levelSmallRest<C>(q, st, i < r.n, ob, os, nidx, nmv, hostTotal);
}

template <int C>
__device__ __forceinline__ void genmovesBlock(const QuadBitBoard *__restrict__ boards, const u8 *__restrict__ states,
                                              const u32 *__restrict__ incl, u64 n, u16 *__restrict__ out, u64 blk,
                                              u16 *buf)
{
    u64 b0 = blk * BS;
    u64 i = b0 + threadIdx.x;
    u64 bEnd = b0 + BS < n ? b0 + BS : n;
    u64 blockStart = b0 ? incl[b0 - 1] : 0;
    u64 total = (u64)incl[bEnd - 1] - blockStart;
    u64 excl = 0, mine = 0;
    if (i < n) {
        excl = i ? (u64)incl[i - 1] : 0;
        mine = (u64)incl[i] - excl;  // 0 for a table hit or a duplicate: nothing to write
    }
    const u64 sh = ((uintptr_t)out & 15) == 0 ? (blockStart & 7) : 0;  // the tile's offset (copyMovesOut)
    if (total + sh <= GEN_LCAP) {
        if (mine) {
            u16 *dst = buf + sh + (excl - blockStart);
            generateMoves<C>(boards[i], states[i], [&](u32 k, u16 m) {
                if (k < mine) dst[k] = m;
            });
        }
        __syncthreads();
        if (((uintptr_t)out & 15) == 0) {
            copyMovesOut(buf, out, blockStart, total);
        } else {
            u16 *o = out + blockStart;
            for (u64 j = threadIdx.x; j < total; j += BS) o[j] = buf[j];
        }
    } else if (mine) {
        u16 *dst = out + excl;
        generateMoves<C>(boards[i], states[i], [&](u32 k, u16 m) {
            if (k < mine) dst[k] = m;
        });
    }
}

template <int C>
__global__ __launch_bounds__(BS) void k_genmoves(const QuadBitBoard *__restrict__ boards, const u8 *__restrict__ states,
                                                 const u32 *__restrict__ incl, u64 n, u16 *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) u16 buf[GEN_LCAP];
    genmovesBlock<C>(boards, states, incl, n, out, blockIdx.x, buf);
}

// expand and genmoves of one level in one launch: `genBlocks` blocks generate
// the moves of 256 boards each and `tiles` blocks write the parent index of
// 2,048 children each, interleaved over the grid (block b generates iff
// floor((b+1)G/T) > floor(bG/T)), so the HBM-bound expand tiles run beside
// the VALU-bound generation.  The expand tile takes the first 8 KB of the
// generation's LDS buffer.
template <int C>
__global__ __launch_bounds__(BS) void k_expand_genmoves(const QuadBitBoard *__restrict__ boards,
                                                        const u8 *__restrict__ states, const u32 *__restrict__ incl,
                                                        u64 n, u16 *__restrict__ out, u64 c,
                                                        const u32 *__restrict__ parts, u32 *__restrict__ parentIdx,
                                                        u64 genBlocks, u64 tiles)
{
    static_assert(GEN_LCAP * sizeof(u16) >= EXPAND_TILE * sizeof(u32), "expand tile in the generation buffer");
    __shared__ __attribute__((aligned(16))) u16 buf[GEN_LCAP];
    __shared__ u32 wmax[BS / 64];
    const u64 b = blockIdx.x, T = genBlocks + tiles;
    const u64 g0 = b * genBlocks / T, g1 = (b + 1) * genBlocks / T;
    if (g1 > g0) genmovesBlock<C>(boards, states, incl, n, out, g0, buf);
    else expandBlock(incl, c, parts, parentIdx, b - g0, (u32 *)buf, wmax);
}

// One-ply leaf (makeMove_and_perft_leaf_kernel, perft_kernels.cu:313-350):
// perft(1) of make(parent, move) per entry.  Persistent grid-stride loop with
// per-thread sums and one u64 atomic per block at the end (one atomic per
// 256 entries serialised on a single address cost more than the counting).
template <int C>
__global__ __launch_bounds__(BS) void k_leaf1(const QuadBitBoard *__restrict__ pb, const u8 *__restrict__ ps,
                                              const u32 *__restrict__ pidx, const u16 *__restrict__ mv, u64 n,
                                              u64 *__restrict__ sum)
{
    u64 acc = 0;
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < n; i += (u64)gridDim.x * BS) {
        u32 p = pidx[i];
        QuadBitBoard q = pb[p];
        u8 st = ps[p];
        makeMove<C>(q, st, mv[i]);
        acc += countMoves<1 - C>(q, st);
    }
    u64 s = blockSum64(acc);
    if (threadIdx.x == 0 && s) atomicAdd((unsigned long long *)sum, (unsigned long long)s);
}

// Fused two-ply leaf (fused_2level_leaf_kernel, perft_kernels.cu:382-480),
// wave-cooperative: each wave takes L2_CHUNK entries, makes the child
// positions (one lane each) and counts their moves, scans the counts across
// the wave, stages every (child, grand-child move) pair in LDS, then all 64
// lanes walk the flat pair list doing make + countMoves.  Every lane gets the
// same number of pairs however uneven the children's move counts are, no
// per-thread move list is kept (the reference's CMove[80] stack array and its
// 80-move cap are gone), and generation and counting live in separate loops
// so the counter's register budget is not shared with the generator's.
// Pair lists longer than the LDS capacity are processed in windows.  A chunk
// of 64 entries keeps every lane busy in the child phase (make, count,
// generate).  A pair takes 3 bytes of LDS (the move, and the lane of its
// child in a separate byte array), and the kernels are held to 128 VGPRs, so
// four waves fit a SIMD (140 VGPRs, 4-byte pairs and three waves per SIMD
// measured 3.6 % slower; five waves need 96 VGPRs, which spill, and measured
// 11 % slower).  The window is as large as four blocks per CU allow: 2,688
// pairs (40,736 B of LDS per block).  64 entries of the config-2 tree have
// 1,753 pairs on average; 15 % of the chunks exceeded the 2,016 pairs of the
// earlier window and took a second one (their straddling lanes generate
// twice, and each window ends in a partial pass), 0.3 % exceed 2,688: 1.35 %
// faster.  The table leaf keeps 2,560 (its per-entry sums take 1 KB more),
// the table-lookup form 2,016 (its LDS bishop table leaves two blocks per CU).
#define L2_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4, 4)))
constexpr int L2_CHUNK = 64;
constexpr int L2_CAP = 2688;
constexpr int L2_CAP_TT = 2560;
constexpr int L2_CAP_MAGIC = 2016;
constexpr int L2_WAVES = 4;

// Fancy magic slider tables (pg_magic.h) for the table-lookup form of the
// fused leaf's counter: square rows and attack sets in device memory (845
// KB, L2-resident), filled once by uploadMagicTables.
static __device__ MagicSquare g_magicSq[128];
static __device__ u64 g_magicAtt[MAGIC_ENTRIES];

// Slider-attack method of the fused leaf: computed fills (the default), table
// lookups from device memory, or lookups with the bishop sets and the square
// rows copied into each block's LDS (46 KB more per block: two blocks per CU
// instead of four)
enum { SL_FILLS = 0, SL_MAGIC = 1, SL_MAGIC_LDS = 2 };
int g_sliderMode = SL_FILLS;

// Chunks are taken either by a static stride over the grid, or (QUEUE) from
// an atomic counter `work` so that no wave idles while others still hold
// several chunks (the counter starts at zero; the grid is the resident one).
template <int C, int CHUNK, int CAP, bool QUEUE, int SLM = SL_FILLS>
__device__ __forceinline__ void leaf2Body(const QuadBitBoard *__restrict__ pb, const u8 *__restrict__ ps,
                                          const u32 *__restrict__ pidx, const u16 *__restrict__ mv, u64 n,
                                          u64 *__restrict__ sum, u64 *work)
{
    static_assert(CHUNK <= 64, "one entry per lane");
    __shared__ QuadBitBoard sb[L2_WAVES][CHUNK];
    __shared__ u8 ss[L2_WAVES][CHUNK];
    __shared__ u16 slm[L2_WAVES][CAP + 1];  // slot CAP takes the moves outside the window
    __shared__ u8 sll[L2_WAVES][CAP + 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    using SL = typename std::conditional<SLM == SL_FILLS, FillSliders, MagicSliders>::type;
    SL sl{};
    if constexpr (SLM == SL_MAGIC) {
        sl = MagicSliders{g_magicSq, g_magicAtt, g_magicAtt + MAGIC_ROOK_ENTRIES};
    } else if constexpr (SLM == SL_MAGIC_LDS) {
        __shared__ u64 lb[MAGIC_BISHOP_ENTRIES];
        __shared__ MagicSquare lsq[128];
        for (int i = threadIdx.x; i < MAGIC_BISHOP_ENTRIES; i += blockDim.x) lb[i] = g_magicAtt[MAGIC_ROOK_ENTRIES + i];
        for (int i = threadIdx.x; i < 128; i += blockDim.x) lsq[i] = g_magicSq[i];
        __syncthreads();
        sl = MagicSliders{lsq, g_magicAtt, lb};
    }
    u64 acc = 0;
    const u64 stride = (u64)gridDim.x * L2_WAVES * CHUNK;
    u64 next = ((u64)blockIdx.x * L2_WAVES + w) * CHUNK;
    // the queue ticket of the next chunk is taken as the current one starts,
    // so the atomic's round trip overlaps a chunk's work (0.7 % faster)
    u64 ticket = 0;
    if constexpr (QUEUE) {
        if (lane == 0) ticket = atomicAdd((unsigned long long *)work, (unsigned long long)CHUNK);
    }
    for (;;) {
        u64 base;
        if constexpr (QUEUE) {
            base = __shfl(ticket, 0);
            if (base >= n) break;
            if (lane == 0) ticket = atomicAdd((unsigned long long *)work, (unsigned long long)CHUNK);
        } else {
            base = next;
            next += stride;
        }
        if (base >= n) break;
        u32 cnt = 0;
        u64 e = base + lane;
        BB catt;  // the child's enemy attacks, from its count, for its generation
        if (lane < CHUNK && e < n) {
            u32 p = pidx[e];
            QuadBitBoard q = pb[p];
            u8 st = ps[p];
            makeMove<C>(q, st, mv[e]);
            sb[w][lane] = q;
            ss[w][lane] = st;
            Frame fr;
            cnt = countChildrenFrame<1 - C>(q, st, sl, fr);
            catt = fr.att;
        }
        u32 incl = waveInclScan(cnt);
        u32 off = incl - cnt;
        u32 total = __shfl(incl, 63);
        waveLdsSync();
        for (u32 wb = 0; wb < total; wb += CAP) {
            if (cnt && off < wb + CAP && off + cnt > wb) {
                QuadBitBoard q = sb[w][lane];
                u8 st = ss[w][lane];
                u16 *dm = slm[w];
                u8 *dl = sll[w];
                Frame fr = framePins<1 - C>(q);
                fr.att = catt;
                generateMovesF<1 - C>(fr, st, [&](u32 k, u16 m) {
                    u32 j = off + k - wb;  // wraps below the window
                    j = j < (u32)CAP ? j : (u32)CAP;  // outside the window: the spare slot, no branch
                    dm[j] = m, dl[j] = (u8)lane;
                });
            }
            waveLdsSync();
            u32 wn = total - wb < (u32)CAP ? total - wb : (u32)CAP;
            u32 wacc = 0;  // at most CAP / 64 pairs of at most 218 moves per lane
            for (u32 j = lane; j < wn; j += 64) {
                u16 x = slm[w][j];
                u32 src = sll[w][j];
                QuadBitBoard c = sb[w][src];
                u8 s2 = ss[w][src];
                makeMove<1 - C>(c, s2, x);
                wacc += countMovesSL<C>(c, s2, sl);
            }
            acc += wacc;
            waveLdsSync();
        }
    }
    u64 s = waveSum64(acc);
    if (lane == 0 && s) atomicAdd((unsigned long long *)sum, (unsigned long long)s);
}

// the queue-fed leaf with the table-lookup counter (pg_set_slider_attacks)
template <int C, int SLM>
__global__ __launch_bounds__(L2_WAVES * 64) L2_OCCUPANCY void k_leaf2q_magic(const QuadBitBoard *__restrict__ pb,
                                                                const u8 *__restrict__ ps,
                                                                const u32 *__restrict__ pidx,
                                                                const u16 *__restrict__ mv, u64 n,
                                                                u64 *__restrict__ sum, u64 *work)
{
    leaf2Body<C, L2_CHUNK, L2_CAP_MAGIC, true, SLM>(pb, ps, pidx, mv, n, sum, work);
}

// Two shapes of the same leaf: 64-entry chunks (default) and 32-entry chunks
// with a 1,536-pair list (half the child phase's lanes idle; measured 8 %
// slower on the config-2 tree).  pg_set_leaf2_variant picks.
template <int C>
__global__ __launch_bounds__(L2_WAVES * 64) L2_OCCUPANCY void k_leaf2(const QuadBitBoard *__restrict__ pb,
                                                         const u8 *__restrict__ ps, const u32 *__restrict__ pidx,
                                                         const u16 *__restrict__ mv, u64 n, u64 *__restrict__ sum)
{
    leaf2Body<C, L2_CHUNK, L2_CAP, false>(pb, ps, pidx, mv, n, sum, nullptr);
}
template <int C>
__global__ __launch_bounds__(L2_WAVES * 64) L2_OCCUPANCY void k_leaf2q(const QuadBitBoard *__restrict__ pb,
                                                          const u8 *__restrict__ ps, const u32 *__restrict__ pidx,
                                                          const u16 *__restrict__ mv, u64 n, u64 *__restrict__ sum,
                                                          u64 *work)
{
    leaf2Body<C, L2_CHUNK, L2_CAP, true>(pb, ps, pidx, mv, n, sum, work);
}
template <int C>
__global__ __launch_bounds__(L2_WAVES * 64) L2_OCCUPANCY void k_leaf2_c32(const QuadBitBoard *__restrict__ pb,
                                                             const u8 *__restrict__ ps, const u32 *__restrict__ pidx,
                                                             const u16 *__restrict__ mv, u64 n,
                                                             u64 *__restrict__ sum)
{
    leaf2Body<C, 32, 1536, false>(pb, ps, pidx, mv, n, sum, nullptr);
}
int g_leaf2Variant = 0;
// grid of the grid-stride leaves, per CU (x2 for the one-ply leaf): with four
// resident two-ply blocks per CU, 16 (four rounds) took the tables' startpos
// perft 11 from 9.78 to 9.52 s and config 4 from 2.21 to 2.16 s against 8
int g_leafBlocksPerCU = 16;

// ------------------------------------------------------------ TT-mode kernels
// With a transposition table each expanded level probes the table at its
// remaining depth (a hit is a known subtree: no children are generated), the
// leaf returns per-entry values, and an upsweep sums every parent's children
// (a contiguous range, so no atomics) and stores the sums in the table
// (reference: makemove_and_count_moves_kernel<useTT>, reduce_by_parent,
// add_tt_hits_and_store, perft_kernels.cu:201-257, 554-599).

// make + table probe (TT mode): a hit is a known subtree.  Grid-stride over a
// capped grid, so that each wave adds its counter tallies once.
template <int C>
__global__ __launch_bounds__(BS) void k_make_probe(const QuadBitBoard *__restrict__ pb, const u8 *__restrict__ ps,
                                                   const u32 *__restrict__ pidx, const u16 *__restrict__ mv, u64 n,
                                                   QuadBitBoard *__restrict__ ob, u8 *__restrict__ os,
                                                   u64 *__restrict__ known, const TTEntry *__restrict__ tt, u64 mask,
                                                   int rd, u64 *ctr)
{
    TTTally tally;
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < n; i += (u64)gridDim.x * BS) {
        u32 p = pidx[i];
        QuadBitBoard q = pb[p];
        u8 st = ps[p];
        makeMove<C>(q, st, mv[i]);
        ob[i] = q;
        os[i] = st;
        u64 v;
        bool hit = ttProbe(tt, mask, hashPosition(q, st, 1 - C), rd, &v);
        known[i] = hit ? v + 1 : 0;
        tally.add(true, hit, false);
    }
    ttFlush(ctr, tally);
}

// Within-level merging (the table and the comparison are in pg_tt.h)
template <int C>  // side to move in these boards
__global__ __launch_bounds__(BS) void k_dedup_write(const QuadBitBoard *__restrict__ b, const u8 *__restrict__ s,
                                                    const u64 *__restrict__ known, u64 n, DedupEntry *__restrict__ tab,
                                                    u64 tmask)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    if (i >= n || known[i]) return;
    Hash128 h = hashPosition(b[i], s[i], C);
    tab[dedupSlot(h, tmask)] = DedupEntry{h.hi, i + 1};
}

template <int C>
__global__ __launch_bounds__(BS) void k_dedup_count(const QuadBitBoard *__restrict__ b, const u8 *__restrict__ s,
                                                    const u64 *__restrict__ known, u64 n,
                                                    const DedupEntry *__restrict__ tab, u64 tmask,
                                                    u32 *__restrict__ redir, u32 *__restrict__ cnt,
                                                    u32 *__restrict__ tileSums)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    u32 c = 0;
    if (i < n) {
        u32 r = (u32)i;
        if (!known[i]) {
            QuadBitBoard q = b[i];
            u8 st = s[i];
            r = dedupTarget(b, s, n, tab, tmask, i, q, st, hashPosition(q, st, C));
            if (r == (u32)i) c = countChildren<C>(q, st);
        }
        redir[i] = r;
        cnt[i] = c;
    }
    u64 t = blockSum64(c);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = (u32)t;
}

__global__ __launch_bounds__(BS) void k_apply_redirects(const u32 *__restrict__ redir, u64 n, u64 *__restrict__ val)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    if (i < n && redir[i] != (u32)i) val[i] = val[redir[i]];
}

// one-ply leaf with per-entry values
template <int C>
__global__ __launch_bounds__(BS) void k_leaf1_val(const QuadBitBoard *__restrict__ pb, const u8 *__restrict__ ps,
                                                  const u32 *__restrict__ pidx, const u16 *__restrict__ mv, u64 n,
                                                  u64 *__restrict__ val)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    u32 p = pidx[i];
    QuadBitBoard q = pb[p];
    u8 st = ps[p];
    makeMove<C>(q, st, mv[i]);
    val[i] = countMoves<1 - C>(q, st);
}

// two-ply leaf with per-entry values and the depth-2 table (the reference's
// leaf TT, perft_kernels.cu:867-877): wave-cooperative like k_leaf2, with each
// entry's partial sum kept in LDS.  Each wave visits chunks by a static stride
// over a grid of up to 16 blocks per CU, through a fixed permutation of the
// level's chunks, so the chunks in flight together are spread over the level.
// In-order chunks put the copies of a position that nearby parents share
// (A-then-B and B-then-A) in flight together, where none of them finds the
// other's stored count: the chunk queue served 9 % fewer requests/s than the
// plain stride, and the permutation took startpos perft 10 with the tables
// from 0.91 to 0.84 s.
template <int C>
__global__ __launch_bounds__(L2_WAVES * 64) L2_OCCUPANCY void k_leaf2_val(const QuadBitBoard *__restrict__ pb,
                                                             const u8 *__restrict__ ps, const u32 *__restrict__ pidx,
                                                             const u16 *__restrict__ mv, u64 n, u64 *__restrict__ val,
                                                             TTEntry *__restrict__ tt, u64 mask, u64 *ctr)
{
    __shared__ QuadBitBoard sb[L2_WAVES][L2_CHUNK];
    __shared__ u8 ss[L2_WAVES][L2_CHUNK];
    __shared__ u32 sacc[L2_WAVES][L2_CHUNK];
    __shared__ u16 slm[L2_WAVES][L2_CAP_TT + 1];  // slot L2_CAP_TT: moves outside the window
    __shared__ u8 sll[L2_WAVES][L2_CAP_TT + 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    TTTally tally;
    // chunk k of each wave's sequence is chunk k * P mod nChunks of the level
    // (P a prime above any chunk count): the chunks in flight together are
    // spread over the whole level
    const u64 nChunks = (n + L2_CHUNK - 1) / L2_CHUNK;
    const u64 P = 1000000007ULL % (nChunks ? nChunks : 1);
    for (u64 k = (u64)blockIdx.x * L2_WAVES + w; k < nChunks; k += (u64)gridDim.x * L2_WAVES) {
        const u64 base = ((k * P) % nChunks) * L2_CHUNK;
        u32 cnt = 0;
        u64 e = base + lane, hv = 0;
        bool valid = lane < L2_CHUNK && e < n, hit = false;
        BB catt;  // the child's enemy attacks, from its count, for its generation
        Hash128 h{};  // for the probe and, on a miss, the store
        if (valid) {
            u32 p = pidx[e];
            QuadBitBoard q = pb[p];
            u8 st = ps[p];
            makeMove<C>(q, st, mv[e]);
            sb[w][lane] = q;
            ss[w][lane] = st;
            h = hashPosition(q, st, 1 - C);
            hit = tt && ttProbe(tt, mask, h, 2, &hv);
            if (!hit) {
                Frame fr;
                cnt = countChildrenFrame<1 - C>(q, st, FillSliders(), fr);
                catt = fr.att;
            }
        }
        if (lane < L2_CHUNK) sacc[w][lane] = 0;
        u32 incl = waveInclScan(cnt);
        u32 off = incl - cnt;
        u32 total = __shfl(incl, 63);
        waveLdsSync();
        for (u32 wb = 0; wb < total; wb += L2_CAP_TT) {
            if (cnt && off < wb + L2_CAP_TT && off + cnt > wb) {
                QuadBitBoard q = sb[w][lane];
                u8 st = ss[w][lane];
                u16 *dm = slm[w];
                u8 *dl = sll[w];
                Frame fr = framePins<1 - C>(q);
                fr.att = catt;
                generateMovesF<1 - C>(fr, st, [&](u32 k, u16 m) {
                    u32 j = off + k - wb;
                    j = j < (u32)L2_CAP_TT ? j : (u32)L2_CAP_TT;
                    dm[j] = m, dl[j] = (u8)lane;
                });
            }
            waveLdsSync();
            u32 wn = total - wb < (u32)L2_CAP_TT ? total - wb : (u32)L2_CAP_TT;
            for (u32 j = lane; j < wn; j += 64) {
                u16 x = slm[w][j];
                u32 src = sll[w][j];
                QuadBitBoard c = sb[w][src];
                u8 s2 = ss[w][src];
                makeMove<1 - C>(c, s2, x);
                atomicAdd(&sacc[w][src], countMoves<C>(c, s2));
            }
            waveLdsSync();
        }
        bool store = false;
        if (valid) {
            u64 v = hit ? hv : sacc[w][lane];
            val[e] = v;
            if (tt && !hit) {
                ttStore(tt, mask, h, 2, v);
                store = true;
            }
        }
        tally.add(valid && tt, hit, store);
        waveLdsSync();
    }
    if (tt) ttFlush(ctr, tally);
}

// upsweep of one level: value = table hit, or the sum of the children's
// values; misses are stored at this level's remaining depth
template <int C>  // side to move in this level's boards
__global__ __launch_bounds__(BS) void k_upsweep(const QuadBitBoard *__restrict__ boards, const u8 *__restrict__ states,
                                                const u32 *__restrict__ incl, const u64 *__restrict__ known,
                                                const u32 *__restrict__ redir, const u64 *__restrict__ childVal,
                                                u64 n, u64 *__restrict__ val, TTEntry *__restrict__ tt, u64 mask,
                                                int rd, u64 *ctr)
{
    TTTally tally;
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < n; i += (u64)gridDim.x * BS) {
        if (redir[i] != (u32)i) continue;  // duplicates are filled in afterwards
        u64 v;
        if (known[i]) {
            v = known[i] - 1;
        } else {
            v = 0;
            u64 b = i ? incl[i - 1] : 0, e = incl[i];
            for (u64 k = b; k < e; k++) v += childVal[k];
            if (tt) {
                ttStore(tt, mask, hashPosition(boards[i], states[i], C), rd, v);
                tally.add(false, false, true);
            }
        }
        val[i] = v;
    }
    if (tt) ttFlush(ctr, tally);
}

// one u64 of device memory to a pinned host slot (the count at the end of a
// call, for the host to poll instead of a readback copy and a synchronise)
__global__ void k_publish_u64(const u64 *__restrict__ src, u64 *slot)
{
    if (threadIdx.x == 0) publishToHost(slot, *src);
}

__global__ __launch_bounds__(BS) void k_sum_u64(const u64 *__restrict__ v, u64 n, u64 *__restrict__ sum)
{
    u64 i = (u64)blockIdx.x * BS + threadIdx.x;
    u64 s = blockSum64(i < n ? v[i] : 0);
    if (threadIdx.x == 0 && s) atomicAdd((unsigned long long *)sum, (unsigned long long)s);
}

// ------------------------------------------------------------ launch helpers
static inline unsigned gridFor(u64 n, int per) { return (unsigned)((n + per - 1) / per); }

unsigned residentGrid(const void *fn, int blockSize)
{
    static std::mutex m;
    static std::unordered_map<const void *, unsigned> cache;
    std::lock_guard<std::mutex> g(m);
    auto it = cache.find(fn);
    if (it != cache.end()) return it->second;
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, blockSize, 0) != hipSuccess || perCU < 1) perCU = 1;
    unsigned grid = (unsigned)perCU * (unsigned)(lib().cuCount ? lib().cuCount : 256);
    cache.emplace(fn, grid);
    return grid;
}

// grid of the grid-stride TT kernels: at most 8 blocks per CU, so that each
// wave's counter tallies are added once for many entries
static unsigned tallyGrid(u64 n)
{
    u64 cap = (u64)(lib().cuCount ? lib().cuCount : 256) * 8;
    u64 g = (n + BS - 1) / BS;
    return (unsigned)(g < cap ? (g ? g : 1) : cap);
}

static int launchMakecount(const QuadBitBoard *pb, const u8 *ps, const u32 *pidx, const u16 *mv, int color, u64 n,
                           QuadBitBoard *ob, u8 *os, u32 *cnt, u32 *tileSums, hipStream_t st)
{
    if (!n) return PG_OK;
    if (color == WHITE)
        hipLaunchKernelGGL(k_makecount<WHITE>, dim3(gridFor(n, BS)), dim3(BS), 0, st, pb, ps, pidx, mv, n, ob, os, cnt,
                           tileSums);
    else
        hipLaunchKernelGGL(k_makecount<BLACK>, dim3(gridFor(n, BS)), dim3(BS), 0, st, pb, ps, pidx, mv, n, ob, os, cnt,
                           tileSums);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// scan: tileSums (per 256) -> group offsets -> in-place inclusive scan.
// Scan buffers of an n-entry level: group offsets (G + 1 u64, the total at
// [G]), in-group tile offsets (T u32) and group sums (G u32), T = ceil(n/256),
// G = ceil(T/64).  With `parts` (capacity expandPartsCap(n)) the scan also
// writes the next expand's partitions, for launchExpandParts.
struct ScanBufs {
    u64 *groupOffs;
    u32 *tileIn, *groupSums;
};
static u64 scanTiles(u64 n) { return (n + BS - 1) / BS; }
static u64 expandPartsCap(u64 n) { return (n * MAX_MOVES + EXPAND_TILE - 1) / EXPAND_TILE + 1; }
static bool allocScanBufs(Bump &A, u64 n, ScanBufs &b)
{
    const u64 T = scanTiles(n), G = scanGroupCount(T);
    b.groupOffs = A.alloc<u64>(G + 1);
    b.tileIn = A.alloc<u32>(T);
    b.groupSums = A.alloc<u32>(G);
    return b.groupOffs && b.tileIn && b.groupSums;
}
// With `hostTotal` (a slot of a fine-grained pinned buffer) the total is
// published there as soon as it is known, before the apply pass runs.
static int launchScan(u32 *d, u64 n, const u32 *tileSums, const ScanBufs &b, hipStream_t st, u32 *parts = nullptr,
                      u64 *hostTotal = nullptr)
{
    if (!n) return PG_OK;
    const u64 T = scanTiles(n), G = scanGroupCount(T);
    if (G <= (u64)SCAN_GROUP) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(TOP_BS), 0, st, tileSums, T, b.tileIn, b.groupOffs,
                           hostTotal);
    } else {
        hipLaunchKernelGGL(k_scan_groups, dim3((unsigned)((G + 15) / 16)), dim3(BS), 0, st, tileSums, T, b.tileIn,
                           b.groupSums);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TOP_BS), 0, st, b.groupSums, G, b.groupOffs, hostTotal);
    }
    hipLaunchKernelGGL(k_scan_apply, dim3(gridFor(n, 4 * BS)), dim3(BS), 0, st, d, n, b.tileIn, b.groupOffs, parts,
                       parts ? expandPartsCap(n) : 0ull);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

static int launchExpand(const u32 *incl, u64 n, u64 c, u32 *parts, u32 *parentIdx, hipStream_t st)
{
    if (!c) return PG_OK;
    u64 tiles = (c + EXPAND_TILE - 1) / EXPAND_TILE;
    hipLaunchKernelGGL(k_partitions, dim3(gridFor(tiles + 1, BS)), dim3(BS), 0, st, incl, n, c, tiles, parts);
    hipLaunchKernelGGL(k_expand, dim3((unsigned)tiles), dim3(BS), 0, st, incl, n, c, parts, parentIdx);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// k_expand on partitions the scan wrote
static int launchExpandParts(const u32 *incl, u64 n, u64 c, const u32 *parts, u32 *parentIdx, hipStream_t st)
{
    if (!c) return PG_OK;
    u64 tiles = (c + EXPAND_TILE - 1) / EXPAND_TILE;
    hipLaunchKernelGGL(k_expand, dim3((unsigned)tiles), dim3(BS), 0, st, incl, n, c, parts, parentIdx);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// Expand (on partitions the scan wrote) and genmoves of a level in one launch,
// for levels of at most FUSE_MAX_CHILDREN children; above, the two run as
// separate launches.  The fused launch saves a launch on small levels
// (Kiwipete perft 5 0.318 -> 0.315 ms, startpos perft 6 0.267 -> 0.263 ms); on
// the largest levels the separate expand keeps its own occupancy (24 VGPRs,
// 8 KB of LDS): fused throughout, config-2 perft 7 with the one-ply leaf took
// 7.24 ms against 6.98 ms.
constexpr u64 FUSE_MAX_CHILDREN = 1ull << 21;
static int launchExpandGenmoves(const QuadBitBoard *boards, const u8 *states, const u32 *incl, int color, u64 n,
                                u64 c, const u32 *parts, u32 *parentIdx, u16 *out, hipStream_t st)
{
    if (!n || !c) return PG_OK;
    const u64 tiles = (c + EXPAND_TILE - 1) / EXPAND_TILE, gen = gridFor(n, BS);
    const dim3 g((unsigned)(gen + tiles)), b(BS);
    if (color == WHITE)
        hipLaunchKernelGGL(k_expand_genmoves<WHITE>, g, b, 0, st, boards, states, incl, n, out, c, parts, parentIdx,
                           gen, tiles);
    else
        hipLaunchKernelGGL(k_expand_genmoves<BLACK>, g, b, 0, st, boards, states, incl, n, out, c, parts, parentIdx,
                           gen, tiles);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

static int launchGenmoves(const QuadBitBoard *boards, const u8 *states, const u32 *incl, int color, u64 n, u16 *out,
                          hipStream_t st)
{
    if (!n) return PG_OK;
    if (color == WHITE)
        hipLaunchKernelGGL(k_genmoves<WHITE>, dim3(gridFor(n, BS)), dim3(BS), 0, st, boards, states, incl, n, out);
    else
        hipLaunchKernelGGL(k_genmoves<BLACK>, dim3(gridFor(n, BS)), dim3(BS), 0, st, boards, states, incl, n, out);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// `work`: a zeroed device u64 for the fused leaf's chunk queue, or null for
// the static-stride form
static int launchLeaf(const QuadBitBoard *pb, const u8 *ps, const u32 *pidx, const u16 *mv, int color, u64 n,
                      int plies, u64 *sum, hipStream_t st, u64 *work = nullptr)
{
    if (!n) return PG_OK;
    // grid-stride leaves: g_leafBlocksPerCU blocks per CU (several rounds of
    // resident blocks; short per-block shares keep the last round's tail
    // small), smaller when the work is smaller
    const u64 cus = (u64)(lib().cuCount ? lib().cuCount : 256);
    if (plies == 1) {
        u64 blocks = (n + BS - 1) / BS;
        u64 cap = cus * 2 * g_leafBlocksPerCU;
        dim3 g((unsigned)(blocks < cap ? blocks : cap)), b(BS);
        if (color == WHITE) hipLaunchKernelGGL(k_leaf1<WHITE>, g, b, 0, st, pb, ps, pidx, mv, n, sum);
        else hipLaunchKernelGGL(k_leaf1<BLACK>, g, b, 0, st, pb, ps, pidx, mv, n, sum);
    } else {
        const u64 chunk = g_leaf2Variant == 1 ? 32 : L2_CHUNK;
        u64 chunks = (n + chunk - 1) / chunk;
        u64 blocks = (chunks + L2_WAVES - 1) / L2_WAVES;
        u64 cap = cus * g_leafBlocksPerCU;
        dim3 g((unsigned)(blocks < cap ? blocks : cap)), b(L2_WAVES * 64);
        if (g_leaf2Variant == 1) {
            if (color == WHITE) hipLaunchKernelGGL(k_leaf2_c32<WHITE>, g, b, 0, st, pb, ps, pidx, mv, n, sum);
            else hipLaunchKernelGGL(k_leaf2_c32<BLACK>, g, b, 0, st, pb, ps, pidx, mv, n, sum);
        } else if (work && g_sliderMode != SL_FILLS) {
            const void *fn = g_sliderMode == SL_MAGIC
                                 ? (color == WHITE ? (const void *)k_leaf2q_magic<WHITE, SL_MAGIC>
                                                   : (const void *)k_leaf2q_magic<BLACK, SL_MAGIC>)
                                 : (color == WHITE ? (const void *)k_leaf2q_magic<WHITE, SL_MAGIC_LDS>
                                                   : (const void *)k_leaf2q_magic<BLACK, SL_MAGIC_LDS>);
            u64 r = residentGrid(fn, L2_WAVES * 64);
            dim3 gq((unsigned)(blocks < r ? blocks : r));
            if (g_sliderMode == SL_MAGIC) {
                if (color == WHITE) hipLaunchKernelGGL((k_leaf2q_magic<WHITE, SL_MAGIC>), gq, b, 0, st, pb, ps, pidx, mv, n, sum, work);
                else hipLaunchKernelGGL((k_leaf2q_magic<BLACK, SL_MAGIC>), gq, b, 0, st, pb, ps, pidx, mv, n, sum, work);
            } else {
                if (color == WHITE) hipLaunchKernelGGL((k_leaf2q_magic<WHITE, SL_MAGIC_LDS>), gq, b, 0, st, pb, ps, pidx, mv, n, sum, work);
                else hipLaunchKernelGGL((k_leaf2q_magic<BLACK, SL_MAGIC_LDS>), gq, b, 0, st, pb, ps, pidx, mv, n, sum, work);
            }
        } else if (work) {
            u64 r = residentGrid((const void *)(color == WHITE ? k_leaf2q<WHITE> : k_leaf2q<BLACK>), L2_WAVES * 64);
            dim3 gq((unsigned)(blocks < r ? blocks : r));
            if (color == WHITE) hipLaunchKernelGGL(k_leaf2q<WHITE>, gq, b, 0, st, pb, ps, pidx, mv, n, sum, work);
            else hipLaunchKernelGGL(k_leaf2q<BLACK>, gq, b, 0, st, pb, ps, pidx, mv, n, sum, work);
        } else {
            if (color == WHITE) hipLaunchKernelGGL(k_leaf2<WHITE>, g, b, 0, st, pb, ps, pidx, mv, n, sum);
            else hipLaunchKernelGGL(k_leaf2<BLACK>, g, b, 0, st, pb, ps, pidx, mv, n, sum);
        }
    }
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// --------------------------------------------------------- TT BFS driver
// Same level loop as bfsPerft with table probes per level, per-entry leaf
// values and an upsweep that stores every computed subtree.
struct TTLevel {
    const QuadBitBoard *boards;
    const u8 *states;
    u32 *incl;
    u64 *known;
    u32 *redir;
    u64 *val;
    u64 n;
};

// A call's inputs in one host-to-device copy, staged in the stream's pinned
// buffer: the root board, its state, `nz` zeroed u64 counters, and the first
// level's parent indices (all the root, 0) and root moves.  The device block
// is [board 32 B | state, padded to 8 | nz x 8 B | n1 x 4 B | n1 x 2 B].
struct RootBlock {
    QuadBitBoard *b0;
    u8 *s0;
    u64 *zeros;
    u32 *idx;
    u16 *mv;
};
static int uploadRoot(Bump &A, const QuadBitBoard &root, u8 state, const u16 *moves, u32 n1, int nz, u64 *pinned,
                      hipStream_t st, RootBlock &r)
{
    const size_t offZ = 40, offIdx = offZ + 8 * (size_t)nz, offMv = offIdx + 4 * (size_t)n1,
                 bytes = offMv + 2 * (size_t)n1;
    static_assert(PIN_STAGE + 40 + 8 * 4 + 6 * MAX_MOVES <= PINNED_BYTES, "staging area too small");
    if (n1 > MAX_MOVES || nz > 4) {
        setError("uploadRoot: %u root moves, %d counters", n1, nz);
        return PG_ERR_INVALID;
    }
    char *d = A.alloc<char>(bytes);
    if (!d) return PG_ERR_ARENA_OOM;
    char *h = (char *)pinned + PIN_STAGE;
    memcpy(h, &root, sizeof(QuadBitBoard));
    memset(h + 32, 0, offMv - 32);  // the state's padding, the counters and the parent indices
    h[32] = (char)state;
    memcpy(h + offMv, moves, 2 * (size_t)n1);
    PG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    r = {(QuadBitBoard *)d, (u8 *)(d + 32), (u64 *)(d + offZ), (u32 *)(d + offIdx), (u16 *)(d + offMv)};
    return PG_OK;
}

static int bfsPerftTT(const QuadBitBoard &root, u8 state, int color, int depth, int leafPlies, const DeviceTT &T,
                      Bump &A, hipStream_t st, u64 *pinned, uint64_t *count)
{
    Lib &L = lib();
    TTEntry *tt = (TTEntry *)T.entries;
    u64 mask = T.mask;
    u16 rootMoves[MAX_MOVES];
    u32 n1 = generateMoves(root, state, color, rootMoves);
    if (n1 == 0) {
        *count = 0;
        return PG_OK;
    }
    RootBlock rb;
    PG_TRY(uploadRoot(A, root, state, rootMoves, n1, 4, pinned, st, rb));  // [leaf sum, probes, hits, stores]
    const QuadBitBoard *b0 = rb.b0;
    const u8 *s0 = rb.s0;
    u64 *dSum = rb.zeros, *ctr = rb.zeros + 1;
    u32 *idx = rb.idx;
    u16 *mv = rb.mv;
    std::vector<TTLevel> lv;
    const QuadBitBoard *pb = b0;
    const u8 *ps = s0;
    u64 n = n1;
    int last = depth - leafPlies;
    for (int lvl = 1; lvl < last; lvl++) {
        int pc = color ^ ((lvl - 1) & 1);
        u64 tiles = (n + BS - 1) / BS;
        QuadBitBoard *ob = A.alloc<QuadBitBoard>(n);
        u8 *os = A.alloc<u8>(n);
        u32 *cnt = A.alloc<u32>(n);
        u64 *known = A.alloc<u64>(n);
        u64 *val = A.alloc<u64>(n);
        u32 *tileSums = A.alloc<u32>(tiles);
        ScanBufs sb;
        const bool sbOk = allocScanBufs(A, n, sb);
        u32 *redir = A.alloc<u32>(n);
        u32 *parts = A.alloc<u32>(expandPartsCap(n));
        u64 tslots = 1;
        while (tslots < 2 * n) tslots <<= 1;
        DedupEntry *tab = A.alloc<DedupEntry>(tslots);
        if (!ob || !os || !cnt || !known || !val || !tileSums || !sbOk || !redir || !parts || !tab)
            return PG_ERR_ARENA_OOM;
        dim3 g(gridFor(n, BS)), gc(tallyGrid(n)), b(BS);
        int rd = depth - lvl;
        // the merge table needs no clearing (every slot read is written first)
        if (pc == WHITE) {
            hipLaunchKernelGGL(k_make_probe<WHITE>, gc, b, 0, st, pb, ps, idx, mv, n, ob, os, known, tt, mask, rd, ctr);
            hipLaunchKernelGGL(k_dedup_write<BLACK>, g, b, 0, st, ob, os, known, n, tab, tslots - 1);
            hipLaunchKernelGGL(k_dedup_count<BLACK>, g, b, 0, st, ob, os, known, n, tab, tslots - 1, redir, cnt,
                               tileSums);
        } else {
            hipLaunchKernelGGL(k_make_probe<BLACK>, gc, b, 0, st, pb, ps, idx, mv, n, ob, os, known, tt, mask, rd, ctr);
            hipLaunchKernelGGL(k_dedup_write<WHITE>, g, b, 0, st, ob, os, known, n, tab, tslots - 1);
            hipLaunchKernelGGL(k_dedup_count<WHITE>, g, b, 0, st, ob, os, known, n, tab, tslots - 1, redir, cnt,
                               tileSums);
        }
        PG_HIP(hipGetLastError());
        volatile u64 *slot = pinned + PIN_POLL;  // the level total, published by the scan
        *slot = HOST_SLOT_EMPTY;
        PG_TRY(launchScan(cnt, n, tileSums, sb, st, parts, pinned + PIN_POLL));
        u64 c;
        PG_TRY(waitPublished(slot, st, c));
        lv.push_back({ob, os, cnt, known, redir, val, n});
        if (c == 0) {
            // every entry was a table hit, a duplicate or terminal: nothing below
            n = 0;
            break;
        }
        if (c >= 0xFFFFFFFFull) return PG_ERR_ARENA_OOM;
        u32 *nidx = A.alloc<u32>(c);
        u16 *nmv = A.alloc<u16>(c);
        if (!nidx || !nmv) return PG_ERR_ARENA_OOM;
        if (c > FUSE_MAX_CHILDREN) {
            PG_TRY(launchExpandParts(cnt, n, c, parts, nidx, st));
            PG_TRY(launchGenmoves(ob, os, cnt, pc ^ 1, n, nmv, st));
        } else if (c) {
            PG_TRY(launchExpandGenmoves(ob, os, cnt, pc ^ 1, n, c, parts, nidx, nmv, st));
        }
        pb = ob;
        ps = os;
        idx = nidx;
        mv = nmv;
        n = c;
    }
    // leaf values (none when a level above came out empty)
    int pc = color ^ ((last - 1) & 1);
    u64 *leafVal = A.alloc<u64>(n);
    if (!leafVal) return PG_ERR_ARENA_OOM;
    if (n && (int)lv.size() == last - 1) {
        if (leafPlies == 1) {
            dim3 g(gridFor(n, BS)), b(BS);
            if (pc == WHITE) hipLaunchKernelGGL(k_leaf1_val<WHITE>, g, b, 0, st, pb, ps, idx, mv, n, leafVal);
            else hipLaunchKernelGGL(k_leaf1_val<BLACK>, g, b, 0, st, pb, ps, idx, mv, n, leafVal);
        } else {
            u64 chunks = (n + L2_CHUNK - 1) / L2_CHUNK;
            u64 blocks = (chunks + L2_WAVES - 1) / L2_WAVES;
            u64 cap = (u64)(L.cuCount ? L.cuCount : 256) * g_leafBlocksPerCU;
            dim3 g((unsigned)(blocks < cap ? blocks : cap)), b(L2_WAVES * 64);
            if (pc == WHITE) hipLaunchKernelGGL(k_leaf2_val<WHITE>, g, b, 0, st, pb, ps, idx, mv, n, leafVal, tt, mask, ctr);
            else hipLaunchKernelGGL(k_leaf2_val<BLACK>, g, b, 0, st, pb, ps, idx, mv, n, leafVal, tt, mask, ctr);
        }
        PG_HIP(hipGetLastError());
    }
    // upsweep: the deepest built level down to 1
    const u64 *childVal = leafVal;
    for (int lvl = (int)lv.size(); lvl >= 1; lvl--) {
        TTLevel &l = lv[lvl - 1];
        int side = color ^ (lvl & 1);
        dim3 g(gridFor(l.n, BS)), gc(tallyGrid(l.n)), b(BS);
        if (side == WHITE)
            hipLaunchKernelGGL(k_upsweep<WHITE>, gc, b, 0, st, l.boards, l.states, l.incl, l.known, l.redir, childVal,
                               l.n, l.val, tt, mask, depth - lvl, ctr);
        else
            hipLaunchKernelGGL(k_upsweep<BLACK>, gc, b, 0, st, l.boards, l.states, l.incl, l.known, l.redir, childVal,
                               l.n, l.val, tt, mask, depth - lvl, ctr);
        hipLaunchKernelGGL(k_apply_redirects, g, b, 0, st, l.redir, l.n, l.val);
        childVal = l.val;
    }
    hipLaunchKernelGGL(k_sum_u64, dim3(gridFor(n1, BS)), dim3(BS), 0, st, childVal, (u64)n1, dSum);
    PG_HIP(hipGetLastError());
    PG_HIP(hipMemcpyAsync(pinned, dSum, 4 * sizeof(u64), hipMemcpyDeviceToHost, st));  // the sum, then ctr[0..2]
    PG_HIP(hipStreamSynchronize(st));
    *count = pinned[0];
    ttCountDevice(pinned[1], pinned[2], pinned[3]);
    return PG_OK;
}

// --------------------------------------------------------------- BFS driver
// perft_gpu_host_bfs (perft_kernels.cu:620-990): root moves on the host, then
// levels 1 .. depth-leafPlies-1 on the device, the host learning each next
// level's size from the total the scan publishes to a pinned slot (the
// reference reads it back with a copy), then the leaf.  Returns
// PG_ERR_ARENA_OOM (and no count) when a level does not fit; the caller
// recomputes one ply down.
static int bfsPerftLevels(const QuadBitBoard &root, u8 state, int color, int depth, int leafPlies, bool useTT,
                          void *arena, size_t arenaBytes, hipStream_t st, u64 *pinned, uint64_t *count,
                          size_t *peakBytes)
{
    if (depth <= 0) {
        *count = 1;
        return PG_OK;
    }
    if (depth == 1) {
        *count = countMoves(root, state, color);
        return PG_OK;
    }
    if (leafPlies < 1) leafPlies = 2;
    if (leafPlies > depth - 1) leafPlies = depth - 1;
    if (leafPlies > 2) leafPlies = 2;

    u16 rootMoves[MAX_MOVES];
    u32 n1 = generateMoves(root, state, color, rootMoves);
    if (n1 == 0) {
        *count = 0;
        return PG_OK;
    }
    Bump A(arena, arenaBytes);
    if (useTT && lib().tt.entries) {
        int s = bfsPerftTT(root, state, color, depth, leafPlies, lib().tt, A, st, pinned, count);
        if (peakBytes) *peakBytes = A.peak;
        return s;
    }
    Lib &L = lib();
    Profiler &P = L.prof;
    const int last = depth - leafPlies;  // level whose entries feed the leaf
    const QuadBitBoard *pb;
    const u8 *ps;
    u32 *idx;
    u16 *mv;
    u64 *dSum;  // [leaf sum, leaf chunk queue]
    u64 n = n1, nParents = 1;
    int lvl = 1;
    if (last > 1) {
        // level 1 from kernel arguments: no upload before the first kernel
        dSum = A.alloc<u64>(2);
        QuadBitBoard *ob = A.alloc<QuadBitBoard>(n1);
        u8 *os = A.alloc<u8>(n1);
        u32 *nidx = A.alloc<u32>((u64)n1 * MAX_MOVES);
        u16 *nmv = A.alloc<u16>((u64)n1 * MAX_MOVES);
        if (!dSum || !ob || !os || !nidx || !nmv) return PG_ERR_ARENA_OOM;
        RootArgs ra{};
        ra.board = root;
        ra.n = n1;
        ra.state = state;
        memcpy(ra.moves, rootMoves, sizeof(u16) * n1);
        volatile u64 *slot = pinned + PIN_POLL;
        *slot = HOST_SLOT_EMPTY;
        if (color == WHITE)
            hipLaunchKernelGGL(k_level_root<WHITE>, dim3(1), dim3(SMALL_LEVEL), 0, st, ra, ob, os, nidx, nmv,
                               pinned + PIN_POLL, dSum);
        else
            hipLaunchKernelGGL(k_level_root<BLACK>, dim3(1), dim3(SMALL_LEVEL), 0, st, ra, ob, os, nidx, nmv,
                               pinned + PIN_POLL, dSum);
        PG_HIP(hipGetLastError());
        u64 c;
        PG_TRY(waitPublished(slot, st, c));
        if (c == 0) {
            PG_HIP(hipStreamSynchronize(st));  // the block may still be running
            *count = 0;
            if (peakBytes) *peakBytes = A.peak;
            P.collect();
            return PG_OK;
        }
        pb = ob;
        ps = os;
        idx = nidx;
        mv = nmv;
        nParents = n1;
        n = c;
        lvl = 2;
    } else {
        RootBlock rb;
        PG_TRY(uploadRoot(A, root, state, rootMoves, n1, 2, pinned, st, rb));
        pb = rb.b0;
        ps = rb.s0;
        dSum = rb.zeros;
        idx = rb.idx;
        mv = rb.mv;
    }
    for (; lvl < last; lvl++) {
        int pc = color ^ ((lvl - 1) & 1);  // side to move in the parent boards
        if (n <= SMALL_LEVEL) {
            // one block does the whole level; the children's arrays take
            // the bound n * MAX_MOVES (at most 1.5 MB)
            QuadBitBoard *ob = A.alloc<QuadBitBoard>(n);
            u8 *os = A.alloc<u8>(n);
            u32 *nidx = A.alloc<u32>(n * MAX_MOVES);
            u16 *nmv = A.alloc<u16>(n * MAX_MOVES);
            if (!ob || !os || !nidx || !nmv) return PG_ERR_ARENA_OOM;
            volatile u64 *slot = pinned + PIN_POLL;
            *slot = HOST_SLOT_EMPTY;
            if (pc == WHITE)
                hipLaunchKernelGGL(k_level_small<WHITE>, dim3(1), dim3(SMALL_LEVEL), 0, st, pb, ps, idx, mv, (u32)n,
                                   ob, os, nidx, nmv, pinned + PIN_POLL);
            else
                hipLaunchKernelGGL(k_level_small<BLACK>, dim3(1), dim3(SMALL_LEVEL), 0, st, pb, ps, idx, mv, (u32)n,
                                   ob, os, nidx, nmv, pinned + PIN_POLL);
            PG_HIP(hipGetLastError());
            u64 c;
            PG_TRY(waitPublished(slot, st, c));
            if (c == 0) {
                PG_HIP(hipStreamSynchronize(st));  // the block may still be running
                *count = 0;
                if (peakBytes) *peakBytes = A.peak;
                P.collect();
                return PG_OK;
            }
            pb = ob;
            ps = os;
            idx = nidx;
            mv = nmv;
            nParents = n;
            n = c;
            continue;
        }
        u64 tiles = (n + BS - 1) / BS;
        QuadBitBoard *ob = A.alloc<QuadBitBoard>(n);
        u8 *os = A.alloc<u8>(n);
        u32 *cnt = A.alloc<u32>(n);
        u32 *tileSums = A.alloc<u32>(tiles);
        ScanBufs sb;
        const bool sbOk = allocScanBufs(A, n, sb);
        u32 *parts = A.alloc<u32>(expandPartsCap(n));
        if (!ob || !os || !cnt || !tileSums || !sbOk || !parts) return PG_ERR_ARENA_OOM;
        // algorithmic bytes: 43N + 33P (make+count), 8N (scan)
        P.begin(K_MAKECOUNT, st);
        PG_TRY(launchMakecount(pb, ps, idx, mv, pc, n, ob, os, cnt, tileSums, st));
        P.end(43.0 * n + 33.0 * nParents, (double)n, st);
        // the scan publishes the level total to the pinned slot, which the
        // host polls (no readback copy, no stream synchronisation)
        volatile u64 *slot = pinned + PIN_POLL;
        *slot = HOST_SLOT_EMPTY;
        P.begin(K_SCAN, st);
        PG_TRY(launchScan(cnt, n, tileSums, sb, st, parts, pinned + PIN_POLL));
        P.end(8.0 * n, (double)n, st);
        u64 c;
        PG_TRY(waitPublished(slot, st, c));
        if (c == 0) {
            PG_HIP(hipStreamSynchronize(st));  // the apply pass may still run
            *count = 0;
            if (peakBytes) *peakBytes = A.peak;
            P.collect();
            return PG_OK;
        }
        if (c >= 0xFFFFFFFFull) return PG_ERR_ARENA_OOM;
        u32 *nidx = A.alloc<u32>(c);
        u16 *nmv = A.alloc<u16>(c);
        if (!nidx || !nmv) return PG_ERR_ARENA_OOM;
        // 4N + 4C (expand), 37N + 2C (genmoves); one launch for small levels
        if (c > FUSE_MAX_CHILDREN) {
            P.begin(K_EXPAND, st);
            PG_TRY(launchExpandParts(cnt, n, c, parts, nidx, st));
            P.end(4.0 * n + 4.0 * c, (double)c, st);
            P.begin(K_GENMOVES, st);
            PG_TRY(launchGenmoves(ob, os, cnt, pc ^ 1, n, nmv, st));
            P.end(37.0 * n + 2.0 * c, (double)n, st);
        } else {
            PG_TRY(launchExpandGenmoves(ob, os, cnt, pc ^ 1, n, c, parts, nidx, nmv, st));
        }
        pb = ob;
        ps = os;
        idx = nidx;
        mv = nmv;
        nParents = n;
        n = c;
    }
    int pc = color ^ ((last - 1) & 1);
    // leaf: 6K + 33P
    P.begin(leafPlies == 1 ? K_LEAF1 : K_LEAF2, st);
    PG_TRY(launchLeaf(pb, ps, idx, mv, pc, n, leafPlies, dSum, st, dSum + 1));
    P.end(6.0 * n + 33.0 * nParents, (double)n, st);
    volatile u64 *slot = pinned + PIN_POLL;
    *slot = HOST_SLOT_EMPTY;
    hipLaunchKernelGGL(k_publish_u64, dim3(1), dim3(64), 0, st, dSum, pinned + PIN_POLL);
    PG_HIP(hipGetLastError());
    u64 total;
    PG_TRY(waitPublished(slot, st, total));
    P.collect();
    *count = total;
    if (peakBytes) *peakBytes = A.peak;
    return PG_OK;
}

// A call that fails (or stops at a level with no children) may leave kernels
// queued on its buffer, which the caller may free or reuse next: such a
// return drains the stream first.  A complete call ends with the host having
// seen the last kernel's count, so nothing of it is left to run.
int bfsPerft(const QuadBitBoard &root, u8 state, int color, int depth, int leafPlies, bool useTT, void *arena,
             size_t arenaBytes, hipStream_t st, u64 *pinned, uint64_t *count, size_t *peakBytes)
{
    int s = bfsPerftLevels(root, state, color, depth, leafPlies, useTT, arena, arenaBytes, st, pinned, count,
                           peakBytes);
    if (s != PG_OK) (void)hipStreamSynchronize(st);
    return s;
}

}  // namespace pg

using namespace pg;

// ------------------------------------------------------- kernel-level C ABI
static hipStream_t toStream(void *s) { return s ? (hipStream_t)s : lib().stream; }

extern "C" int pg_level_makecount(const void *parentBoards, const uint8_t *parentStates, const uint32_t *parentIdx,
                                  const uint16_t *moves, int color, uint64_t n, void *boards, uint8_t *states,
                                  uint32_t *counts, uint32_t *tileSums, void *stream)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if ((color & ~1) || (n && (!parentBoards || !parentStates || !parentIdx || !moves || !boards || !states ||
                               !counts || !tileSums))) {
        setError("pg_level_makecount: bad argument");
        return PG_ERR_INVALID;
    }
    return launchMakecount((const QuadBitBoard *)parentBoards, parentStates, parentIdx, moves, color, n,
                           (QuadBitBoard *)boards, states, counts, tileSums, toStream(stream));
}

extern "C" int pg_scan_u32(uint32_t *data, uint64_t n, const uint32_t *tileSums, void *scratch, uint64_t *total,
                           void *stream)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if (n && (!data || !scratch)) {
        setError("pg_scan_u32: bad argument");
        return PG_ERR_INVALID;
    }
    hipStream_t st = toStream(stream);
    if (!n) {
        if (total) *total = 0;
        return PG_OK;
    }
    // scratch: group offsets (G + 1 u64), in-group tile offsets (T u32),
    // group sums (G u32), then the tile sums when the caller has none (T u32)
    const u64 T = scanTiles(n), G = scanGroupCount(T);
    ScanBufs sb;
    sb.groupOffs = (u64 *)scratch;
    sb.tileIn = (u32 *)(sb.groupOffs + G + 1);
    sb.groupSums = sb.tileIn + T;
    const u32 *ts = tileSums;
    if (!ts) {
        u32 *own = sb.groupSums + G;
        hipLaunchKernelGGL(k_tile_sums, dim3(gridFor(n, BS)), dim3(BS), 0, st, data, n, own);
        ts = own;
    }
    PG_TRY(launchScan(data, n, ts, sb, st));
    if (total) {
        PG_HIP(hipMemcpyAsync(lib().hostPinned, sb.groupOffs + G, sizeof(u64), hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        *total = lib().hostPinned[0];
    }
    return PG_OK;
}

// ------------------------------------------- scan without a caller workspace
// The scan's workspace of pg_scan_u32_static lives in this file-scope device
// array, one copy per GPU, held until the process ends.  Host code never uses
// the array's name (in host code it is a host-side stand-in, not a device
// address): initStaticScratch looks up each device's address once, with that
// device current, and every launch takes that address.
static constexpr uint64_t STATIC_SCAN_MAX_N = 1ull << 28;
static constexpr uint64_t STATIC_SCRATCH_BYTES = 9ull << 20;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];
static constexpr int MAX_DEVICES = 64;
static void *g_scratchPtr[MAX_DEVICES];

// bytes of pg_scan_u32's workspace for n entries when the tile sums are
// computed too: group offsets (G + 1 u64), in-group tile offsets (T u32),
// group sums (G u32), tile sums (T u32)
static uint64_t scanScratchBytes(uint64_t n)
{
    const u64 T = scanTiles(n), G = scanGroupCount(T);
    return 8 * (G + 1) + 4 * T + 4 * G + 4 * T;
}
static_assert(8 * (((1ull << 28) / 256 + 63) / 64 + 1) + 8 * ((1ull << 28) / 256) +
                      4 * (((1ull << 28) / 256 + 63) / 64) <= (9ull << 20),
              "the static scratch holds the workspace of the largest scan it accepts");

namespace pg {
int initStaticScratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) {
        setError("device %d beyond the static scratch table", device);
        return PG_ERR_INVALID;
    }
    if (g_scratchPtr[device]) return PG_OK;
    void *p = nullptr;
    PG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchPtr[device] = p;
    return PG_OK;
}
}  // namespace pg

extern "C" int pg_scan_u32_static(uint32_t *data, uint64_t n, const uint32_t *tileSums, uint64_t *total,
                                  void *stream)
{
    Lib &L = lib();
    if (!L.init) return PG_ERR_NOT_INIT;
    if (n > STATIC_SCAN_MAX_N) {
        setError("pg_scan_u32_static: n = %llu exceeds the static workspace's maximum of %llu entries",
                 (unsigned long long)n, (unsigned long long)STATIC_SCAN_MAX_N);
        return PG_ERR_INVALID;
    }
    if (n && !data) {
        setError("pg_scan_u32_static: bad argument");
        return PG_ERR_INVALID;
    }
    void *ws = (L.device >= 0 && L.device < MAX_DEVICES) ? g_scratchPtr[L.device] : nullptr;
    if (!ws) {
        setError("pg_scan_u32_static: no static workspace on device %d", L.device);
        return PG_ERR_NOT_INIT;
    }
    if (scanScratchBytes(n) > STATIC_SCRATCH_BYTES) {
        setError("pg_scan_u32_static: workspace of %llu bytes exceeds %llu", (unsigned long long)scanScratchBytes(n),
                 (unsigned long long)STATIC_SCRATCH_BYTES);
        return PG_ERR_INVALID;
    }
    return pg_scan_u32(data, n, tileSums, ws, total, stream);
}

extern "C" int pg_expand(const uint32_t *incl, uint64_t n, uint64_t c, uint32_t *parentIdx, void *scratch,
                         void *stream)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if (c && (!incl || !n || !parentIdx || !scratch)) {
        setError("pg_expand: bad argument");
        return PG_ERR_INVALID;
    }
    return launchExpand(incl, n, c, (u32 *)scratch, parentIdx, toStream(stream));
}

extern "C" int pg_genmoves(const void *boards, const uint8_t *states, const uint32_t *incl, int color, uint64_t n,
                           uint16_t *childMoves, void *stream)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if ((color & ~1) || (n && (!boards || !states || !incl || !childMoves))) {
        setError("pg_genmoves: bad argument");
        return PG_ERR_INVALID;
    }
    return launchGenmoves((const QuadBitBoard *)boards, states, incl, color, n, childMoves, toStream(stream));
}

extern "C" int pg_leaf(const void *parentBoards, const uint8_t *parentStates, const uint32_t *parentIdx,
                       const uint16_t *moves, int color, uint64_t n, int plies, uint64_t *sum, void *stream)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if ((color & ~1) || (plies != 1 && plies != 2) || !sum ||
        (n && (!parentBoards || !parentStates || !parentIdx || !moves))) {
        setError("pg_leaf: bad argument");
        return PG_ERR_INVALID;
    }
    return launchLeaf((const QuadBitBoard *)parentBoards, parentStates, parentIdx, moves, color, n, plies, sum,
                      toStream(stream));
}

extern "C" int pg_profile_enable(int on)
{
    lib().prof.on = on != 0;
    return PG_OK;
}

extern "C" int pg_profile_kernels(uint32_t mask)
{
    lib().prof.mask = mask;
    return PG_OK;
}

extern "C" int pg_profile_reset(void)
{
    lib().prof.reset();
    return PG_OK;
}

extern "C" int pg_profile_get(pg_kernel_stat *out, int maxOut, int *nOut)
{
    if (!out || !nOut) return PG_ERR_INVALID;
    Profiler &P = lib().prof;
    int n = 0;
    for (int k = 0; k < K_COUNT && n < maxOut; k++) {
        if (!P.launches[k]) continue;
        pg_kernel_stat &s = out[n++];
        memset(&s, 0, sizeof s);
        strncpy(s.name, KERNEL_NAMES[k], sizeof s.name - 1);
        s.launches = P.launches[k];
        s.ms = P.ms[k];
        s.bytes = P.bytes[k];
        s.items = P.items[k];
        s.big_bytes = P.bigBytes[k];
        s.big_ms = P.bigN[k] ? P.bigMs[k] / P.bigN[k] : 0;
        s.big_launches = P.bigN[k];
    }
    *nOut = n;
    return PG_OK;
}

extern "C" int pg_device_synchronize(void)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    PG_HIP(hipDeviceSynchronize());
    return PG_OK;
}

// device memory helpers for callers without a HIP runtime of their own
// (the kernel-level entry points take device pointers)
extern "C" int pg_device_malloc(uint64_t bytes, void **ptr)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if (!ptr) return PG_ERR_INVALID;
    PG_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return PG_OK;
}

extern "C" int pg_device_free(void *ptr)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    PG_HIP(hipFree(ptr));
    return PG_OK;
}

extern "C" int pg_memcpy(void *dst, const void *src, uint64_t bytes, int kind)
{
    if (!lib().init) return PG_ERR_NOT_INIT;
    if (kind < 0 || kind > 2 || (bytes && (!dst || !src))) return PG_ERR_INVALID;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    PG_HIP(hipMemcpy(dst, src, bytes, k));
    return PG_OK;
}

extern "C" void pg_set_leaf2_variant(int v) { g_leaf2Variant = v == 1 ? 1 : 0; }
extern "C" void pg_set_leaf_grid(int blocksPerCU) { g_leafBlocksPerCU = blocksPerCU >= 1 && blocksPerCU <= 256 ? blocksPerCU : 16; }

// slider-attack method of the fused two-ply leaf: 0 = computed fills (the
// default), 1 = fancy magic tables in device memory, 2 = the same with the
// bishop sets and the square rows in each block's LDS.  The tables are built
// (and checked against a ray walk for every blocker subset) and uploaded to
// the library's device by the first call there that asks for them; the mode
// applies to the launches that follow (a call made while a server runs is
// refused, so no kernel in flight sees it change).
extern "C" int pg_set_slider_attacks(int mode)
{
    if (mode < SL_FILLS || mode > SL_MAGIC_LDS) {
        setError("pg_set_slider_attacks: mode %d is not 0, 1 or 2", mode);
        return PG_ERR_INVALID;
    }
    Lib &L = lib();
    if (!L.init) {
        setError("pg_init has not run");
        return PG_ERR_NOT_INIT;
    }
    std::lock_guard<std::mutex> lock(L.mu);
    PG_TRY(refuseWhileServing("pg_set_slider_attacks"));
    if (mode != SL_FILLS) {
        static int uploadedTo = -1;  // the device that holds the tables
        if (uploadedTo != L.device) {
            std::unique_ptr<MagicTables> t(new MagicTables);
            if (!buildMagicTables(*t)) {
                setError("pg_set_slider_attacks: the magic search failed");
                return PG_ERR_INVALID;
            }
            PG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_magicSq), t->sq, sizeof t->sq));
            PG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_magicAtt), t->att, sizeof t->att));
            uploadedTo = L.device;
        }
    }
    g_sliderMode = mode;
    return PG_OK;
}
