// pg_batch.hip -- batched perft: thousands of positions of uneven size in one
// device-resident pass (no reference counterpart; the reference perfts one
// position per call, perft_kernels.cu:620).
//
// Design:
//  * every root is normalised to white-to-move by a colour flip (vertical
//    mirror + colour swap, perft-invariant), so a BFS level has one side to
//    move and the kernels stay colour-templated;
//  * roots are grouped by depth; a group is one multi-root BFS: level 0 = the
//    roots, each level's entries carry (parent, move, root id), children keep
//    their parents' order so a root's entries stay contiguous;
//  * level sizes never leave the device: after each scan a one-thread kernel
//    bump-allocates the next level inside the arena (setting an OOM flag when
//    it does not fit) and every kernel runs a fixed persistent grid that
//    reads its size from the level descriptor.  With fixed launch shapes the
//    whole level sequence of a group is captured once in a hipGraph and
//    replayed for every batch of the same shape;
//  * the leaf is persistent: waves pull 64-entry chunks from a device work
//    queue (an atomic cursor), so roots with huge subtrees and roots with tiny
//    ones share the GPU without a tail; results go to per-root u64 counters
//    with one atomic per wave and root;
//  * with the transposition table every level after the roots probes it,
//    merges identical positions across all roots and counts only what is left
//    to expand; the TT leaf visits its chunks in a fixed permuted order and an
//    upsweep stores every merged subtree once;
//  * an autotuner picks the leaf plies (1 or 2), with the table on whether a
//    group uses it and how much of it, and the persistent grid, per (depth,
//    group size), by timing each choice on a sample of the group, and caches
//    the choice.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <tuple>
#include <vector>

#include "pg_device.h"
#include "pg_launcher.h"
#include "pg_tt.h"

namespace pg {

constexpr int MAXLV = 24;
constexpr int B_CHUNK = 64;  // leaf entries per work-queue pop (one per lane)

struct LvlDev {
    u64 n;                 // entries (level 0: roots)
    u32 *idx;              // parent index into the previous level (level >= 1)
    u16 *mv;               // move applied to the parent (level >= 1)
    u32 *root;             // root id (level >= 1)
    QuadBitBoard *boards;  // positions of this level (level 0: the roots)
    u8 *states;
    u32 *cnt;              // child counts, then their inclusive prefix
    u32 *tileSums;
    u64 *tileOffs;
    u32 *parts;            // expand partitions of the next level
    u64 *known;            // TT mode: table hit + 1 (0 = miss)
    u64 *val;              // TT mode: subtree count of each entry
    u32 *redir;            // TT mode, levels >= 1: entry an identical position merges into
    DedupEntry *dtab;      // TT mode, levels >= 1: merge table (dmask + 1 slots, >= 2n)
    u64 dmask;
};

struct BatchDev {
    char *arena;
    u64 arenaCap, arenaOff;  // device-side bump allocator over the arena
    u32 oom, leafLevel;
    u64 work;                // leaf work-queue cursor
    u64 *rootCounts;
    TTEntry *tt;             // shared device table (TT mode)
    u64 ttMask;
    u32 depth, useTT;
    u64 ttSalt;              // XORed into every table key (0 = the shared key space)
    u64 ttCtr[3];            // probes, hits, stores
    LvlDev lv[MAXLV];
};

__device__ __forceinline__ void *dalloc(BatchDev *d, u64 bytes)
{
    bytes = (bytes + 255) & ~(u64)255;
    if (d->arenaOff + bytes > d->arenaCap) {
        d->oom = 1;
        return nullptr;
    }
    void *p = d->arena + d->arenaOff;
    d->arenaOff += bytes;
    return p;
}

// A table key in the group's key space.  The autotuner gives each trial its
// own salt, so that a trial never finds what an earlier trial stored for the
// same sample; a salted entry has its real depth and count, and no unsalted
// probe can match it.
__device__ __forceinline__ Hash128 salted(Hash128 h, u64 salt)
{
    h.lo ^= salt;
    h.hi ^= salt;
    return h;
}

// ---------------------------------------------------------------- kernels
// zero the per-root counters and the work cursor (a kernel, not a memset
// node, so the captured graph replays correctly)
__global__ void k_b_init(BatchDev *d)
{
    const u64 nRoots = d->lv[0].n;
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nRoots; i += (u64)gridDim.x * blockDim.x) d->rootCounts[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d->work = 0;
        d->ttCtr[0] = d->ttCtr[1] = d->ttCtr[2] = 0;
    }
}

__global__ __launch_bounds__(BS) void k_b_count0(BatchDev *d)
{
    if (d->oom) return;
    const LvlDev a = d->lv[0];
    const bool useTT = d->useTT;
    TTEntry *const tt = d->tt;
    const u64 ttMask = d->ttMask;
    const int depth = (int)d->depth;
    const u64 salt = d->ttSalt;
    TTTally tally;
    u64 tiles = (a.n + BS - 1) / BS;
    for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
        u64 i = t * BS + threadIdx.x;
        u32 c = 0;
        bool hit = false;
        if (i < a.n) {
            if (useTT) {
                u64 v;
                hit = ttProbe(tt, ttMask, salted(hashPosition(a.boards[i], a.states[i], WHITE), salt), depth, &v);
                a.known[i] = hit ? v + 1 : 0;
                tally.add(true, hit, false);
            }
            c = hit ? 0 : countChildren<WHITE>(a.boards[i], a.states[i]);
            a.cnt[i] = c;
        }
        u64 s = blockSum64(c);
        if (threadIdx.x == 0) a.tileSums[t] = (u32)s;
        __syncthreads();
    }
    if (useTT) ttFlush(d->ttCtr, tally);
}

template <int C>  // side to move in the parent boards
__global__ __launch_bounds__(BS) void k_b_makecount(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    const LvlDev p = d->lv[L - 1];
    u64 tiles = (a.n + BS - 1) / BS;
    if (d->useTT) {
        // make + probe; the counts follow the level's merge (k_b_dedup_count)
        TTEntry *const tt = d->tt;
        const u64 ttMask = d->ttMask;
        const int rd = (int)d->depth - L;
        const u64 salt = d->ttSalt;
        TTTally tally;
        for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < a.n; i += (u64)gridDim.x * BS) {
            u32 pi = a.idx[i];
            QuadBitBoard q = p.boards[pi];
            u8 st = p.states[pi];
            makeMove<C>(q, st, a.mv[i]);
            a.boards[i] = q;
            a.states[i] = st;
            u64 v;
            bool hit = ttProbe(tt, ttMask, salted(hashPosition(q, st, 1 - C), salt), rd, &v);
            a.known[i] = hit ? v + 1 : 0;
            tally.add(true, hit, false);
        }
        ttFlush(d->ttCtr, tally);
        return;
    }
    for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
        u64 i = t * BS + threadIdx.x;
        u32 c = 0;
        if (i < a.n) {
            u32 pi = a.idx[i];
            QuadBitBoard q = p.boards[pi];
            u8 st = p.states[pi];
            makeMove<C>(q, st, a.mv[i]);
            a.boards[i] = q;
            a.states[i] = st;
            c = countChildren<1 - C>(q, st);
            a.cnt[i] = c;
        }
        u64 s = blockSum64(c);
        if (threadIdx.x == 0) a.tileSums[t] = (u32)s;
        __syncthreads();
    }
}

// TT mode, levels >= 1: merge identical positions of level L (the scheme of
// the plain engine's k_dedup_*), then count the children of the entries that
// remain to be expanded.  The table is not cleared: every slot that
// k_b_dedup_count reads was written by its reader in k_b_dedup_write.
template <int C>  // side to move in level L's boards
__global__ __launch_bounds__(BS) void k_b_dedup_write(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < a.n; i += (u64)gridDim.x * BS) {
        if (a.known[i]) continue;
        Hash128 h = hashPosition(a.boards[i], a.states[i], C);
        a.dtab[dedupSlot(h, a.dmask)] = DedupEntry{h.hi, i + 1};
    }
}

template <int C>  // side to move in level L's boards
__global__ __launch_bounds__(BS) void k_b_dedup_count(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    u64 tiles = (a.n + BS - 1) / BS;
    for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
        u64 i = t * BS + threadIdx.x;
        u32 c = 0;
        if (i < a.n) {
            u32 r = (u32)i;
            if (!a.known[i]) {
                QuadBitBoard q = a.boards[i];
                u8 st = a.states[i];
                r = dedupTarget(a.boards, a.states, a.n, a.dtab, a.dmask, i, q, st, hashPosition(q, st, C));
                if (r == (u32)i) c = countChildren<C>(q, st);
            }
            a.redir[i] = r;
            a.cnt[i] = c;
        }
        u64 s = blockSum64(c);
        if (threadIdx.x == 0) a.tileSums[t] = (u32)s;
        __syncthreads();
    }
}

// duplicates take the value of the entry they merged into
__global__ __launch_bounds__(BS) void k_b_redirect(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < a.n; i += (u64)gridDim.x * BS)
        if (a.redir[i] != (u32)i) a.val[i] = a.val[a.redir[i]];
}

__global__ __launch_bounds__(TOP_BS) void k_b_scan_top(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    scanTileSums(a.tileSums, (a.n + BS - 1) / BS, a.tileOffs);
}

__global__ __launch_bounds__(BS) void k_b_scan_apply(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    const u64 tiles = (a.n + BS - 1) / BS;
    // one wave per 256-entry tile (no block barrier inside)
    for (u64 t = (u64)blockIdx.x * (BS / 64) + (threadIdx.x >> 6); t < tiles; t += (u64)gridDim.x * (BS / 64))
        scanApplyTile(a.cnt, a.n, a.tileOffs, t);
}

// one thread: size the next level and carve its arrays out of the arena
__global__ void k_b_alloc(BatchDev *d, int L)
{
    if (d->oom) return;
    LvlDev &a = d->lv[L];
    LvlDev &b = d->lv[L + 1];
    u64 tiles = (a.n + BS - 1) / BS;
    u64 c = a.n ? a.tileOffs[tiles] : 0;
    b.n = 0;
    if (c >= 0xFFFFFFFFull) {
        d->oom = 1;
        return;
    }
    a.parts = (u32 *)dalloc(d, 4 * (c / EXPAND_TILE + 2));
    b.idx = (u32 *)dalloc(d, 4 * c + 4);
    b.mv = (u16 *)dalloc(d, 2 * c + 2);
    b.root = (u32 *)dalloc(d, 4 * c + 4);
    if ((u32)(L + 1) < d->leafLevel) {
        u64 bt = (c + BS - 1) / BS;
        b.boards = (QuadBitBoard *)dalloc(d, 32 * c + 32);
        b.states = (u8 *)dalloc(d, c + 1);
        b.cnt = (u32 *)dalloc(d, 4 * c + 4);
        b.tileSums = (u32 *)dalloc(d, 4 * bt + 4);
        b.tileOffs = (u64 *)dalloc(d, 8 * (bt + 1));
        if (d->useTT) {
            u64 slots = 256;
            while (slots < 2 * c) slots <<= 1;
            b.known = (u64 *)dalloc(d, 8 * c + 8);
            b.redir = (u32 *)dalloc(d, 4 * c + 4);
            b.dtab = (DedupEntry *)dalloc(d, sizeof(DedupEntry) * slots);
            b.dmask = slots - 1;
        }
    }
    if (d->useTT) b.val = (u64 *)dalloc(d, 8 * c + 8);
    if (!d->oom) b.n = c;
}

__global__ __launch_bounds__(BS) void k_b_partitions(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    u64 n = a.n, c = d->lv[L + 1].n;
    if (!c) return;
    u64 tiles = (c + EXPAND_TILE - 1) / EXPAND_TILE;
    for (u64 t = (u64)blockIdx.x * BS + threadIdx.x; t <= tiles; t += (u64)gridDim.x * BS) {
        u64 child = t * EXPAND_TILE;
        if (child > c) child = c;
        u64 p = upperBound(a.cnt, n, child);
        a.parts[t] = (u32)(p < n ? p : n - 1);
    }
}

// parent index (and root id) of every child of level L+1: one 2048-child tile
// per block iteration, parents' first children scattered into an LDS tile and
// filled forward by a max-scan (the scheme of k_expand)
__global__ __launch_bounds__(BS) void k_b_expand(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    const LvlDev b = d->lv[L + 1];
    __shared__ __attribute__((aligned(16))) u32 tile[EXPAND_TILE];
    __shared__ u32 wmax[BS / 64];
    const u64 c = b.n;
    const u64 tiles = (c + EXPAND_TILE - 1) / EXPAND_TILE;
    for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
        const u64 c0 = t * EXPAND_TILE;
        for (int k = threadIdx.x; k < EXPAND_TILE; k += BS) tile[k] = 0;
        __syncthreads();
        u32 o[EXPAND_PER], r[EXPAND_PER];
        expandTileParents(a.cnt, a.parts[t], a.parts[t + 1], c0, tile, wmax, o);
        // each child's root: the parent's root (level 1 entries are roots' children)
        const u64 base = c0 + (u64)threadIdx.x * EXPAND_PER;
        if (L == 0) {
#pragma unroll
            for (int k = 0; k < EXPAND_PER; k++) r[k] = o[k];
        } else {
            // (a select between o[k] and a.root[o[k]] made the compiler spill
            // o to scratch and load through a selected pointer)
#pragma unroll
            for (int k = 0; k < EXPAND_PER; k++) {
                u32 x = 0;
                if (base + k < c) x = a.root[o[k]];
                r[k] = x;
            }
        }
        storeTileStreaming(tile, o, b.idx, c0, c);
        __syncthreads();
        storeTileStreaming(tile, r, b.root, c0, c);
        __syncthreads();  // the tile is cleared again by the next iteration
    }
}

template <int C>  // side to move in this level's boards
__global__ __launch_bounds__(BS) void k_b_genmoves(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    u16 *out = d->lv[L + 1].mv;
    u64 n = a.n;
    const u32 *incl = a.cnt;
    __shared__ __attribute__((aligned(16))) u16 buf[GEN_LCAP];
    u64 tiles = (n + BS - 1) / BS;
    const bool aligned = ((uintptr_t)out & 15) == 0;  // arena carvings are 256-byte aligned
    for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
        u64 b0 = t * BS, i = b0 + threadIdx.x;
        u64 bEnd = b0 + BS < n ? b0 + BS : n;
        u64 blockStart = b0 ? incl[b0 - 1] : 0;
        u64 total = (u64)incl[bEnd - 1] - blockStart;
        u64 excl = (i < n && i) ? (u64)incl[i - 1] : 0;
        u64 mine = i < n ? (u64)incl[i] - excl : 0;  // 0 for a table hit: nothing to write
        const u64 sh = aligned ? (blockStart & 7) : 0;  // the tile's offset (copyMovesOut)
        if (total + sh <= GEN_LCAP) {
            if (mine) {
                u16 *dst = buf + sh + (excl - blockStart);
                generateMoves<C>(a.boards[i], a.states[i], [&](u32 k, u16 m) {
                    if (k < mine) dst[k] = m;
                });
            }
            __syncthreads();
            if (aligned) {
                copyMovesOut(buf, out, blockStart, total);
            } else {
                u16 *o = out + blockStart;
                for (u64 j = threadIdx.x; j < total; j += BS) o[j] = buf[j];
            }
        } else if (mine) {
            u16 *dst = out + excl;
            generateMoves<C>(a.boards[i], a.states[i], [&](u32 k, u16 m) {
                if (k < mine) dst[k] = m;
            });
        }
        __syncthreads();
    }
}

// add a per-lane value to its root's counter: one atomic per wave when the
// lanes share a root (entries are root-ordered), else one per lane
__device__ __forceinline__ void addToRoots(u64 *rootCounts, bool valid, u32 root, u64 v)
{
    u32 r0 = __shfl(root, __builtin_ctzll(__ballot(valid) | (1ULL << 63)));
    bool same = valid ? root == r0 : true;
    if (__all(same)) {
        u64 s = waveSum64(valid ? v : 0);
        if ((threadIdx.x & 63) == 0 && s) atomicAdd((unsigned long long *)&rootCounts[r0], (unsigned long long)s);
    } else if (valid && v) {
        atomicAdd((unsigned long long *)&rootCounts[root], (unsigned long long)v);
    }
}

// persistent leaf, one ply: pop 64-entry chunks from the work queue
template <int C>
__global__ __launch_bounds__(BS) void k_b_leaf1(BatchDev *d)
{
    if (d->oom) return;
    const u32 L = d->leafLevel;
    const LvlDev a = d->lv[L];
    const LvlDev p = d->lv[L - 1];
    const u64 n = a.n;
    const int lane = threadIdx.x & 63;
    // next chunk's ticket taken as the current one starts (see k_b_leaf2)
    u64 ticket = 0;
    if (lane == 0) ticket = atomicAdd((unsigned long long *)&d->work, 64ull);
    for (;;) {
        const u64 base = __shfl(ticket, 0);
        if (base >= n) break;
        if (lane == 0) ticket = atomicAdd((unsigned long long *)&d->work, 64ull);
        u64 e = base + lane;
        bool valid = e < n;
        u32 v = 0, root = 0;
        if (valid) {
            u32 pi = a.idx[e];
            QuadBitBoard q = p.boards[pi];
            u8 st = p.states[pi];
            makeMove<C>(q, st, a.mv[e]);
            v = countMoves<1 - C>(q, st);
            root = a.root[e];
        }
        addToRoots(d->rootCounts, valid, root, v);
    }
}

// persistent fused two-ply leaf: the wave-cooperative scheme of k_leaf2
// (child boards and the flat pair list in LDS) with work-queue chunks.  A
// chunk whose entries all belong to one root (nearly all: roots own millions
// of leaf entries and entries are root-ordered) sums its pair counts in
// registers and adds one wave total to the root; a chunk that straddles roots
// keeps per-entry partial sums in LDS (atomics) and attributes them per lane.
// The choice is wave-uniform, so the pair loop does not diverge.  Pairs take
// 3 bytes of LDS and the kernels 128 VGPRs at most, so four blocks fit a CU
// and four waves a SIMD, as for k_leaf2q; a window of 2,560 pairs is what
// four blocks per CU leave beside the per-entry sums.
constexpr int BL2_CAP = 2560;
constexpr int BL2_WAVES = 4;
#define BL2_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4, 4)))
template <int C>
__global__ __launch_bounds__(BL2_WAVES * 64) BL2_OCCUPANCY void k_b_leaf2(BatchDev *d)
{
    if (d->oom) return;
    __shared__ QuadBitBoard sb[BL2_WAVES][B_CHUNK];
    __shared__ u8 ss[BL2_WAVES][B_CHUNK];
    __shared__ u32 sacc[BL2_WAVES][B_CHUNK];
    __shared__ u16 slm[BL2_WAVES][BL2_CAP + 1];  // slot BL2_CAP: moves outside the window
    __shared__ u8 sll[BL2_WAVES][BL2_CAP + 1];
    const u32 L = d->leafLevel;
    // descriptor fields into registers once (the atomics below would otherwise
    // force them to be re-read from the descriptor)
    const u32 *__restrict__ aIdx = d->lv[L].idx;
    const u16 *__restrict__ aMv = d->lv[L].mv;
    const u32 *__restrict__ aRoot = d->lv[L].root;
    const QuadBitBoard *__restrict__ pBoards = d->lv[L - 1].boards;
    const u8 *__restrict__ pStates = d->lv[L - 1].states;
    u64 *rootCounts = d->rootCounts;
    const u64 n = d->lv[L].n;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // the next chunk's queue ticket is taken as the current one starts, so
    // the atomic's round trip overlaps a chunk's work
    u64 ticket = 0;
    if (lane == 0) ticket = atomicAdd((unsigned long long *)&d->work, (unsigned long long)B_CHUNK);
    for (;;) {
        const u64 base = __shfl(ticket, 0);
        if (base >= n) break;
        if (lane == 0) ticket = atomicAdd((unsigned long long *)&d->work, (unsigned long long)B_CHUNK);
        u32 cnt = 0, root = 0;
        u64 e = base + lane;
        bool valid = lane < B_CHUNK && e < n;
        BB catt;  // the child's enemy attacks, from its count, for its generation
        if (valid) {
            u32 pi = aIdx[e];
            QuadBitBoard q = pBoards[pi];
            u8 st = pStates[pi];
            makeMove<C>(q, st, aMv[e]);
            sb[w][lane] = q;
            ss[w][lane] = st;
            Frame fr;
            cnt = countChildrenFrame<1 - C>(q, st, FillSliders(), fr);
            catt = fr.att;
            root = aRoot[e];
        }
        const u32 r0 = __shfl(root, 0);  // lane 0 is always valid
        const bool oneRoot = __all(valid ? root == r0 : true);
        if (!oneRoot && lane < B_CHUNK) sacc[w][lane] = 0;
        u32 incl = waveInclScan(cnt);
        u32 off = incl - cnt;
        u32 total = __shfl(incl, 63);
        u32 acc = 0;  // a lane's pairs of one chunk: at most 218 x 218 moves
        waveLdsSync();
        for (u32 wb = 0; wb < total; wb += BL2_CAP) {
            if (cnt && off < wb + BL2_CAP && off + cnt > wb) {
                QuadBitBoard q = sb[w][lane];
                u8 st = ss[w][lane];
                u16 *dm = slm[w];
                u8 *dl = sll[w];
                Frame fr = framePins<1 - C>(q);
                fr.att = catt;
                generateMovesF<1 - C>(fr, st, [&](u32 k, u16 m) {
                    u32 j = off + k - wb;
                    j = j < (u32)BL2_CAP ? j : (u32)BL2_CAP;
                    dm[j] = m, dl[j] = (u8)lane;
                });
            }
            waveLdsSync();
            u32 wn = total - wb < (u32)BL2_CAP ? total - wb : (u32)BL2_CAP;
            for (u32 j = lane; j < wn; j += 64) {
                u16 x = slm[w][j];
                u32 src = sll[w][j];
                QuadBitBoard c = sb[w][src];
                u8 s2 = ss[w][src];
                makeMove<1 - C>(c, s2, x);
                u32 cm = countMoves<C>(c, s2);
                if (oneRoot) acc += cm;
                else atomicAdd(&sacc[w][src], cm);  // LDS atomic
            }
            waveLdsSync();
        }
        if (oneRoot) {
            u64 sum = waveSum64(acc);
            if (lane == 0 && sum) atomicAdd((unsigned long long *)&rootCounts[r0], (unsigned long long)sum);
        } else {
            u64 v = valid ? sacc[w][lane] : 0;
            addToRoots(rootCounts, valid, root, v);
        }
        waveLdsSync();
    }
}
// ---------------------------------------------------------- TT-mode kernels
// Leaf values per entry (instead of per-root atomics), then an upsweep that
// sums each parent's contiguous child range, stores the subtree in the table,
// and finally the roots' sums.  All sizes come from the descriptor.
template <int C>
__global__ __launch_bounds__(BS) void k_b_leaf1v(BatchDev *d)
{
    if (d->oom) return;
    const u32 L = d->leafLevel;
    const LvlDev a = d->lv[L];
    const LvlDev p = d->lv[L - 1];
    for (u64 e = (u64)blockIdx.x * BS + threadIdx.x; e < a.n; e += (u64)gridDim.x * BS) {
        u32 pi = a.idx[e];
        QuadBitBoard q = p.boards[pi];
        u8 st = p.states[pi];
        makeMove<C>(q, st, a.mv[e]);
        a.val[e] = countMoves<1 - C>(q, st);
    }
}

// TT-mode two-ply leaf with per-entry values and the depth-2 table.  Chunks
// are visited in a fixed permuted order rather than taken in order from the
// queue: the chunks in flight together are then spread over the level,
// whereas in-order chunks put the copies of a position that nearby parents
// share (A-then-B and B-then-A) in flight together, where none of them finds
// the other's stored count (a single-root perft 7 made 25 % more leaf misses
// in order).
template <int C>
__global__ __launch_bounds__(BL2_WAVES * 64) BL2_OCCUPANCY void k_b_leaf2v(BatchDev *d)
{
    if (d->oom) return;
    __shared__ QuadBitBoard sb[BL2_WAVES][B_CHUNK];
    __shared__ u8 ss[BL2_WAVES][B_CHUNK];
    __shared__ u32 sacc[BL2_WAVES][B_CHUNK];
    __shared__ u16 slm[BL2_WAVES][BL2_CAP + 1];  // slot BL2_CAP: moves outside the window
    __shared__ u8 sll[BL2_WAVES][BL2_CAP + 1];
    const u32 L = d->leafLevel;
    const LvlDev a = d->lv[L];
    const LvlDev p = d->lv[L - 1];
    const u64 n = a.n;
    TTEntry *const tt = d->tt;
    const u64 ttMask = d->ttMask;
    const u64 salt = d->ttSalt;
    TTTally tally;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // chunk k of each wave's sequence is chunk k * P mod nChunks of the level
    // (P a prime above any chunk count): the chunks in flight together are
    // spread over the whole level
    const u64 nChunks = (n + B_CHUNK - 1) / B_CHUNK;
    const u64 P = 1000000007ULL % (nChunks ? nChunks : 1);
    for (u64 k = (u64)blockIdx.x * BL2_WAVES + w; k < nChunks; k += (u64)gridDim.x * BL2_WAVES) {
        const u64 base = ((k * P) % nChunks) * B_CHUNK;
        u32 cnt = 0;
        u64 e = base + lane, hv = 0;
        bool valid = lane < B_CHUNK && e < n, hit = false;
        BB catt;  // the child's enemy attacks, from its count, for its generation
        Hash128 h{};  // for the probe and, on a miss, the store
        if (valid) {
            u32 pi = a.idx[e];
            QuadBitBoard q = p.boards[pi];
            u8 st = p.states[pi];
            makeMove<C>(q, st, a.mv[e]);
            sb[w][lane] = q;
            ss[w][lane] = st;
            h = salted(hashPosition(q, st, 1 - C), salt);
            hit = ttProbe(tt, ttMask, h, 2, &hv);
            if (!hit) {
                Frame fr;
                cnt = countChildrenFrame<1 - C>(q, st, FillSliders(), fr);
                catt = fr.att;
            }
        }
        if (lane < B_CHUNK) sacc[w][lane] = 0;
        u32 incl = waveInclScan(cnt);
        u32 off = incl - cnt;
        u32 total = __shfl(incl, 63);
        waveLdsSync();
        for (u32 wb = 0; wb < total; wb += BL2_CAP) {
            if (cnt && off < wb + BL2_CAP && off + cnt > wb) {
                QuadBitBoard q = sb[w][lane];
                u8 st = ss[w][lane];
                u16 *dm = slm[w];
                u8 *dl = sll[w];
                Frame fr = framePins<1 - C>(q);
                fr.att = catt;
                generateMovesF<1 - C>(fr, st, [&](u32 k, u16 m) {
                    u32 j = off + k - wb;
                    j = j < (u32)BL2_CAP ? j : (u32)BL2_CAP;
                    dm[j] = m, dl[j] = (u8)lane;
                });
            }
            waveLdsSync();
            u32 wn = total - wb < (u32)BL2_CAP ? total - wb : (u32)BL2_CAP;
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
        bool store = valid && !hit;
        if (valid) {
            u64 v = hit ? hv : sacc[w][lane];
            a.val[e] = v;
            if (store) ttStore(tt, ttMask, h, 2, v);
        }
        tally.add(valid, hit, store);
        waveLdsSync();
    }
    ttFlush(d->ttCtr, tally);
}

template <int C>  // side to move in level-L boards
__global__ __launch_bounds__(BS) void k_b_upsweep(BatchDev *d, int L)
{
    if (d->oom) return;
    const LvlDev a = d->lv[L];
    const u64 *cv = d->lv[L + 1].val;
    const int rd = (int)d->depth - L;
    TTEntry *const tt = d->tt;
    const u64 ttMask = d->ttMask;
    const u64 salt = d->ttSalt;
    TTTally tally;
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < a.n; i += (u64)gridDim.x * BS) {
        if (a.redir[i] != (u32)i) continue;  // duplicates are filled in by k_b_redirect
        u64 v;
        if (a.known[i]) {
            v = a.known[i] - 1;
        } else {
            v = 0;
            u64 b = i ? a.cnt[i - 1] : 0, e = a.cnt[i];
            for (u64 k = b; k < e; k++) v += cv[k];
            ttStore(tt, ttMask, salted(hashPosition(a.boards[i], a.states[i], C), salt), rd, v);
            tally.add(false, false, true);
        }
        a.val[i] = v;
    }
    ttFlush(d->ttCtr, tally);
}

__global__ __launch_bounds__(BS) void k_b_rootsum(BatchDev *d)
{
    if (d->oom) return;
    const LvlDev a = d->lv[0];
    const u64 *cv = d->lv[1].val;
    TTEntry *const tt = d->tt;
    const u64 ttMask = d->ttMask;
    const int depth = (int)d->depth;
    const u64 salt = d->ttSalt;
    u64 *const rootCounts = d->rootCounts;
    for (u64 i = (u64)blockIdx.x * BS + threadIdx.x; i < a.n; i += (u64)gridDim.x * BS) {
        u64 v;
        if (a.known[i]) {
            v = a.known[i] - 1;
        } else {
            v = 0;
            u64 b = i ? a.cnt[i - 1] : 0, e = a.cnt[i];
            for (u64 k = b; k < e; k++) v += cv[k];
            ttStore(tt, ttMask, salted(hashPosition(a.boards[i], a.states[i], WHITE), salt), depth, v);
        }
        rootCounts[i] = v;
    }
}

// ------------------------------------------------------------------ host
struct GroupPlan {
    int depth, plies, leafLevel;
    bool useTT;
    int gridPerCU = 8;  // blocks per CU of the persistent level kernels and the TT leaf
    int ttLog2 = 0;     // the table's first 2^ttLog2 entries, 0 = the whole table
    u64 ttSalt = 0;     // key space of the group's table entries (autotuner trials)
};

// a tuned choice: leaf plies, table use, persistent grid, table span
struct TunedPlan {
    int plies;
    bool useTT;
    int gridPerCU;
    int ttLog2 = 0;
};

// mask of the table span a group probes and stores in: its first 2^log2
// entries (every entry is keyed by the full hash, so any span is exact)
static u64 ttSpanMask(int log2)
{
    const u64 full = lib().tt.mask;
    if (log2 <= 0 || log2 >= 63) return full;
    const u64 m = (1ull << log2) - 1;
    return m < full ? m : full;
}
static int log2Floor(u64 v)
{
    int r = 0;
    while (v > 1) v >>= 1, r++;
    return r;
}

// One batched engine: its descriptor and graphs, and where it runs.  The
// library's own (g_batch) runs on the library arena and stream under the
// library mutex; a serving slot owns one over its arena partition, stream and
// pinned buffer (pg_serve.hip).
struct BatchState {
    BatchDev *dev = nullptr;  // device descriptor
    BatchDev host{};          // host copy used to initialise it
    // (depth, plies, tt, grid per CU) -> graph
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;
    // (depth, size bucket, tt requested) -> choice
    std::map<std::tuple<int, int, int>, TunedPlan> tuned;
    char *arena = nullptr;  // null: the library arena (usableArena bytes)
    size_t arenaBytes = 0;
    hipStream_t stream = nullptr;  // null: the library stream
    u64 *pinned = nullptr;         // null: the library's pinned buffer
};
static BatchState g_batch;

static unsigned persistentBlocks(int perCU)
{
    int cu = lib().cuCount ? lib().cuCount : 256;
    return (unsigned)(cu * perCU);
}

// enqueue the whole level sequence of one group (capturable: fixed shapes,
// sizes read on the device)
static int enqueueGroup(BatchState &B, const GroupPlan &g, hipStream_t st)
{
    BatchDev *d = B.dev;
    unsigned G = persistentBlocks(g.gridPerCU);
    hipLaunchKernelGGL(k_b_init, dim3(64), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_b_count0, dim3(G), dim3(BS), 0, st, d);
    for (int L = 0; L < g.leafLevel; L++) {
        int side = L & 1;  // level-L boards: white to move on even levels
        if (L > 0) {
            if (side == BLACK) hipLaunchKernelGGL(k_b_makecount<WHITE>, dim3(G), dim3(BS), 0, st, d, L);
            else hipLaunchKernelGGL(k_b_makecount<BLACK>, dim3(G), dim3(BS), 0, st, d, L);
            if (g.useTT) {
                if (side == WHITE) {
                    hipLaunchKernelGGL(k_b_dedup_write<WHITE>, dim3(G), dim3(BS), 0, st, d, L);
                    hipLaunchKernelGGL(k_b_dedup_count<WHITE>, dim3(G), dim3(BS), 0, st, d, L);
                } else {
                    hipLaunchKernelGGL(k_b_dedup_write<BLACK>, dim3(G), dim3(BS), 0, st, d, L);
                    hipLaunchKernelGGL(k_b_dedup_count<BLACK>, dim3(G), dim3(BS), 0, st, d, L);
                }
            }
        }
        hipLaunchKernelGGL(k_b_scan_top, dim3(1), dim3(TOP_BS), 0, st, d, L);
        hipLaunchKernelGGL(k_b_scan_apply, dim3(G), dim3(BS), 0, st, d, L);
        hipLaunchKernelGGL(k_b_alloc, dim3(1), dim3(1), 0, st, d, L);
        hipLaunchKernelGGL(k_b_partitions, dim3(G), dim3(BS), 0, st, d, L);
        hipLaunchKernelGGL(k_b_expand, dim3(G), dim3(BS), 0, st, d, L);
        if (side == WHITE) hipLaunchKernelGGL(k_b_genmoves<WHITE>, dim3(G), dim3(BS), 0, st, d, L);
        else hipLaunchKernelGGL(k_b_genmoves<BLACK>, dim3(G), dim3(BS), 0, st, d, L);
    }
    int pside = (g.leafLevel - 1) & 1;  // side to move in the leaf entries' parents
    if (g.useTT) {
        if (g.plies == 1) {
            if (pside == WHITE) hipLaunchKernelGGL(k_b_leaf1v<WHITE>, dim3(G), dim3(BS), 0, st, d);
            else hipLaunchKernelGGL(k_b_leaf1v<BLACK>, dim3(G), dim3(BS), 0, st, d);
        } else {
            if (pside == WHITE) hipLaunchKernelGGL(k_b_leaf2v<WHITE>, dim3(G), dim3(BL2_WAVES * 64), 0, st, d);
            else hipLaunchKernelGGL(k_b_leaf2v<BLACK>, dim3(G), dim3(BL2_WAVES * 64), 0, st, d);
        }
        for (int L = g.leafLevel - 1; L >= 1; L--) {
            if ((L & 1) == WHITE) hipLaunchKernelGGL(k_b_upsweep<WHITE>, dim3(G), dim3(BS), 0, st, d, L);
            else hipLaunchKernelGGL(k_b_upsweep<BLACK>, dim3(G), dim3(BS), 0, st, d, L);
            hipLaunchKernelGGL(k_b_redirect, dim3(G), dim3(BS), 0, st, d, L);
        }
        hipLaunchKernelGGL(k_b_rootsum, dim3(G), dim3(BS), 0, st, d);
    } else if (g.plies == 1) {
        // work-queue leaves: exactly the resident blocks, so none starts late to an empty queue
        unsigned R = residentGrid((const void *)(pside == WHITE ? k_b_leaf1<WHITE> : k_b_leaf1<BLACK>), BS);
        if (pside == WHITE) hipLaunchKernelGGL(k_b_leaf1<WHITE>, dim3(R), dim3(BS), 0, st, d);
        else hipLaunchKernelGGL(k_b_leaf1<BLACK>, dim3(R), dim3(BS), 0, st, d);
    } else {
        unsigned R = residentGrid((const void *)(pside == WHITE ? k_b_leaf2<WHITE> : k_b_leaf2<BLACK>), BL2_WAVES * 64);
        if (pside == WHITE) hipLaunchKernelGGL(k_b_leaf2<WHITE>, dim3(R), dim3(BL2_WAVES * 64), 0, st, d);
        else hipLaunchKernelGGL(k_b_leaf2<BLACK>, dim3(R), dim3(BL2_WAVES * 64), 0, st, d);
    }
    PG_HIP(hipGetLastError());
    return PG_OK;
}

static void flipColors(QuadBitBoard &q, u8 &state)
{
    u64 occ = q.bb[1] | q.bb[2] | q.bb[3];
    u64 white = occ & ~q.bb[0];
    q.bb[0] = __builtin_bswap64(white);
    q.bb[1] = __builtin_bswap64(q.bb[1]);
    q.bb[2] = __builtin_bswap64(q.bb[2]);
    q.bb[3] = __builtin_bswap64(q.bb[3]);
    state = (u8)((state & 0xF0) | ((state & 3) << 2) | ((state >> 2) & 3));
}

// run one group (roots already white to move) on the arena; OOM splits it
static int runGroup(BatchState &B, const GroupPlan &g, const QuadBitBoard *roots, const u8 *states, u64 nRoots,
                    bool useGraph, u64 *counts)
{
    Lib &L = lib();
    hipStream_t st = B.stream ? B.stream : L.stream;
    char *arena = B.arena ? B.arena : (char *)L.arena;
    const size_t arenaBytes = B.arena ? B.arenaBytes : usableArena();
    u64 *pinned = B.pinned ? B.pinned : L.hostPinned;
    // arena layout: [roots | states | level-0 arrays | rootCounts | device bump region]
    Bump A(arena, arenaBytes);
    QuadBitBoard *dRoots = A.alloc<QuadBitBoard>(nRoots);
    u8 *dStates = A.alloc<u8>(nRoots);
    u32 *cnt0 = A.alloc<u32>(nRoots);
    u64 tiles0 = (nRoots + BS - 1) / BS;
    u32 *ts0 = A.alloc<u32>(tiles0);
    u64 *to0 = A.alloc<u64>(tiles0 + 1);
    u64 *rc = A.alloc<u64>(nRoots);
    u64 *known0 = g.useTT ? A.alloc<u64>(nRoots) : nullptr;
    if (!dRoots || !dStates || !cnt0 || !ts0 || !to0 || !rc || (g.useTT && !known0)) return PG_ERR_ARENA_OOM;
    PG_HIP(hipMemcpyAsync(dRoots, roots, nRoots * sizeof(QuadBitBoard), hipMemcpyHostToDevice, st));
    PG_HIP(hipMemcpyAsync(dStates, states, nRoots, hipMemcpyHostToDevice, st));
    BatchDev &h = B.host;
    memset(&h, 0, sizeof h);
    h.arena = arena + A.off;
    h.arenaCap = arenaBytes - A.off;
    h.leafLevel = (u32)g.leafLevel;
    h.rootCounts = rc;
    h.lv[0].n = nRoots;
    h.lv[0].boards = dRoots;
    h.lv[0].states = dStates;
    h.lv[0].cnt = cnt0;
    h.lv[0].tileSums = ts0;
    h.lv[0].tileOffs = to0;
    h.lv[0].known = known0;
    h.depth = (u32)g.depth;
    h.useTT = g.useTT ? 1 : 0;
    h.tt = (TTEntry *)L.tt.entries;
    h.ttMask = ttSpanMask(g.ttLog2);
    h.ttSalt = g.ttSalt;
    PG_HIP(hipMemcpyAsync(B.dev, &h, sizeof h, hipMemcpyHostToDevice, st));
    // the event profiler is single-threaded: only the library's engine records
    Profiler &P = L.prof;
    const bool prof = &B == &g_batch;
    if (prof) P.begin(K_BATCH_GROUP, st);
    if (useGraph) {
        auto key = std::make_tuple(g.depth, g.plies, (int)g.useTT, g.gridPerCU);
        auto it = B.graphs.find(key);
        if (it == B.graphs.end()) {
            hipGraph_t graph;
            PG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int s = enqueueGroup(B, g, st);
            PG_HIP(hipStreamEndCapture(st, &graph));
            if (s != PG_OK) return s;
            hipGraphExec_t exec;
            PG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            it = B.graphs.emplace(key, exec).first;
        }
        PG_HIP(hipGraphLaunch(it->second, st));
    } else {
        PG_TRY(enqueueGroup(B, g, st));
    }
    if (prof) P.end(0, (double)nRoots, st);
    // one readback for the whole group: the OOM flag, then the counts
    PG_HIP(hipMemcpyAsync(pinned, &B.dev->oom, sizeof(u32), hipMemcpyDeviceToHost, st));
    PG_HIP(hipMemcpyAsync(counts, rc, nRoots * sizeof(u64), hipMemcpyDeviceToHost, st));
    if (g.useTT) PG_HIP(hipMemcpyAsync(pinned + 1, B.dev->ttCtr, 3 * sizeof(u64), hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
    if (prof) P.collect();
    if (g.useTT) ttCountDevice(pinned[1], pinned[2], pinned[3]);
    if (*(u32 *)pinned) return PG_ERR_ARENA_OOM;
    return PG_OK;
}

static int runGroupSplit(BatchState &B, const GroupPlan &g, const QuadBitBoard *roots, const u8 *states, u64 nRoots,
                         bool useGraph, u64 *counts)
{
    int s = runGroup(B, g, roots, states, nRoots, useGraph, counts);
    if (s != PG_ERR_ARENA_OOM) return s;
    if (nRoots == 1) {
        // a single root too large for one pass: its children become a batch
        // one ply shallower (each flipped back to white to move)
        if (g.depth <= 2) return PG_ERR_ARENA_OOM;
        u16 mv[MAX_MOVES];
        u32 nc = generateMoves<WHITE>(roots[0], states[0], mv);
        std::vector<QuadBitBoard> kids(nc);
        std::vector<u8> ks(nc);
        for (u32 i = 0; i < nc; i++) {
            kids[i] = roots[0];
            ks[i] = states[0];
            makeMove<WHITE>(kids[i], ks[i], mv[i]);
            flipColors(kids[i], ks[i]);
        }
        int plies = g.plies < g.depth - 1 ? g.plies : g.depth - 2;
        GroupPlan sub{g.depth - 1, plies, g.depth - 1 - plies, g.useTT, g.gridPerCU, g.ttLog2, g.ttSalt};
        std::vector<u64> kc(nc);
        if (nc) PG_TRY(runGroupSplit(B, sub, kids.data(), ks.data(), nc, useGraph, kc.data()));
        u64 total = 0;
        for (u64 v : kc) total += v;
        counts[0] = total;
        return PG_OK;
    }
    u64 h = nRoots / 2;
    PG_TRY(runGroupSplit(B, g, roots, states, h, useGraph, counts));
    return runGroupSplit(B, g, roots + h, states + h, nRoots - h, useGraph, counts + h);
}

// Autotuner: picks a group's leaf plies (1 or 2: where the BFS hands over to
// the fused leaf), when the table is on whether the group uses it (its probes
// and stores cost more than they save on shallow groups), then the
// persistent grid of its level kernels and TT leaf (4, 8 or 16 blocks per
// CU), then how much of the device table it uses: the whole table, or its
// first 1 GiB or 128 MiB (a span that the 256 MB last-level cache can hold
// makes probes cheaper, and a span too small for the group's stores loses
// hits), by timing each choice on a sample of the group.  Each choice first
// runs on a quarter of one sample (graph capture and first launches: the
// graph does not depend on the number of roots, and each trial's table keys
// are salted, so a larger warm-up buys nothing) and is timed on a second,
// disjoint sample when the group has room, so a timed run does not hit its
// own entries.  A span is tried only when it holds at least
// eight entries per store the whole group is expected to make (the timed
// sample's stores scaled to the group).  The sample shrinks 4x per ply above
// depth 4 so that tuning a deep group stays cheap.  Choices are cached per
// (depth, log2 size, table requested).
constexpr int TUNE_MAX_DEPTH = 7;  // groups this deep are not tuned
// Deeper groups than this are leaf-bound: the fused two-ply leaf won every
// trial there (config 3 with and without the table), so the one-ply leaf is
// not tried on them, and without the table they get the fixed plan at once,
// since the grid made no difference either (config 3's tuned pass took
// 0.5466 / 0.5457 / 0.5457 s at 4 / 8 / 16 blocks per CU).  With the table
// their grid and table span are still tried (about 3 % on config 3's table
// pass).  The one-ply trials at depths 5 and 6 were most of the first call:
// config 3's first call took 0.95 s with them and 0.57 s without, the tuned
// pass 0.547 s either way.
constexpr int TUNE_ONE_PLY_MAX_DEPTH = 4;
static const int TUNE_GRIDS[] = {4, 8, 16};
static const int TUNE_SPANS[] = {26, 23};  // log2 entries: 1 GiB and 128 MiB of 16-B entries
static u64 g_tuneTrials = 0;               // trials run so far (their key salts)
static TunedPlan chooseGroupPlan(int depth, const std::vector<QuadBitBoard> &roots, const std::vector<u8> &states,
                                 bool useGraph, bool useTT, int forcedGrid, int forcedSpan)
{
    int bucket = 0;
    for (u64 n = roots.size(); n > 1; n >>= 1) bucket++;
    auto key = std::make_tuple(depth, bucket, (int)useTT);
    auto it = g_batch.tuned.find(key);
    if (it != g_batch.tuned.end()) return it->second;
    // Deep groups are leaf-bound, where the fused two-ply leaf always wins
    // and the table always pays; a one-ply trial there would cost more than
    // a whole shallow batch (startpos children at depth 9: minutes).
    if (depth >= TUNE_MAX_DEPTH || (!useTT && depth > TUNE_ONE_PLY_MAX_DEPTH))
        return g_batch.tuned[key] = {2, useTT, forcedGrid ? forcedGrid : 8, forcedSpan};
    u64 ns = 256;
    for (int d = depth; d > 4 && ns > 8; d--) ns >>= 2;
    ns = std::min<u64>(roots.size(), ns);
    const u64 t0i = roots.size() >= 2 * ns ? ns : 0;  // timed sample [t0i, t0i + ns)
    const u64 nw = ns >= 4 ? ns / 4 : 1;               // warm-up sample [0, nw)
    std::vector<u64> tmp(ns);
    u64 stores = 0;  // table stores of the last timed run
    auto trial = [&](GroupPlan g) {
        // a key space of its own: no trial finds what another stored
        g.ttSalt = 0x9E3779B97F4A7C15ULL * ++g_tuneTrials;
        double t = 1e30;
        if (runGroupSplit(g_batch, g, roots.data(), states.data(), nw, useGraph, tmp.data()) == PG_OK) {
            uint64_t before[6], after[6];
            ttStats(before);
            auto a = std::chrono::steady_clock::now();
            if (runGroupSplit(g_batch, g, roots.data() + t0i, states.data() + t0i, ns, useGraph, tmp.data()) == PG_OK)
                t = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
            ttStats(after);
            stores = after[2] - before[2];
        }
        return t;
    };
    const int grid0 = forcedGrid ? forcedGrid : 8;
    double best = 1e30;
    TunedPlan choice{2, useTT, grid0, forcedSpan};
    u64 choiceStores = 0;
    for (int tt = useTT ? 1 : 0; tt >= 0; tt--) {
        for (int p = depth > TUNE_ONE_PLY_MAX_DEPTH ? 2 : 1; p <= 2 && p < depth; p++) {
            double t = trial(GroupPlan{depth, p, depth - p, tt != 0, grid0, forcedSpan});
            if (t < best) best = t, choice = {p, tt != 0, grid0, forcedSpan}, choiceStores = stores;
        }
    }
    if (!forcedGrid) {
        for (int gpc : TUNE_GRIDS) {
            if (gpc == grid0) continue;
            double t = trial(GroupPlan{depth, choice.plies, depth - choice.plies, choice.useTT, gpc, forcedSpan});
            if (t < best) best = t, choice.gridPerCU = gpc;
        }
    }
    if (choice.useTT && !forcedSpan) {
        const double groupStores = (double)choiceStores * (double)roots.size() / (double)ns;
        const int full = log2Floor(lib().tt.mask + 1);
        for (int span : TUNE_SPANS) {
            if (span >= full || (double)(1ull << span) < 8.0 * groupStores) continue;
            double t = trial(GroupPlan{depth, choice.plies, depth - choice.plies, true, choice.gridPerCU, span});
            if (t < best) best = t, choice.ttLog2 = span;
        }
    }
    g_batch.tuned[key] = choice;
    return choice;
}

// ------------------------------------------------ engines of serving slots
BatchState *batchSlotCreate(void *arena, size_t bytes, hipStream_t st, u64 *pinned)
{
    BatchState *B = new BatchState();
    if (hipMalloc((void **)&B->dev, sizeof(BatchDev)) != hipSuccess) {
        (void)hipGetLastError();
        delete B;
        return nullptr;
    }
    B->arena = (char *)arena;
    B->arenaBytes = bytes;
    B->stream = st;
    B->pinned = pinned;
    return B;
}

void batchSlotDestroy(BatchState *B)
{
    if (!B) return;
    for (auto &kv : B->graphs) (void)hipGraphExecDestroy(kv.second);
    if (B->dev) (void)hipFree(B->dev);
    delete B;
}

// Every position at its depth (2..MAXLV-2) as multi-root groups on the slot's
// engine, with a fixed plan (no trials in a serving loop): the two-ply leaf,
// the table when useTT and one exists, 8 blocks per CU; graphs captured per
// slot on first use.
int batchSlotRun(BatchState *B, int n, const pg_position *pos, const uint8_t *depths, bool useTT, uint64_t *counts)
{
    std::map<int, std::vector<int>> groups;
    for (int i = 0; i < n; i++) {
        if (depths[i] < 2 || depths[i] > MAXLV - 2) {
            setError("batchSlotRun: depth %d outside 2..%d", depths[i], MAXLV - 2);
            return PG_ERR_INVALID;
        }
        groups[depths[i]].push_back(i);
    }
    const bool tt = useTT && lib().tt.entries;
    for (auto &kv : groups) {
        const int depth = kv.first;
        const std::vector<int> &ids = kv.second;
        std::vector<QuadBitBoard> roots(ids.size());
        std::vector<u8> states(ids.size());
        for (size_t k = 0; k < ids.size(); k++) {
            memcpy(roots[k].bb, pos[ids[k]].bb, 32);
            states[k] = pos[ids[k]].state;
            if (pos[ids[k]].color == BLACK) flipColors(roots[k], states[k]);
        }
        const int plies = depth > 2 ? 2 : 1;
        GroupPlan g{depth, plies, depth - plies, tt, 8};
        std::vector<u64> out(ids.size());
        PG_TRY(runGroupSplit(*B, g, roots.data(), states.data(), ids.size(), true, out.data()));
        for (size_t k = 0; k < ids.size(); k++) counts[ids[k]] = out[k];
    }
    return PG_OK;
}

int perftBatch(int n, const pg_position *pos, const uint8_t *depths, const pg_batch_options *opt, uint64_t *counts)
{
    Lib &L = lib();
    if (!L.init) {
        setError("pg_init has not run");
        return PG_ERR_NOT_INIT;
    }
    std::lock_guard<std::mutex> lock(L.mu);
    PG_TRY(refuseWhileServing("pg_perft_batch"));
    if (!g_batch.dev) PG_HIP(hipMalloc((void **)&g_batch.dev, sizeof(BatchDev)));
    bool useGraph = !opt || opt->useGraph;
    bool useTT = opt && opt->useTT && L.tt.entries;
    int forcedPlies = opt ? opt->leafPlies : 0;
    const int forcedGrid = opt && opt->gridPerCU > 0 ? (opt->gridPerCU > 32 ? 32 : opt->gridPerCU) : 0;
    // a table budget in MB: the largest power-of-two span of 16-B entries within it
    const int forcedSpan = opt && opt->ttBudgetMB > 0 ? log2Floor((u64)opt->ttBudgetMB) + 16 : 0;
    // group by depth; depth <= 1 on the host as the reference does (perft_kernels.cu:623-626)
    std::map<int, std::vector<int>> groups;
    for (int i = 0; i < n; i++) {
        int d = depths[i];
        if (d > MAXLV - 2) {
            setError("depth %d too large for the batched path", d);
            return PG_ERR_INVALID;
        }
        if (d <= 1) {
            QuadBitBoard q;
            memcpy(q.bb, pos[i].bb, 32);
            counts[i] = d == 0 ? 1 : countMoves(q, pos[i].state, pos[i].color);
        } else {
            groups[d].push_back(i);
        }
    }
    for (auto &kv : groups) {
        int depth = kv.first;
        const std::vector<int> &ids = kv.second;
        std::vector<QuadBitBoard> roots(ids.size());
        std::vector<u8> states(ids.size());
        for (size_t k = 0; k < ids.size(); k++) {
            memcpy(roots[k].bb, pos[ids[k]].bb, 32);
            states[k] = pos[ids[k]].state;
            if (pos[ids[k]].color == BLACK) flipColors(roots[k], states[k]);
        }
        int plies = forcedPlies;
        bool groupTT = useTT;
        int grid = forcedGrid ? forcedGrid : 8;
        int span = forcedSpan;
        if (plies != 1 && plies != 2) {
            TunedPlan c = chooseGroupPlan(depth, roots, states, useGraph, useTT, forcedGrid, forcedSpan);
            plies = c.plies;
            groupTT = c.useTT;
            grid = forcedGrid ? forcedGrid : c.gridPerCU;
            span = forcedSpan ? forcedSpan : c.ttLog2;
        }
        if (plies >= depth) plies = depth - 1;
        GroupPlan g{depth, plies, depth - plies, groupTT, grid, span};
        std::vector<u64> out(ids.size());
        PG_TRY(runGroupSplit(g_batch, g, roots.data(), states.data(), ids.size(), useGraph, out.data()));
        for (size_t k = 0; k < ids.size(); k++) counts[ids[k]] = out[k];
    }
    return PG_OK;
}

}  // namespace pg

using namespace pg;

extern "C" {
void pg_set_leaf_plies(int plies)
{
    if (plies == 1 || plies == 2) g_leafPlies = plies;
}

int pg_batch_tuned(int *rows, int maxRows, int *nOut)
{
    if (!nOut || (maxRows > 0 && !rows)) return PG_ERR_INVALID;
    std::lock_guard<std::mutex> lock(lib().mu);
    int n = 0;
    for (auto &kv : g_batch.tuned) {
        if (n >= maxRows) break;
        int *r = rows + 7 * n++;
        r[0] = std::get<0>(kv.first);
        r[1] = std::get<1>(kv.first);
        r[2] = std::get<2>(kv.first);
        r[3] = kv.second.plies;
        r[4] = kv.second.useTT ? 1 : 0;
        r[5] = kv.second.gridPerCU;
        r[6] = kv.second.useTT && lib().tt.entries ? log2Floor(ttSpanMask(kv.second.ttLog2) + 1) : 0;
    }
    *nOut = n;
    return PG_OK;
}

int pg_perft_batch(int n, const pg_position *positions, const uint8_t *depths, const pg_batch_options *opt,
                   uint64_t *counts)
{
    if (n < 0 || (n && (!positions || !depths || !counts))) return PG_ERR_INVALID;
    for (int i = 0; i < n; i++) PG_TRY(validatePosition(&positions[i]));
    if (n == 0) return PG_OK;
    return perftBatch(n, positions, depths, opt, counts);
}

}
