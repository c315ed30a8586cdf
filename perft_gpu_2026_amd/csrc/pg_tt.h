// pg_tt.h -- position hash and transposition-table formats.
//
// Hash: a keyed 128-bit hash of the 33-byte position + side to move, computed
// from the board itself (each bitboard word through an independent
// murmur3 fmix64 finaliser, summed).  The reference keeps an incremental
// 128-bit Zobrist hash beside every BFS entry (zobrist.h:81-227, 16 B per
// entry per level); hashing from the board instead stores nothing per entry
// and costs ~1 % of a leaf entry's arithmetic.  Any hash gives identical
// counts; only collisions matter, and verification uses the full other 64 bits.
//
// Device TT (tt.h:16-70 and :118-155 redesigned): ONE table for all depths,
// 16-B entries {lock, data}, data = count << 8 | depth, lock = key ^ data
// (XOR lock: a torn 16-B read fails the check, so it reads as a miss; 128-bit
// accesses are single dwordx4 instructions but not architecturally atomic).
// Entries pair up in 32-B buckets, bucket = mix(hash.lo, depth); key =
// hash.hi.  The bucket's first entry keeps the deeper of what it holds and
// what is stored (its own key and ties replace it), the second takes every
// other store, so the count of a deep subtree survives the flood of depth-2
// leaf stores that an always-replace slot loses.  Depth-tagged so requests of
// different depths share the table (serving).  Counts >= 2^56 are not stored.
//
// Host TT (tt.h:165-232): lossless chained table -- a bump pool of entries and
// CAS-prepended bucket chains; never evicts, drops new stores once the pool is
// full.  Thread-safe.
#pragma once
#include "pg_chess.h"
#include "pg_internal.h"

namespace pg {

PG_HD u64 fmix64(u64 k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

PG_HD Hash128 hashPosition(const QuadBitBoard &q, u8 state, int color)
{
    u64 tag = (u64)state | ((u64)color << 8);
    Hash128 h;
    h.lo = fmix64(q.bb[0] ^ 0x243F6A8885A308D3ULL) + fmix64(q.bb[1] ^ 0x13198A2E03707344ULL) +
           fmix64(q.bb[2] ^ 0xA4093822299F31D0ULL) + fmix64(q.bb[3] ^ 0x082EFA98EC4E6C89ULL) +
           fmix64(tag ^ 0x452821E638D01377ULL);
    h.hi = fmix64(q.bb[0] ^ 0xBE5466CF34E90C6CULL) + fmix64(q.bb[1] ^ 0xC0AC29B7C97C50DDULL) +
           fmix64(q.bb[2] ^ 0x3F84D5B5B5470917ULL) + fmix64(q.bb[3] ^ 0x9216D5D98979FB1BULL) +
           fmix64(tag ^ 0xD1310BA698DFB5ACULL);
    return h;
}

// ---------------------------------------------------------------- device TT
static constexpr u64 TT_MAX_COUNT = (1ULL << 56) - 1;

// first entry of the bucket (mask = entries - 1, entries a power of two >= 2)
PG_HD u64 ttBucket(const Hash128 &h, int depth, u64 mask)
{
    return (h.lo ^ ((u64)depth * 0x9E3779B97F4A7C15ULL)) & mask & ~1ULL;
}

#if defined(__HIPCC__)
struct alignas(16) TTEntry {
    u64 lock, data;
};

__device__ __forceinline__ bool ttProbe(const TTEntry *tt, u64 mask, const Hash128 &h, int depth, u64 *count)
{
    if (!tt) return false;
    const TTEntry *b = tt + ttBucket(h, depth, mask);
    TTEntry e0 = b[0], e1 = b[1];  // one 32-B bucket
    bool h0 = ((e0.lock ^ e0.data) == h.hi) && ((e0.data & 0xFF) == (u64)depth);
    bool h1 = ((e1.lock ^ e1.data) == h.hi) && ((e1.data & 0xFF) == (u64)depth);
    *count = (h0 ? e0.data : e1.data) >> 8;
    return h0 || h1;
}

__device__ __forceinline__ void ttStore(TTEntry *tt, u64 mask, const Hash128 &h, int depth, u64 count)
{
    if (!tt || count > TT_MAX_COUNT) return;
    u64 data = (count << 8) | (u64)depth;
    TTEntry e{h.hi ^ data, data};
    TTEntry *b = tt + ttBucket(h, depth, mask);
    TTEntry e0 = b[0];
    u64 d0 = e0.data & 0xFF;
    bool own0 = ((e0.lock ^ e0.data) == h.hi) && d0 == (u64)depth;
    if (own0 || (u64)depth >= d0) b[0] = e;  // one 16-B store
    else b[1] = e;
}

// ----------------------------------------------------- within-level merging
// (write_dedup_table / check_duplicates / apply_redirects,
// perft_kernels.cu:486-552): every unknown entry of a level writes
// {hash.hi, index + 1} to a table of at least 2n slots, then (in the next
// kernel) reads its slot back.  The table is never cleared: every slot that
// is read was written by the entry reading it, so it holds one of this
// level's entries (both 8-B halves come from this level's writers), never
// what an earlier level or call left there.  An entry that finds another
// index with its key AND an identical
// position (boards compared, so a hash collision or a torn 16-B slot can
// never merge two different positions) becomes a duplicate with no children,
// and takes the original's value after the upsweep.
struct DedupEntry {
    u64 key, idx;
};

__device__ __forceinline__ u64 dedupSlot(const Hash128 &h, u64 tmask) { return (h.lo >> 17) & tmask; }

// the entry of the level that position (q, st) at index i merges into: i
// itself unless the table names an identical other entry
__device__ __forceinline__ u32 dedupTarget(const QuadBitBoard *__restrict__ b, const u8 *__restrict__ s, u64 n,
                                           const DedupEntry *__restrict__ tab, u64 tmask, u64 i,
                                           const QuadBitBoard &q, u8 st, const Hash128 &h)
{
    DedupEntry e = tab[dedupSlot(h, tmask)];
    u64 j = e.idx - 1;
    bool dup = e.key == h.hi && e.idx && j != i && j < n && s[j] == st && b[j].bb[0] == q.bb[0] &&
               b[j].bb[1] == q.bb[1] && b[j].bb[2] == q.bb[2] && b[j].bb[3] == q.bb[3];
    return dup ? (u32)j : (u32)i;
}
#endif

// ------------------------------------------------------------------ host TT
bool hostTTProbe(const Hash128 &h, int depth, uint64_t *count);
void hostTTStore(const Hash128 &h, int depth, uint64_t count);
int ttInit(uint64_t deviceMB, uint64_t hostMB);
int ttFree();
int ttClear();
int ttClearDevice();  // the device table only (clearDeviceTTs)
int ttStats(uint64_t out[6]);
void ttCountDevice(u64 probes, u64 hits, u64 stores);

}  // namespace pg
