// pg_host.hip -- host side of the engine: library state, FEN, workload
// generator, the launcher (host recursion above the launch depth with the
// host TT tier and OOM fallback, launcher.cu:533-737), the -cpu path
// (launcher.cu:739-816) and the per-position batch baseline.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pg_internal.h"
#include "pg_launcher.h"
#include "pg_tt.h"

namespace pg {

static thread_local std::string g_err;
void setError(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

Lib &lib()
{
    static Lib L;
    return L;
}

// Polls a pinned slot until a kernel of stream `st` publishes a value there.
// After 20 ms it synchronises the stream instead, which reports a failed
// kernel rather than spinning on a value that never comes.
int waitPublished(volatile uint64_t *slot, hipStream_t st, uint64_t &v)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 1;; i++) {
        v = *slot;
        if (v != HOST_SLOT_EMPTY) return PG_OK;
        if (!(i & 255) && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
        __builtin_ia32_pause();
    }
    PG_HIP(hipStreamSynchronize(st));
    v = *slot;
    if (v == HOST_SLOT_EMPTY) {
        setError("a kernel did not publish its level total");
        return PG_ERR_HIP;
    }
    return PG_OK;
}

int refuseWhileServing(const char *what)
{
    if (!lib().serving.load()) return PG_OK;
    setError("%s: a server owns the arena and the device table until pg_serve_destroy", what);
    return PG_ERR_BUSY;
}

int validatePosition(const pg_position *p)
{
    if (!p) {
        setError("position is NULL");
        return PG_ERR_INVALID;
    }
    const u64 b1 = p->bb[1], b2 = p->bb[2], b3 = p->bb[3], occ = b1 | b2 | b3;
    const u64 kings = b2 & b3 & ~b1;
    const char *why = nullptr;
    if (p->color > 1) why = "side to move is not 0 or 1";
    else if (b1 & b2 & b3) why = "piece code 7 on a square";
    else if (p->bb[0] & ~occ) why = "black pieces outside the occupied squares";
    else if (popc(kings & p->bb[0]) != 1 || popc(kings & ~p->bb[0]) != 1) why = "not exactly one king per side";
    if (!why) return PG_OK;
    setError("invalid position: %s", why);
    return PG_ERR_INVALID;
}

// launcher state (launcher.cu:565-568)
static int g_effectiveLD = 0;
static std::atomic<uint64_t> g_oomCount{0};
static size_t g_lastPeak = 0;
static size_t g_peakSinceReset = 0;  // largest g_lastPeak since resetPeakSinceReset
static uint64_t g_arenaLimit = 0;  // 0 = whole arena; a test hook to force OOM
int g_leafPlies = 2;  // fused two-ply leaf (config 2 perft 7: 5.93 ms, 6.95 ms with the one-ply leaf)
static bool g_frontierBatch = true;  // launcher: batch the frontier above the launch depth
// launcher with the tables: the reference's serial recursion, the unique
// frontier over parallel host threads, or the unique frontier through the
// batched engine with the device table (pg_set_tt_frontier)
enum { TT_FRONTIER_SERIAL = 0, TT_FRONTIER_PARALLEL = 1, TT_FRONTIER_BATCH = 2 };
static int g_ttFrontier = TT_FRONTIER_SERIAL;
static const int FRONTIER_MAX_PLIES = 5;
static const int FRONTIER_AUTO_MIN_DEPTH = 7;     // see perftLaunch
static const double FRONTIER_AUTO_MAX_POSITIONS = 8e6;  // bf^plies bound of an automatic frontier

static inline QuadBitBoard toQBB(const pg_position *p)
{
    QuadBitBoard q;
    memcpy(q.bb, p->bb, 32);
    return q;
}

// ------------------------------------------------------------------ FEN
// readFENString semantics (util.cpp:7-95): '/' or '\' separate ranks, the
// clocks are ignored, an unknown piece letter is skipped.
static int parseFen(const char *fen, pg_position *pos)
{
    if (!fen || !pos) return PG_ERR_INVALID;
    memset(pos, 0, sizeof *pos);
    int i = 0, row = 0, col = 0;
    bool done = false;
    for (; fen[i]; i++) {
        char c = fen[i];
        if (c == '/' || c == '\\') {
            row++;
            col = 0;
        } else if (c >= '1' && c <= '8') {
            col += c - '0';
        } else {
            static const char *P = "pnbrqk";
            const char *f = strchr(P, c >= 'A' && c <= 'Z' ? c + 32 : c);
            if (f && c != ' ' && *f) {
                int piece = (int)(f - P) + 1;
                int sq = (7 - row) * 8 + col;
                if (sq < 0 || sq > 63) return PG_ERR_INVALID;
                u64 b = 1ULL << sq;
                if (c >= 'a') pos->bb[0] |= b;
                if (piece & 1) pos->bb[1] |= b;
                if (piece & 2) pos->bb[2] |= b;
                if (piece & 4) pos->bb[3] |= b;
                col++;
            }
        }
        if (row >= 7 && col == 8) {
            done = true;
            break;
        }
    }
    if (!done) return PG_ERR_INVALID;
    i++;
    while (fen[i] == ' ') i++;
    pos->color = (fen[i] == 'b' || fen[i] == 'B') ? BLACK : WHITE;
    if (fen[i]) i++;
    while (fen[i] == ' ') i++;
    while (fen[i] && fen[i] != ' ') {
        switch (fen[i]) {
        case 'K': pos->state |= 1; break;
        case 'Q': pos->state |= 2; break;
        case 'k': pos->state |= 4; break;
        case 'q': pos->state |= 8; break;
        }
        i++;
    }
    while (fen[i] == ' ') i++;
    if (fen[i] >= 'a' && fen[i] <= 'h') pos->state |= (u8)((fen[i] - 'a' + 1) << 4);
    // exactly one king per side is required by the move generator
    return validatePosition(pos);
}

static int writeFen(const pg_position *pos, char *out, int outLen)
{
    std::string s;
    QuadBitBoard q = toQBB(pos);
    for (int r = 7; r >= 0; r--) {
        int run = 0;
        for (int f = 0; f < 8; f++) {
            int sq = r * 8 + f;
            int pc = pieceAt(q, sq);
            if (!pc) {
                run++;
                continue;
            }
            if (run) s += (char)('0' + run), run = 0;
            char ch = ".PNBRQK"[pc];
            if (q.bb[0] >> sq & 1) ch = (char)(ch + 32);
            s += ch;
        }
        if (run) s += (char)('0' + run);
        if (r) s += '/';
    }
    s += pos->color == WHITE ? " w " : " b ";
    if (!(pos->state & 15)) s += '-';
    if (pos->state & 1) s += 'K';
    if (pos->state & 2) s += 'Q';
    if (pos->state & 4) s += 'k';
    if (pos->state & 8) s += 'q';
    s += ' ';
    int ep = pos->state >> 4;
    bool wrote = false;
    if (ep) {
        int f = ep - 1;
        u64 cap = 1ULL << (pos->color == WHITE ? 32 + f : 24 + f);
        u64 adj = ((cap << 1) & NOT_A) | ((cap >> 1) & NOT_H);
        u64 occ = q.bb[1] | q.bb[2] | q.bb[3];
        u64 mine = pos->color == WHITE ? occ & ~q.bb[0] : occ & q.bb[0];
        u64 pawns = q.bb[1] & ~q.bb[2] & ~q.bb[3];
        if (adj & pawns & mine) {
            s += (char)('a' + f);
            s += pos->color == WHITE ? '6' : '3';
            wrote = true;
        }
    }
    if (!wrote) s += '-';
    s += " 0 1";
    if ((int)s.size() + 1 > outLen) return PG_ERR_INVALID;
    memcpy(out, s.c_str(), s.size() + 1);
    return PG_OK;
}

// ------------------------------------------------------- workload generator
struct SplitMix {
    u64 s;
    u64 next()
    {
        u64 z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // Lemire, "Fast random integer generation in an interval" (2019)
    u32 below(u32 n)
    {
        unsigned __int128 m = (unsigned __int128)next() * n;
        u64 l = (u64)m;
        if (l < n) {
            u64 t = (0 - (u64)n) % n;
            while (l < t) {
                m = (unsigned __int128)next() * n;
                l = (u64)m;
            }
        }
        return (u32)(m >> 64);
    }
};

static int canonKey(u16 m)
{
    int fl = m >> 12;
    return ((m & 63) * 64 + ((m >> 6) & 63)) * 8 + ((fl & 8) ? (fl & 3) + 1 : 0);
}

static const char *START_FEN = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1";

static void walk(SplitMix &rng, int plies, int requireNonTerminal, pg_position *pos)
{
    for (;;) {
        parseFen(START_FEN, pos);
        bool ok = true;
        for (int p = 0; p < plies; p++) {
            QuadBitBoard q = toQBB(pos);
            u16 mv[MAX_MOVES];
            u32 n = generateMoves(q, pos->state, pos->color, mv);
            if (!n) {
                ok = false;
                break;
            }
            std::sort(mv, mv + n, [](u16 a, u16 b) { return canonKey(a) < canonKey(b); });
            u16 m = mv[rng.below(n)];
            makeMove(q, pos->state, pos->color, m);
            memcpy(pos->bb, q.bb, 32);
            pos->color ^= 1;
        }
        if (ok && requireNonTerminal && countMoves(toQBB(pos), pos->state, pos->color) == 0) ok = false;
        if (ok) return;
    }
}

// ---------------------------------------------------------------- init
int initLib(int device, uint64_t arenaBytes)
{
    Lib &L = lib();
    if (L.init) return PG_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        setError("no GPU visible");
        return PG_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        setError("device %d out of range (%d visible)", device, ndev);
        return PG_ERR_INVALID;
    }
    PG_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PG_HIP(hipGetDeviceProperties(&prop, device));
    L.cuCount = prop.multiProcessorCount;
    size_t freeB = 0, totalB = 0;
    PG_HIP(hipMemGetInfo(&freeB, &totalB));
    if (arenaBytes == 0) {
        arenaBytes = (uint64_t)(freeB * 0.6);
        const uint64_t cap = 192ull << 30;
        if (arenaBytes > cap) arenaBytes = cap;
    }
    arenaBytes &= ~(uint64_t)255;
    PG_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    PG_HIP(hipMalloc(&L.arena, arenaBytes));
    PG_HIP(hipHostMalloc((void **)&L.hostPinned, PINNED_BYTES, PINNED_FLAGS));
    // the device address of this GPU's static scan workspace (pg_scan_u32_static)
    PG_TRY(initStaticScratch(device));
    L.arenaBytes = arenaBytes;
    L.device = device;
    L.init = true;
    return PG_OK;
}

int shutdownLib()
{
    Lib &L = lib();
    if (!L.init) return PG_OK;
    ttFree();
    (void)hipStreamSynchronize(L.stream);
    (void)hipFree(L.arena);
    (void)hipHostFree(L.hostPinned);
    (void)hipStreamDestroy(L.stream);
    L.arena = nullptr;
    L.hostPinned = nullptr;
    L.stream = nullptr;
    L.init = false;
    return PG_OK;
}

// ------------------------------------------------------- GPU BFS (one call)
// Drives bfsPerft on the library arena (or a caller buffer).  The GPU path
// never computes on the host except depth <= 1 and the root move list.
int gpuBfs(const pg_position *pos, int depth, bool useTT, void *buf, uint64_t bytes, uint64_t *count)
{
    Lib &L = lib();
    if (!L.init) {
        setError("pg_init has not run");
        return PG_ERR_NOT_INIT;
    }
    std::lock_guard<std::mutex> g(L.mu);
    // checked under the lock: pg_serve_create sets the flag, then takes the
    // lock, so a call that holds the lock first finishes before any slot runs
    PG_TRY(refuseWhileServing("GPU perft"));
    size_t peak = 0;
    size_t avail = L.arenaBytes;
    if (g_arenaLimit && g_arenaLimit < avail) avail = g_arenaLimit;
    int s = bfsPerft(toQBB(pos), pos->state, pos->color, depth, g_leafPlies, useTT, buf ? buf : L.arena,
                     buf ? bytes : avail, L.stream, L.hostPinned, count, &peak);
    g_lastPeak = peak;
    if (peak > g_peakSinceReset) g_peakSinceReset = peak;
    return s;
}

// -------------------------------------------------------------- launcher
// perft_cpu_recurse (launcher.cu:572-715): host TT probe, GPU call at or
// below the effective launch depth, CPU split one ply down on OOM (which
// also lowers the effective launch depth), host TT store of 64-bit results.
// Host frontier feeding the device (the reference recurses serially above the
// launch depth and makes one GPU call per node, launcher.cu:572-715): expand
// `depth - ld` plies on the host, merge identical positions with their
// multiplicities (exact: the whole position is the key), and run every unique
// frontier position through the batched engine in one pass.  This is the
// default without the tables.  With them the serial recursion stays the
// default: each GPU call finds the subtrees the calls before it stored, and
// that beat both frontier forms on the startpos TT runs (launch depth 9;
// perft 10 / 11: serial 0.685 / 8.47 s, the frontier over four parallel slots
// 0.770 / 8.77 s, the frontier through the batched engine with the table
// 0.831 / 11.11 s).
struct PosKey {
    u64 bb[4];
    u8 state, color;
    bool operator==(const PosKey &o) const
    {
        return bb[0] == o.bb[0] && bb[1] == o.bb[1] && bb[2] == o.bb[2] && bb[3] == o.bb[3] && state == o.state &&
               color == o.color;
    }
};
struct PosKeyHash {
    size_t operator()(const PosKey &k) const
    {
        u64 h = (u64)k.state | ((u64)k.color << 8);
        for (int i = 0; i < 4; i++) h = (h ^ k.bb[i]) * 0x9E3779B97F4A7C15ULL + (h >> 29);
        return (size_t)(h ^ (h >> 31));
    }
};

using PosMap = std::unordered_map<PosKey, u64, PosKeyHash>;

// the children of `parents[b, e)` (with the parents' multiplicities)
template <class Sink>
static void expandRange(const std::vector<std::pair<PosKey, u64>> &parents, size_t b, size_t e, Sink &&sink)
{
    for (size_t i = b; i < e; i++) {
        const PosKey &k = parents[i].first;
        QuadBitBoard q;
        memcpy(q.bb, k.bb, 32);
        u16 mv[MAX_MOVES];
        u32 n = generateMoves(q, k.state, k.color, mv);
        for (u32 j = 0; j < n; j++) {
            PosKey c;
            QuadBitBoard cq = q;
            c.state = k.state;
            makeMove(cq, c.state, k.color, mv[j]);
            memcpy(c.bb, cq.bb, 32);
            c.color = k.color ^ 1;
            sink(c, parents[i].second);
        }
    }
}

// One ply of the frontier over the host threads: each thread expands a slice
// of the parents into per-shard buffers (the shard from the key's hash), then
// each thread merges one shard.  Startpos's fifth ply (4,865,609 children of
// 72,078 parents, 822,518 unique) took 0.78 s in one hash map on one thread.
static std::vector<std::pair<PosKey, u64>> expandPlyParallel(const std::vector<std::pair<PosKey, u64>> &parents)
{
    const unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<std::vector<std::pair<PosKey, u64>>>> buf(T, std::vector<std::vector<std::pair<PosKey, u64>>>(T));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const size_t b = parents.size() * t / T, e = parents.size() * (t + 1) / T;
            expandRange(parents, b, e, [&](const PosKey &c, u64 m) {
                buf[t][(PosKeyHash()(c) >> 40) % T].emplace_back(c, m);
            });
        });
    for (auto &x : th) x.join();
    th.clear();
    std::vector<std::vector<std::pair<PosKey, u64>>> out(T);
    for (unsigned sh = 0; sh < T; sh++)
        th.emplace_back([&, sh] {
            size_t n = 0;
            for (unsigned t = 0; t < T; t++) n += buf[t][sh].size();
            PosMap m;
            m.reserve(n);
            for (unsigned t = 0; t < T; t++) {
                for (auto &kv : buf[t][sh]) m[kv.first] += kv.second;
                std::vector<std::pair<PosKey, u64>>().swap(buf[t][sh]);
            }
            out[sh].assign(m.begin(), m.end());
        });
    for (auto &x : th) x.join();
    std::vector<std::pair<PosKey, u64>> all;
    size_t n = 0;
    for (auto &o : out) n += o.size();
    all.reserve(n);
    for (auto &o : out) all.insert(all.end(), o.begin(), o.end());
    return all;
}

// The unique positions `plies` plies below `pos` and how many move sequences
// reach each (exact: the whole position is the key).  Plies of more than 2^16
// parents run over the host threads.
static void expandFrontier(const pg_position *pos, int plies, std::vector<pg_position> &ps, std::vector<u64> &mult)
{
    PosKey k0;
    memcpy(k0.bb, pos->bb, 32);
    k0.state = pos->state;
    k0.color = pos->color;
    std::vector<std::pair<PosKey, u64>> cur{{k0, 1}};
    for (int ply = 0; ply < plies; ply++) {
        if (cur.size() > (1u << 16)) {
            cur = expandPlyParallel(cur);
            continue;
        }
        PosMap next;
        next.reserve(cur.size() * 32);
        expandRange(cur, 0, cur.size(), [&](const PosKey &c, u64 m) { next[c] += m; });
        cur.assign(next.begin(), next.end());
    }
    ps.clear();
    mult.clear();
    ps.reserve(cur.size());
    mult.reserve(cur.size());
    for (auto &kv : cur) {
        pg_position p{};
        memcpy(p.bb, kv.first.bb, 32);
        p.state = kv.first.state;
        p.color = kv.first.color;
        ps.push_back(p);
        mult.push_back(kv.second);
    }
}

static void sumFrontier(const std::vector<u64> &mult, const std::vector<u64> &counts, u128 *out)
{
    unsigned __int128 total = 0;
    for (size_t i = 0; i < mult.size(); i++) total += (unsigned __int128)mult[i] * counts[i];
    out->lo = (u64)total;
    out->hi = (u64)(total >> 64);
}

// the unique frontier through the batched engine in one pass (with the
// device table when useTT)
static int frontierPerft(const pg_position *pos, int depth, int ld, bool useTT, u128 *out)
{
    std::vector<pg_position> ps;
    std::vector<u64> mult;
    expandFrontier(pos, depth - ld, ps, mult);
    std::vector<u8> ds(ps.size(), (u8)ld);
    std::vector<u64> counts(ps.size());
    pg_batch_options opt{};
    opt.useGraph = 1;
    opt.useTT = useTT ? 1 : 0;
    const size_t CH = 1 << 20;
    for (size_t b = 0; b < ps.size(); b += CH) {
        int n = (int)std::min(CH, ps.size() - b);
        PG_TRY(perftBatch(n, ps.data() + b, ds.data() + b, &opt, counts.data() + b));
    }
    sumFrontier(mult, counts, out);
    return PG_OK;
}

// With the tables: the unique frontier at the launch depth (host tier probed
// first) spread over TT_SLOTS host threads, each with its own HIP stream,
// pinned readback buffer and 1/TT_SLOTS of the arena, all sharing the device
// table and the host tier, so the GPU calls of several frontier nodes overlap
// on the device.  A node that does not fit its slot is recomputed afterwards
// through the serial recursion on the whole arena (which lowers the launch
// depth as the reference does).
static const int TT_SLOTS = 4;
struct FrontierSlot {
    hipStream_t stream = nullptr;
    u64 *pinned = nullptr;
};
static FrontierSlot g_ttSlots[TT_SLOTS];

static int recurse(const pg_position *pos, int depth, bool useTT, u128 *out);

static int frontierPerftParallel(const pg_position *pos, int depth, int ld, u128 *out)
{
    Lib &L = lib();
    std::vector<pg_position> ps;
    std::vector<u64> mult;
    expandFrontier(pos, depth - ld, ps, mult);
    const size_t n = ps.size();
    std::vector<u64> counts(n, 0);
    std::vector<u8> state(n, 0);  // 0 = to compute, 1 = done, 2 = did not fit its slot
    std::vector<Hash128> hs(n);
    for (size_t i = 0; i < n; i++) {
        hs[i] = hashPosition(toQBB(&ps[i]), ps[i].state, ps[i].color);
        uint64_t v;
        if (ld >= 2 && hostTTProbe(hs[i], ld, &v)) counts[i] = v, state[i] = 1;
    }
    {
        std::lock_guard<std::mutex> g(L.mu);
        PG_TRY(refuseWhileServing("GPU perft"));
        for (auto &s : g_ttSlots) {
            if (!s.stream) PG_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            if (!s.pinned) PG_HIP(hipHostMalloc((void **)&s.pinned, PINNED_BYTES, PINNED_FLAGS));
        }
        const bool profOn = L.prof.on;
        L.prof.on = false;  // the event profiler is single-threaded
        const size_t per = (usableArena() / TT_SLOTS) & ~(size_t)255;
        std::atomic<size_t> next{0};
        std::atomic<int> err{PG_OK};
        auto work = [&](int k) {
            (void)hipSetDevice(L.device);
            char *arena = (char *)L.arena + (size_t)k * per;
            for (size_t i; (i = next.fetch_add(1)) < n && err.load() == PG_OK;) {
                if (state[i]) continue;
                const pg_position &p = ps[i];
                size_t peak = 0;
                uint64_t c = 0;
                int s = bfsPerft(toQBB(&p), p.state, p.color, ld, g_leafPlies, true, arena, per, g_ttSlots[k].stream,
                                 g_ttSlots[k].pinned, &c, &peak);
                if (s == PG_ERR_ARENA_OOM) {
                    state[i] = 2;
                } else if (s != PG_OK) {
                    int expect = PG_OK;
                    err.compare_exchange_strong(expect, s);
                } else {
                    counts[i] = c;
                    state[i] = 1;
                    if (ld >= 2) hostTTStore(hs[i], ld, c);
                }
            }
        };
        std::vector<std::thread> th;
        for (int k = 0; k < TT_SLOTS; k++) th.emplace_back(work, k);
        for (auto &t : th) t.join();
        L.prof.on = profOn;
        if (err.load() != PG_OK) return err.load();
    }
    for (size_t i = 0; i < n; i++) {
        if (state[i] != 2) continue;
        g_oomCount++;
        u128 r;
        PG_TRY(recurse(&ps[i], ld, true, &r));
        counts[i] = r.lo;
    }
    sumFrontier(mult, counts, out);
    return PG_OK;
}

static int recurse(const pg_position *pos, int depth, bool useTT, u128 *out)
{
    Hash128 h{0, 0};
    if (useTT && depth >= 2) {
        h = hashPosition(toQBB(pos), pos->state, pos->color);
        uint64_t v;
        if (hostTTProbe(h, depth, &v)) {
            *out = u128(v);
            return PG_OK;
        }
    }
    u128 total;
    bool split = depth > g_effectiveLD;
    if (!split) {
        uint64_t c = 0;
        int s = gpuBfs(pos, depth, useTT, nullptr, 0, &c);
        if (s == PG_ERR_ARENA_OOM) {
            g_oomCount++;
            if (g_effectiveLD > 2 && depth >= g_effectiveLD) {
                printf("  >> OOM at GPU depth %d, decreasing effective LD: %d -> %d\n", depth, g_effectiveLD,
                       g_effectiveLD - 1);
                g_effectiveLD--;
            }
            split = true;
        } else if (s != PG_OK) {
            return s;
        } else {
            total = u128(c);
        }
    }
    const int fld = std::max(1, std::min(g_effectiveLD, depth - 1));
    const bool frontier = split && depth > 1 && depth - fld <= FRONTIER_MAX_PLIES;
    if (frontier && !useTT && g_frontierBatch) {
        PG_TRY(frontierPerft(pos, depth, fld, false, &total));
    } else if (frontier && useTT && g_ttFrontier == TT_FRONTIER_PARALLEL && fld >= 2) {
        PG_TRY(frontierPerftParallel(pos, depth, fld, &total));
    } else if (frontier && useTT && g_ttFrontier == TT_FRONTIER_BATCH) {
        PG_TRY(frontierPerft(pos, depth, fld, true, &total));
    } else if (split) {
        if (depth <= 1) {
            // a GPU call of depth 1 cannot run out of memory; depth 0 is a leaf
            total = u128(depth == 0 ? 1 : countMoves(toQBB(pos), pos->state, pos->color));
        } else {
            u16 mv[MAX_MOVES];
            QuadBitBoard q = toQBB(pos);
            u32 n = generateMoves(q, pos->state, pos->color, mv);
            for (u32 i = 0; i < n; i++) {
                pg_position c = *pos;
                QuadBitBoard cq = q;
                makeMove(cq, c.state, pos->color, mv[i]);
                memcpy(c.bb, cq.bb, 32);
                c.color ^= 1;
                u128 sub;
                PG_TRY(recurse(&c, depth - 1, useTT, &sub));
                total += sub;
            }
        }
    }
    if (useTT && depth >= 2 && total.hi == 0) hostTTStore(h, depth, total.lo);
    *out = total;
    return PG_OK;
}

int perftLaunch(const pg_position *pos, int depth, int launchDepth, bool useTT, u128 *out)
{
    if (depth < 0 || depth > 32) {
        setError("depth %d out of range", depth);
        return PG_ERR_INVALID;
    }
    if (launchDepth <= 0) {
        float bf;
        PG_TRY(estimateLD(pos, &bf, &launchDepth));
        // Without the tables the nodes above the launch depth are expanded on
        // the host, identical positions merged (with their multiplicities) and
        // the rest run through the batched engine: a deeper frontier merges
        // more and leaves less to compute.  From depth 7 the automatic launch
        // depth leaves a frontier of four plies, five from depth 10 (launch
        // depth max(4, depth - plies)), fewer where bf^plies would pass 8
        // million positions: startpos perft 9 / 10 took 0.47 / 6.35 s against
        // 1.18 / 35.3 s at the memory-based launch depth (13.5 s for perft 10
        // with four plies), Kiwipete perft 7 / 8 0.079 / 2.5 s against
        // 0.122 / 5.8 s, Position 3 perft 10 0.31 s against 0.74 s; a
        // four-ply frontier at depth 7 was slower than three.
        if (!useTT && g_frontierBatch && depth >= FRONTIER_AUTO_MIN_DEPTH) {
            int plies = depth >= 10 ? 5 : 4;
            while (plies > 1 && pow((double)bf, plies) > FRONTIER_AUTO_MAX_POSITIONS) plies--;
            launchDepth = std::min(launchDepth, std::max(4, depth - plies));
        }
    }
    g_effectiveLD = std::min(launchDepth, depth);
    if (g_effectiveLD < 1) g_effectiveLD = 1;
    if (depth == 0) {
        *out = u128(1);
        return PG_OK;
    }
    return recurse(pos, depth, useTT, out);
}

// estimateLaunchDepth (launcher.cu:533-562): branching factor from perft(1..3)
// on the host, "unstable" positions get 5.  The memory model is this engine's:
// a 2-ply-leaf BFS of depth d holds ~perft(d-2) leaf entries of 6 B plus the
// boards of the level above (~38 B / bf per leaf entry), and must fit half of
// the arena.
int estimateLD(const pg_position *pos, float *bfOut, int *ldOut)
{
    uint64_t p1, p2, p3;
    PG_TRY(perftCpu(pos, 1, false, &p1));
    PG_TRY(perftCpu(pos, 2, false, &p2));
    PG_TRY(perftCpu(pos, 3, false, &p3));
    if (!p1 || !p2 || !p3) {
        *bfOut = 1;
        *ldOut = 5;
        return PG_OK;
    }
    double geo = sqrt(((double)p3 / p2) * ((double)p2 / p1));
    double ari = (((double)p3 / p2) + ((double)p2 / p1)) / 2;
    double bf = (geo + ari) / 2;
    *bfOut = (float)bf;
    if (ari / geo > 2.0 || bf < 2.0) {
        *ldOut = 5;
        return PG_OK;
    }
    double arena = lib().arenaBytes ? (double)lib().arenaBytes : (double)(16ull << 30);
    double perEntry = 6.0 + 38.0 / bf;
    int d = (int)(log(arena / 2.0 / perEntry) / log(bf)) + 2;
    *ldOut = d < 1 ? 1 : d;
    return PG_OK;
}

// -------------------------------------------------------------- -cpu path
// perft_cpu<> (launcher.cu:762-805): generate/make/recurse, countMoves at
// depth 1, host TT for depth >= 2 when enabled.  Single-threaded.
template <int C>
static uint64_t cpuRec(const QuadBitBoard &q, u8 st, int depth, bool useTT)
{
    if (depth == 1) return countMoves<C>(q, st);
    Hash128 h{0, 0};
    if (useTT) {
        h = hashPosition(q, st, C);
        uint64_t v;
        if (hostTTProbe(h, depth, &v)) return v;
    }
    uint64_t total = 0;
    auto child = [&](u32, u16 m) {
        QuadBitBoard c = q;
        u8 s2 = st;
        makeMove<C>(c, s2, m);
        total += cpuRec<1 - C>(c, s2, depth - 1, useTT);
    };
    generateMoves<C>(q, st, child);
    if (useTT) hostTTStore(h, depth, total);
    return total;
}

int perftCpu(const pg_position *pos, int depth, bool useTT, uint64_t *out)
{
    if (depth < 0 || depth > 32) return PG_ERR_INVALID;
    if (depth == 0) {
        *out = 1;
        return PG_OK;
    }
    QuadBitBoard q = toQBB(pos);
    *out = pos->color == WHITE ? cpuRec<WHITE>(q, pos->state, depth, useTT) : cpuRec<BLACK>(q, pos->state, depth, useTT);
    return PG_OK;
}

int getEffectiveLD() { return g_effectiveLD; }
void setEffectiveLD(int ld) { g_effectiveLD = ld; }
uint64_t getOomFallbackCount() { return g_oomCount.load(); }
size_t lastBfsPeak() { return g_lastPeak; }
size_t peakSinceReset() { return g_peakSinceReset; }
void resetPeakSinceReset() { g_peakSinceReset = 0; }
size_t usableArena()
{
    size_t a = lib().arenaBytes;
    return (g_arenaLimit && g_arenaLimit < a) ? g_arenaLimit : a;
}

}  // namespace pg

using namespace pg;

// ================================================================ C ABI
extern "C" {

const char *pg_last_error(void) { return g_err.c_str(); }

int pg_init(int device, uint64_t arenaBytes) { return initLib(device, arenaBytes); }
int pg_shutdown(void)
{
    PG_TRY(refuseWhileServing("pg_shutdown"));
    return shutdownLib();
}
uint64_t pg_arena_bytes(void) { return lib().arenaBytes; }
void pg_set_arena_limit(uint64_t bytes) { g_arenaLimit = bytes; }
void pg_set_frontier_batch(int on) { g_frontierBatch = on != 0; }
int pg_set_tt_frontier(int mode)
{
    if (mode < TT_FRONTIER_SERIAL || mode > TT_FRONTIER_BATCH) {
        setError("pg_set_tt_frontier: mode %d is not 0, 1 or 2", mode);
        return PG_ERR_INVALID;
    }
    g_ttFrontier = mode;
    return PG_OK;
}

int pg_parse_fen(const char *fen, pg_position *pos) { return parseFen(fen, pos); }
int pg_write_fen(const pg_position *pos, char *out, int outLen)
{
    if (!out) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    return writeFen(pos, out, outLen);
}

int pg_gen_moves(const pg_position *pos, uint16_t *moves, int *n)
{
    if (!moves || !n) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    *n = (int)generateMoves(toQBB(pos), pos->state, pos->color, moves);
    return PG_OK;
}

int pg_count_moves(const pg_position *pos, int *n)
{
    if (!n) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    *n = (int)countMoves(toQBB(pos), pos->state, pos->color);
    return PG_OK;
}

int pg_make_move(pg_position *pos, uint16_t move)
{
    PG_TRY(validatePosition(pos));
    QuadBitBoard q = toQBB(pos);
    makeMove(q, pos->state, pos->color, move);
    memcpy(pos->bb, q.bb, 32);
    pos->color ^= 1;
    return PG_OK;
}

int pg_random_position(uint64_t seed, int plies, int requireNonTerminal, pg_position *pos)
{
    if (!pos || plies < 0 || plies > 1000) return PG_ERR_INVALID;
    SplitMix rng{seed};
    walk(rng, plies, requireNonTerminal, pos);
    return PG_OK;
}

int pg_random_request(uint64_t seed, int pliesLo, int pliesHi, int depthLo, int depthHi, pg_position *pos, int *plies,
                      int *depth)
{
    if (!pos || !plies || !depth || pliesLo < 0 || pliesHi < pliesLo || depthHi < depthLo) return PG_ERR_INVALID;
    SplitMix rng{seed};
    *plies = pliesLo + (int)rng.below((u32)(pliesHi - pliesLo + 1));
    *depth = depthLo + (int)rng.below((u32)(depthHi - depthLo + 1));
    walk(rng, *plies, 1, pos);
    return PG_OK;
}

int pg_perft_bfs(const pg_position *pos, int depth, int useTT, void *gpuBuffer, uint64_t bufferBytes, uint64_t *count)
{
    if (!count || depth < 0 || depth > 20) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    return gpuBfs(pos, depth, useTT != 0, gpuBuffer, bufferBytes, count);
}

int pg_perft(const pg_position *pos, int depth, int launchDepth, int useTT, pg_u128 *count)
{
    if (!count) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    if (!lib().init) return PG_ERR_NOT_INIT;
    PG_TRY(refuseWhileServing("pg_perft"));
    u128 r;
    int s = perftLaunch(pos, depth, launchDepth, useTT != 0, &r);
    count->lo = r.lo;
    count->hi = r.hi;
    return s;
}

int pg_perft_cpu(const pg_position *pos, int depth, int useTT, uint64_t *count)
{
    if (!count) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    return perftCpu(pos, depth, useTT != 0, count);
}

int pg_estimate_launch_depth(const pg_position *pos, float *bf, int *ld)
{
    if (!bf || !ld) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    return estimateLD(pos, bf, ld);
}

int pg_get_effective_ld(void) { return getEffectiveLD(); }
void pg_set_effective_ld(int ld) { setEffectiveLD(ld); }
uint64_t pg_get_oom_fallback_count(void) { return getOomFallbackCount(); }
uint64_t pg_last_bfs_peak_bytes(void) { return lastBfsPeak(); }

int pg_perft_batch_plain(int n, const pg_position *positions, const uint8_t *depths, int useTT, uint64_t *counts)
{
    if (n < 0 || (n && (!positions || !depths || !counts))) return PG_ERR_INVALID;
    for (int i = 0; i < n; i++) {
        PG_TRY(validatePosition(&positions[i]));
        if (depths[i] > 20) {
            setError("pg_perft_batch_plain: depth %d out of range", depths[i]);
            return PG_ERR_INVALID;
        }
    }
    if (n) PG_TRY(refuseWhileServing("pg_perft_batch_plain"));
    for (int i = 0; i < n; i++) {
        u128 r;
        // one GPU BFS call per position; a position too big for the arena
        // is split on the host like the launcher does
        int s = gpuBfs(&positions[i], depths[i], useTT != 0, nullptr, 0, &counts[i]);
        if (s == PG_ERR_ARENA_OOM) {
            int saved = g_effectiveLD;
            g_effectiveLD = depths[i] - 1;
            s = recurse(&positions[i], depths[i], useTT != 0, &r);
            g_effectiveLD = saved;
            counts[i] = r.lo;
        }
        if (s != PG_OK) return s;
    }
    return PG_OK;
}

}  // extern "C"
