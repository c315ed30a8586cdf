// pg_tt.hip -- transposition-table storage: the device table in HBM and the
// host lossless tier (initTT / freeTT / clearDeviceTTs / printTTStats,
// launcher.cu:128-531).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "pg_launcher.h"
#include "pg_tt.h"

namespace pg {

struct HostEntry {
    u64 key;    // hash.hi
    u64 lo;     // hash.lo (full verification)
    u64 count;
    u32 next;   // pool index, 0 = end of chain
    u32 depth;
};

struct HostTT {
    std::atomic<u32> *heads = nullptr;
    u64 mask = 0;
    HostEntry *pool = nullptr;
    u64 poolSize = 0;
    std::atomic<u64> used{1};
    std::atomic<u64> probes{0}, hits{0}, stores{0};
};
static HostTT g_host;
static std::atomic<u64> g_devProbes{0}, g_devHits{0}, g_devStores{0};

static inline u64 hostBucket(const Hash128 &h, int depth) { return (h.lo ^ ((u64)depth * 0xC2B2AE3D27D4EB4FULL)) & g_host.mask; }

bool hostTTProbe(const Hash128 &h, int depth, uint64_t *count)
{
    if (!g_host.heads) return false;
    g_host.probes.fetch_add(1, std::memory_order_relaxed);
    for (u32 i = g_host.heads[hostBucket(h, depth)].load(std::memory_order_acquire); i; i = g_host.pool[i].next) {
        const HostEntry &e = g_host.pool[i];
        if (e.key == h.hi && e.lo == h.lo && e.depth == (u32)depth) {
            *count = e.count;
            g_host.hits.fetch_add(1, std::memory_order_relaxed);
            return true;
        }
    }
    return false;
}

void hostTTStore(const Hash128 &h, int depth, uint64_t count)
{
    if (!g_host.heads) return;
    u64 idx = g_host.used.fetch_add(1, std::memory_order_relaxed);
    if (idx >= g_host.poolSize) return;  // pool full: drop (tt.h:211)
    HostEntry &e = g_host.pool[idx];
    e.key = h.hi;
    e.lo = h.lo;
    e.count = count;
    e.depth = (u32)depth;
    std::atomic<u32> &head = g_host.heads[hostBucket(h, depth)];
    u32 old = head.load(std::memory_order_relaxed);
    do {
        e.next = old;
    } while (!head.compare_exchange_weak(old, (u32)idx, std::memory_order_release, std::memory_order_relaxed));
    g_host.stores.fetch_add(1, std::memory_order_relaxed);
}

static void hostFree()
{
    free(g_host.heads);
    free(g_host.pool);
    g_host.heads = nullptr;
    g_host.pool = nullptr;
    g_host.mask = 0;
    g_host.poolSize = 0;
    g_host.used = 1;
}

static int hostInit(uint64_t hostMB)
{
    hostFree();
    if (hostMB == 0) return PG_OK;
    u64 bytes = hostMB << 20;
    // 1/8 of the budget for bucket heads, the rest for entries; pool index is u32
    u64 buckets = 1;
    while (buckets * 2 * sizeof(u32) <= bytes / 8) buckets *= 2;
    u64 entries = (bytes - buckets * sizeof(u32)) / sizeof(HostEntry);
    if (entries > 0xFFFFFFF0ull) entries = 0xFFFFFFF0ull;
    g_host.heads = (std::atomic<u32> *)calloc(buckets, sizeof(u32));
    g_host.pool = (HostEntry *)malloc(entries * sizeof(HostEntry));
    if (!g_host.heads || !g_host.pool) {
        hostFree();
        setError("host TT allocation of %llu MB failed", (unsigned long long)hostMB);
        fprintf(stderr, "warning: %s, host TT disabled\n", pg_last_error());
        return PG_OK;  // launcher.cu:339-345: warn and run without the table
    }
    g_host.mask = buckets - 1;
    g_host.poolSize = entries;
    return PG_OK;
}

static int deviceInit(uint64_t deviceMB)
{
    Lib &L = lib();
    if (L.tt.entries) {
        (void)hipFree(L.tt.entries);
        L.tt = DeviceTT{};
    }
    if (deviceMB == 0 || !L.init) return PG_OK;
    u64 bytes = deviceMB << 20;
    u64 slots = 1;
    while (slots * 2 * 16 <= bytes) slots *= 2;
    void *p = nullptr;
    // a table that does not fit beside the arena is halved until it does
    // (down to 1/16 of the request), as the reference's auto sizing takes
    // what is free (launcher.cu:146-160)
    const u64 want = slots;
    while (hipMalloc(&p, slots * 16) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        if (slots <= want / 16 || slots <= 65536) break;
        slots >>= 1;
    }
    if (!p) {
        fprintf(stderr, "warning: device TT allocation of %llu MB failed, device TT disabled\n",
                (unsigned long long)deviceMB);
        return PG_OK;
    }
    if (slots != want)
        fprintf(stderr, "warning: device TT of %llu MB does not fit, using %llu MB\n", (unsigned long long)deviceMB,
                (unsigned long long)(slots * 16 >> 20));
    PG_HIP(hipMemsetAsync(p, 0, slots * 16, L.stream));
    PG_HIP(hipStreamSynchronize(L.stream));
    L.tt.entries = p;
    L.tt.mask = slots - 1;
    return PG_OK;
}

int ttInit(uint64_t deviceMB, uint64_t hostMB)
{
    PG_TRY(hostInit(hostMB));
    return deviceInit(deviceMB);
}

int ttFree()
{
    hostFree();
    Lib &L = lib();
    if (L.tt.entries) (void)hipFree(L.tt.entries);
    L.tt = DeviceTT{};
    return PG_OK;
}

int ttClearDevice()
{
    Lib &L = lib();
    if (L.tt.entries) {
        PG_HIP(hipMemsetAsync(L.tt.entries, 0, (L.tt.mask + 1) * 16, L.stream));
        PG_HIP(hipStreamSynchronize(L.stream));
    }
    return PG_OK;
}

// The host tier is cleared by handing its bucket heads back and taking fresh
// zeroed ones: a large calloc maps zero pages that the OS fills on first
// touch, so a 32 GB head array (a 256 GB tier) clears in microseconds instead
// of a memset of every page, and untouched buckets cost no resident memory.
int ttClear()
{
    PG_TRY(ttClearDevice());
    if (g_host.heads) {
        const u64 buckets = g_host.mask + 1;
        free(g_host.heads);
        g_host.heads = (std::atomic<u32> *)calloc(buckets, sizeof(u32));
        if (!g_host.heads) {
            hostFree();
            setError("host TT: its cleared bucket heads could not be allocated");
            fprintf(stderr, "warning: %s, host TT disabled\n", pg_last_error());  // as hostInit
            return PG_OK;
        }
        g_host.used = 1;
    }
    return PG_OK;
}

void ttCountDevice(u64 probes, u64 hits, u64 stores)
{
    g_devProbes += probes;
    g_devHits += hits;
    g_devStores += stores;
}

int ttStats(uint64_t out[6])
{
    out[0] = g_devProbes;
    out[1] = g_devHits;
    out[2] = g_devStores;
    out[3] = g_host.probes;
    out[4] = g_host.hits;
    out[5] = g_host.stores;
    return PG_OK;
}

// ----------------------------------------------------------- snapshot/resume
// File: magic "PGTT0001", device slot count, host bucket count, host pool
// entries in use, then the device table, the host heads and the host pool.
// A long run can save its tables and a later process resume from them.
static const char TT_MAGIC[8] = {'P', 'G', 'T', 'T', '0', '0', '0', '1'};

int ttSave(const char *path)
{
    Lib &L = lib();
    FILE *f = fopen(path, "wb");
    if (!f) {
        setError("cannot open %s for writing", path);
        return PG_ERR_INVALID;
    }
    u64 devSlots = L.tt.entries ? L.tt.mask + 1 : 0;
    u64 buckets = g_host.heads ? g_host.mask + 1 : 0;
    u64 used = g_host.heads ? std::min<u64>(g_host.used.load(), g_host.poolSize) : 0;
    u64 hdr[3] = {devSlots, buckets, used};
    bool ok = fwrite(TT_MAGIC, 1, 8, f) == 8 && fwrite(hdr, 8, 3, f) == 3;
    const u64 CH = 64ull << 20;
    std::vector<char> buf;
    if (ok && devSlots) {
        buf.resize(CH);
        for (u64 off = 0; ok && off < devSlots * 16; off += CH) {
            u64 len = std::min(CH, devSlots * 16 - off);
            if (hipMemcpy(buf.data(), (char *)L.tt.entries + off, len, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
            else ok = fwrite(buf.data(), 1, len, f) == len;
        }
    }
    if (ok && buckets) ok = fwrite((const void *)g_host.heads, sizeof(u32), buckets, f) == buckets;
    if (ok && used) ok = fwrite(g_host.pool, sizeof(HostEntry), used, f) == used;
    fclose(f);
    if (!ok) {
        setError("writing %s failed", path);
        return PG_ERR_INVALID;
    }
    return PG_OK;
}

int ttLoad(const char *path)
{
    Lib &L = lib();
    FILE *f = fopen(path, "rb");
    if (!f) {
        setError("cannot open %s", path);
        return PG_ERR_INVALID;
    }
    char magic[8];
    u64 hdr[3];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, TT_MAGIC, 8) || fread(hdr, 8, 3, f) != 3) {
        fclose(f);
        setError("%s is not a TT snapshot", path);
        return PG_ERR_INVALID;
    }
    u64 devSlots = hdr[0], buckets = hdr[1], used = hdr[2];
    // the tables must already exist with the same geometry (pg_tt_init with the same sizes)
    u64 haveDev = L.tt.entries ? L.tt.mask + 1 : 0, haveBuckets = g_host.heads ? g_host.mask + 1 : 0;
    if (devSlots != haveDev || buckets != haveBuckets || used > g_host.poolSize) {
        fclose(f);
        setError("snapshot geometry (%llu slots, %llu buckets) differs from the current tables (%llu, %llu)",
                 (unsigned long long)devSlots, (unsigned long long)buckets, (unsigned long long)haveDev,
                 (unsigned long long)haveBuckets);
        return PG_ERR_INVALID;
    }
    bool ok = true;
    const u64 CH = 64ull << 20;
    std::vector<char> buf;
    if (devSlots) {
        buf.resize(CH);
        for (u64 off = 0; ok && off < devSlots * 16; off += CH) {
            u64 len = std::min(CH, devSlots * 16 - off);
            ok = fread(buf.data(), 1, len, f) == len &&
                 hipMemcpy((char *)L.tt.entries + off, buf.data(), len, hipMemcpyHostToDevice) == hipSuccess;
        }
    }
    if (ok && buckets) ok = fread((void *)g_host.heads, sizeof(u32), buckets, f) == buckets;
    if (ok && used) ok = fread(g_host.pool, sizeof(HostEntry), used, f) == used;
    fclose(f);
    // every chain link must stay inside the entries in use (index 0 ends a
    // chain): the probe follows them, so a corrupt or edited file could send
    // it outside the pool
    bool linksOk = true;
    if (ok && buckets) {
        const u64 lim = used ? used : 1;
        for (u64 b = 0; b < buckets && linksOk; b++) linksOk = g_host.heads[b].load(std::memory_order_relaxed) < lim;
        for (u64 i = 1; i < used && linksOk; i++) linksOk = g_host.pool[i].next < lim;
    }
    if (!ok || !linksOk) {
        // a part-filled table would answer with wrong counts: start empty
        ttClear();
        setError(ok ? "%s holds chain links outside its entries (tables cleared)" : "reading %s failed (tables cleared)",
                 path);
        return PG_ERR_INVALID;
    }
    if (buckets) g_host.used = used ? used : 1;
    return PG_OK;
}

void u128::toDecimal(char *buf, int n) const
{
    if (hi == 0) {
        snprintf(buf, n, "%llu", (unsigned long long)lo);
        return;
    }
    u32 limbs[4] = {(u32)lo, (u32)(lo >> 32), (u32)hi, (u32)(hi >> 32)};
    u32 groups[8];
    int ng = 0;
    for (;;) {
        bool zero = true;
        u64 rem = 0;
        for (int i = 3; i >= 0; i--) {
            u64 cur = (rem << 32) | limbs[i];
            limbs[i] = (u32)(cur / 1000000000u);
            rem = cur % 1000000000u;
            if (limbs[i]) zero = false;
        }
        groups[ng++] = (u32)rem;
        if (zero) break;
    }
    int off = snprintf(buf, n, "%u", groups[ng - 1]);
    for (int i = ng - 2; i >= 0 && off < n; i--) off += snprintf(buf + off, n - off, "%09u", groups[i]);
}

}  // namespace pg

using namespace pg;

extern "C" {
// the tables are shared with a server's slots while one exists: creating,
// clearing, freeing, saving or loading them then returns PG_ERR_BUSY
int pg_tt_init(uint64_t deviceMB, uint64_t hostMB)
{
    PG_TRY(refuseWhileServing("pg_tt_init"));
    return ttInit(deviceMB, hostMB);
}
int pg_tt_clear(void)
{
    PG_TRY(refuseWhileServing("pg_tt_clear"));
    return ttClear();
}
int pg_tt_free(void)
{
    PG_TRY(refuseWhileServing("pg_tt_free"));
    return ttFree();
}
int pg_tt_stats(uint64_t out[6])
{
    if (!out) return PG_ERR_INVALID;
    return ttStats(out);
}
int pg_tt_save(const char *path)
{
    if (!path) return PG_ERR_INVALID;
    PG_TRY(refuseWhileServing("pg_tt_save"));
    return ttSave(path);
}
int pg_tt_load(const char *path)
{
    if (!path) return PG_ERR_INVALID;
    PG_TRY(refuseWhileServing("pg_tt_load"));
    return ttLoad(path);
}
int pg_u128_to_string(uint64_t lo, uint64_t hi, char *buf, int n)
{
    if (!buf || n < 2) return PG_ERR_INVALID;
    u128 v;
    v.lo = lo;
    v.hi = hi;
    v.toDecimal(buf, n);
    return PG_OK;
}
}
