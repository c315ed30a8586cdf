// pg_internal.h -- library state shared by the host code and the kernel
// launchers (not part of the public ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/perft_gpu.h"
#include "pg_chess.h"

namespace pg {

// thread-local error text returned by pg_last_error
void setError(const char *fmt, ...);

#define PG_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::pg::setError("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return PG_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)

#define PG_TRY(call)              \
    do {                          \
        int s_ = (call);          \
        if (s_ != PG_OK) return s_; \
    } while (0)

// Bump allocator over a device byte range (GpuBumpAllocator,
// perft_kernels.cu:175-195): 256-B aligned, reset per call.
struct Bump {
    char *base = nullptr;
    size_t size = 0, off = 0, peak = 0;
    Bump() = default;
    Bump(void *b, size_t s) : base((char *)b), size(s) {}
    template <typename T>
    T *alloc(size_t n)
    {
        size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        if (off + bytes > size) return nullptr;
        T *p = (T *)(base + off);
        off += bytes;
        if (off > peak) peak = off;
        return p;
    }
    size_t mark() const { return off; }
    void release(size_t m) { off = m; }
};

// Device transposition table: one depth-tagged table of 16-B entries
// {key ^ data, data}, data = count << 8 | depth (see pg_tt.h).
struct DeviceTT {
    void *entries = nullptr;
    uint64_t mask = 0;  // slots - 1 (power of two)
};

// Live per-kernel timing with HIP events on the stream the kernels run on,
// enabled by pg_profile_enable; accumulates launches, milliseconds and the
// algorithmic bytes / items of every launch.
enum KernelId { K_MAKECOUNT, K_SCAN, K_EXPAND, K_GENMOVES, K_LEAF1, K_LEAF2, K_BATCH_GROUP, K_TT, K_COUNT };
static const char *const KERNEL_NAMES[K_COUNT] = {"makecount", "scan", "expand", "genmoves",
                                                  "leaf1",     "leaf2", "batch_group", "tt"};
struct Profiler {
    bool on = false;
    uint32_t mask = ~0u;  // kernels timed when on (bit = KernelId)
    struct Rec {
        int id;
        hipEvent_t a, b;
        double bytes, items;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[K_COUNT] = {}, bytes[K_COUNT] = {}, items[K_COUNT] = {};
    uint64_t launches[K_COUNT] = {};
    // the launches with the most algorithmic bytes (the largest level of a
    // repeated call): their bytes, summed time and number
    double bigBytes[K_COUNT] = {}, bigMs[K_COUNT] = {};
    uint64_t bigN[K_COUNT] = {};
    int cur = -1;
    hipEvent_t curA = nullptr;

    hipEvent_t get()
    {
        if (pool.empty()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    void begin(int id, hipStream_t st)
    {
        if (!on || !((mask >> id) & 1)) return;
        cur = id;
        curA = get();
        (void)hipEventRecord(curA, st);
    }
    void end(double b, double it, hipStream_t st)
    {
        if (!on || cur < 0) return;
        hipEvent_t e = get();
        (void)hipEventRecord(e, st);
        pending.push_back({cur, curA, e, b, it});
        cur = -1;
    }
    // waits for each recorded end event, then adds up the intervals
    void collect()
    {
        for (auto &r : pending) {
            float t = 0;
            (void)hipEventSynchronize(r.b);  // the caller may not have synchronised the stream
            (void)hipEventElapsedTime(&t, r.a, r.b);
            ms[r.id] += t;
            bytes[r.id] += r.bytes;
            items[r.id] += r.items;
            launches[r.id]++;
            if (r.bytes > bigBytes[r.id]) bigBytes[r.id] = r.bytes, bigMs[r.id] = 0, bigN[r.id] = 0;
            if (r.bytes == bigBytes[r.id]) bigMs[r.id] += t, bigN[r.id]++;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    void reset()
    {
        for (int i = 0; i < K_COUNT; i++)
            ms[i] = bytes[i] = items[i] = bigBytes[i] = bigMs[i] = 0, launches[i] = bigN[i] = 0;
    }
};

// The pinned host buffer of every stream that runs GPU BFS calls (the
// library's, each serving and frontier slot's), allocated fine-grained
// (hipHostMallocCoherent) so that a kernel's system-scope store reaches the
// host while the stream still runs: its first bytes take the readbacks, u64
// slot PIN_POLL takes the level totals the scan publishes for the host to
// poll, and bytes [PIN_STAGE, PINNED_BYTES) stage a call's inputs for its one
// host-to-device copy.
constexpr size_t PINNED_BYTES = 4096, PIN_STAGE = 1024, PIN_POLL = 8;
constexpr unsigned PINNED_FLAGS = hipHostMallocCoherent;
// what a polled slot holds until a kernel publishes (no level total reaches it)
constexpr uint64_t HOST_SLOT_EMPTY = ~0ull;
int waitPublished(volatile uint64_t *slot, hipStream_t st, uint64_t &v);

struct Lib {
    Profiler prof;
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    void *arena = nullptr;
    size_t arenaBytes = 0;
    uint64_t *hostPinned = nullptr;  // small pinned readback buffer
    DeviceTT tt;
    int cuCount = 0;
    std::mutex mu;  // serialises use of the arena / stream by host threads
    // set while a server exists: its slots own the arena and share the device
    // table, so calls that use either return PG_ERR_BUSY until it is destroyed
    std::atomic<bool> serving{false};
};
Lib &lib();
// PG_ERR_BUSY (with the error text naming `what`) while a server exists, else PG_OK
int refuseWhileServing(const char *what);
// PG_OK if the position is one the move generator can take: exactly one king
// per side, black pieces a subset of the occupied squares, no piece code 7,
// side to move 0 or 1; else PG_ERR_INVALID with the reason as the error text
int validatePosition(const pg_position *p);

// kernels (pg_kernels.hip)
// looks up, with `device` current, that device's address of the static scan
// workspace of pg_scan_u32_static (once per device)
int initStaticScratch(int device);
// Blocks of kernel `fn` (block size `blockSize`) resident on the whole device
// at once: occupancy per CU x CUs.  A persistent grid of exactly this size
// runs in one round; a larger one leaves a partly filled last round.
unsigned residentGrid(const void *fn, int blockSize);
// `pinned`: a host-pinned buffer of at least 8 u64 owned by the caller for
// the call's readbacks (each concurrent caller brings its own)
int bfsPerft(const QuadBitBoard &root, u8 state, int color, int depth, int leafPlies, bool useTT, void *arena,
             size_t arenaBytes, hipStream_t stream, u64 *pinned, uint64_t *count, size_t *peakBytes);

// host TT tier (pg_host.cpp)
struct Hash128 {
    uint64_t lo, hi;
};

}  // namespace pg
