// pg_serve.hip -- perft serving loop (no reference counterpart).
//
// Requests (position, depth) arrive at any time through pg_serve_submit.  The
// server splits its HBM budget into the shared device transposition table and
// `maxInflight` slots; each slot owns an arena partition, a HIP stream, a
// pinned readback buffer and a worker thread.  Admission is continuous: a
// queued request is taken by the first slot that frees up (arrival order), so
// requests that arrive while others are on the GPU start as soon as a slot
// and its memory are free, and the slots' kernels overlap on the device.  A
// request runs as one BFS of its whole depth in its slot (table probes and
// stores at every level, shared by all requests); one that does not fit its
// slot is split one ply down on the host, as the launcher does on OOM
// (launcher.cu:600-610), so results never depend on memory pressure.
// Counts are exact, so a request's result is identical alone and in the mix.
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <vector>

#include "pg_launcher.h"
#include "pg_internal.h"
#include "pg_tt.h"

namespace pg {

static double nowSec()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct ServeRequest {
    uint64_t id;
    pg_position pos;
    int depth;
    double submit;
};

struct Slot {
    hipStream_t stream = nullptr;
    void *arena = nullptr;
    size_t bytes = 0;
    u64 *pinned = nullptr;
    BatchState *batch = nullptr;  // the slot's batched engine (maxBatch > 1)
    std::thread th;
};

// requests deeper than this run alone (a whole-depth BFS fills the GPU)
static const int SERVE_MERGE_MAX_DEPTH = 5;

struct Server {
    pg_serve_config cfg{};
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<ServeRequest> queue;
    std::deque<pg_serve_result> done;
    std::vector<Slot> slots;
    bool stop = false;
    bool ownsTT = false;
    int leafPlies = 2;

    // one request in one slot: whole-depth BFS, split on the host when it does not fit
    int runOne(Slot &s, const pg_position &p, int depth, uint64_t *out)
    {
        QuadBitBoard q;
        memcpy(q.bb, p.bb, 32);
        size_t peak = 0;
        int st = bfsPerft(q, p.state, p.color, depth, leafPlies, true, s.arena, s.bytes, s.stream, s.pinned, out, &peak);
        if (st != PG_ERR_ARENA_OOM) return st;
        if (depth <= 1) return PG_ERR_ARENA_OOM;
        u16 mv[MAX_MOVES];
        u32 n = generateMoves(q, p.state, p.color, mv);
        uint64_t total = 0;
        for (u32 i = 0; i < n; i++) {
            pg_position c = p;
            QuadBitBoard cq = q;
            makeMove(cq, c.state, p.color, mv[i]);
            memcpy(c.bb, cq.bb, 32);
            c.color ^= 1;
            uint64_t sub = 0;
            int r = runOne(s, c, depth - 1, &sub);
            if (r != PG_OK) return r;
            total += sub;
        }
        *out = total;
        return PG_OK;
    }

    // Takes the request at the head of the queue and, when it is shallow and
    // maxBatch > 1, the other waiting requests of the same depth up to
    // maxBatch in all (they run in one multi-root pass on the slot's batched
    // engine); every request taken keeps its own submission time.
    void worker(int k)
    {
        Slot &s = slots[k];
        (void)hipSetDevice(lib().device);
        std::vector<ServeRequest> take;
        for (;;) {
            take.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                cvWork.wait(lk, [&] { return stop || !queue.empty(); });
                if (stop && queue.empty()) return;
                take.push_back(queue.front());
                queue.pop_front();
                const int d = take[0].depth;
                if (s.batch && cfg.maxBatch > 1 && d >= 2 && d <= SERVE_MERGE_MAX_DEPTH) {
                    for (auto it = queue.begin(); it != queue.end() && (int)take.size() < cfg.maxBatch;) {
                        if (it->depth == d) {
                            take.push_back(*it);
                            it = queue.erase(it);
                        } else {
                            ++it;
                        }
                    }
                }
            }
            const double start = nowSec();
            std::vector<uint64_t> counts(take.size(), 0);
            std::vector<int> status(take.size(), PG_OK);
            if (take.size() > 1) {
                std::vector<pg_position> ps(take.size());
                std::vector<uint8_t> ds(take.size());
                for (size_t i = 0; i < take.size(); i++) ps[i] = take[i].pos, ds[i] = (uint8_t)take[i].depth;
                if (batchSlotRun(s.batch, (int)take.size(), ps.data(), ds.data(), true, counts.data()) != PG_OK) {
                    // not expected (OOM splits inside); each request alone instead
                    for (size_t i = 0; i < take.size(); i++) status[i] = runOne(s, take[i].pos, take[i].depth, &counts[i]);
                }
            } else {
                status[0] = runOne(s, take[0].pos, take[0].depth, &counts[0]);
            }
            const double end = nowSec();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t i = 0; i < take.size(); i++) {
                    pg_serve_result res{};
                    res.id = take[i].id;
                    res.submitTime = take[i].submit;
                    res.startTime = start;
                    res.endTime = end;
                    res.count = counts[i];
                    res.status = status[i];
                    done.push_back(res);
                }
            }
            cvDone.notify_all();
        }
    }
};

static int serveCreateLocked(const pg_serve_config *cfg, void **handle);
static int serveCreate(const pg_serve_config *cfg, void **handle)
{
    Lib &L = lib();
    if (!L.init) {
        setError("pg_init has not run");
        return PG_ERR_NOT_INIT;
    }
    // one server at a time; from here until pg_serve_destroy the slots own
    // the arena and the table, and every other call that uses them is refused
    bool expected = false;
    if (!L.serving.compare_exchange_strong(expected, true)) {
        setError("a server already exists");
        return PG_ERR_BUSY;
    }
    // wait for a GPU call of another thread that started before the flag was set
    std::lock_guard<std::mutex> drain(L.mu);
    int s = serveCreateLocked(cfg, handle);
    if (s != PG_OK) L.serving = false;
    return s;
}

static int serveCreateLocked(const pg_serve_config *cfg, void **handle)
{
    Lib &L = lib();
    Server *S = new Server();
    S->cfg = *cfg;
    L.prof.on = false;  // the event profiler is single-threaded; slots run concurrently
    int slots = cfg->maxInflight > 0 ? cfg->maxInflight : 4;
    if (slots > 16) slots = 16;
    u64 budget = cfg->hbmBudgetBytes ? cfg->hbmBudgetBytes : L.arenaBytes;
    u64 ttBytes = cfg->ttBytes;
    if (ttBytes >= budget) ttBytes = budget / 4;
    // the shared table: reuse the library's if one exists, else create it in the budget
    if (!L.tt.entries && ttBytes) {
        if (ttInit(ttBytes >> 20, 0) != PG_OK) {
            delete S;
            return PG_ERR_HIP;
        }
        S->ownsTT = L.tt.entries != nullptr;
    }
    u64 ttUsed = L.tt.entries ? (L.tt.mask + 1) * 16 : 0;
    u64 forSlots = budget > ttUsed ? budget - ttUsed : 0;
    // slots are carved out of the library arena (already resident in HBM)
    if (forSlots > L.arenaBytes) forSlots = L.arenaBytes;
    u64 per = (forSlots / slots) & ~(u64)255;
    if (per < (64ull << 20)) {
        setError("serving budget too small: %llu bytes per slot", (unsigned long long)per);
        delete S;
        return PG_ERR_INVALID;
    }
    S->slots.resize(slots);
    for (int k = 0; k < slots; k++) {
        Slot &s = S->slots[k];
        s.arena = (char *)L.arena + k * per;
        s.bytes = per;
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc((void **)&s.pinned, PINNED_BYTES, PINNED_FLAGS) != hipSuccess) {
            setError("serving slot %d: stream/pinned allocation failed", k);
            delete S;
            return PG_ERR_HIP;
        }
        if (cfg->maxBatch > 1 && !(s.batch = batchSlotCreate(s.arena, s.bytes, s.stream, s.pinned))) {
            setError("serving slot %d: batched engine allocation failed", k);
            delete S;
            return PG_ERR_HIP;
        }
    }
    for (int k = 0; k < slots; k++) S->slots[k].th = std::thread(&Server::worker, S, k);
    *handle = S;
    return PG_OK;
}

static int serveDestroy(Server *S)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cvWork.notify_all();
    for (auto &s : S->slots)
        if (s.th.joinable()) s.th.join();
    for (auto &s : S->slots) {
        if (s.stream) {
            (void)hipStreamSynchronize(s.stream);
            (void)hipStreamDestroy(s.stream);
        }
        if (s.pinned) (void)hipHostFree(s.pinned);
        batchSlotDestroy(s.batch);
    }
    if (S->ownsTT) {
        Lib &L = lib();
        (void)hipFree(L.tt.entries);
        L.tt = DeviceTT{};
    }
    delete S;
    lib().serving = false;
    return PG_OK;
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_serve_create(const pg_serve_config *cfg, void **handle)
{
    if (!cfg || !handle) return PG_ERR_INVALID;
    return serveCreate(cfg, handle);
}

int pg_serve_submit(void *handle, uint64_t id, const pg_position *pos, int depth)
{
    if (!handle || depth < 0 || depth > 16) return PG_ERR_INVALID;
    PG_TRY(validatePosition(pos));
    Server *S = (Server *)handle;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            setError("the server is shutting down");
            return PG_ERR_BUSY;
        }
        if (S->cfg.maxQueued > 0 && S->queue.size() >= (size_t)S->cfg.maxQueued) {
            setError("%d requests already wait (maxQueued)", S->cfg.maxQueued);
            return PG_ERR_BUSY;
        }
        S->queue.push_back({id, *pos, depth, nowSec()});
    }
    S->cvWork.notify_one();
    return PG_OK;
}

int pg_serve_poll(void *handle, pg_serve_result *results, int maxResults, int timeoutMs, int *nOut)
{
    if (!handle || !nOut || (maxResults > 0 && !results)) return PG_ERR_INVALID;
    Server *S = (Server *)handle;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && timeoutMs > 0)
        S->cvDone.wait_for(lk, std::chrono::milliseconds(timeoutMs), [&] { return !S->done.empty(); });
    int n = 0;
    while (n < maxResults && !S->done.empty()) {
        results[n++] = S->done.front();
        S->done.pop_front();
    }
    *nOut = n;
    return PG_OK;
}

int pg_serve_destroy(void *handle)
{
    if (!handle) return PG_ERR_INVALID;
    return serveDestroy((Server *)handle);
}

double pg_clock(void) { return nowSec(); }
}
