// pg_launcher_api.hip -- the reference's C++ launcher interface (launcher.h:
// 1-47, utils.h:7), declared in include/perft_launcher.h, on top of this
// library's launcher, GPU BFS, -cpu path and tables.  Host code only.
//
// The reference's fatal errors exit(1) (launcher.cu:64-67, 80-83); so do
// these functions when the library reports anything but an arena OOM, which
// perft_gpu_host_bfs reports through g_lastBfsOom as the reference does.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <chrono>

#include "../../include/perft_launcher.h"
#include "pg_internal.h"
#include "pg_launcher.h"
#include "pg_tt.h"

// ---- globals (launcher.h:6-11, 37-47; perft.cu:7-8)
void *preAllocatedBufferHost = nullptr;
uint64 g_preAllocatedMemorySize = 0;
bool g_useTT = true;
size_t g_lastBfsPeakMemory = 0;
bool g_lastBfsOom = false;
int g_deviceTTBudgetMB = 0;
int g_hostTTBudgetMB = 0;

namespace {

uint64_t g_oomBase = 0;       // pg_get_oom_fallback_count at the last resetCallStats
uint64_t g_statsBase[6] = {};  // pg_tt_stats at the last resetCallStats

void fatal(const char *what)
{
    fprintf(stderr, "%s failed: %s\n", what, pg_last_error());
    exit(1);
}

pg_position toPosition(const QuadBitBoard *pos, const GameState *gs, uint8 color)
{
    pg_position p{};
    memcpy(p.bb, pos->bb, sizeof p.bb);
    p.state = gs->raw;
    p.color = color;
    return p;
}

double seconds()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// "Perft(%02d): <count>, time: <s> seconds, nps: <n>" (launcher.cu:732-735, 750-753)
void printPerftLine(uint32 depth, uint64_t lo, uint64_t hi, double secs)
{
    char buf[64];
    pg_u128_to_string(lo, hi, buf, sizeof buf);
    printf("\nPerft(%02u): %s, time: %g seconds", depth, buf, secs);
    if (secs > 0) printf(", nps: %llu", (unsigned long long)(((double)hi * 18446744073709551616.0 + (double)lo) / secs));
    printf("\n");
    fflush(stdout);
}

bool tablesOn() { return g_useTT; }

}  // namespace

void initGPU(int gpu)
{
    size_t freeB = 0, totalB = 0;
    if (hipSetDevice(gpu) != hipSuccess || hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
        fprintf(stderr, "hipSetDevice(%d) failed\n", gpu);
        exit(1);
    }
    printf("\ngpu: %d, memory total: %llu, free: %llu\n", gpu, (unsigned long long)totalB,
           (unsigned long long)freeB);
    // with the tables on, a third of the device for the arena leaves the rest
    // to initTT (launcher.cu:74-77); without them the library's own default
    uint64_t arena = g_useTT ? (uint64_t)totalB / 3 : 0;
    if (pg_init(gpu, arena) != PG_OK) fatal("initGPU");
    preAllocatedBufferHost = pg::lib().arena;
    g_preAllocatedMemorySize = pg_arena_bytes();
    printf("Allocated BFS buffer: %llu MB (%.0f%% of %llu MB total)\n",
           (unsigned long long)(g_preAllocatedMemorySize >> 20), 100.0 * g_preAllocatedMemorySize / totalB,
           (unsigned long long)(totalB >> 20));
}

void initMoveGen() {}

uint32 estimateLaunchDepth(QuadBitBoard *pos, GameState *gs, uint8 rootColor, float *outBranchingFactor)
{
    pg_position p = toPosition(pos, gs, rootColor);
    float bf = 0;
    int ld = 0;
    if (pg_estimate_launch_depth(&p, &bf, &ld) != PG_OK) fatal("estimateLaunchDepth");
    if (outBranchingFactor) *outBranchingFactor = bf;
    return (uint32)ld;
}

void perftLauncher(QuadBitBoard *pos, GameState *gs, uint8 rootColor, uint32 depth, int launchDepth)
{
    pg_position p = toPosition(pos, gs, rootColor);
    pg_u128 r{0, 0};
    double t0 = seconds();
    int s = pg_perft(&p, (int)depth, launchDepth, tablesOn(), &r);
    double t1 = seconds();
    if (s != PG_OK) fatal("perftLauncher");
    printPerftLine(depth, r.lo, r.hi, t1 - t0);
}

uint64 perft_gpu_host_bfs(QuadBitBoard *pos, GameState *gs, uint8 rootColor, int depth, void *gpuBuffer,
                          size_t bufferSize)
{
    pg_position p = toPosition(pos, gs, rootColor);
    uint64_t c = 0;
    int s = pg_perft_bfs(&p, depth, tablesOn(), gpuBuffer, gpuBuffer ? bufferSize : 0, &c);
    g_lastBfsOom = s == PG_ERR_ARENA_OOM;
    g_lastBfsPeakMemory = (size_t)pg_last_bfs_peak_bytes();
    if (g_lastBfsOom) return 0;
    if (s != PG_OK) fatal("perft_gpu_host_bfs");
    return c;
}

void perftCPU(QuadBitBoard *pos, GameState *gs, uint8 rootColor, uint32 depth)
{
    double t0 = seconds();
    uint64 c = perft_cpu_dispatch(pos, gs, rootColor, depth);
    printPerftLine(depth, c, 0, seconds() - t0);
}

uint64 perft_cpu_dispatch(QuadBitBoard *pos, GameState *gs, uint8 color, uint32 depth)
{
    pg_position p = toPosition(pos, gs, color);
    uint64_t c = 0;
    if (pg_perft_cpu(&p, (int)depth, tablesOn(), &c) != PG_OK) fatal("perft_cpu_dispatch");
    return c;
}

void initTT(int launchDepth, int maxLaunchDepth, int maxDepth, float branchingFactor)
{
    (void)branchingFactor;
    if (!g_useTT) return;
    // device depths 3..maxLaunchDepth-1 and host depths launchDepth..maxDepth
    // (launcher.cu:137-141, 278-280) share one table per tier here
    uint64_t devMB = 0, hostMB = 0;
    if (maxLaunchDepth > 3) {
        if (g_deviceTTBudgetMB > 0) {
            devMB = (uint64_t)g_deviceTTBudgetMB;
        } else {
            size_t freeB = 0, totalB = 0;
            if (hipMemGetInfo(&freeB, &totalB) == hipSuccess) devMB = (uint64_t)(freeB * 0.9) >> 20;
            printf("Auto device TT budget: %llu MB (90%% of %llu MB free)\n", (unsigned long long)devMB,
                   (unsigned long long)(freeB >> 20));
        }
    }
    if (launchDepth <= maxDepth) {
        if (g_hostTTBudgetMB > 0) {
            hostMB = (uint64_t)g_hostTTBudgetMB;
        } else {
            uint64_t ram = (uint64_t)sysconf(_SC_PHYS_PAGES) * (uint64_t)sysconf(_SC_PAGE_SIZE);
            hostMB = (uint64_t)(ram * 0.9) >> 20;
            if (hostMB > (64ull << 10)) hostMB = 64ull << 10;
            printf("Host TT budget: auto %llu MB (90%% of %llu MB system RAM, at most 64 GiB)\n",
                   (unsigned long long)hostMB, (unsigned long long)(ram >> 20));
        }
    }
    // a table that cannot be allocated is reported and left out, as the
    // reference warns and runs without it (launcher.cu:339-345)
    if (pg_tt_init(devMB, hostMB) != PG_OK) fprintf(stderr, "warning: initTT: %s\n", pg_last_error());
}

void clearDeviceTTs()
{
    if (!g_useTT) return;
    if (pg::ttClearDevice() != PG_OK) fatal("clearDeviceTTs");
}

void freeTT()
{
    if (!g_useTT) return;
    pg_tt_free();
}

void printTTStats()
{
    uint64_t st[6] = {};
    pg_tt_stats(st);
    uint64_t d[6];
    for (int i = 0; i < 6; i++) d[i] = st[i] - g_statsBase[i];
    if (d[0])
        printf("  Device TT probes: %llu hits / %llu total (%.1f%%), %llu stores\n", (unsigned long long)d[1],
               (unsigned long long)d[0], 100.0 * d[1] / d[0], (unsigned long long)d[2]);
    if (d[3])
        printf("  Host TT probes: %llu hits / %llu total (%.1f%%)\n", (unsigned long long)d[4],
               (unsigned long long)d[3], 100.0 * d[4] / d[3]);
    size_t peak = pg::peakSinceReset();
    if (peak && g_preAllocatedMemorySize)
        printf("  Peak BFS memory: %llu MB / %llu MB (%.1f%%)\n", (unsigned long long)(peak >> 20),
               (unsigned long long)(g_preAllocatedMemorySize >> 20), 100.0 * peak / g_preAllocatedMemorySize);
    uint64 ooms = getOomFallbackCount();
    if (ooms) printf("  OOM fallbacks: %llu\n", (unsigned long long)ooms);
    fflush(stdout);
}

void resetCallStats()
{
    g_oomBase = pg_get_oom_fallback_count();
    pg_tt_stats(g_statsBase);
    pg::resetPeakSinceReset();
}

uint64 getOomFallbackCount() { return pg_get_oom_fallback_count() - g_oomBase; }
int getEffectiveLD() { return pg_get_effective_ld(); }
void setEffectiveLD(int ld) { pg_set_effective_ld(ld); }

void readFENString(const char fen[], QuadBitBoard *qbb, GameState *gs, uint8 *color)
{
    pg_position p{};
    if (pg_parse_fen(fen, &p) != PG_OK) {
        fprintf(stderr, "readFENString: %s: %s\n", fen ? fen : "(null)", pg_last_error());
        memset(&p, 0, sizeof p);
    }
    memcpy(qbb->bb, p.bb, sizeof qbb->bb);
    gs->raw = p.state;
    *color = p.color;
}
