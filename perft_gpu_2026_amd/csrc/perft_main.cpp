// perft_main.cpp -- the `perft_gpu` command line, a drop-in for perft.cu:13-186:
//   perft_gpu <fen> <depth> [<launchdepth>] [-nott] [-dtt <MB>] [-htt <MB>] [-ld <N>] [-cpu]
// Flags may appear anywhere; a positional third argument is the launch depth;
// with no arguments it runs the start position to depth 10.  Output lines keep
// the reference's form: "Perft(%02d): <count>, time: <s> seconds, nps: <n>".
// Extra flags of this engine: -leaf <1|2> (leaf plies of the GPU BFS),
// -arena <MB> (BFS arena size), -divide (per-root-move counts at <depth>),
// -suite <file> (EPD perft suite: "fen ;D1 n1 ;D2 n2 ..." per line),
// -ttsave <file> (write both table tiers to <file> after every finished depth)
// and -ttload <file> (fill the tables from such a file before the first
// depth; the tables must have the sizes the file was written with).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/perft_gpu.h"

extern "C" int pg_u128_to_string(uint64_t lo, uint64_t hi, char *buf, int n);
extern "C" void pg_set_leaf_plies(int plies);

static double now()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static void printResult(int depth, uint64_t lo, uint64_t hi, double secs)
{
    char buf[64];
    pg_u128_to_string(lo, hi, buf, sizeof buf);
    printf("\nPerft(%02d): %s, time: %g seconds", depth, buf, secs);
    if (secs > 0) printf(", nps: %llu", (unsigned long long)(((double)hi * 18446744073709551616.0 + (double)lo) / secs));
    printf("\n");
    fflush(stdout);
}

static void moveText(uint16_t m, char *buf)
{
    int f = m & 63, t = (m >> 6) & 63, fl = m >> 12;
    buf[0] = (char)('a' + (f & 7));
    buf[1] = (char)('1' + (f >> 3));
    buf[2] = (char)('a' + (t & 7));
    buf[3] = (char)('1' + (t >> 3));
    buf[4] = (fl & 8) ? "nbrq"[fl & 3] : 0;
    buf[5] = 0;
}

// one perft through the launcher (GPU) or the host recursion (-cpu)
static int perftOne(const pg_position *p, int depth, int ld, bool cpu, bool tt, pg_u128 *out)
{
    if (cpu) {
        uint64_t c = 0;
        int s = pg_perft_cpu(p, depth, tt, &c);
        out->lo = c;
        out->hi = 0;
        return s;
    }
    return pg_perft(p, depth, ld, tt, out);
}

static int runDivide(const pg_position *pos, int depth, int ld, bool cpu, bool tt)
{
    uint16_t mv[256];
    int n = 0;
    pg_gen_moves(pos, mv, &n);
    unsigned __int128 total = 0;
    double t0 = now();
    for (int i = 0; i < n; i++) {
        pg_position c = *pos;
        pg_make_move(&c, mv[i]);
        pg_u128 r{1, 0};
        if (depth > 1 && perftOne(&c, depth - 1, ld < depth - 1 ? ld : depth - 1, cpu, tt, &r) != PG_OK) {
            printf("perft failed: %s\n", pg_last_error());
            return 1;
        }
        char mt[8], num[64];
        moveText(mv[i], mt);
        pg_u128_to_string(r.lo, r.hi, num, sizeof num);
        printf("%s: %s\n", mt, num);
        total += ((unsigned __int128)r.hi << 64) | r.lo;
    }
    char num[64];
    pg_u128_to_string((uint64_t)total, (uint64_t)(total >> 64), num, sizeof num);
    printf("\nMoves: %d\nTotal: %s, time: %g seconds\n", n, num, now() - t0);
    return 0;
}

// EPD suite: "<fen> ;D1 20 ;D2 400 ..." per line; depths above maxDepth are skipped
static int runSuite(const char *path, int maxDepth, bool cpu, bool tt)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        printf("cannot open %s\n", path);
        return 1;
    }
    char line[4096];
    int failures = 0, checks = 0;
    while (fgets(line, sizeof line, f)) {
        std::string l(line);
        size_t semi = l.find(';');
        if (semi == std::string::npos) continue;
        std::string fen = l.substr(0, semi);
        pg_position p;
        if (pg_parse_fen(fen.c_str(), &p) != PG_OK) {
            printf("bad FEN: %s\n", fen.c_str());
            failures++;
            continue;
        }
        size_t pos = semi;
        while (pos != std::string::npos) {
            int d = 0;
            unsigned long long want = 0;
            if (sscanf(l.c_str() + pos + 1, " D%d %llu", &d, &want) == 2 && d >= 1 && d <= maxDepth) {
                pg_u128 r{0, 0};
                int s = perftOne(&p, d, d, cpu, tt, &r);
                bool ok = s == PG_OK && r.hi == 0 && r.lo == want;
                printf("%s D%d %llu %s\n", ok ? "ok  " : "FAIL", d, (unsigned long long)r.lo, fen.c_str());
                failures += !ok;
                checks++;
            }
            pos = l.find(';', pos + 1);
        }
    }
    fclose(f);
    printf("\nsuite: %d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}

// Table snapshots for long runs (no reference counterpart: its perft 14 ran
// 27.8 h with no resume, README.md:63).  A failed load leaves the tables
// empty, so the run continues from scratch.
static void loadSnapshot(const char *path)
{
    if (!path) return;
    if (pg_tt_load(path) == PG_OK) printf("TT snapshot loaded: %s\n", path);
    else printf("warning: TT snapshot not loaded: %s\n", pg_last_error());
    fflush(stdout);
}
static void saveSnapshot(const char *path)
{
    if (path && pg_tt_save(path) != PG_OK) printf("warning: TT snapshot not saved: %s\n", pg_last_error());
}

int main(int argc, char *argv[])
{
    bool cpuMode = false, useTT = true, divide = false;
    const char *suite = nullptr, *ttSave = nullptr, *ttLoad = nullptr;
    long ldOverride = 0, dttMB = 0, httMB = 0, arenaMB = 0, leaf = 0;
    // flags anywhere (perft.cu:17-57): consume them and compact argv
    for (int i = 1; i < argc; i++) {
        long *valueFlag = nullptr;
        if (!strcmp(argv[i], "-cpu")) cpuMode = true;
        else if (!strcmp(argv[i], "-divide")) divide = true;
        else if ((!strcmp(argv[i], "-suite") || !strcmp(argv[i], "-ttsave") || !strcmp(argv[i], "-ttload")) &&
                 i + 1 < argc) {
            const char *&dst = argv[i][3] == 'i' ? suite : argv[i][3] == 's' ? ttSave : ttLoad;
            dst = argv[i + 1];
            for (int j = i; j + 2 < argc; j++) argv[j] = argv[j + 2];
            argc -= 2;
            i--;
            continue;
        }
        else if (!strcmp(argv[i], "-nott")) useTT = false;
        else if (!strcmp(argv[i], "-ld")) valueFlag = &ldOverride;
        else if (!strcmp(argv[i], "-dtt")) valueFlag = &dttMB;
        else if (!strcmp(argv[i], "-htt")) valueFlag = &httMB;
        else if (!strcmp(argv[i], "-arena")) valueFlag = &arenaMB;
        else if (!strcmp(argv[i], "-leaf")) valueFlag = &leaf;
        else continue;
        int take = 1;
        if (valueFlag) {
            if (i + 1 >= argc) break;
            *valueFlag = atol(argv[i + 1]);
            take = 2;
        }
        for (int j = i; j + take < argc; j++) argv[j] = argv[j + take];
        argc -= take;
        i--;
    }

    if (!cpuMode) {
        int s = pg_init(0, (uint64_t)arenaMB << 20);
        if (s == PG_ERR_NO_DEVICE) {
            printf("No of GPUs detected: 0\n");
            printf("No GPUs found. Exiting.\n");
            return 1;
        }
        printf("No of GPUs detected: 1\n");
        if (s != PG_OK) {
            printf("GPU init failed: %s\n", pg_last_error());
            return 1;
        }
        printf("BFS arena: %.1f GiB\n", pg_arena_bytes() / 1073741824.0);
    }
    if (leaf == 1 || leaf == 2) pg_set_leaf_plies((int)leaf);

    if (suite) {
        int maxD = argc >= 2 ? atoi(argv[1]) : 6;
        if (useTT) pg_tt_init(cpuMode ? 0 : 4096, 1024);
        int rc = runSuite(suite, maxD > 0 ? maxD : 6, cpuMode, useTT);
        pg_tt_free();
        if (!cpuMode) pg_shutdown();
        return rc;
    }

    const char *fen = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1";
    int maxDepth = 10;
    if (argc >= 3) {
        fen = argv[1];
        maxDepth = atoi(argv[2]);
    } else {
        printf("\nUsage: perft_gpu <fen> <depth> [<launchdepth>] [-nott] [-dtt <MB>] [-htt <MB>] [-ld <N>] [-cpu]\n");
        printf("\nAs no parameters were provided... running default test\n");
    }
    pg_position pos;
    if (pg_parse_fen(fen, &pos) != PG_OK) {
        printf("invalid FEN: %s\n", fen);
        return 1;
    }
    if (maxDepth < 0 || maxDepth > 20) {
        printf("depth %d out of range\n", maxDepth);
        return 1;
    }

    if (divide) {
        if (useTT) pg_tt_init(cpuMode ? 0 : 4096, 1024);
        int rc = runDivide(&pos, maxDepth, ldOverride > 0 ? (int)ldOverride : maxDepth, cpuMode, useTT);
        pg_tt_free();
        if (!cpuMode) pg_shutdown();
        return rc;
    }

    if (cpuMode) {
        printf("CPU mode\n");
        fflush(stdout);
        if (useTT) {
            pg_tt_init(0, httMB ? (uint64_t)httMB : 1024);
            loadSnapshot(ttLoad);
        }
        for (int d = 1; d <= maxDepth; d++) {
            uint64_t c = 0;
            double t0 = now();
            pg_perft_cpu(&pos, d, useTT, &c);
            printResult(d, c, 0, now() - t0);
            if (useTT) saveSnapshot(ttSave);
        }
        pg_tt_free();
        return 0;
    }

    float bf = 30.0f;
    int ld = 0;
    pg_estimate_launch_depth(&pos, &bf, &ld);
    // launch-depth policy, perft.cu:119-154
    if (ld > 11) ld = 11;
    if (ld < 6) ld = 6;
    if (maxDepth >= 10 && ld < 9) ld = 9;
    if (ldOverride > 0) ld = (int)ldOverride;
    else if (argc >= 4) ld = atoi(argv[3]);
    if (maxDepth < ld) ld = maxDepth;
    int currentLD = ld;
    printf("Launch depth: %d\n", currentLD);
    fflush(stdout);
    if (useTT) {
        // the reference gives its device tables most of the free memory
        // (launcher.cu:146-160); 64 GB fits beside the default arena on a
        // 288 GB device (the library halves it where it does not fit)
        uint64_t dMB = dttMB ? (uint64_t)dttMB : 65536, hMB = httMB ? (uint64_t)httMB : 4096;
        if (pg_tt_init(dMB, hMB) != PG_OK) printf("warning: TT init failed: %s\n", pg_last_error());
        loadSnapshot(ttLoad);
    }
    int rc = 0;
    for (int d = 1; d <= maxDepth; d++) {
        int eff = currentLD < d ? currentLD : d;
        pg_u128 r{0, 0};
        uint64_t oomBefore = pg_get_oom_fallback_count();
        double t0 = now();
        int s = pg_perft(&pos, d, eff, useTT, &r);
        double t1 = now();
        if (s != PG_OK) {
            printf("perft(%d) failed: %s\n", d, pg_last_error());
            rc = 1;
            break;
        }
        printResult(d, r.lo, r.hi, t1 - t0);
        if (useTT) saveSnapshot(ttSave);
        uint64_t ooms = pg_get_oom_fallback_count() - oomBefore;
        if (ooms > 0 && d >= ld) {
            int newLD = pg_get_effective_ld();
            if (newLD != currentLD) {
                printf("  >> Launch depth adjusted: %d -> %d (OOM fallbacks: %llu)\n", currentLD, newLD,
                       (unsigned long long)ooms);
                currentLD = newLD;
            }
        }
    }
    pg_tt_free();
    pg_shutdown();
    return rc;
}
