// launcher_caller.cpp -- a C++ program written only against the reference's
// launcher interface (include/perft_launcher.h, the names and signatures of
// launcher.h:1-47 and utils.h:7), linked with -lperft_gpu.  It runs the
// per-depth loop of perft.cu:100-176 and a few direct calls, printing one
// "key value" line per check for the test that runs it:
//   launcher_caller <fen> <depth> <launchDepth> [-nott]
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "perft_launcher.h"

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: launcher_caller <fen> <depth> <launchDepth> [-nott]\n");
        return 2;
    }
    const char *fen = argv[1];
    int maxDepth = atoi(argv[2]), launchDepth = atoi(argv[3]);
    if (argc > 4 && !strcmp(argv[4], "-nott")) g_useTT = false;
    g_deviceTTBudgetMB = 4096;  // as perft.cu's -dtt / -htt set them
    g_hostTTBudgetMB = 256;

    initGPU(0);
    initMoveGen();
    QuadBitBoard bb;
    GameState gs;
    uint8 color;
    readFENString(fen, &bb, &gs, &color);

    // the whole depth in one GPU call, in the arena, then in 64 KB of it
    uint64 c = perft_gpu_host_bfs(&bb, &gs, color, maxDepth, preAllocatedBufferHost, g_preAllocatedMemorySize);
    printf("bfs %llu oom %d peak %llu\n", c, (int)g_lastBfsOom, (unsigned long long)g_lastBfsPeakMemory);
    c = perft_gpu_host_bfs(&bb, &gs, color, maxDepth, preAllocatedBufferHost, 64 * 1024);
    printf("bfs_small %llu oom %d\n", c, (int)g_lastBfsOom);
    printf("cpu3 %llu\n", perft_cpu_dispatch(&bb, &gs, color, 3));
    float bf = 0;
    uint32 est = estimateLaunchDepth(&bb, &gs, color, &bf);
    printf("estimate %u bf %.2f\n", est, bf);

    // perft.cu's GPU loop with the given launch depth
    initTT(launchDepth, launchDepth, maxDepth, bf);
    for (int depth = 1; depth <= maxDepth; depth++) {
        int effectiveLD = launchDepth < depth ? launchDepth : depth;
        resetCallStats();
        setEffectiveLD(effectiveLD);
        perftLauncher(&bb, &gs, color, depth, effectiveLD);
        printTTStats();
        printf("oom_fallbacks %llu ld %d\n", getOomFallbackCount(), getEffectiveLD());
    }
    // a cleared device table gives the same count again
    clearDeviceTTs();
    perftLauncher(&bb, &gs, color, maxDepth, launchDepth);
    freeTT();

    // perft.cu's -cpu loop
    initTT(2, 2, 3, 30.0f);
    for (int depth = 1; depth <= 3; depth++) perftCPU(&bb, &gs, color, depth);
    freeTT();
    printf("done\n");
    return 0;
}
