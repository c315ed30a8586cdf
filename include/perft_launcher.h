/*
 * perft_launcher.h -- the reference's C++ launcher interface, implemented by
 * libperft_gpu.so on top of the C ABI in perft_gpu.h.
 *
 * A C++ caller written against the reference's launcher.h:1-47 (and
 * utils.h:7 for readFENString) compiles against this header unchanged and
 * links with -lperft_gpu: the types, names and signatures below are the
 * reference's, so the symbols match (C++ linkage, global namespace).  Each
 * declaration names the reference function it replaces; the behaviour notes
 * say where this engine differs.  Include either this header or the
 * reference's launcher.h/chess.h, not both.
 *
 * Threading: like the reference, a single host thread drives these calls.
 */
#ifndef PERFT_LAUNCHER_H
#define PERFT_LAUNCHER_H
#include <stddef.h>

/* chess.h:16-19 */
typedef unsigned char uint8;
typedef unsigned short uint16;
typedef unsigned int uint32;
typedef unsigned long long uint64;

/* chess.h:66-76: bb[0] = black occupancy, bb[1..3] = bits 0..2 of the piece
 * code (P=1 N=2 B=3 R=4 Q=5 K=6) */
struct QuadBitBoard {
    uint64 bb[4];
};

/* chess.h:78-92: bits 1:0 white castle, 3:2 black castle, 7:4 en-passant
 * file + 1; the side to move is passed separately */
struct GameState {
    union {
        struct {
            uint8 whiteCastle : 2;
            uint8 blackCastle : 2;
            uint8 enPassent : 4;
        };
        uint8 raw;
    };
};

/* ---- globals (launcher.h:6-11, 37-47; perft.cu:7-8) ---- */
/* the BFS arena: set by initGPU to the library's device arena and its size */
extern void *preAllocatedBufferHost;
extern uint64 g_preAllocatedMemorySize;
/* tables on/off (default on); -nott clears it */
extern bool g_useTT;
/* peak arena bytes of the last perft_gpu_host_bfs call */
extern size_t g_lastBfsPeakMemory;
/* set when the last perft_gpu_host_bfs call ran out of arena (count 0) */
extern bool g_lastBfsOom;
/* table budgets in MB for initTT (0 = automatic), perft.cu:7-8 */
extern int g_deviceTTBudgetMB;
extern int g_hostTTBudgetMB;

/* initGPU, launcher.cu:59-88: selects the device and allocates the BFS arena
 * (with the tables on, a third of the device memory, leaving the rest to
 * initTT; without them 60 % of free memory, at most 192 GiB -- the reference
 * caps its arena at 16 GiB).  Prints the reference's two lines; exit(1) on
 * failure, as the reference does. */
void initGPU(int gpu);
/* initMoveGen, launcher.cu:823-826: nothing to build (the move generator's
 * only table is compiled in) */
void initMoveGen();

/* estimateLaunchDepth, launcher.cu:533-562 */
uint32 estimateLaunchDepth(QuadBitBoard *pos, GameState *gs, uint8 rootColor, float *outBranchingFactor = nullptr);

/* perftLauncher, launcher.cu:718-737: host recursion above launchDepth (host
 * table tier when g_useTT), GPU BFS at it, OOM fallback one ply down; prints
 * "Perft(%02d): <count>, time: <s> seconds, nps: <n>" */
void perftLauncher(QuadBitBoard *pos, GameState *gs, uint8 rootColor, uint32 depth, int launchDepth);
/* perft_gpu_host_bfs, perft_kernels.cu:620-990: the whole depth on the GPU in
 * [gpuBuffer, gpuBuffer + bufferSize) (the arena when gpuBuffer is NULL);
 * returns 0 and sets g_lastBfsOom when a level does not fit */
uint64 perft_gpu_host_bfs(QuadBitBoard *pos, GameState *gs, uint8 rootColor, int depth, void *gpuBuffer,
                          size_t bufferSize);

/* perftCPU, launcher.cu:739-759: timed -cpu perft with its Perft line */
void perftCPU(QuadBitBoard *pos, GameState *gs, uint8 rootColor, uint32 depth);
/* perft_cpu_dispatch, launcher.cu:807-816: single-threaded host perft (host
 * table when g_useTT and initTT made one) */
uint64 perft_cpu_dispatch(QuadBitBoard *pos, GameState *gs, uint8 color, uint32 depth);

/* initTT, launcher.cu:128-351: the reference makes one device table per depth
 * 3..maxLaunchDepth-1 and one host table per depth launchDepth..maxDepth;
 * this engine makes one depth-tagged device table (when maxLaunchDepth > 3)
 * and one host tier (when any host depth exists), sized by
 * g_deviceTTBudgetMB / g_hostTTBudgetMB, or automatically: 90 % of the free
 * device memory and min(90 % of RAM, 64 GiB).  No-op when !g_useTT. */
void initTT(int launchDepth, int maxLaunchDepth, int maxDepth, float branchingFactor);
/* clearDeviceTTs, launcher.cu:489-503: zero the device table (the host tier
 * keeps its entries) */
void clearDeviceTTs();
/* freeTT, launcher.cu:505-531 */
void freeTT();
/* printTTStats, launcher.cu:382-487: table probes and hits and the peak
 * arena use since resetCallStats (the reference prints these only in its
 * verbose build) */
void printTTStats();
/* resetCallStats, launcher.cu:353-376: zero the OOM fallback count and the
 * counters printTTStats reports */
void resetCallStats();
/* getOomFallbackCount, getEffectiveLD, setEffectiveLD, launcher.cu:380, 566-568 */
uint64 getOomFallbackCount();
int getEffectiveLD();
void setEffectiveLD(int ld);

/* readFENString, util.cpp:7-95 (utils.h:7): placement, side, castling, EP;
 * the clocks are ignored; an unreadable FEN gives an empty board, which every
 * perft call above refuses */
void readFENString(const char fen[], QuadBitBoard *qbb, GameState *gs, uint8 *color);

#endif
