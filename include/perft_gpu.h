/*
 * perft_gpu.h -- C ABI of the MI355X perft engine (libperft_gpu.so).
 *
 * Plain pointers, sizes and integer status codes; no exceptions cross this
 * boundary and no HIP or torch types appear in it (streams are passed as
 * `void *` holding a hipStream_t, NULL = the library's own stream).
 *
 * The reference exposes its engine as C++ functions in launcher.h:1-47 plus the
 * `perft_gpu` command line (perft.cu:13-186).  Each entry point below names the
 * reference interface it replaces.  The C++ interface with the reference's own
 * names and signatures (initGPU, perftLauncher, perft_gpu_host_bfs, the
 * g_lastBfsOom global, ...) is declared in include/perft_launcher.h and is
 * compiled into the same library on top of these entry points.
 *
 * Position encoding (chess.h:66-92): bb[0] = black occupancy, bb[1..3] = bits
 * 0..2 of the piece code (P=1 N=2 B=3 R=4 Q=5 K=6); state = white castle in
 * bits 1:0, black castle in 3:2, en-passant file+1 in 7:4; color 0 = white to
 * move, 1 = black.  Moves are u16: from | to << 6 | flags << 12
 * (chess.h:96-144).
 */
#ifndef PERFT_GPU_H
#define PERFT_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference has none: it exit(1)s on fatal errors and
 *      reports BFS out-of-memory through the global g_lastBfsOom,
 *      launcher.h:46; here OOM is the distinct PG_ERR_ARENA_OOM) ---- */
#define PG_OK             0
#define PG_ERR_ARENA_OOM  1   /* BFS level did not fit the arena: no count produced */
#define PG_ERR_INVALID    2   /* bad argument (depth, FEN, sizes) */
#define PG_ERR_HIP        3   /* a HIP runtime call failed */
#define PG_ERR_NOT_INIT   4   /* pg_init has not run */
#define PG_ERR_NO_DEVICE  5   /* no GPU visible */
#define PG_ERR_BUSY       6   /* serving: admission refused (budget / queue full) */

typedef struct pg_position {
    uint64_t bb[4];
    uint8_t  state;
    uint8_t  color;
    uint8_t  _pad[6];
} pg_position;

typedef struct pg_u128 {   /* uint128.h:7-109: host sums above the launch depth */
    uint64_t lo, hi;
} pg_u128;

/* ===================== initialisation (launcher.h:12-13) ===================== */

/* initGPU(int) + initMoveGen(), launcher.cu:59-88 and perft_kernels.cu:995-1134.
 * Selects `device`, allocates the BFS arena of `arenaBytes` (0 = automatic:
 * 60 % of free HBM, at most 192 GiB) and creates the library stream.
 * The move generator needs no tables, so there is nothing to upload. */
int pg_init(int device, uint64_t arenaBytes);
/* frees the arena, the TTs and the stream (perft.cu:179-182) */
int pg_shutdown(void);
/* last error text (thread-local), never NULL */
const char *pg_last_error(void);
uint64_t pg_arena_bytes(void);
/* caps the bytes of the arena a GPU call may use (0 = all of it); an injectable
 * arena size that lets a test force the OOM fallback (launcher.cu:600-610) */
void pg_set_arena_limit(uint64_t bytes);
/* launcher above the launch depth without the TT: 1 (default) expands the
 * frontier on the host, merges identical positions and runs them through the
 * batched engine; 0 keeps the reference's serial host recursion with one GPU
 * call per node (launcher.cu:572-715).  With the TT the recursion is always
 * serial: each GPU call merges identical positions per level and the table
 * carries subtrees between calls. */
void pg_set_frontier_batch(int on);
/* the launcher above the launch depth WITH the tables: 0 = the reference's
 * serial host recursion, one GPU call per node (launcher.cu:572-715);
 * 1 = the unique frontier at the launch depth (host tier probed first) over
 * four host threads, each with its own stream and quarter of the arena,
 * sharing both table tiers, so the GPU calls overlap on the device; 2 = the
 * unique frontier through the batched engine with the device table.  The
 * frontier is expanded on the host when it lies at most 4 plies down. */
int pg_set_tt_frontier(int mode);

/* ========================= host helpers (util.cpp) ========================== */

/* readFENString, util.cpp:7-95: placement, side, castling, EP; clocks ignored */
int pg_parse_fen(const char *fen, pg_position *pos);
/* FEN writer (no reference counterpart; EP square written only when a pawn of
 * the side to move stands next to the double-pushed pawn) */
int pg_write_fen(const pg_position *pos, char *out, int outLen);
/* legal moves of a position, host side (generateMoves, MoveGeneratorBitboard.h:2072) */
int pg_gen_moves(const pg_position *pos, uint16_t *moves, int *n);
/* perft(1) as the reference counts it (countMoves, MoveGeneratorBitboard.h:
 * 1563-1943): the number of legal moves, except that a side in check from a
 * pawn that has just double-pushed counts its en-passant captures once even
 * when two pawns can take (:1639-1649) */
int pg_count_moves(const pg_position *pos, int *n);
int pg_make_move(pg_position *pos, uint16_t move);
/* seeded random walk used by the workloads (no reference counterpart): SplitMix64
 * stream, moves sorted by (from, to, promotion), Lemire-unbiased pick; a walk that
 * reaches a terminal position, or ends on one when requireNonTerminal, is
 * redrawn from the same stream */
int pg_random_position(uint64_t seed, int plies, int requireNonTerminal, pg_position *pos);
/* plies U[pliesLo,pliesHi], then depth U[depthLo,depthHi], then the walk, all
 * from the stream of `seed` */
int pg_random_request(uint64_t seed, int pliesLo, int pliesHi, int depthLo, int depthHi,
                      pg_position *pos, int *plies, int *depth);

/* =========================== perft entry points ============================= */

/* perft_gpu_host_bfs, launcher.h:22 / perft_kernels.cu:620-990: one position,
 * all of `depth` on the GPU in one call (BFS levels + fused leaf).  Uses the
 * library arena, or [gpuBuffer, gpuBuffer+bufferBytes) when gpuBuffer != NULL
 * (a device pointer; lets a caller force OOM with a tiny buffer, as the
 * reference's bufferSize parameter does).  Returns PG_ERR_ARENA_OOM and no count
 * when a level does not fit.  useTT != 0 probes/stores the device TT.  It
 * returns only after the device is done with the buffer: the host learns each
 * level's size and the count from values the kernels publish to pinned host
 * memory, and a call that fails or ends early drains its stream. */
int pg_perft_bfs(const pg_position *pos, int depth, int useTT,
                 void *gpuBuffer, uint64_t bufferBytes, uint64_t *count);

/* perftLauncher, launcher.h:18 / launcher.cu:572-737: host recursion above the
 * launch depth (host TT tier when enabled), GPU BFS at the launch depth, and on
 * PG_ERR_ARENA_OOM a fallback one ply down that lowers the effective launch
 * depth.  launchDepth <= 0 picks it with pg_estimate_launch_depth; then,
 * without the tables and from depth 7, it is lowered to max(4, depth - 4)
 * (depth - 5 from depth 10; fewer plies where the frontier would pass 8
 * million positions):
 * the plies above it are expanded on the host, identical positions merged
 * (each counted once, times its multiplicity) and the rest run through the
 * batched engine, which computes fewer nodes than the tree holds.  An explicit
 * launchDepth is used as given; the perft_gpu command line passes the
 * reference's launch-depth policy, whose frontier of at most one ply repeats
 * no position, so every node is computed. */
int pg_perft(const pg_position *pos, int depth, int launchDepth, int useTT, pg_u128 *count);

/* perft_cpu_dispatch, launcher.h:22-23 / launcher.cu:762-816: the -cpu path,
 * host recursion only (host TT when enabled) */
int pg_perft_cpu(const pg_position *pos, int depth, int useTT, uint64_t *count);

/* estimateLaunchDepth, launcher.h:16 / launcher.cu:533-562 */
int pg_estimate_launch_depth(const pg_position *pos, float *branchingFactor, int *launchDepth);
/* getEffectiveLD / setEffectiveLD / getOomFallbackCount, launcher.h:33-35 */
int pg_get_effective_ld(void);
void pg_set_effective_ld(int ld);
uint64_t pg_get_oom_fallback_count(void);
/* g_lastBfsPeakMemory, launcher.h:38 */
uint64_t pg_last_bfs_peak_bytes(void);
/* leaf plies of the GPU BFS: 2 = fused two-ply leaf (default), 1 = one-ply leaf
 * over a materialised last level (switches.h:21 USE_FUSED_2LEVEL_LEAF) */
void pg_set_leaf_plies(int plies);
/* shape of the fused two-ply leaf: 0 = 64 entries per wave (default),
 * 1 = 32 entries per wave */
void pg_set_leaf2_variant(int variant);
/* grid of the grid-stride leaf kernels in blocks per CU (default 16; the one-ply
 * leaf uses twice as many) -- a tuning knob of this engine, no reference
 * counterpart */
void pg_set_leaf_grid(int blocksPerCU);
/* slider attacks of the fused two-ply leaf's move counter: 0 = computed
 * fills (default), 1 = fancy magic lookups (the reference's method,
 * MoveGeneratorBitboard.h:589-648) with the tables in device memory, 2 = the
 * same with the bishop table in each block's LDS.  The first call with 1 or 2
 * builds the tables (own multipliers, checked for every blocker subset) and
 * uploads them; needs pg_init. */
int pg_set_slider_attacks(int mode);
/* host-only check of the magic tables (no GPU needed): builds them and looks
 * up `samples` random occupancies per square and slider against a ray walk;
 * *tries = multipliers tried by the search.  PG_OK or PG_ERR_INVALID. */
int pg_magic_check(uint64_t samples, uint64_t *tries);
/* decimal text of a 128-bit count (uint128.h:46-108) */
int pg_u128_to_string(uint64_t lo, uint64_t hi, char *buf, int n);

/* ====================== transposition tables (tt.h) ========================= */

/* initTT, launcher.h:26 / launcher.cu:128-351: one depth-tagged device table of
 * 16-B entries in HBM (power-of-two slots in deviceMB) and the host lossless
 * tier of hostMB.  0 = automatic sizes. */
int pg_tt_init(uint64_t deviceMB, uint64_t hostMB);
int pg_tt_clear(void);          /* clearDeviceTTs, launcher.cu:489 */
int pg_tt_free(void);           /* freeTT */
/* printTTStats-style counters: device probes/hits/stores, host probes/hits/stores */
int pg_tt_stats(uint64_t out[6]);
/* TT snapshot / resume (no reference counterpart; its 27.8 h perft 14 run had
 * no resume, README.md:63): write both tiers to a file, or load a file into
 * tables created by pg_tt_init with the same sizes (a file that fails part
 * way through leaves the tables cleared) */
int pg_tt_save(const char *path);
int pg_tt_load(const char *path);

/* ===================== kernel-level ABI (device pointers) =================== */
/* One BFS level of the reference's host-driven loop (perft_kernels.cu:681-835),
 * callable on its own.  All pointers are device pointers; n/c are entry counts;
 * `color` is the side to move in the PARENT boards. */

/* makemove_and_count_moves_kernel, perft_kernels.cu:201-257: child board i =
 * make(parentBoards[parentIdx[i]], moves[i]); counts[i] = legal moves of it.
 * tileSums (ceil(n/256) u32) receives per-256-entry sums for pg_scan_u32. */
int pg_level_makecount(const void *parentBoards, const uint8_t *parentStates, const uint32_t *parentIdx,
                       const uint16_t *moves, int color, uint64_t n, void *boards, uint8_t *states,
                       uint32_t *counts, uint32_t *tileSums, void *stream);
/* cub::DeviceScan::InclusiveSum replacement, perft_kernels.cu:786: in-place
 * inclusive scan of n u32 with a wave64 block scan.  tileSums[t] is the sum
 * of data[256t .. 256t+255]; if it is NULL the sums are computed here.  scratch
 * must hold 8*(ceil(n/256)+4) bytes, plus 4*ceil(n/256) more when tileSums is
 * NULL; *total (host pointer, may be NULL) receives the 64-bit sum
 * (synchronises the stream). */
int pg_scan_u32(uint32_t *data, uint64_t n, const uint32_t *tileSums, void *scratch, uint64_t *total, void *stream);
/* No counterpart in the reference: pg_scan_u32 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a 9 MiB static device
 * array per GPU, held until the process ends; n may be at most 2^28
 * (268,435,456) entries, larger n returns PG_ERR_INVALID.  Calls no hipMalloc
 * or hipFree.  Every call on one GPU shares that GPU's workspace, so calls
 * that overlap on one GPU (other streams or threads) race. */
int pg_scan_u32_static(uint32_t *data, uint64_t n, const uint32_t *tileSums, uint64_t *total, void *stream);
/* merge_path_partitions + mp_interval_expand_kernel, perft_kernels.cu:77-173:
 * parentIdx[j] = #{p < n : incl[p] <= j} for j < c.  scratch: 4*(ceil(c/2048)+1) B */
int pg_expand(const uint32_t *incl, uint64_t n, uint64_t c, uint32_t *parentIdx, void *scratch, void *stream);
/* generate_moves_kernel, perft_kernels.cu:263-311: moves of boards[i] written
 * to childMoves[incl[i-1] .. incl[i]) (LDS-staged, coalesced) */
int pg_genmoves(const void *boards, const uint8_t *states, const uint32_t *incl, int color, uint64_t n,
                uint16_t *childMoves, void *stream);
/* makeMove_and_perft_leaf_kernel (:313-350, plies = 1) and
 * fused_2level_leaf_kernel (:382-480, plies = 2): sums perft(plies) of
 * make(parentBoards[parentIdx[i]], moves[i]) over i < n into *sum (device u64,
 * accumulated, not cleared). */
int pg_leaf(const void *parentBoards, const uint8_t *parentStates, const uint32_t *parentIdx,
            const uint16_t *moves, int color, uint64_t n, int plies, uint64_t *sum, void *stream);

/* ====================== live kernel profiling (no reference counterpart;
 * the reference's diagnostics are compiled out, switches.h:53) ============== */
typedef struct pg_kernel_stat {
    char     name[32];
    uint64_t launches;
    double   ms;      /* summed HIP-event time of the launches */
    double   bytes;   /* summed algorithmic bytes of the launches */
    double   items;   /* summed entries processed */
    /* the launches with the most algorithmic bytes (a repeated call's largest
     * level): their bytes each, their mean time and their number */
    double   big_bytes;
    double   big_ms;
    uint64_t big_launches;
} pg_kernel_stat;
/* VALU calibration (no reference counterpart): a kernel whose loop is exactly
 * 32 VALU instructions per iteration in 8 independent chains, on wavesPerSimd
 * (1-8) waves per SIMD of every CU, `iters` iterations.  kind 0 v_xor_b32,
 * 1 v_bitop3_b32, 2 v_lshlrev_b64, 3 v_bcnt_u32_b32, 4 v_lshl_add_u64, 5 the
 * move counter's mix (bitwise, 64-bit shift, select, add, popcount),
 * 6 v_lshlrev_b32, 7 v_alignbit_b32, 8 v_cndmask_b32, 9 v_add_u32,
 * 10 v_and_or_b32, 11 v_sub_co_u32, 12 v_and_b32, 13 v_or_b32, 14 v_mov_b32,
 * 15 v_bfrev_b32, 16 v_or3_b32, 17 v_ffbl_b32, 18 v_add3_u32,
 * 19 v_cmp_eq_u32, 20 v_xor_b32 in its VOP3 encoding.  *ms is
 * the launch's time, *waveInsts the VALU instructions it issued: their
 * quotient is the VALU issue rate the leaf kernels are measured against. */
int pg_valu_probe(int kind, int wavesPerSimd, uint64_t iters, double *ms, uint64_t *waveInsts);
int pg_profile_enable(int on);
/* which kernels an enabled profiler times: bit k for the k-th kernel that
 * pg_profile_get names (makecount, scan, expand, genmoves, leaf1, leaf2,
 * batch_group, tt); all by default.  Each timed launch adds two stream
 * events, so timing only the dominant kernel keeps a timed run closest to
 * an untimed one. */
int pg_profile_kernels(uint32_t mask);
int pg_profile_reset(void);
int pg_profile_get(pg_kernel_stat *out, int maxOut, int *nOut);
int pg_device_synchronize(void);
/* device memory for callers that have no HIP runtime of their own (a second
 * runtime in the same process cannot share the device): kind 0 = host to
 * device, 1 = device to host, 2 = device to device; synchronous */
int pg_device_malloc(uint64_t bytes, void **ptr);
int pg_device_free(void *ptr);
int pg_memcpy(void *dst, const void *src, uint64_t bytes, int kind);

/* ============================ batched perft ================================= */
/* No reference counterpart (the reference perfts one position per call).
 * counts[i] = perft(position i, depths[i]).  pg_perft_batch runs the whole batch
 * as one multi-root BFS whose leaf phase is a persistent work-queue kernel;
 * pg_perft_batch_plain calls pg_perft_bfs once per position (the baseline). */
typedef struct pg_batch_options {
    int gridPerCU;        /* blocks per CU of the persistent level kernels and the
                             table leaf (1-32), 0 = autotuned (4, 8 or 16); the
                             blocks are 256 threads, the scan's tile size */
    int leafPlies;        /* 1 or 2, 0 = autotuned */
    int useGraph;         /* capture the level sequence in a hipGraph */
    int useTT;
    int ttBudgetMB;       /* with useTT: the part of the device table a group
                             uses (its first power-of-two entries within this
                             many MB, at least 1 MB, at most the table); 0 =
                             autotuned (the whole table, 1 GiB or 128 MiB) */
    int reserved[3];
} pg_batch_options;
int pg_perft_batch(int n, const pg_position *positions, const uint8_t *depths, const pg_batch_options *opt,
                   uint64_t *counts);
int pg_perft_batch_plain(int n, const pg_position *positions, const uint8_t *depths, int useTT, uint64_t *counts);
/* the autotuner's cached choices, seven ints per row: depth, log2 of the group
 * size, table requested, leaf plies, table used, grid blocks per CU, log2 of
 * the device-table entries used (0 when the table is not used) */
int pg_batch_tuned(int *rows, int maxRows, int *nOut);

/* ================================ serving =================================== */
/* No reference counterpart.  A request loop that admits perft queries as they
 * arrive, gates admission by an HBM byte budget, keeps BFS buffers and the TT
 * resident, and overlaps requests on the device (continuous batching). */
typedef struct pg_serve_config {
    uint64_t hbmBudgetBytes;  /* BFS buffers + TT */
    uint64_t ttBytes;         /* part of the budget given to the shared device TT */
    int maxInflight;          /* requests on the device at once (slots, at most 16) */
    int maxBatch;             /* queued requests of one depth a free slot takes at
                                 once and runs as one multi-root device pass (1 =
                                 one request per pass) */
    int maxQueued;            /* requests waiting beyond which pg_serve_submit
                                 returns PG_ERR_BUSY (0 = no limit) */
    int reserved[3];
} pg_serve_config;
typedef struct pg_serve_result {
    uint64_t id;
    uint64_t count;
    double   submitTime, startTime, endTime;  /* seconds, library clock */
    int      status;
    int      _pad;
} pg_serve_result;
/* One server at a time: while it exists its slots own the library arena and
 * share the device table, so pg_perft_bfs, pg_perft, pg_perft_batch,
 * pg_perft_batch_plain, pg_shutdown, the pg_tt_* management calls and a
 * second pg_serve_create return PG_ERR_BUSY until pg_serve_destroy. */
int pg_serve_create(const pg_serve_config *cfg, void **handle);
/* PG_ERR_BUSY when maxQueued requests already wait, PG_ERR_INVALID for a bad
 * position or a depth outside 0..16 */
int pg_serve_submit(void *handle, uint64_t id, const pg_position *pos, int depth);
/* copies up to maxResults finished results; waits up to timeoutMs for one */
int pg_serve_poll(void *handle, pg_serve_result *results, int maxResults, int timeoutMs, int *nOut);
int pg_serve_destroy(void *handle);
double pg_clock(void);

#ifdef __cplusplus
}
#endif
#endif
