/*
 * oracle.h -- CPU oracle for the perft engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry and the cpu_baseline leg of bench.py may load it; the product
 * (perft_gpu_2026_amd) never links, loads or calls it.
 *
 * It is a plain-C restatement of the reference's -cpu path:
 *   perft recursion          launcher.cu:762-816   (perft_cpu<>, perft_cpu_dispatch)
 *   legal move counting      MoveGeneratorBitboard.h:1563-1943 (countMoves*)
 *   legal move generation    MoveGeneratorBitboard.h:977-1554  (generateMoves*)
 *   make move                MoveGeneratorBitboard.h:1945-2067 (makeMove, host flavour)
 *   FEN reading              util.cpp:7-95
 *   position layout          chess.h:66-144 (QuadBitBoard, GameState, CMove)
 * written from the semantics, with classical ray-scan slider attacks instead
 * of the reference's magic tables, plus an independent brute-force
 * (pseudo-legal + king-in-check filter) generator used to pin it.
 *
 * Pinning (what fixes the results, see tests/test_oracle.py):
 *   - published perft counts of six standard positions (start, Kiwipete,
 *     positions 3-6) and the reference README's own counts (README.md:51-103);
 *   - count == |generate| on every node of full trees, except where the
 *     reference's own count differs from its generator (below);
 *   - fast path == brute-force path on seeded random positions;
 *   - the 218-move position of chess.h:177.
 *
 * Counting (perft(1)) follows the reference's countMovesOutOfCheck, which
 * counts the en-passant evasions of a side in check as one move even when two
 * pawns can take (MoveGeneratorBitboard.h:1639-1649; its generator emits both,
 * :1085-1090).  Position 3 at depth 10 shows it: 860,322,381,070 (the
 * reference's README.md count) against 860,322,602,309 legal leaves.
 */
#ifndef PG_ORACLE_H
#define PG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Position: quad-bitboard bb[4] (bb0 = black occupancy, bb1..3 = bits of the
 * piece code P=1 N=2 B=3 R=4 Q=5 K=6) + state byte (bits 1:0 white castle,
 * 3:2 black castle, 7:4 en-passant file+1) + side to move (0 white, 1 black). */

void     oracle_init(void);
int      oracle_parse_fen(const char *fen, uint64_t bb[4], uint8_t *state, uint8_t *color);
int      oracle_write_fen(const uint64_t bb[4], uint8_t state, uint8_t color, char *out, int outLen);

uint32_t oracle_count_moves(const uint64_t bb[4], uint8_t state, uint8_t color);
uint32_t oracle_gen_moves(const uint64_t bb[4], uint8_t state, uint8_t color, uint16_t *moves);
void     oracle_make_move(uint64_t bb[4], uint8_t *state, uint8_t color, uint16_t move);

uint64_t oracle_perft(const uint64_t bb[4], uint8_t state, uint8_t color, uint32_t depth);
uint64_t oracle_perft_mt(const uint64_t bb[4], uint8_t state, uint8_t color, uint32_t depth, int threads);
/* perft of many positions on `threads` threads; counts[i] = perft(pos i, depths[i]) */
void     oracle_perft_batch(int n, const uint64_t *bbs, const uint8_t *states, const uint8_t *colors,
                            const uint32_t *depths, uint64_t *counts, int threads);

/* brute force: pseudo-legal moves filtered by "own king not attacked after make" */
uint32_t oracle_brute_gen_moves(const uint64_t bb[4], uint8_t state, uint8_t color, uint16_t *moves);
/* the reference's perft(1) count from the brute generator: in check, all
 * en-passant captures together count once (MoveGeneratorBitboard.h:1639-1649) */
uint32_t oracle_brute_count_moves(const uint64_t bb[4], uint8_t state, uint8_t color);
uint64_t oracle_perft_brute(const uint64_t bb[4], uint8_t state, uint8_t color, uint32_t depth);

/* seeded random walk: SplitMix64(seed), legal moves sorted by (from,to,promo),
 * Lemire-unbiased pick; a walk that hits a terminal position (or ends on one
 * when requireNonTerminal) is redrawn from the same stream. returns redraws. */
int      oracle_random_position(uint64_t seed, int plies, int requireNonTerminal,
                                uint64_t bb[4], uint8_t *state, uint8_t *color);
/* batched/serving draw: plies U[pliesLo,pliesHi] then depth U[depthLo,depthHi]
 * from SplitMix64(seed), then a non-terminal walk on the same stream */
int      oracle_random_request(uint64_t seed, int pliesLo, int pliesHi, int depthLo, int depthHi,
                               uint64_t bb[4], uint8_t *state, uint8_t *color, int *plies, int *depth);

#ifdef __cplusplus
}
#endif
#endif
