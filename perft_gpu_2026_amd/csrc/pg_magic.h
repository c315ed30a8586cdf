// pg_magic.h -- fancy magic slider-attack tables (the reference's
// MoveGeneratorBitboard.h:589-648 / GlobalVars.cpp:4-161 lookup scheme), an
// alternative to the computed fills of pg_chess.h for the fused leaf's
// counter, selected with pg_set_slider_attacks.
//
// attacks(sq, occ) = table[offset[sq] + (((occ & mask[sq]) * magic[sq]) >> shift[sq])]
// with mask = the squares of the slider's lines from sq that can block (board
// edges and sq excluded), shift = 64 - popcount(mask).  Rook squares take
// 2^10..2^12 entries each (102,400 in all, 800 KB), bishop squares 2^5..2^9
// (5,248, 41 KB).  The magic multipliers are this build's own: found at
// set-up by a seeded search (SplitMix64 candidates with few bits set), and
// each accepted one is checked against a ray walk for every blocker subset
// of its square, so a lookup is exact by construction.
#pragma once
#include <stdint.h>

namespace pg {

struct MagicSquare {
    uint64_t mask, magic;
    uint32_t offset, shift;
    uint64_t pad;  // 32-B rows
};

constexpr int MAGIC_ROOK_ENTRIES = 102400;
constexpr int MAGIC_BISHOP_ENTRIES = 5248;
constexpr int MAGIC_ENTRIES = MAGIC_ROOK_ENTRIES + MAGIC_BISHOP_ENTRIES;

struct MagicTables {
    MagicSquare sq[2][64];        // [0] rook, [1] bishop
    uint64_t att[MAGIC_ENTRIES];  // rook entries, then bishop entries (bishop offsets count from there)
};

// Builds the tables (deterministic: a fixed seed), checks every entry, and
// returns the number of candidate multipliers tried, or 0 if a square found
// none (never seen; the search is bounded).
uint64_t buildMagicTables(MagicTables &t);

// attacks by ray walk from sq over occ (board edges stop the rays; the first
// occupied square is included), the check the tables are built against
uint64_t rayWalkAttacks(int sq, uint64_t occ, bool bishop);

}  // namespace pg
