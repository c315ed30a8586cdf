// pg_magic.hip -- host construction of the fancy magic slider tables
// (pg_magic.h).  The search follows the published fancy-magic method
// (chessprogramming.org "Looking for Magics"): for each square, enumerate
// every subset of the blocker mask (carry-rippler), compute its attack set by
// a ray walk, and try sparse random multipliers until no two subsets with
// different attack sets share an index.  The reference ships its own
// multipliers (GlobalVars.cpp); none of them are used here.
#include <string.h>

#include <vector>

#include "pg_magic.h"

namespace pg {

uint64_t rayWalkAttacks(int sq, uint64_t occ, bool bishop)
{
    static const int DR[2][4] = {{1, -1, 0, 0}, {1, 1, -1, -1}}, DF[2][4] = {{0, 0, 1, -1}, {1, -1, 1, -1}};
    const int k = bishop ? 1 : 0;
    uint64_t a = 0;
    for (int d = 0; d < 4; d++) {
        int r = sq / 8 + DR[k][d], f = sq % 8 + DF[k][d];
        while (r >= 0 && r < 8 && f >= 0 && f < 8) {
            uint64_t b = 1ULL << (r * 8 + f);
            a |= b;
            if (occ & b) break;
            r += DR[k][d];
            f += DF[k][d];
        }
    }
    return a;
}

// squares whose occupancy can change the attack set: the ray squares short of
// the board edge in each direction
static uint64_t blockerMask(int sq, bool bishop)
{
    static const int DR[2][4] = {{1, -1, 0, 0}, {1, 1, -1, -1}}, DF[2][4] = {{0, 0, 1, -1}, {1, -1, 1, -1}};
    const int k = bishop ? 1 : 0;
    uint64_t m = 0;
    for (int d = 0; d < 4; d++) {
        int r = sq / 8 + DR[k][d], f = sq % 8 + DF[k][d];
        while (r + DR[k][d] >= 0 && r + DR[k][d] < 8 && f + DF[k][d] >= 0 && f + DF[k][d] < 8) {
            m |= 1ULL << (r * 8 + f);
            r += DR[k][d];
            f += DF[k][d];
        }
    }
    return m;
}

static uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t buildMagicTables(MagicTables &t)
{
    memset(&t, 0, sizeof t);
    uint64_t rng = 0x6D61676963ULL, tries = 0;
    uint32_t offset = 0;
    std::vector<uint64_t> occ, att, used;
    std::vector<uint32_t> epoch;
    for (int k = 0; k < 2; k++) {
        const bool bishop = k == 1;
        for (int sq = 0; sq < 64; sq++) {
            const uint64_t mask = blockerMask(sq, bishop);
            const int bits = __builtin_popcountll(mask);
            const uint32_t n = 1u << bits;
            occ.assign(n, 0);
            att.assign(n, 0);
            uint64_t sub = 0;
            for (uint32_t i = 0; i < n; i++) {  // every subset of the mask
                occ[i] = sub;
                att[i] = rayWalkAttacks(sq, sub, bishop);
                sub = (sub - mask) & mask;
            }
            used.assign(n, 0);
            epoch.assign(n, 0);
            uint64_t magic = 0;
            for (uint32_t e = 1; e < 100000000u; e++) {
                tries++;
                const uint64_t m = splitmix(rng) & splitmix(rng) & splitmix(rng);
                if (__builtin_popcountll((mask * m) >> 56) < 6) continue;
                bool ok = true;
                for (uint32_t i = 0; i < n && ok; i++) {
                    const uint32_t idx = (uint32_t)((occ[i] * m) >> (64 - bits));
                    if (epoch[idx] != e) epoch[idx] = e, used[idx] = att[i];
                    else if (used[idx] != att[i]) ok = false;
                }
                if (ok) {
                    magic = m;
                    break;
                }
            }
            if (!magic) return 0;
            MagicSquare &q = t.sq[k][sq];
            q.mask = mask;
            q.magic = magic;
            // bishop offsets count from the start of the bishop entries
            q.offset = bishop ? offset - (uint32_t)MAGIC_ROOK_ENTRIES : offset;
            q.shift = (uint32_t)(64 - bits);
            for (uint32_t i = 0; i < n; i++) t.att[offset + ((occ[i] * magic) >> (64 - bits))] = att[i];
            // every subset through the finished table against the ray walk
            const uint64_t *base = bishop ? t.att + MAGIC_ROOK_ENTRIES : t.att;
            for (uint32_t i = 0; i < n; i++)
                if (base[q.offset + ((occ[i] * magic) >> q.shift)] != att[i]) return 0;
            offset += n;
        }
        if (k == 0 && offset != (uint32_t)MAGIC_ROOK_ENTRIES) return 0;
    }
    return offset == (uint32_t)MAGIC_ENTRIES ? tries : 0;
}

}  // namespace pg

// host-only check of the table construction (no GPU needed): builds the
// tables, then looks up `samples` pseudo-random occupancies per square and
// slider against the ray walk; *tries = candidate multipliers tried.
extern "C" int pg_magic_check(uint64_t samples, uint64_t *tries)
{
    std::vector<pg::MagicTables> t(1);
    const uint64_t n = pg::buildMagicTables(t[0]);
    if (tries) *tries = n;
    if (!n) return 2;  // PG_ERR_INVALID
    uint64_t x = 0x853C49E6748FEA9BULL;
    for (uint64_t i = 0; i < samples; i++) {
        for (int k = 0; k < 2; k++) {
            for (int sq = 0; sq < 64; sq++) {
                x ^= x << 13, x ^= x >> 7, x ^= x << 17;
                const uint64_t occ = x & (x >> 5);
                const pg::MagicSquare &q = t[0].sq[k][sq];
                const uint64_t *base = k ? t[0].att + pg::MAGIC_ROOK_ENTRIES : t[0].att;
                if (base[q.offset + (((occ & q.mask) * q.magic) >> q.shift)] != pg::rayWalkAttacks(sq, occ, k == 1))
                    return 2;
            }
        }
    }
    return 0;
}
