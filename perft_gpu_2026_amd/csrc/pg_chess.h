// pg_chess.h -- position layout and legal move arithmetic, shared by the host
// (recursion above the launch depth, root moves, -cpu mode) and the gfx950
// kernels.
//
// Layout (kept from the reference, chess.h:66-144):
//   QuadBitBoard: bb[0] = black occupancy, bb[1..3] = bits 0..2 of the piece
//     code P=1 N=2 B=3 R=4 Q=5 K=6 (32 B, 32-B aligned: one 2x dwordx4 load).
//   GameState: one byte, bits 1:0 white castle (1 = king side, 2 = queen
//     side), bits 3:2 black castle, bits 7:4 en-passant file + 1.
//   The side to move is not stored; it is a template parameter.
//   Move: u16 = from | to << 6 | flags << 12, flags as chess.h:120-144.
//
// Design for CDNA4 (own design, not the reference's): every routine is
// branch-light so a 64-lane wavefront stays converged.
//   * slider attacks are computed, not looked up: per-direction Kogge-Stone
//     occluded fills for whole piece sets, the rank-wise subtraction trick for
//     east/west, and hyperbola quintessence with a full 64-bit bit reverse for
//     single-square lines; the only table is 64 rows of per-square masks.
//   * set arithmetic runs on 32-bit halves (BB below): gfx950's
//     v_bitop3_b32 evaluates any three-input bitwise function in one
//     instruction, and the compiler forms it only for 32-bit operations;
//     64-bit shifts stay single v_lshlrev_b64 / v_lshrrev_b64 instructions.
//   * pins and checks come from the few enemy sliders on the king's lines;
//     pinned pieces are handled by per-line source masks instead of per-piece
//     loops (a piece pinned on a line may move only along it).
//   * legal-move counting is set-wise: slider and pawn counts are popcounts
//     of per-direction target sets, knights one table row each; only en
//     passant (rare) branches.
// The semantics follow the reference's countMoves/makeMove
// (MoveGeneratorBitboard.h:1563-2067), including its en-passant rules
// (rank discovered check by rook/queen only; EP evasion only when the captured
// pawn is the checker) and castling by rights + empty, unattacked path.
#pragma once
#include <stdint.h>

#include "pg_magic.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PG_HD __host__ __device__ __forceinline__
#else
#define PG_HD inline __attribute__((always_inline))
#endif

namespace pg {

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t u8;

enum { WHITE = 0, BLACK = 1 };
enum { PAWN = 1, KNIGHT = 2, BISHOP = 3, ROOK = 4, QUEEN = 5, KING = 6 };
enum {
    MF_QUIET = 0, MF_DPUSH = 1, MF_KCASTLE = 2, MF_QCASTLE = 3, MF_CAPTURE = 4, MF_EP = 5,
    MF_PROMO = 8  // 8 N, 9 B, 10 R, 11 Q
};
enum { MAX_MOVES = 256 };

struct alignas(32) QuadBitBoard {
    u64 bb[4];
};
struct GameState {
    u8 raw;
};

static constexpr u64 FILE_A = 0x0101010101010101ULL;
static constexpr u64 FILE_H = 0x8080808080808080ULL;
static constexpr u64 NOT_A = ~FILE_A;
static constexpr u64 NOT_H = ~FILE_H;
static constexpr u64 NOT_AB = ~(FILE_A | (FILE_A << 1));
static constexpr u64 NOT_GH = ~(FILE_H | (FILE_H >> 1));
static constexpr u64 RANK_1 = 0xFFULL;
static constexpr u64 RANK_3 = 0xFF0000ULL;
static constexpr u64 RANK_6 = 0xFF0000000000ULL;
static constexpr u64 RANK_8 = 0xFF00000000000000ULL;
static constexpr u64 DIAG_A1H8 = 0x8040201008040201ULL;
static constexpr u64 DIAG_A8H1 = 0x0102040810204080ULL;

PG_HD int popc(u64 x) { return __builtin_popcountll(x); }
PG_HD int lsb(u64 x) { return __builtin_ctzll(x); }

// ------------------------------------------------------------ 64-bit sets
// A set of squares as two 32-bit halves.  Bitwise operators work per half,
// so chains of them fuse into v_bitop3_b32 / v_and_or_b32 / v_or3_b32 (the
// compiler keeps 64-bit bitwise operations whole and never fuses those);
// add, subtract, popcount and find-first go through the 64-bit value (the
// same v_add/sub_co pairs, v_bcnt and v_ffbl either way).  Built with
// -fno-slp-vectorize so the halves are not merged back into 64-bit vectors.
struct BB {
    u32 lo, hi;
    PG_HD constexpr BB() : lo(0), hi(0) {}
    PG_HD constexpr BB(u64 x) : lo((u32)x), hi((u32)(x >> 32)) {}
    PG_HD constexpr BB(u32 l, u32 h, int) : lo(l), hi(h) {}
    PG_HD constexpr u64 v() const { return ((u64)hi << 32) | lo; }
    PG_HD constexpr explicit operator bool() const { return v() != 0; }
};
PG_HD constexpr BB operator&(BB a, BB b) { return BB(a.lo & b.lo, a.hi & b.hi, 0); }
PG_HD constexpr BB operator|(BB a, BB b) { return BB(a.lo | b.lo, a.hi | b.hi, 0); }
PG_HD constexpr BB operator^(BB a, BB b) { return BB(a.lo ^ b.lo, a.hi ^ b.hi, 0); }
PG_HD constexpr BB operator~(BB a) { return BB(~a.lo, ~a.hi, 0); }
PG_HD constexpr BB &operator&=(BB &a, BB b) { return a = a & b; }
PG_HD constexpr BB &operator|=(BB &a, BB b) { return a = a | b; }
PG_HD constexpr BB &operator^=(BB &a, BB b) { return a = a ^ b; }
PG_HD constexpr BB operator-(BB a, BB b) { return BB(a.v() - b.v()); }
PG_HD constexpr bool operator==(BB a, BB b) { return a.v() == b.v(); }
PG_HD constexpr bool operator!=(BB a, BB b) { return a.v() != b.v(); }
PG_HD constexpr bool operator!(BB a) { return a.v() == 0; }
PG_HD int popc(BB x) { return __builtin_popcountll(x.v()); }
PG_HD int lsb(BB x) { return __builtin_ctzll(x.v()); }
PG_HD BB lowBit(BB x) { return BB(x.v() & (0 - x.v())); }
PG_HD BB clearLow(BB x) { return BB(x.v() & (x.v() - 1)); }
PG_HD BB bit(int sq) { return BB(1ULL << sq); }

// (a & m) | c.  On gfx950 one v_bitop3_b32 per half: the compiler would pick
// v_and_or_b32, which never issues together with another wave's instruction,
// while v_bitop3_b32 (like v_and/v_or/v_xor) can (measured with the VALU
// calibration kernel: 0.42 against 0.23 instructions per SIMD and cycle).
PG_HD BB andOr(BB a, BB m, BB c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return BB(__builtin_amdgcn_bitop3_b32(a.lo, m.lo, c.lo, 0xEA), __builtin_amdgcn_bitop3_b32(a.hi, m.hi, c.hi, 0xEA),
              0);
#else
    return (a & m) | c;
#endif
}
// a | b | c, one v_bitop3_b32 per half instead of v_or3_b32 (which never pairs)
PG_HD BB or3(BB a, BB b, BB c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return BB(__builtin_amdgcn_bitop3_b32(a.lo, b.lo, c.lo, 0xFE), __builtin_amdgcn_bitop3_b32(a.hi, b.hi, c.hi, 0xFE),
              0);
#else
    return a | b | c;
#endif
}
// (a & b) != 0: the halves' intersections are or-ed by one v_bitop3_b32
// (the compiler would form v_and_or_b32, which does not pair)
PG_HD bool meets(BB a, BB b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a.lo, b.lo, a.hi & b.hi, 0xEA) != 0;
#else
    return (bool)(a & b);
#endif
}
PG_HD BB rev64(BB x) { return BB(__builtin_bitreverse32(x.hi), __builtin_bitreverse32(x.lo), 0); }

// Board shift by S (positive: towards h8).  On the device one 64-bit shift
// instruction; a shift by 32 is a move of halves.  (The compiler would split
// it into a 32-bit shift and an align per half, 10 % more VALU instructions
// in the fused leaf.)  Each shift is followed by `s_nop 1`: the wave waits two
// cycles before its next instruction, which mostly reads the shift's result,
// and the other waves of the SIMD issue meanwhile.  On the config-2 leaf that
// took 5.91 to 5.81 ms with the same VALU instructions; `s_nop 0` measured the
// same, `s_nop 3` 5.95 ms and `s_nop 7` 6.36 ms.
template <int S>
PG_HD constexpr u64 shift(u64 b)
{
    if constexpr (S > 0) return b << S;
    else return b >> (-S);
}
template <int S>
PG_HD BB shift(BB b)
{
    static_assert(S != 0 && S > -64 && S < 64, "shift amount");
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (S == 32) {
        return BB(0u, b.lo, 0);
    } else if constexpr (S == -32) {
        return BB(b.hi, 0u, 0);
    } else {
        u64 r;
        if constexpr (S > 0) asm("v_lshlrev_b64 %0, %2, %1\n\ts_nop 1" : "=v"(r) : "v"(b.v()), "n"(S));
        else asm("v_lshrrev_b64 %0, %2, %1\n\ts_nop 1" : "=v"(r) : "v"(b.v()), "n"(-S));
        return BB(r);
    }
#else
    return BB(shift<S>(b.v()));
#endif
}

// ---------------------------------------------------------------- piece sets
struct Sets {
    BB occ, us, them, pawns, knights, diag, orth, kings;
};

template <int C>
PG_HD Sets derive(const QuadBitBoard &q)
{
    Sets s;
    BB b1 = q.bb[1], b2 = q.bb[2], b3 = q.bb[3];
    s.occ = b1 | b2 | b3;
    BB black = BB(q.bb[0]) & s.occ;
    s.us = C == WHITE ? (s.occ ^ black) : black;
    s.them = s.occ ^ s.us;
    s.pawns = b1 & ~(b2 | b3);
    s.knights = b2 & ~(b1 | b3);
    s.diag = b1 & (b2 ^ b3);  // B = 011, Q = 101
    s.orth = b3 & ~b2;        // R = 100, Q = 101
    s.kings = b2 & b3;        // K = 110 (111 unused)
    return s;
}

PG_HD int pieceAt(const QuadBitBoard &q, int sq)
{
    return (int)(((q.bb[1] >> sq) & 1) | (((q.bb[2] >> sq) & 1) << 1) | (((q.bb[3] >> sq) & 1) << 2));
}

// ------------------------------------------------------------ leaper attacks
// Written once for plain 64-bit words (compile-time tables) and for BB.
template <class T>
PG_HD constexpr T knightAttacks(T n)
{
    if constexpr (__is_same(T, BB)) {
        BB h1 = andOr(shift<-1>(n), BB(NOT_H), shift<1>(n) & NOT_A);
        BB h2 = andOr(shift<-2>(n), BB(NOT_GH), shift<2>(n) & NOT_AB);
        return or3(shift<16>(h1), shift<-16>(h1), shift<8>(h2)) | shift<-8>(h2);
    } else {
        T l1 = shift<-1>(n) & NOT_H, l2 = shift<-2>(n) & NOT_GH;
        T r1 = shift<1>(n) & NOT_A, r2 = shift<2>(n) & NOT_AB;
        T h1 = l1 | r1, h2 = l2 | r2;
        return shift<16>(h1) | shift<-16>(h1) | shift<8>(h2) | shift<-8>(h2);
    }
}

template <class T>
PG_HD constexpr T kingAttacks(T k)
{
    T a = (shift<1>(k) & NOT_A) | (shift<-1>(k) & NOT_H);
    T r = k | a;
    return a | shift<8>(r) | shift<-8>(r);
}

template <int C, class T>  // squares attacked by pawns of colour C
PG_HD constexpr T pawnAttacks(T p)
{
    if constexpr (__is_same(T, BB)) {
        if constexpr (C == WHITE) return andOr(shift<9>(p), BB(NOT_A), shift<7>(p) & NOT_H);
        else return andOr(shift<-7>(p), BB(NOT_A), shift<-9>(p) & NOT_H);
    } else if constexpr (C == WHITE) {
        return (shift<9>(p) & NOT_A) | (shift<7>(p) & NOT_H);
    } else {
        return (shift<-7>(p) & NOT_A) | (shift<-9>(p) & NOT_H);
    }
}

// ------------------------------------------------------ per-square masks
// Line masks through `sq`, excluding sq itself.
PG_HD constexpr u64 fileMaskEx(int sq) { return (FILE_A << (sq & 7)) ^ (1ULL << sq); }
PG_HD constexpr u64 rankMaskEx(int sq) { return (0xFFULL << (sq & 56)) ^ (1ULL << sq); }
PG_HD constexpr u64 diagMaskEx(int sq)  // a1-h8 direction (NE/SW)
{
    int d = (sq >> 3) - (sq & 7);  // rank - file
    u64 m = d >= 0 ? (DIAG_A1H8 << (8 * d)) : (DIAG_A1H8 >> (-8 * d));
    return m ^ (1ULL << sq);
}
PG_HD constexpr u64 antiMaskEx(int sq)  // a8-h1 direction (NW/SE)
{
    int d = (sq >> 3) + (sq & 7) - 7;
    u64 m = d >= 0 ? (DIAG_A8H1 << (8 * d)) : (DIAG_A8H1 >> (-8 * d));
    return m ^ (1ULL << sq);
}

// One 64-byte row per square, built at compile time: the four lines through
// the square (square excluded), knight and king attacks, and pawn attacks by
// colour.  One constexpr table serves host and device: on the device it is
// read-only data of the code object (4 KB, cached), addressed relative to the
// program counter.  (A __constant__ table is registered with the runtime for
// host access and is reached through the GOT: a scalar load, and a wait on
// it, inside every loop that reads it; this form took the fused leaf 0.25 %
// lower.)  A countMoves reads the king's row (lines, king, knight and pawn
// attacks), the enemy king's king attacks and one row per own knight.
struct SqMasks {
    u64 file, rank, diag, anti, knight, king, pawn[2];
};
struct SqTable {
    SqMasks t[64];
};
constexpr SqTable buildSqTable()
{
    SqTable T{};
    for (int sq = 0; sq < 64; sq++) {
        u64 b = 1ULL << sq;
        SqMasks &m = T.t[sq];
        m.file = fileMaskEx(sq);
        m.rank = rankMaskEx(sq);
        m.diag = diagMaskEx(sq);
        m.anti = antiMaskEx(sq);
        m.knight = knightAttacks<u64>(b);
        m.king = kingAttacks<u64>(b);
        m.pawn[WHITE] = pawnAttacks<WHITE, u64>(b);
        m.pawn[BLACK] = pawnAttacks<BLACK, u64>(b);
    }
    return T;
}
static constexpr SqTable kSqHost = buildSqTable();

PG_HD const SqMasks &sqMasks(int sq)
{
    return kSqHost.t[sq];
}

// ------------------------------------------------------- Kogge-Stone fills
// Squares attacked along direction S (board shift) by every slider in `gen`,
// given the set of empty (transparent) squares.  WRAP masks file wrap-around.
template <int S, u64 WRAP>
PG_HD BB rayAttacks(BB gen, BB empty)
{
    BB pro = empty & WRAP;
    gen |= pro & shift<S>(gen);
    pro &= shift<S>(pro);
    gen |= pro & shift<2 * S>(gen);
    pro &= shift<2 * S>(pro);
    gen |= pro & shift<4 * S>(gen);
    return shift<S>(gen) & WRAP;
}

// East attacks of every slider in s, rank-wise subtraction (o ^ (o - 2s)):
// the h-file is treated as occupied so a borrow never leaves its rank, and
// h-file sliders (no east moves) are dropped.
PG_HD BB eastAttacks(BB s, BB occ)
{
    BB o = occ | FILE_H;
    BB g = s & NOT_H;
    return o ^ (o - shift<1>(g));
}
// West attacks: east attacks in the bit-reversed board
PG_HD BB westAttacks(BB s, BB occ) { return rev64(eastAttacks(rev64(s), rev64(occ))); }

PG_HD BB orthAttacksSet(BB s, BB empty)
{
    return rayAttacks<8, ~0ULL>(s, empty) | rayAttacks<-8, ~0ULL>(s, empty) | eastAttacks(s, ~empty) |
           westAttacks(s, ~empty);
}
PG_HD BB diagAttacksSet(BB s, BB empty)
{
    return rayAttacks<9, NOT_A>(s, empty) | rayAttacks<7, NOT_H>(s, empty) | rayAttacks<-7, NOT_A>(s, empty) |
           rayAttacks<-9, NOT_H>(s, empty);
}

// Hyperbola quintessence with a full bit reverse: attacks from the square with
// bit `b` along the line `m` (mask excluding the square).  Works for ranks too
// because reversing all 64 bits reverses the order inside each rank.
PG_HD BB lineAttacks(BB occ, BB m, BB b)
{
    BB o = occ & m;
    BB f = o - b;
    BB r = rev64(rev64(o) - rev64(b));
    return (f ^ r) & m;
}

PG_HD BB rookAttacks(int sq, BB occ)
{
    const SqMasks &M = sqMasks(sq);
    BB b = bit(sq);
    return lineAttacks(occ, M.file, b) | lineAttacks(occ, M.rank, b);
}
PG_HD BB bishopAttacks(int sq, BB occ)
{
    const SqMasks &M = sqMasks(sq);
    BB b = bit(sq);
    return lineAttacks(occ, M.diag, b) | lineAttacks(occ, M.anti, b);
}

// ------------------------------------------------------------ legality frame
// Everything the counter and the generator need about checks and pins.
struct Frame {
    Sets s;
    int k;           // own king square
    BB kBit;
    BB kAtt;         // king-move squares from the own king square
    BB att;          // squares attacked by the enemy (x-ray through the king on check lines)
    BB checkers;
    BB checkMask;    // landing squares that resolve the check (all if none, none if double)
    BB pinFile, pinRank, pinDiag, pinAnti;  // pin rays (king..pinner] per line
    BB pinned;       // own pieces on any pin ray
    int nCheck;
};

// Pins and checks from the "snipers": enemy sliders on the king's four
// empty-board lines.  For each sniper, the squares strictly between it and the
// king decide: none occupied -> it gives check (check ray = between + sniper,
// and the square behind the king on that line is attacked through the king);
// exactly one occupied and ours -> that piece is pinned to the line.  Snipers
// are few (usually 0-3), so this beats scanning all four lines both ways.
// `att` is left to the caller (countMoves fuses it with its own fills).
template <int C>
PG_HD Frame framePins(const QuadBitBoard &q)
{
    Frame f;
    f.s = derive<C>(q);
    const Sets &s = f.s;
    f.kBit = s.kings & s.us;
    f.k = lsb(f.kBit);
    const int k = f.k;
    const SqMasks &K = sqMasks(k);
    BB eDiag = s.diag & s.them, eOrth = s.orth & s.them;
    BB mFile = K.file, mRank = K.rank, mDiag = K.diag, mAnti = K.anti;
    BB snipers = andOr(mFile | mRank, eOrth, (mDiag | mAnti) & eDiag);
    BB below = f.kBit - BB(1), above = ~below & ~f.kBit;
    f.kAtt = K.king;
    BB chkRays, behind, pinAll;
    while (snipers) {
        BB sb = lowBit(snipers);
        snipers ^= sb;
        int sq = lsb(sb);
        // the sniper's line, selected per 32-bit half (a select of whole BB
        // values here is lowered to an indexed private array)
        bool inF = meets(mFile, sb), inR = meets(mRank, sb), inD = meets(mDiag, sb);
        BB line(inF ? mFile.lo : inR ? mRank.lo : inD ? mDiag.lo : mAnti.lo,
                inF ? mFile.hi : inR ? mRank.hi : inD ? mDiag.hi : mAnti.hi, 0);
        bool up = sq > k;
        BB btw = line & (up ? (sb - shift<1>(f.kBit)) : (f.kBit - shift<1>(sb)));
        BB blk = btw & s.occ;
        BB ray = btw | sb;
        bool check = !blk;
        bool pin = blk && !clearLow(blk) && meets(blk, s.us);
        chkRays |= check ? ray : BB();
        behind |= check ? (f.kAtt & line & (up ? below : above)) : BB();
        pinAll |= pin ? ray : BB();
    }
    // pin rays of different lines meet only at the king, so the lines split them
    f.pinFile = pinAll & mFile;
    f.pinRank = pinAll & mRank;
    f.pinDiag = pinAll & mDiag;
    f.pinAnti = pinAll & mAnti;
    BB leaperCheckers = andOr(BB(K.knight), s.knights & s.them, BB(K.pawn[C]) & s.pawns & s.them);
    f.checkers = andOr(chkRays, eOrth | eDiag, leaperCheckers);
    f.nCheck = popc(f.checkers);
    f.checkMask = f.nCheck == 0 ? BB(~0ULL) : (f.nCheck == 1 ? (chkRays | leaperCheckers) : BB());
    f.pinned = pinAll & s.us;  // every pin ray lies on one of the four lines
    // leaper and king attacks here; slider attacks are added by the caller
    BB eK = s.kings & s.them;
    BB eKAtt = eK ? BB(sqMasks(lsb(eK)).king) : BB();
    f.att = or3(pawnAttacks<1 - C>(s.pawns & s.them), knightAttacks(s.knights & s.them), eKAtt) | behind;
    return f;
}

template <int C>
PG_HD Frame frame(const QuadBitBoard &q)
{
    Frame f = framePins<C>(q);
    BB empty = ~f.s.occ;
    f.att |= orthAttacksSet(f.s.orth & f.s.them, empty) | diagAttacksSet(f.s.diag & f.s.them, empty);
    return f;
}

// One direction of Kogge-Stone fills for two slider sets over the same empty
// squares: the propagator chain (empty & wrap, shifted and-ed twice) is
// computed once and shared by the enemy-attack fill and our move fill.
template <int S, u64 WRAP>
PG_HD void rayPair(BB enemy, BB own, BB empty, BB &att, BB &ownAtt)
{
    BB p0 = empty & WRAP;
    BB p1 = p0 & shift<S>(p0);
    BB p2 = p1 & shift<2 * S>(p1);
    enemy = andOr(p0, shift<S>(enemy), enemy);
    own = andOr(p0, shift<S>(own), own);
    enemy = andOr(p1, shift<2 * S>(enemy), enemy);
    own = andOr(p1, shift<2 * S>(own), own);
    enemy = andOr(p2, shift<4 * S>(enemy), enemy);
    own = andOr(p2, shift<4 * S>(own), own);
    att = andOr(shift<S>(enemy), BB(WRAP), att);
    ownAtt = shift<S>(own) & WRAP;
}

// En-passant moves, reference semantics (MoveGeneratorBitboard.h:1732-1770,
// 1639-1650).  Returns a mask of legal EP source pawns.  Set-wise except for
// the rare discovered rank check (the king on the capture rank with an enemy
// rook or queen on it), which tests each source pawn.
template <int C>
PG_HD BB epSources(const Frame &f, u8 state)
{
    int ep = state >> 4;
    if (!ep) return BB();
    const Sets &s = f.s;
    BB target = bit((ep - 1) + (C == WHITE ? 40 : 16));
    BB capSq = C == WHITE ? shift<-8>(target) : shift<8>(target);
    BB src = ((shift<1>(capSq) & NOT_A) | (shift<-1>(capSq) & NOT_H)) & s.pawns & s.us;
    if (f.nCheck) {
        // only when the captured pawn is the single checker, unpinned pawns only
        return (f.nCheck == 1 && capSq == f.checkers) ? (src & ~f.pinned) : BB();
    }
    // a pinned pawn captures only along its own pin line: the target must lie
    // on the pin rays of the diagonal (or anti-diagonal) the pawn is pinned on
    BB onLine = (meets(target, f.pinDiag) ? f.pinDiag : BB()) | (meets(target, f.pinAnti) ? f.pinAnti : BB());
    BB legal = src & f.pinned & onLine;
    BB freeSrc = src & ~f.pinned;
    // removing both pawns must not open the king's rank to a rook/queen
    BB kRank = sqMasks(f.k).rank;
    if ((kRank & capSq) && (kRank & s.orth & s.them)) {
        for (BB b = freeSrc; b; b = clearLow(b)) {
            BB p = lowBit(b);
            BB r = lineAttacks(s.occ ^ capSq ^ p, kRank, f.kBit);
            if (!(r & s.orth & s.them)) legal |= p;
        }
    } else {
        legal |= freeSrc;
    }
    return legal;
}

// ------------------------------------------------------------- move counting
// Number of legal moves.  Set-wise: the eight slider directions are filled
// once for the enemy's attacks and our moves together; pawns are counted by
// shifted sets, knights by one table row each.
//
// EXACT = true gives |generateMoves| (used to size a BFS level or a pair
// list).  EXACT = false is the reference's countMoves, the perft(1) every
// count of the reference's -cpu path ends in: out of check it counts at most
// ONE en-passant evasion however many pawns can capture, because it ORs both
// capture directions onto the single target square before the popcount
// (countMovesOutOfCheck, MoveGeneratorBitboard.h:1639-1649), while its
// generator emits one evasion per capturing pawn (:1085-1090).  The two differ
// only when the side to move is in check from a pawn that just double-pushed
// between two of its pawns (221,239 times among the perft(10) leaves of
// Position 3, whose published count includes it).
// Slider attacks by table lookup (pg_magic.h): `sq` holds the 64 rook rows
// then the 64 bishop rows, `ratt` / `batt` the rook and bishop attack sets
// (device memory, or a block's LDS copy of the bishop sets).
struct MagicSliders {
    const MagicSquare *sq;
    const u64 *ratt, *batt;
    PG_HD BB rook(int s, BB occ) const
    {
        const MagicSquare &m = sq[s];
        return BB(ratt[m.offset + (((occ.v() & m.mask) * m.magic) >> m.shift)]);
    }
    PG_HD BB bishop(int s, BB occ) const
    {
        const MagicSquare &m = sq[64 + s];
        return BB(batt[m.offset + (((occ.v() & m.mask) * m.magic) >> m.shift)]);
    }
};
struct FillSliders {};  // the computed fills below (default)

// With `frameOut` non-null the legality frame, its enemy attacks complete,
// is handed back for generateMovesF on the same position.
template <int C, bool EXACT, class SL = FillSliders>
PG_HD u32 countMovesT(const QuadBitBoard &q, u8 state, const SL &sl = SL(), Frame *frameOut = nullptr)
{
    Frame f = framePins<C>(q);
    const Sets &s = f.s;
    BB empty = ~s.occ;
    BB targets = ~s.us & f.checkMask;
    BB freeP = ~f.pinned;

    BB att = f.att;
    u32 n = 0;
    if constexpr (__is_same(SL, MagicSliders)) {
        // one lookup per slider: the enemy's through our king (its squares
        // behind the king are attacked), ours limited to a pin line when pinned
        const BB occNoK = s.occ ^ f.kBit;
        for (BB b = s.diag & s.them; b; b = clearLow(b)) att |= sl.bishop(lsb(b), occNoK);
        for (BB b = s.orth & s.them; b; b = clearLow(b)) att |= sl.rook(lsb(b), occNoK);
        for (BB b = s.diag & s.us & ~(f.pinFile | f.pinRank); b; b = clearLow(b)) {
            const int from = lsb(b);
            const BB fb = bit(from);
            BB t = sl.bishop(from, s.occ) & targets;
            t &= (fb & f.pinDiag) ? f.pinDiag : ((fb & f.pinAnti) ? f.pinAnti : BB(~0ULL));
            n += popc(t);
        }
        for (BB b = s.orth & s.us & ~(f.pinDiag | f.pinAnti); b; b = clearLow(b)) {
            const int from = lsb(b);
            const BB fb = bit(from);
            BB t = sl.rook(from, s.occ) & targets;
            t &= (fb & f.pinFile) ? f.pinFile : ((fb & f.pinRank) ? f.pinRank : BB(~0ULL));
            n += popc(t);
        }
    } else {
        // sliders: a direction's fill reaches each target from exactly one
        // slider, so per-direction popcounts count moves exactly; a pinned
        // slider moves only along its pin line
        BB D = s.diag & s.us, O = s.orth & s.us;
        BB eD = s.diag & s.them, eO = s.orth & s.them;
        BB dA1H8 = D & (freeP | f.pinDiag), dA8H1 = D & (freeP | f.pinAnti);
        BB oFile = O & (freeP | f.pinFile), oRank = O & (freeP | f.pinRank);
        BB r;
        rayPair<8, ~0ULL>(eO, oFile, empty, att, r);
        n += popc(r & targets);
        rayPair<-8, ~0ULL>(eO, oFile, empty, att, r);
        n += popc(r & targets);
        att = or3(att, eastAttacks(eO, s.occ), westAttacks(eO, s.occ));
        n += popc(eastAttacks(oRank, s.occ) & targets) + popc(westAttacks(oRank, s.occ) & targets);
        rayPair<9, NOT_A>(eD, dA1H8, empty, att, r);
        n += popc(r & targets);
        rayPair<-9, NOT_H>(eD, dA1H8, empty, att, r);
        n += popc(r & targets);
        rayPair<7, NOT_H>(eD, dA8H1, empty, att, r);
        n += popc(r & targets);
        rayPair<-7, NOT_A>(eD, dA8H1, empty, att, r);
        n += popc(r & targets);
    }

    n += popc(f.kAtt & ~s.us & ~att);
    if (frameOut) {
        *frameOut = f;
        frameOut->att = att;
    }

    // knights: pinned knights never move
    for (BB N = s.knights & s.us & freeP; N; N = clearLow(N)) n += popc(BB(sqMasks(lsb(N)).knight) & targets);

    // pawns
    BB P = s.pawns & s.us;
    const u64 promo = C == WHITE ? RANK_8 : RANK_1;
    BB push = P & (freeP | f.pinFile);
    BB s1 = (C == WHITE ? shift<8>(push) : shift<-8>(push)) & empty;
    BB s2 = (C == WHITE ? shift<8>(s1 & RANK_3) : shift<-8>(s1 & RANK_6)) & empty & f.checkMask;
    s1 &= f.checkMask;
    // capture towards the h-file: white NE (+9, a1-h8 line), black SE (-7, a8-h1 line)
    BB pe = P & (freeP | (C == WHITE ? f.pinDiag : f.pinAnti));
    BB pw = P & (freeP | (C == WHITE ? f.pinAnti : f.pinDiag));
    BB ce = (C == WHITE ? shift<9>(pe) : shift<-7>(pe)) & NOT_A & s.them & f.checkMask;
    BB cw = (C == WHITE ? shift<7>(pw) : shift<-9>(pw)) & NOT_H & s.them & f.checkMask;
    // pushes land on empty squares and captures on enemy ones, and a square
    // is reached by a single push or by a double push, never both (the double
    // push needs the single-push square empty): s1, s2 and ce are disjoint,
    // so one popcount covers them; cw can share squares with ce.  A promotion
    // counts four moves: three more for each promoting target.
    n += popc(or3(s1, s2, ce)) + popc(cw) + 3 * (popc((s1 | ce) & promo) + popc(cw & promo));

    // castling: rights, empty path, unattacked path, not in check
    u32 cr = C == WHITE ? (state & 3) : ((state >> 2) & 3);
    const u64 ksEmpty = C == WHITE ? 0x60ULL : 0x6000000000000000ULL;
    const u64 qsEmpty = C == WHITE ? 0x0EULL : 0x0E00000000000000ULL;
    const u64 qsSafe = C == WHITE ? 0x0CULL : 0x0C00000000000000ULL;
    bool noCheck = f.nCheck == 0;
    n += (noCheck && (cr & 1) && !((s.occ | att) & ksEmpty)) ? 1 : 0;
    n += (noCheck && (cr & 2) && !(s.occ & qsEmpty) && !(att & qsSafe)) ? 1 : 0;

    if (state >> 4) {
        BB eps = epSources<C>(f, state);
        n += (!EXACT && f.nCheck) ? (eps ? 1u : 0u) : (u32)popc(eps);
    }
    return n;
}

// the reference's perft(1) count (leaves of every perft)
template <int C>
PG_HD u32 countMoves(const QuadBitBoard &q, u8 state)
{
    return countMovesT<C, false>(q, state);
}
// the number of moves generateMoves emits (sizes of BFS levels and pair lists)
template <int C>
PG_HD u32 countChildren(const QuadBitBoard &q, u8 state)
{
    return countMovesT<C, true>(q, state);
}

// countChildren that also hands back the position's frame, so that
// generateMovesF can emit the same moves without recomputing it
template <int C, class SL>
PG_HD u32 countChildrenFrame(const QuadBitBoard &q, u8 state, const SL &sl, Frame &f)
{
    return countMovesT<C, true, SL>(q, state, sl, &f);
}

// the same two counts with a chosen slider-attack method (FillSliders or MagicSliders)
template <int C, class SL>
PG_HD u32 countMovesSL(const QuadBitBoard &q, u8 state, const SL &sl)
{
    return countMovesT<C, false, SL>(q, state, sl);
}
template <int C, class SL>
PG_HD u32 countChildrenSL(const QuadBitBoard &q, u8 state, const SL &sl)
{
    return countMovesT<C, true, SL>(q, state, sl);
}

PG_HD u32 countMoves(const QuadBitBoard &q, u8 state, int color)
{
    return color == WHITE ? countMoves<WHITE>(q, state) : countMoves<BLACK>(q, state);
}
PG_HD u32 countChildren(const QuadBitBoard &q, u8 state, int color)
{
    return color == WHITE ? countChildren<WHITE>(q, state) : countChildren<BLACK>(q, state);
}

// --------------------------------------------------------- move generation
// Target-set emission.  A set's squares are visited in ascending order, the
// low 32-bit half first: a move then costs one 32-bit find-first and one
// clear-lowest (the 64-bit forms take four and four instructions), and the
// move is one shift-or (pieces: base | to << 6, base = from | flags << 12)
// or one multiply-add (pawns: from = to - d, so move = 65 to - d + flags).
PG_HD u32 ctz32(u32 x) { return (u32)__builtin_ctz(x); }

template <typename Emit>
PG_HD void emitTo(BB t, u32 base, u32 &n, Emit &emit)
{
    for (u32 x = t.lo; x; x &= x - 1) emit(n++, (u16)(base | (ctz32(x) << 6)));
    for (u32 x = t.hi; x; x &= x - 1) emit(n++, (u16)((base | 0x800u) | (ctz32(x) << 6)));
}
// pawn moves by d squares (from = to - d), flags already in `add` = flags << 12 - d
template <typename Emit>
PG_HD void emitPawnTo(BB t, u32 add, u32 &n, Emit &emit)
{
    for (u32 x = t.lo; x; x &= x - 1) emit(n++, (u16)(ctz32(x) * 65u + add));
    for (u32 x = t.hi; x; x &= x - 1) emit(n++, (u16)((ctz32(x) + 32u) * 65u + add));
}
// promotions by d squares, in the order N, B, Q, R (MoveGeneratorBitboard.h:920-923)
template <typename Emit>
PG_HD void emitPromoTo(BB t, u32 minusD, u32 &n, Emit &emit)
{
    for (BB r = t; r; r = clearLow(r)) {
        const u32 m = (u32)lsb(r) * 65u + minusD;
        emit(n, (u16)(m | ((MF_PROMO + 0) << 12)));
        emit(n + 1, (u16)(m | ((MF_PROMO + 1) << 12)));
        emit(n + 2, (u16)(m | ((MF_PROMO + 3) << 12)));
        emit(n + 3, (u16)(m | ((MF_PROMO + 2) << 12)));
        n += 4;
    }
}
// pawn moves of t by d squares, promotions included, in ascending order of
// the target: white's promotions (rank 8) follow its other moves, black's
// (rank 1) come first
template <int C, typename Emit>
PG_HD void emitPawnMoves(BB t, int d, u32 &n, Emit &emit)
{
    const u64 promo = C == WHITE ? RANK_8 : RANK_1;
    const BB plain = t & ~BB(promo), pr = t & promo;
    if constexpr (C == BLACK)
        if (pr) emitPromoTo(pr, (u32)(-d), n, emit);
    emitPawnTo(plain, (u32)(-d), n, emit);
    if constexpr (C == WHITE)
        if (pr) emitPromoTo(pr, (u32)(-d), n, emit);
}

// Emits every legal move into `out` (if non-null) and returns the count.
// Order: king, castling, knights, diagonal sliders, orthogonal sliders,
// pawn pushes, double pushes, captures, en passant; each piece's or pawn
// set's targets in ascending order.  A double push carries MF_DPUSH only when
// an enemy pawn stands beside its target (otherwise no EP state can follow);
// captures carry no capture flag.
template <int C, typename Emit>
PG_HD u32 generateMovesF(const Frame &f, u8 state, Emit &&emit)
{
    const Sets &s = f.s;
    BB empty = ~s.occ;
    BB targets = ~s.us & f.checkMask;
    u32 n = 0;
    emitTo(f.kAtt & ~s.us & ~f.att, (u32)f.k, n, emit);
    if (f.nCheck == 0) {
        u32 cr = C == WHITE ? (state & 3) : ((state >> 2) & 3);
        const u64 ksEmpty = C == WHITE ? 0x60ULL : 0x6000000000000000ULL;
        const u64 qsEmpty = C == WHITE ? 0x0EULL : 0x0E00000000000000ULL;
        const u64 qsSafe = C == WHITE ? 0x0CULL : 0x0C00000000000000ULL;
        if ((cr & 1) && !((s.occ | f.att) & ksEmpty)) {
            emit(n, (u16)(f.k | ((f.k + 2) << 6) | (MF_KCASTLE << 12)));
            n++;
        }
        if ((cr & 2) && !(s.occ & qsEmpty) && !(f.att & qsSafe)) {
            emit(n, (u16)(f.k | ((f.k - 2) << 6) | (MF_QCASTLE << 12)));
            n++;
        }
    }
    BB freeP = ~f.pinned;
    for (BB b = s.knights & s.us & freeP; b; b = clearLow(b)) {
        int from = lsb(b);
        emitTo(BB(sqMasks(from).knight) & targets, (u32)from, n, emit);
    }
    for (BB b = s.diag & s.us & ~(f.pinFile | f.pinRank); b; b = clearLow(b)) {
        int from = lsb(b);
        BB fb = bit(from);
        BB t = bishopAttacks(from, s.occ) & targets;
        t &= (fb & f.pinDiag) ? f.pinDiag : ((fb & f.pinAnti) ? f.pinAnti : BB(~0ULL));
        emitTo(t, (u32)from, n, emit);
    }
    for (BB b = s.orth & s.us & ~(f.pinDiag | f.pinAnti); b; b = clearLow(b)) {
        int from = lsb(b);
        BB fb = bit(from);
        BB t = rookAttacks(from, s.occ) & targets;
        t &= (fb & f.pinFile) ? f.pinFile : ((fb & f.pinRank) ? f.pinRank : BB(~0ULL));
        emitTo(t, (u32)from, n, emit);
    }
    BB P = s.pawns & s.us;
    const int fwd = C == WHITE ? 8 : -8;
    BB push = P & (freeP | f.pinFile);
    BB s1 = (C == WHITE ? shift<8>(push) : shift<-8>(push)) & empty;
    BB s2 = (C == WHITE ? shift<8>(s1 & RANK_3) : shift<-8>(s1 & RANK_6)) & empty & f.checkMask;
    s1 &= f.checkMask;
    emitPawnMoves<C>(s1, fwd, n, emit);
    // double pushes: MF_DPUSH where an enemy pawn stands beside the target
    {
        BB ePawns = s.pawns & s.them;
        BB dp = s2 & ((shift<1>(ePawns) & NOT_A) | (shift<-1>(ePawns) & NOT_H));
        const u32 add = (u32)(-2 * fwd);
        for (u32 x = s2.lo; x; x &= x - 1) {
            const u32 to = ctz32(x);
            emit(n++, (u16)(to * 65u + add + (((dp.lo >> to) & 1u) << 12)));
        }
        for (u32 x = s2.hi; x; x &= x - 1) {
            const u32 to = ctz32(x);
            emit(n++, (u16)((to + 32u) * 65u + add + (((dp.hi >> to) & 1u) << 12)));
        }
    }
    BB pe = P & (freeP | (C == WHITE ? f.pinDiag : f.pinAnti));
    BB pw = P & (freeP | (C == WHITE ? f.pinAnti : f.pinDiag));
    const int de = C == WHITE ? 9 : -7, dw = C == WHITE ? 7 : -9;
    BB ce = (C == WHITE ? shift<9>(pe) : shift<-7>(pe)) & NOT_A & s.them & f.checkMask;
    BB cw = (C == WHITE ? shift<7>(pw) : shift<-9>(pw)) & NOT_H & s.them & f.checkMask;
    emitPawnMoves<C>(ce, de, n, emit);
    emitPawnMoves<C>(cw, dw, n, emit);
    if (state >> 4) {
        int to = ((state >> 4) - 1) + (C == WHITE ? 40 : 16);
        for (BB b = epSources<C>(f, state); b; b = clearLow(b)) {
            emit(n, (u16)(lsb(b) | (to << 6) | (MF_EP << 12)));
            n++;
        }
    }
    return n;
}

template <int C, typename Emit>
PG_HD u32 generateMoves(const QuadBitBoard &q, u8 state, Emit &&emit)
{
    return generateMovesF<C>(frame<C>(q), state, emit);
}

struct MoveListSink {
    u16 *out;
    PG_HD void operator()(u32 i, u16 m) const { out[i] = m; }
};

template <int C>
PG_HD u32 generateMoves(const QuadBitBoard &q, u8 state, u16 *out)
{
    return generateMoves<C>(q, state, MoveListSink{out});
}

PG_HD u32 generateMoves(const QuadBitBoard &q, u8 state, int color, u16 *out)
{
    return color == WHITE ? generateMoves<WHITE>(q, state, out) : generateMoves<BLACK>(q, state, out);
}

// ----------------------------------------------------------------- make move
// Castle rights clear when a move touches a1/e1/h1/a8/e8/h8 (from or to);
// the EP file is set only by MF_DPUSH.  Equivalent to the reference's device
// flavour (MoveGeneratorBitboard.h:2021-2040) on consistent positions.
template <int C>
PG_HD void makeMove(QuadBitBoard &q, u8 &state, u16 m)
{
    u32 from = m & 63, to = (m >> 6) & 63, flags = m >> 12;
    u64 fb64 = 1ULL << from, tb64 = 1ULL << to;
    BB fb = fb64, tb = tb64;
    u32 piece = (u32)(((q.bb[1] >> from) & 1) | (((q.bb[2] >> from) & 1) << 1) | (((q.bb[3] >> from) & 1) << 2));
    piece = (flags & MF_PROMO) ? (flags & 3) + KNIGHT : piece;
    // en passant: the captured pawn sits behind the target
    BB epCap = (flags == MF_EP) ? (C == WHITE ? shift<-8>(tb) : shift<8>(tb)) : BB();
    // castling: the rook jumps to the square the king passes
    const u32 base = C == WHITE ? 0 : 56;
    BB rookFrom = flags == MF_KCASTLE ? BB(1ULL << (base + 7)) : (flags == MF_QCASTLE ? BB(1ULL << base) : BB());
    BB rookTo = flags == MF_KCASTLE ? BB(1ULL << (base + 5)) : (flags == MF_QCASTLE ? BB(1ULL << (base + 3)) : BB());
    BB clr = ~(fb | tb | epCap | rookFrom);
    BB b0 = q.bb[0], b1 = q.bb[1], b2 = q.bb[2], b3 = q.bb[3];
    q.bb[0] = andOr(b0, clr, C == BLACK ? (tb | rookTo) : BB()).v();
    q.bb[1] = andOr(b1, clr, (piece & 1) ? tb : BB()).v();
    q.bb[2] = andOr(b2, clr, (piece & 2) ? tb : BB()).v();
    q.bb[3] = andOr(b3, clr, ((piece & 4) ? tb : BB()) | rookTo).v();

    // castling rights lost when a move touches a1/e1/h1/a8/e8/h8: each right's
    // two squares shifted onto its bit (K = 1: e1, h1; Q = 2: a1, e1;
    // k = 4: e8, h8; q = 8: a8, e8), squares 56-63 being bits 24-31 of hi
    const BB touch = fb | tb;
    const u32 lo = touch.lo, hi = touch.hi;
    const u32 c = (((lo >> 4) | (lo >> 7)) & 1u) | (((lo << 1) | (lo >> 3)) & 2u) | (((hi >> 26) | (hi >> 29)) & 4u) |
                  (((hi >> 21) | (hi >> 25)) & 8u);
    u32 st = state & 0x0F & ~c;
    st |= (flags == MF_DPUSH) ? (((from & 7) + 1) << 4) : 0u;
    state = (u8)st;
}

PG_HD void makeMove(QuadBitBoard &q, u8 &state, int color, u16 m)
{
    if (color == WHITE) makeMove<WHITE>(q, state, m);
    else makeMove<BLACK>(q, state, m);
}

// perft(2) of a position: sum over its legal moves of the child's move count
// (the body of the fused two-ply leaf).
template <int C>
PG_HD u32 perft2(const QuadBitBoard &q, u8 state)
{
    u32 total = 0;
    auto child = [&](u32, u16 m) {
        QuadBitBoard c = q;
        u8 st = state;
        makeMove<C>(c, st, m);
        total += countMoves<1 - C>(c, st);
    };
    generateMoves<C>(q, state, child);
    return total;
}

}  // namespace pg
