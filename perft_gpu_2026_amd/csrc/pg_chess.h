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
//     occluded fills for whole piece sets, and hyperbola quintessence with a
//     full 64-bit bit reverse (v_bfrev_b32 x2) for single-square lines;
//     no table reads, so the hot loop is pure VALU.
//   * pins and checks come from the few enemy sliders on the king's lines;
//     pinned pieces are handled by per-line source masks instead of per-piece
//     loops (a piece pinned on a line may move only along it).
//   * legal-move counting is set-wise: knight, slider and pawn counts are
//     popcounts of per-direction target sets; only en passant (rare) branches.
// The semantics follow the reference's countMoves/makeMove
// (MoveGeneratorBitboard.h:1563-2067), including its en-passant rules
// (rank discovered check by rook/queen only; EP evasion only when the captured
// pawn is the checker) and castling by rights + empty, unattacked path.
#pragma once
#include <stdint.h>

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
PG_HD u64 rev64(u64 x) { return __builtin_bitreverse64(x); }
PG_HD u64 sel(bool c, u64 a, u64 b) { return c ? a : b; }

template <int S>
PG_HD u64 shift(u64 b)
{
    if constexpr (S > 0) return b << S;
    else return b >> (-S);
}

// ---------------------------------------------------------------- piece sets
struct Sets {
    u64 occ, us, them, pawns, knights, diag, orth, kings;
};

template <int C>
PG_HD Sets derive(const QuadBitBoard &q)
{
    Sets s;
    u64 b1 = q.bb[1], b2 = q.bb[2], b3 = q.bb[3];
    s.occ = b1 | b2 | b3;
    u64 black = q.bb[0] & s.occ;
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
PG_HD u64 knightAttacks(u64 n)
{
    u64 l1 = (n >> 1) & NOT_H, l2 = (n >> 2) & NOT_GH;
    u64 r1 = (n << 1) & NOT_A, r2 = (n << 2) & NOT_AB;
    u64 h1 = l1 | r1, h2 = l2 | r2;
    return (h1 << 16) | (h1 >> 16) | (h2 << 8) | (h2 >> 8);
}

PG_HD u64 kingAttacks(u64 k)
{
    u64 a = ((k << 1) & NOT_A) | ((k >> 1) & NOT_H);
    u64 r = k | a;
    return a | (r << 8) | (r >> 8);
}

template <int C>  // squares attacked by pawns of colour C
PG_HD u64 pawnAttacks(u64 p)
{
    if constexpr (C == WHITE) return ((p << 9) & NOT_A) | ((p << 7) & NOT_H);
    else return ((p >> 7) & NOT_A) | ((p >> 9) & NOT_H);
}

// ------------------------------------------------------- Kogge-Stone fills
// Squares attacked along direction S (board shift) by every slider in `gen`,
// given the set of empty (transparent) squares.  WRAP masks file wrap-around.
template <int S, u64 WRAP>
PG_HD u64 rayAttacks(u64 gen, u64 empty)
{
    u64 pro = empty & WRAP;
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
PG_HD u64 eastAttacks(u64 s, u64 occ)
{
    u64 o = occ | FILE_H;
    u64 g = s & NOT_H;
    return o ^ (o - (g << 1));
}
// West attacks: east attacks in the bit-reversed board
PG_HD u64 westAttacks(u64 s, u64 occ) { return rev64(eastAttacks(rev64(s), rev64(occ))); }

PG_HD u64 orthAttacksSet(u64 s, u64 empty)
{
    return rayAttacks<8, ~0ULL>(s, empty) | rayAttacks<-8, ~0ULL>(s, empty) | eastAttacks(s, ~empty) |
           westAttacks(s, ~empty);
}
PG_HD u64 diagAttacksSet(u64 s, u64 empty)
{
    return rayAttacks<9, NOT_A>(s, empty) | rayAttacks<7, NOT_H>(s, empty) |
           rayAttacks<-7, NOT_A>(s, empty) | rayAttacks<-9, NOT_H>(s, empty);
}

// ------------------------------------------------- single-square line attacks
// Line masks through `sq`, excluding sq itself.
PG_HD u64 fileMaskEx(int sq) { return (FILE_A << (sq & 7)) ^ (1ULL << sq); }
PG_HD u64 rankMaskEx(int sq) { return (0xFFULL << (sq & 56)) ^ (1ULL << sq); }
PG_HD u64 diagMaskEx(int sq)  // a1-h8 direction (NE/SW)
{
    int d = (sq >> 3) - (sq & 7);  // rank - file
    u64 m = d >= 0 ? (DIAG_A1H8 << (8 * d)) : (DIAG_A1H8 >> (-8 * d));
    return m ^ (1ULL << sq);
}
PG_HD u64 antiMaskEx(int sq)  // a8-h1 direction (NW/SE)
{
    int d = (sq >> 3) + (sq & 7) - 7;
    u64 m = d >= 0 ? (DIAG_A8H1 << (8 * d)) : (DIAG_A8H1 >> (-8 * d));
    return m ^ (1ULL << sq);
}

// Hyperbola quintessence with a full bit reverse: attacks from the square with
// bit `b` along the line `m` (mask excluding the square).  Works for ranks too
// because reversing all 64 bits reverses the order inside each rank.
PG_HD u64 lineAttacks(u64 occ, u64 m, u64 b)
{
    u64 o = occ & m;
    u64 f = o - b;
    u64 r = rev64(rev64(o) - rev64(b));
    return (f ^ r) & m;
}

PG_HD u64 rookAttacks(int sq, u64 occ)
{
    u64 b = 1ULL << sq;
    return lineAttacks(occ, fileMaskEx(sq), b) | lineAttacks(occ, rankMaskEx(sq), b);
}
PG_HD u64 bishopAttacks(int sq, u64 occ)
{
    u64 b = 1ULL << sq;
    return lineAttacks(occ, diagMaskEx(sq), b) | lineAttacks(occ, antiMaskEx(sq), b);
}

// ------------------------------------------------------------ legality frame
// Everything the counter and the generator need about checks and pins.
struct Frame {
    Sets s;
    int k;           // own king square
    u64 kBit;
    u64 att;         // squares attacked by the enemy (x-ray through the king on check lines)
    u64 checkers;
    u64 checkMask;   // landing squares that resolve the check (all if none, none if double)
    u64 pinFile, pinRank, pinDiag, pinAnti;  // pin rays (king..pinner] per line
    u64 pinned;      // own pieces on any pin ray
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
    u64 eDiag = s.diag & s.them, eOrth = s.orth & s.them;
    u64 mFile = fileMaskEx(k), mRank = rankMaskEx(k), mDiag = diagMaskEx(k), mAnti = antiMaskEx(k);
    u64 snipers = ((mFile | mRank) & eOrth) | ((mDiag | mAnti) & eDiag);
    u64 below = f.kBit - 1, above = ~below & ~f.kBit;
    u64 kAtt = kingAttacks(f.kBit);
    u64 chkRays = 0, behind = 0, pinAll = 0;
    while (snipers) {
        u64 sb = snipers & (0 - snipers);
        snipers ^= sb;
        int sq = lsb(sb);
        u64 line = (mFile & sb) ? mFile : ((mRank & sb) ? mRank : ((mDiag & sb) ? mDiag : mAnti));
        bool up = sq > k;
        u64 btw = line & (up ? (sb - (f.kBit << 1)) : (f.kBit - (sb << 1)));
        u64 blk = btw & s.occ;
        u64 ray = btw | sb;
        bool check = blk == 0;
        bool pin = blk && !(blk & (blk - 1)) && (blk & s.us);
        chkRays |= check ? ray : 0;
        behind |= check ? (kAtt & line & (up ? below : above)) : 0;
        pinAll |= pin ? ray : 0;
    }
    // pin rays of different lines meet only at the king, so the lines split them
    f.pinFile = pinAll & mFile;
    f.pinRank = pinAll & mRank;
    f.pinDiag = pinAll & mDiag;
    f.pinAnti = pinAll & mAnti;
    u64 leaperCheckers = (knightAttacks(f.kBit) & s.knights & s.them) | (pawnAttacks<C>(f.kBit) & s.pawns & s.them);
    f.checkers = leaperCheckers | (chkRays & (eOrth | eDiag));
    f.nCheck = popc(f.checkers);
    f.checkMask = f.nCheck == 0 ? ~0ULL : (f.nCheck == 1 ? (chkRays | leaperCheckers) : 0ULL);
    f.pinned = (f.pinFile | f.pinRank | f.pinDiag | f.pinAnti) & s.us;
    // leaper and king attacks here; slider attacks are added by the caller
    f.att = pawnAttacks<1 - C>(s.pawns & s.them) | knightAttacks(s.knights & s.them) | kingAttacks(s.kings & s.them) |
            behind;
    return f;
}

template <int C>
PG_HD Frame frame(const QuadBitBoard &q)
{
    Frame f = framePins<C>(q);
    u64 empty = ~f.s.occ;
    f.att |= orthAttacksSet(f.s.orth & f.s.them, empty) | diagAttacksSet(f.s.diag & f.s.them, empty);
    return f;
}

// One direction of Kogge-Stone fills for two slider sets over the same empty
// squares: the propagator chain (empty & wrap, shifted and-ed twice) is
// computed once and shared by the enemy-attack fill and our move fill.
template <int S, u64 WRAP>
PG_HD void rayPair(u64 enemy, u64 own, u64 empty, u64 &att, u64 &ownAtt)
{
    u64 p0 = empty & WRAP;
    u64 p1 = p0 & shift<S>(p0);
    u64 p2 = p1 & shift<2 * S>(p1);
    enemy |= p0 & shift<S>(enemy);
    own |= p0 & shift<S>(own);
    enemy |= p1 & shift<2 * S>(enemy);
    own |= p1 & shift<2 * S>(own);
    enemy |= p2 & shift<4 * S>(enemy);
    own |= p2 & shift<4 * S>(own);
    att |= shift<S>(enemy) & WRAP;
    ownAtt = shift<S>(own) & WRAP;
}

// En-passant moves, reference semantics (MoveGeneratorBitboard.h:1732-1770,
// 1639-1650).  Returns a mask of legal EP source pawns.
template <int C>
PG_HD u64 epSources(const Frame &f, u8 state)
{
    int ep = state >> 4;
    if (!ep) return 0;
    const Sets &s = f.s;
    u64 target = 1ULL << ((ep - 1) + (C == WHITE ? 40 : 16));
    u64 capSq = C == WHITE ? (target >> 8) : (target << 8);
    u64 myPawns = s.pawns & s.us;
    u64 src = (((capSq << 1) & NOT_A) | ((capSq >> 1) & NOT_H)) & myPawns;
    if (!src) return 0;
    if (f.nCheck) {
        // only when the captured pawn is the single checker, unpinned pawns only
        return (f.nCheck == 1 && capSq == f.checkers) ? (src & ~f.pinned) : 0;
    }
    u64 legal = 0;
    while (src) {
        u64 p = src & (0 - src);
        src ^= p;
        if (p & f.pinned) {
            // a pinned pawn may capture along its pin line only
            u64 line = (p & f.pinDiag) ? f.pinDiag : ((p & f.pinAnti) ? f.pinAnti : 0);
            if (target & line) legal |= p;
        } else {
            // removing both pawns must not open the king's rank to a rook/queen
            u64 occ2 = s.occ ^ capSq ^ p;
            u64 r = lineAttacks(occ2, rankMaskEx(f.k), f.kBit);
            if (!(r & s.orth & s.them)) legal |= p;
        }
    }
    return legal;
}

// ------------------------------------------------------------- move counting
// Number of legal moves (perft(1)).  Set-wise, no per-piece loops: the eight
// slider directions are filled once for the enemy's attacks and our moves
// together; knights and pawns are counted by shifted sets.
template <int C>
PG_HD u32 countMoves(const QuadBitBoard &q, u8 state)
{
    Frame f = framePins<C>(q);
    const Sets &s = f.s;
    u64 empty = ~s.occ;
    u64 targets = ~s.us & f.checkMask;
    u64 freeP = ~f.pinned;

    // sliders: a direction's fill reaches each target from exactly one slider,
    // so per-direction popcounts count moves exactly; a pinned slider moves
    // only along its pin line
    u64 D = s.diag & s.us, O = s.orth & s.us;
    u64 eD = s.diag & s.them, eO = s.orth & s.them;
    u64 dA1H8 = D & (freeP | f.pinDiag), dA8H1 = D & (freeP | f.pinAnti);
    u64 oFile = O & (freeP | f.pinFile), oRank = O & (freeP | f.pinRank);
    u64 att = f.att, r;
    u32 n = 0;
    rayPair<8, ~0ULL>(eO, oFile, empty, att, r);
    n += popc(r & targets);
    rayPair<-8, ~0ULL>(eO, oFile, empty, att, r);
    n += popc(r & targets);
    att |= eastAttacks(eO, s.occ) | westAttacks(eO, s.occ);
    n += popc(eastAttacks(oRank, s.occ) & targets) + popc(westAttacks(oRank, s.occ) & targets);
    rayPair<9, NOT_A>(eD, dA1H8, empty, att, r);
    n += popc(r & targets);
    rayPair<-9, NOT_H>(eD, dA1H8, empty, att, r);
    n += popc(r & targets);
    rayPair<7, NOT_H>(eD, dA8H1, empty, att, r);
    n += popc(r & targets);
    rayPair<-7, NOT_A>(eD, dA8H1, empty, att, r);
    n += popc(r & targets);

    n += popc(kingAttacks(f.kBit) & ~s.us & ~att);

    // knights: pinned knights never move
    u64 N = s.knights & s.us & freeP;
    {
        u64 l1 = (N >> 1) & NOT_H, l2 = (N >> 2) & NOT_GH;
        u64 r1 = (N << 1) & NOT_A, r2 = (N << 2) & NOT_AB;
        n += popc((l1 << 16) & targets) + popc((l1 >> 16) & targets) + popc((r1 << 16) & targets) +
             popc((r1 >> 16) & targets) + popc((l2 << 8) & targets) + popc((l2 >> 8) & targets) +
             popc((r2 << 8) & targets) + popc((r2 >> 8) & targets);
    }

    // pawns
    u64 P = s.pawns & s.us;
    u64 promo = C == WHITE ? RANK_8 : RANK_1;
    u64 push = P & (freeP | f.pinFile);
    u64 s1 = (C == WHITE ? push << 8 : push >> 8) & empty;
    u64 s2 = (C == WHITE ? (s1 & RANK_3) << 8 : (s1 & RANK_6) >> 8) & empty & f.checkMask;
    s1 &= f.checkMask;
    // capture towards the h-file: white NE (+9, a1-h8 line), black SE (-7, a8-h1 line)
    u64 pe = P & (freeP | (C == WHITE ? f.pinDiag : f.pinAnti));
    u64 pw = P & (freeP | (C == WHITE ? f.pinAnti : f.pinDiag));
    u64 ce = (C == WHITE ? (pe << 9) : (pe >> 7)) & NOT_A & s.them & f.checkMask;
    u64 cw = (C == WHITE ? (pw << 7) : (pw >> 9)) & NOT_H & s.them & f.checkMask;
    n += popc(s1 & ~promo) + popc(s2) + popc(ce & ~promo) + popc(cw & ~promo) +
         4 * (popc(s1 & promo) + popc(ce & promo) + popc(cw & promo));

    // castling: rights, empty path, unattacked path, not in check
    u32 cr = C == WHITE ? (state & 3) : ((state >> 2) & 3);
    const u64 ksEmpty = C == WHITE ? 0x60ULL : 0x6000000000000000ULL;
    const u64 qsEmpty = C == WHITE ? 0x0EULL : 0x0E00000000000000ULL;
    const u64 qsSafe = C == WHITE ? 0x0CULL : 0x0C00000000000000ULL;
    bool noCheck = f.nCheck == 0;
    n += (noCheck && (cr & 1) && !((s.occ | att) & ksEmpty)) ? 1 : 0;
    n += (noCheck && (cr & 2) && !(s.occ & qsEmpty) && !(att & qsSafe)) ? 1 : 0;

    if (state >> 4) n += popc(epSources<C>(f, state));
    return n;
}

PG_HD u32 countMoves(const QuadBitBoard &q, u8 state, int color)
{
    return color == WHITE ? countMoves<WHITE>(q, state) : countMoves<BLACK>(q, state);
}

// --------------------------------------------------------- move generation
// Emits every legal move into `out` (if non-null) and returns the count.
// Order: king, castling, knights, diagonal sliders, orthogonal sliders,
// pawn pushes, double pushes, captures, en passant.  A double push carries
// MF_DPUSH only when an enemy pawn stands beside its target (otherwise no EP
// state can follow); captures carry no capture flag.
template <int C, typename Emit>
PG_HD u32 generateMoves(const QuadBitBoard &q, u8 state, Emit &&emit)
{
    Frame f = frame<C>(q);
    const Sets &s = f.s;
    u64 empty = ~s.occ;
    u64 targets = ~s.us & f.checkMask;
    u32 n = 0;
    for (u64 t = kingAttacks(f.kBit) & ~s.us & ~f.att; t; t &= t - 1) {
        emit(n, (u16)(f.k | (lsb(t) << 6)));
        n++;
    }
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
    u64 freeP = ~f.pinned;
    for (u64 b = s.knights & s.us & freeP; b; b &= b - 1) {
        int from = lsb(b);
        for (u64 t = knightAttacks(1ULL << from) & targets; t; t &= t - 1) {
            emit(n, (u16)(from | (lsb(t) << 6)));
            n++;
        }
    }
    for (u64 b = s.diag & s.us & ~(f.pinFile | f.pinRank); b; b &= b - 1) {
        int from = lsb(b);
        u64 fb = 1ULL << from;
        u64 t = bishopAttacks(from, s.occ) & targets;
        t &= (fb & f.pinDiag) ? f.pinDiag : ((fb & f.pinAnti) ? f.pinAnti : ~0ULL);
        for (; t; t &= t - 1) {
            emit(n, (u16)(from | (lsb(t) << 6)));
            n++;
        }
    }
    for (u64 b = s.orth & s.us & ~(f.pinDiag | f.pinAnti); b; b &= b - 1) {
        int from = lsb(b);
        u64 fb = 1ULL << from;
        u64 t = rookAttacks(from, s.occ) & targets;
        t &= (fb & f.pinFile) ? f.pinFile : ((fb & f.pinRank) ? f.pinRank : ~0ULL);
        for (; t; t &= t - 1) {
            emit(n, (u16)(from | (lsb(t) << 6)));
            n++;
        }
    }
    u64 P = s.pawns & s.us;
    u64 promo = C == WHITE ? RANK_8 : RANK_1;
    const int fwd = C == WHITE ? 8 : -8;
    u64 push = P & (freeP | f.pinFile);
    u64 s1 = (C == WHITE ? push << 8 : push >> 8) & empty;
    u64 s2 = (C == WHITE ? (s1 & RANK_3) << 8 : (s1 & RANK_6) >> 8) & empty & f.checkMask;
    s1 &= f.checkMask;
    u64 ePawns = s.pawns & s.them;
    auto emitPawn = [&](int from, int to) {
        if ((1ULL << to) & promo) {
            // promotion order N, B, Q, R (MoveGeneratorBitboard.h:920-923)
            emit(n, (u16)(from | (to << 6) | ((MF_PROMO + 0) << 12)));
            emit(n + 1, (u16)(from | (to << 6) | ((MF_PROMO + 1) << 12)));
            emit(n + 2, (u16)(from | (to << 6) | ((MF_PROMO + 3) << 12)));
            emit(n + 3, (u16)(from | (to << 6) | ((MF_PROMO + 2) << 12)));
            n += 4;
        } else {
            emit(n, (u16)(from | (to << 6)));
            n++;
        }
    };
    for (u64 t = s1; t; t &= t - 1) {
        int to = lsb(t);
        emitPawn(to - fwd, to);
    }
    for (u64 t = s2; t; t &= t - 1) {
        int to = lsb(t);
        u64 tb = 1ULL << to;
        bool epPossible = ((((tb << 1) & NOT_A) | ((tb >> 1) & NOT_H)) & ePawns) != 0;
        emit(n, (u16)((to - 2 * fwd) | (to << 6) | ((epPossible ? MF_DPUSH : MF_QUIET) << 12)));
        n++;
    }
    u64 pe = P & (freeP | (C == WHITE ? f.pinDiag : f.pinAnti));
    u64 pw = P & (freeP | (C == WHITE ? f.pinAnti : f.pinDiag));
    const int de = C == WHITE ? 9 : -7, dw = C == WHITE ? 7 : -9;
    u64 ce = (C == WHITE ? (pe << 9) : (pe >> 7)) & NOT_A & s.them & f.checkMask;
    u64 cw = (C == WHITE ? (pw << 7) : (pw >> 9)) & NOT_H & s.them & f.checkMask;
    for (u64 t = ce; t; t &= t - 1) {
        int to = lsb(t);
        emitPawn(to - de, to);
    }
    for (u64 t = cw; t; t &= t - 1) {
        int to = lsb(t);
        emitPawn(to - dw, to);
    }
    if (state >> 4) {
        int to = ((state >> 4) - 1) + (C == WHITE ? 40 : 16);
        for (u64 b = epSources<C>(f, state); b; b &= b - 1) {
            emit(n, (u16)(lsb(b) | (to << 6) | (MF_EP << 12)));
            n++;
        }
    }
    return n;
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
    u64 fb = 1ULL << from, tb = 1ULL << to;
    u32 piece = (u32)(((q.bb[1] >> from) & 1) | (((q.bb[2] >> from) & 1) << 1) | (((q.bb[3] >> from) & 1) << 2));
    piece = (flags & MF_PROMO) ? (flags & 3) + KNIGHT : piece;
    u64 clr = ~(fb | tb);
    // en passant: the captured pawn sits behind the target
    u64 epCap = (flags == MF_EP) ? (C == WHITE ? tb >> 8 : tb << 8) : 0;
    // castling: the rook jumps to the square the king passes
    const u32 base = C == WHITE ? 0 : 56;
    u64 rookFrom = flags == MF_KCASTLE ? (1ULL << (base + 7)) : (flags == MF_QCASTLE ? (1ULL << base) : 0);
    u64 rookTo = flags == MF_KCASTLE ? (1ULL << (base + 5)) : (flags == MF_QCASTLE ? (1ULL << (base + 3)) : 0);
    clr &= ~(epCap | rookFrom);
    q.bb[0] = (q.bb[0] & clr) | (C == BLACK ? (tb | rookTo) : 0);
    q.bb[1] = (q.bb[1] & clr) | ((piece & 1) ? tb : 0);
    q.bb[2] = (q.bb[2] & clr) | ((piece & 2) ? tb : 0);
    q.bb[3] = (q.bb[3] & clr) | ((piece & 4) ? tb : 0) | rookTo;

    u64 touch = fb | tb;
    u32 c = 0;
    c |= (touch & (1ULL << 0)) ? 2u : 0u;
    c |= (touch & (1ULL << 7)) ? 1u : 0u;
    c |= (touch & (1ULL << 4)) ? 3u : 0u;
    c |= (touch & (1ULL << 56)) ? 8u : 0u;
    c |= (touch & (1ULL << 63)) ? 4u : 0u;
    c |= (touch & (1ULL << 60)) ? 12u : 0u;
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
