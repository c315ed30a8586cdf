/*
 * oracle.c -- CPU oracle (test infrastructure only; never part of the product).
 *
 * Plain-C restatement of the reference's -cpu perft path.  Each function cites
 * the reference file:line whose semantics it follows.  Slider attacks use the
 * classical ray-scan method (ray tables + first-blocker bit scan) rather than
 * the reference's fancy-magic tables (MoveGeneratorBitboard.h:589-648): both
 * compute the same attack sets, and an independent method keeps the oracle
 * from sharing a bug with any magic-table implementation.
 *
 * Pins: see oracle.h.  Every function below is exercised by tests/test_oracle.py
 * against published perft counts and the brute-force generator.
 */
#include "oracle.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t  u8;

enum { WHITE = 0, BLACK = 1 };
enum { PAWN = 1, KNIGHT = 2, BISHOP = 3, ROOK = 4, QUEEN = 5, KING = 6 };
/* move flags, chess.h:120-144 */
enum { F_QUIET = 0, F_DPUSH = 1, F_KCASTLE = 2, F_QCASTLE = 3, F_CAPTURE = 4, F_EP = 5,
       F_PROMO = 8 /* 8 N, 9 B, 10 R, 11 Q; +4 = capture */ };

#define BIT(s) (1ULL << (s))
#define FILE_A 0x0101010101010101ULL
#define FILE_H 0x8080808080808080ULL
#define RANK_1 0x00000000000000FFULL
#define RANK_3 0x0000000000FF0000ULL
#define RANK_6 0x0000FF0000000000ULL
#define RANK_8 0xFF00000000000000ULL

/* direction index: 0 N, 1 E, 2 NE, 3 NW (positive), 4 S, 5 W, 6 SW, 7 SE (negative) */
static const int DR[8] = { 1, 0, 1, 1, -1, 0, -1, -1 };
static const int DF[8] = { 0, 1, 1, -1, 0, -1, -1, 1 };

static u64 RAY[8][64];
static u64 KNIGHT_ATT[64], KING_ATT[64];
static u64 BETWEEN[64][64], LINE[64][64];
static int g_init = 0;

static inline int popc(u64 x) { return __builtin_popcountll(x); }
static inline int lsb(u64 x) { return __builtin_ctzll(x); }
static inline int msb(u64 x) { return 63 - __builtin_clzll(x); }

/* tables by ray walks (the reference builds the same sets with Kogge-Stone
 * fills at init, perft_kernels.cu:998-1044) */
void oracle_init(void)
{
    if (g_init) return;
    for (int s = 0; s < 64; s++) {
        int r = s >> 3, f = s & 7;
        for (int d = 0; d < 8; d++) {
            u64 ray = 0;
            for (int rr = r + DR[d], ff = f + DF[d]; rr >= 0 && rr < 8 && ff >= 0 && ff < 8; rr += DR[d], ff += DF[d])
                ray |= BIT(rr * 8 + ff);
            RAY[d][s] = ray;
        }
        static const int KN[8][2] = {{1,2},{2,1},{2,-1},{1,-2},{-1,-2},{-2,-1},{-2,1},{-1,2}};
        u64 kn = 0, kg = 0;
        for (int i = 0; i < 8; i++) {
            int rr = r + KN[i][0], ff = f + KN[i][1];
            if (rr >= 0 && rr < 8 && ff >= 0 && ff < 8) kn |= BIT(rr * 8 + ff);
        }
        for (int dr = -1; dr <= 1; dr++)
            for (int df = -1; df <= 1; df++) {
                int rr = r + dr, ff = f + df;
                if ((dr || df) && rr >= 0 && rr < 8 && ff >= 0 && ff < 8) kg |= BIT(rr * 8 + ff);
            }
        KNIGHT_ATT[s] = kn;
        KING_ATT[s] = kg;
    }
    for (int a = 0; a < 64; a++)
        for (int b = 0; b < 64; b++) {
            BETWEEN[a][b] = 0;
            LINE[a][b] = 0;
            for (int d = 0; d < 8; d++) {
                if (RAY[d][a] & BIT(b)) {
                    int od = (d + 4) & 7;
                    BETWEEN[a][b] = RAY[d][a] & RAY[od][b];
                    LINE[a][b] = RAY[d][a] | RAY[od][a] | BIT(a);
                }
            }
        }
    g_init = 1;
}

/* classical ray attack: the ray up to and including the first blocker */
static inline u64 ray_att(int d, int s, u64 occ)
{
    u64 a = RAY[d][s], bl = a & occ;
    if (bl) a ^= RAY[d][d < 4 ? lsb(bl) : msb(bl)];
    return a;
}
static inline u64 rook_att(int s, u64 occ)
{
    return ray_att(0, s, occ) | ray_att(1, s, occ) | ray_att(4, s, occ) | ray_att(5, s, occ);
}
static inline u64 bishop_att(int s, u64 occ)
{
    return ray_att(2, s, occ) | ray_att(3, s, occ) | ray_att(6, s, occ) | ray_att(7, s, occ);
}

/* -------- position helpers (chess.h:66-92) -------- */
typedef struct {
    u64 occ, white, black, pawns, knights, bishopQueens, rookQueens, kings;
} Pieces;

static inline void derive(const u64 bb[4], Pieces *p)
{
    u64 b1 = bb[1], b2 = bb[2], b3 = bb[3];
    p->occ = b1 | b2 | b3;
    p->black = bb[0] & p->occ;
    p->white = p->occ & ~bb[0];
    p->pawns = b1 & ~b2 & ~b3;
    p->knights = ~b1 & b2 & ~b3;
    p->bishopQueens = (b1 & b2 & ~b3) | (b1 & ~b2 & b3);
    p->rookQueens = (~b1 & ~b2 & b3) | (b1 & ~b2 & b3);
    p->kings = ~b1 & b2 & b3;
}

static inline int piece_at(const u64 bb[4], int s)
{
    return (int)(((bb[1] >> s) & 1) | (((bb[2] >> s) & 1) << 1) | (((bb[3] >> s) & 1) << 2));
}

static inline u64 north(u64 b) { return b << 8; }
static inline u64 south(u64 b) { return b >> 8; }
static inline u64 east(u64 b) { return (b << 1) & ~FILE_A; }
static inline u64 west(u64 b) { return (b >> 1) & ~FILE_H; }

/* squares attacked by `color` (king of the other side made transparent for
 * sliders) -- MoveGeneratorBitboard.h:835-875 */
static u64 attacked_by(const Pieces *p, int color, u64 transparentKing)
{
    u64 mine = color == WHITE ? p->white : p->black;
    u64 occ = p->occ & ~transparentKing;
    u64 pw = p->pawns & mine, a = 0;
    if (color == WHITE) a |= east(north(pw)) | west(north(pw));
    else a |= east(south(pw)) | west(south(pw));
    for (u64 b = p->knights & mine; b; b &= b - 1) a |= KNIGHT_ATT[lsb(b)];
    for (u64 b = p->bishopQueens & mine; b; b &= b - 1) a |= bishop_att(lsb(b), occ);
    for (u64 b = p->rookQueens & mine; b; b &= b - 1) a |= rook_att(lsb(b), occ);
    for (u64 b = p->kings & mine; b; b &= b - 1) a |= KING_ATT[lsb(b)];
    return a;
}

/* pieces singly blocking an enemy slider's line to the king --
 * MoveGeneratorBitboard.h:803-831 */
static u64 pinned_pieces(const Pieces *p, int k, u64 enemy)
{
    u64 att = (bishop_att(k, 0) & p->bishopQueens & enemy) | (rook_att(k, 0) & p->rookQueens & enemy);
    u64 pinned = 0;
    for (; att; att &= att - 1) {
        u64 btw = BETWEEN[lsb(att)][k] & p->occ;
        if (btw && !(btw & (btw - 1))) pinned |= btw;
    }
    return pinned;
}

/* -------- move list sink: count-only or write -------- */
typedef struct { u16 *out; u32 n; } Sink;
static inline void emit(Sink *s, int from, int to, int flags)
{
    if (s->out) s->out[s->n] = (u16)(from | (to << 6) | (flags << 12));
    s->n++;
}
static inline void emit_targets(Sink *s, int from, u64 t, u64 enemy)
{
    for (; t; t &= t - 1) {
        int to = lsb(t);
        emit(s, from, to, (BIT(to) & enemy) ? F_CAPTURE : F_QUIET);
    }
}
static inline void emit_pawn(Sink *s, int from, int to, int isCap)
{
    if (BIT(to) & (RANK_1 | RANK_8)) {
        int c = isCap ? F_CAPTURE : 0;
        /* promotion order N, B, Q, R as MoveGeneratorBitboard.h:920-923 */
        emit(s, from, to, F_PROMO + 0 + c);
        emit(s, from, to, F_PROMO + 1 + c);
        emit(s, from, to, F_PROMO + 3 + c);
        emit(s, from, to, F_PROMO + 2 + c);
    } else {
        emit(s, from, to, isCap ? F_CAPTURE : F_QUIET);
    }
}

/*
 * Legal move generation/counting.  One routine serves both: with s->out NULL
 * it only counts.  Semantics follow countMoves (MoveGeneratorBitboard.h:
 * 1700-1943) and countMovesOutOfCheck (:1563-1694), including the reference's
 * en-passant rules (pinned EP along the pin line :1746-1753, rank discovered
 * check by rook/queen only :1756-1767, EP evasion only when the captured pawn
 * is the checker :1639-1650) and castling on rights + empty path + unattacked
 * path without a rook-presence test (:1896-1913).
 *
 * One difference between counting and generating is the reference's own: out
 * of check, countMovesOutOfCheck ORs the two en-passant capture directions
 * onto the single target square and popcounts that (:1646-1648), so it counts
 * at most ONE en-passant evasion when two pawns can take, while
 * generateMovesOutOfCheck emits one per pawn (:1085-1090).  perft(1) is that
 * count (launcher.cu:765-766, 809-810), so the count mode here does the same.
 */
static u32 legal_moves(const u64 bb[4], u8 state, int color, Sink *s)
{
    Pieces p;
    derive(bb, &p);
    u64 mine = color == WHITE ? p.white : p.black;
    u64 enemy = p.occ & ~mine;
    u64 myKing = p.kings & mine;
    int k = lsb(myKing);
    u64 empty = ~p.occ;
    u64 pinned = pinned_pieces(&p, k, enemy);
    u64 threatened = attacked_by(&p, !color, myKing);
    u64 myPawns = p.pawns & mine;
    u64 dpRank = color == WHITE ? RANK_3 : RANK_6;

    /* king moves (both paths) */
    emit_targets(s, k, KING_ATT[k] & ~(threatened | mine), enemy);

    u64 epTarget = 0, epCaptured = 0;
    if (state >> 4) {
        int f = (state >> 4) - 1;
        epTarget = BIT(f + (color == WHITE ? 40 : 16));
        epCaptured = color == WHITE ? south(epTarget) : north(epTarget);
    }

    if (threatened & myKing) {
        /* out of check, MoveGeneratorBitboard.h:1563-1694 */
        u64 attackers = 0;
        attackers |= (color == WHITE ? (east(north(myKing)) | west(north(myKing)))
                                     : (east(south(myKing)) | west(south(myKing)))) & p.pawns & enemy;
        attackers |= KNIGHT_ATT[k] & p.knights & enemy;
        attackers |= bishop_att(k, p.occ) & p.bishopQueens & enemy;
        attackers |= rook_att(k, p.occ) & p.rookQueens & enemy;
        if (attackers & (attackers - 1)) return s->n;   /* double check: king only */
        u64 safe = attackers | BETWEEN[k][lsb(attackers)];
        u64 movers = mine & ~pinned;
        int epCounted = 0;
        u64 pw = myPawns & ~pinned;
        for (u64 b = pw; b; b &= b - 1) {
            int from = lsb(b);
            u64 pb = BIT(from);
            u64 one = (color == WHITE ? north(pb) : south(pb)) & empty;
            if (one & safe) emit_pawn(s, from, lsb(one), 0);
            u64 two = (color == WHITE ? north(one & dpRank) : south(one & dpRank)) & empty & safe;
            if (two) emit(s, from, lsb(two), F_DPUSH);
            u64 caps = (color == WHITE ? (east(north(pb)) | west(north(pb))) : (east(south(pb)) | west(south(pb))))
                       & enemy & safe;
            for (; caps; caps &= caps - 1) emit_pawn(s, from, lsb(caps), 1);
            if (epTarget && epCaptured == attackers) {
                u64 epc = (color == WHITE ? (east(north(pb)) | west(north(pb))) : (east(south(pb)) | west(south(pb))))
                          & epTarget;
                if (epc && (s->out || !epCounted)) emit(s, from, lsb(epc), F_EP);
                if (epc) epCounted = 1;
            }
        }
        for (u64 b = p.knights & movers; b; b &= b - 1) emit_targets(s, lsb(b), KNIGHT_ATT[lsb(b)] & safe, enemy);
        for (u64 b = p.bishopQueens & movers; b; b &= b - 1)
            emit_targets(s, lsb(b), bishop_att(lsb(b), p.occ) & safe, enemy);
        for (u64 b = p.rookQueens & movers; b; b &= b - 1)
            emit_targets(s, lsb(b), rook_att(lsb(b), p.occ) & safe, enemy);
        return s->n;
    }

    /* en passant, MoveGeneratorBitboard.h:1732-1770 */
    if (epTarget) {
        for (u64 src = (east(epCaptured) | west(epCaptured)) & myPawns; src; src &= src - 1) {
            int from = lsb(src);
            if (BIT(from) & pinned) {
                if (epTarget & LINE[from][k]) emit(s, from, lsb(epTarget), F_EP);
            } else {
                u64 occ2 = p.occ ^ epCaptured ^ BIT(from);
                u64 rankMask = RANK_1 << (k & 56);
                if (!(rook_att(k, occ2) & p.rookQueens & enemy & rankMask)) emit(s, from, lsb(epTarget), F_EP);
            }
        }
    }

    /* pawns */
    for (u64 b = myPawns; b; b &= b - 1) {
        int from = lsb(b);
        u64 pb = BIT(from);
        u64 allowed = (pb & pinned) ? LINE[from][k] : ~0ULL;
        u64 one = (color == WHITE ? north(pb) : south(pb)) & empty;
        if (one & allowed) {
            emit_pawn(s, from, lsb(one), 0);
            u64 two = (color == WHITE ? north(one & dpRank) : south(one & dpRank)) & empty;
            if (two) emit(s, from, lsb(two), F_DPUSH);
        }
        u64 caps = (color == WHITE ? (east(north(pb)) | west(north(pb))) : (east(south(pb)) | west(south(pb))))
                   & enemy & allowed;
        for (; caps; caps &= caps - 1) emit_pawn(s, from, lsb(caps), 1);
    }

    /* castling, MoveGeneratorBitboard.h:1896-1913 */
    int castle = color == WHITE ? (state & 3) : ((state >> 2) & 3);
    if (color == WHITE) {
        if ((castle & 1) && !(0x60ULL & p.occ) && !(0x60ULL & threatened)) emit(s, 4, 6, F_KCASTLE);
        if ((castle & 2) && !(0x0EULL & p.occ) && !(0x0CULL & threatened)) emit(s, 4, 2, F_QCASTLE);
    } else {
        if ((castle & 1) && !(0x6000000000000000ULL & p.occ) && !(0x6000000000000000ULL & threatened))
            emit(s, 60, 62, F_KCASTLE);
        if ((castle & 2) && !(0x0E00000000000000ULL & p.occ) && !(0x0C00000000000000ULL & threatened))
            emit(s, 60, 58, F_QCASTLE);
    }

    /* knights: pinned knights never move */
    for (u64 b = p.knights & mine & ~pinned; b; b &= b - 1)
        emit_targets(s, lsb(b), KNIGHT_ATT[lsb(b)] & ~mine, enemy);
    /* sliders: pinned ones stay on the king line */
    for (u64 b = p.bishopQueens & mine; b; b &= b - 1) {
        int from = lsb(b);
        u64 t = bishop_att(from, p.occ) & ~mine;
        if (BIT(from) & pinned) t &= LINE[from][k];
        emit_targets(s, from, t, enemy);
    }
    for (u64 b = p.rookQueens & mine; b; b &= b - 1) {
        int from = lsb(b);
        u64 t = rook_att(from, p.occ) & ~mine;
        if (BIT(from) & pinned) t &= LINE[from][k];
        emit_targets(s, from, t, enemy);
    }
    return s->n;
}

uint32_t oracle_count_moves(const uint64_t bb[4], uint8_t state, uint8_t color)
{
    Sink s = { 0, 0 };
    return legal_moves(bb, state, color, &s);
}

uint32_t oracle_gen_moves(const uint64_t bb[4], uint8_t state, uint8_t color, uint16_t *moves)
{
    Sink s = { moves, 0 };
    return legal_moves(bb, state, color, &s);
}

/* makeMove, host flavour: MoveGeneratorBitboard.h:1945-2067 (castle rights by
 * king/rook moves and captures on rook corners, EP set on every flagged double
 * push) */
void oracle_make_move(uint64_t bb[4], uint8_t *state, uint8_t color, uint16_t move)
{
    int from = move & 63, to = (move >> 6) & 63, flags = move >> 12;
    u64 src = BIT(from), dst = BIT(to);
    int piece = piece_at(bb, from);
    if (flags & F_PROMO) piece = (flags & 3) + KNIGHT;
    u64 clr = ~(src | dst);
    for (int i = 0; i < 4; i++) bb[i] &= clr;
    if (color == BLACK) bb[0] |= dst;
    if (piece & 1) bb[1] |= dst;
    if (piece & 2) bb[2] |= dst;
    if (piece & 4) bb[3] |= dst;
    if (flags == F_EP) {
        u64 cap = color == WHITE ? south(dst) : north(dst);
        for (int i = 0; i < 4; i++) bb[i] &= ~cap;
    }
    if (flags == F_KCASTLE || flags == F_QCASTLE) {
        int base = color == WHITE ? 0 : 56;
        int rf = flags == F_KCASTLE ? base + 7 : base;
        int rt = flags == F_KCASTLE ? base + 5 : base + 3;
        for (int i = 0; i < 4; i++) bb[i] &= ~BIT(rf);
        bb[3] |= BIT(rt);
        if (color == BLACK) bb[0] |= BIT(rt);
    }
    u8 st = *state & 0x0F;
    /* a capture on the enemy rook corner removes that right */
    if (color == WHITE) {
        if (dst & BIT(63)) st &= ~(1 << 2);
        if (dst & BIT(56)) st &= ~(2 << 2);
    } else {
        if (dst & BIT(7)) st &= ~1;
        if (dst & BIT(0)) st &= ~2;
    }
    if (piece == KING) st &= color == WHITE ? ~3 : ~(3 << 2);
    if (piece == ROOK) {
        if (color == WHITE) {
            if (from == 7) st &= ~1;
            if (from == 0) st &= ~2;
        } else {
            if (from == 63) st &= ~(1 << 2);
            if (from == 56) st &= ~(2 << 2);
        }
    }
    if (flags == F_DPUSH) st |= (u8)(((from & 7) + 1) << 4);
    *state = st;
}

/* -------- perft, launcher.cu:762-816 -------- */
static uint64_t perft_rec(u64 bb[4], u8 state, int color, u32 depth)
{
    if (depth == 1) return oracle_count_moves(bb, state, color);
    u16 moves[256];
    u32 n = oracle_gen_moves(bb, state, color, moves);
    uint64_t total = 0;
    for (u32 i = 0; i < n; i++) {
        u64 c[4] = { bb[0], bb[1], bb[2], bb[3] };
        u8 st = state;
        oracle_make_move(c, &st, color, moves[i]);
        total += perft_rec(c, st, !color, depth - 1);
    }
    return total;
}

uint64_t oracle_perft(const uint64_t bb[4], uint8_t state, uint8_t color, uint32_t depth)
{
    oracle_init();
    if (depth == 0) return 1;
    u64 c[4] = { bb[0], bb[1], bb[2], bb[3] };
    return perft_rec(c, state, color, depth);
}

/* -------- threaded drivers (root-move / position parallel) -------- */
typedef struct {
    int n;
    const u64 *bbs; const u8 *states; const u8 *colors; const u32 *depths;
    u64 *counts;
    int next;
    pthread_mutex_t mu;
} Work;

static void *worker(void *arg)
{
    Work *w = (Work *)arg;
    for (;;) {
        pthread_mutex_lock(&w->mu);
        int i = w->next++;
        pthread_mutex_unlock(&w->mu);
        if (i >= w->n) break;
        w->counts[i] = oracle_perft(&w->bbs[4 * i], w->states[i], w->colors[i], w->depths[i]);
    }
    return NULL;
}

void oracle_perft_batch(int n, const uint64_t *bbs, const uint8_t *states, const uint8_t *colors,
                        const uint32_t *depths, uint64_t *counts, int threads)
{
    oracle_init();
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    Work w;
    memset(&w, 0, sizeof w);
    w.n = n; w.bbs = bbs; w.states = states; w.colors = colors; w.depths = depths; w.counts = counts;
    pthread_mutex_init(&w.mu, NULL);
    pthread_t th[256];
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &w);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&w.mu);
}

uint64_t oracle_perft_mt(const uint64_t bb[4], uint8_t state, uint8_t color, uint32_t depth, int threads)
{
    oracle_init();
    if (depth <= 2) return oracle_perft(bb, state, color, depth);
    /* split at depth 2 for balance */
    u16 m1[256], m2[256];
    u32 n1 = oracle_gen_moves(bb, state, color, m1);
    int cap = 256 * 256;
    u64 *bbs = (u64 *)malloc(sizeof(u64) * 4 * cap);
    u8 *sts = (u8 *)malloc(cap), *cols = (u8 *)malloc(cap);
    u32 *deps = (u32 *)malloc(sizeof(u32) * cap);
    u64 *cnt = (u64 *)malloc(sizeof(u64) * cap);
    int n = 0;
    for (u32 i = 0; i < n1; i++) {
        u64 c[4] = { bb[0], bb[1], bb[2], bb[3] };
        u8 st = state;
        oracle_make_move(c, &st, color, m1[i]);
        u32 n2 = oracle_gen_moves(c, st, !color, m2);
        for (u32 j = 0; j < n2; j++) {
            u64 *d = &bbs[4 * n];
            memcpy(d, c, 32);
            u8 st2 = st;
            oracle_make_move(d, &st2, !color, m2[j]);
            sts[n] = st2; cols[n] = color; deps[n] = depth - 2;
            n++;
        }
    }
    oracle_perft_batch(n, bbs, sts, cols, deps, cnt, threads);
    u64 total = 0;
    for (int i = 0; i < n; i++) total += cnt[i];
    free(bbs); free(sts); free(cols); free(deps); free(cnt);
    return total;
}

/* -------- brute force: pseudo-legal + king-safety filter (independent) -------- */
static int king_attacked(const u64 bb[4], int color)
{
    Pieces p;
    derive(bb, &p);
    u64 mine = color == WHITE ? p.white : p.black;
    u64 kb = p.kings & mine;
    if (!kb) return 1;
    int k = lsb(kb);
    u64 enemy = p.occ & ~mine;
    if (KNIGHT_ATT[k] & p.knights & enemy) return 1;
    if (KING_ATT[k] & p.kings & enemy) return 1;
    if (bishop_att(k, p.occ) & p.bishopQueens & enemy) return 1;
    if (rook_att(k, p.occ) & p.rookQueens & enemy) return 1;
    u64 pa = color == WHITE ? (east(north(kb)) | west(north(kb))) : (east(south(kb)) | west(south(kb)));
    return (pa & p.pawns & enemy) != 0;
}

static int square_attacked(const u64 bb[4], int sq, int byColor)
{
    Pieces p;
    derive(bb, &p);
    u64 them = byColor == WHITE ? p.white : p.black;
    if (KNIGHT_ATT[sq] & p.knights & them) return 1;
    if (KING_ATT[sq] & p.kings & them) return 1;
    if (bishop_att(sq, p.occ) & p.bishopQueens & them) return 1;
    if (rook_att(sq, p.occ) & p.rookQueens & them) return 1;
    u64 sb = BIT(sq);
    /* pawns of byColor attacking sq sit diagonally behind it from byColor's view */
    u64 pa = byColor == WHITE ? (east(south(sb)) | west(south(sb))) : (east(north(sb)) | west(north(sb)));
    return (pa & p.pawns & them) != 0;
}

uint32_t oracle_brute_gen_moves(const uint64_t bb[4], uint8_t state, uint8_t color, uint16_t *moves)
{
    oracle_init();
    Pieces p;
    derive(bb, &p);
    u64 mine = color == WHITE ? p.white : p.black;
    u64 enemy = p.occ & ~mine;
    u16 pseudo[512];
    Sink s = { pseudo, 0 };
    for (u64 b = mine; b; b &= b - 1) {
        int from = lsb(b);
        int pc = piece_at(bb, from);
        u64 pb = BIT(from), t = 0;
        switch (pc) {
        case PAWN: {
            u64 one = (color == WHITE ? north(pb) : south(pb)) & ~p.occ;
            if (one) {
                emit_pawn(&s, from, lsb(one), 0);
                int startRank = color == WHITE ? 1 : 6;
                if ((from >> 3) == startRank) {
                    u64 two = (color == WHITE ? north(one) : south(one)) & ~p.occ;
                    if (two) emit(&s, from, lsb(two), F_DPUSH);
                }
            }
            u64 diag = color == WHITE ? (east(north(pb)) | west(north(pb))) : (east(south(pb)) | west(south(pb)));
            for (u64 c = diag & enemy; c; c &= c - 1) emit_pawn(&s, from, lsb(c), 1);
            if (state >> 4) {
                int f = (state >> 4) - 1;
                u64 ept = BIT(f + (color == WHITE ? 40 : 16));
                u64 capSq = color == WHITE ? south(ept) : north(ept);
                /* EP needs an enemy pawn on the passed square and an empty target */
                if ((diag & ept) && (capSq & p.pawns & enemy) && !(ept & p.occ)) emit(&s, from, lsb(ept), F_EP);
            }
            continue;
        }
        case KNIGHT: t = KNIGHT_ATT[from]; break;
        case BISHOP: t = bishop_att(from, p.occ); break;
        case ROOK: t = rook_att(from, p.occ); break;
        case QUEEN: t = bishop_att(from, p.occ) | rook_att(from, p.occ); break;
        case KING: t = KING_ATT[from]; break;
        }
        emit_targets(&s, from, t & ~mine, enemy);
    }
    /* castling: rights, empty path, king start and transit squares not attacked */
    int castle = color == WHITE ? (state & 3) : ((state >> 2) & 3);
    int base = color == WHITE ? 0 : 56;
    if ((castle & 1) && !(p.occ & (BIT(base + 5) | BIT(base + 6))) &&
        !square_attacked(bb, base + 4, !color) && !square_attacked(bb, base + 5, !color) &&
        !square_attacked(bb, base + 6, !color))
        emit(&s, base + 4, base + 6, F_KCASTLE);
    if ((castle & 2) && !(p.occ & (BIT(base + 1) | BIT(base + 2) | BIT(base + 3))) &&
        !square_attacked(bb, base + 4, !color) && !square_attacked(bb, base + 3, !color) &&
        !square_attacked(bb, base + 2, !color))
        emit(&s, base + 4, base + 2, F_QCASTLE);

    u32 n = 0;
    for (u32 i = 0; i < s.n; i++) {
        u64 c[4] = { bb[0], bb[1], bb[2], bb[3] };
        u8 st = state;
        oracle_make_move(c, &st, color, pseudo[i]);
        if (!king_attacked(c, color)) {
            if (moves) moves[n] = pseudo[i];
            n++;
        }
    }
    return n;
}

/* The reference's perft(1) convention on top of the brute generator: every
 * legal move counts, except that a side in check counts its en-passant
 * captures as one move however many pawns can take (countMovesOutOfCheck,
 * MoveGeneratorBitboard.h:1639-1649). */
uint32_t oracle_brute_count_moves(const uint64_t bb[4], uint8_t state, uint8_t color)
{
    u16 moves[256];
    u32 n = oracle_brute_gen_moves(bb, state, color, moves), nEp = 0;
    for (u32 i = 0; i < n; i++) nEp += (moves[i] >> 12) == F_EP;
    if (nEp > 1 && king_attacked(bb, color)) n -= nEp - 1;
    return n;
}

uint64_t oracle_perft_brute(const uint64_t bb[4], uint8_t state, uint8_t color, uint32_t depth)
{
    oracle_init();
    if (depth == 0) return 1;
    if (depth == 1) return oracle_brute_count_moves(bb, state, color);
    u16 moves[256];
    u32 n = oracle_brute_gen_moves(bb, state, color, moves);
    uint64_t total = 0;
    for (u32 i = 0; i < n; i++) {
        u64 c[4] = { bb[0], bb[1], bb[2], bb[3] };
        u8 st = state;
        oracle_make_move(c, &st, color, moves[i]);
        total += oracle_perft_brute(c, st, !color, depth - 1);
    }
    return total;
}

/* -------- FEN, util.cpp:7-95 (placement, side, castling, EP; clocks ignored) -------- */
int oracle_parse_fen(const char *fen, uint64_t bb[4], uint8_t *state, uint8_t *color)
{
    oracle_init();
    memset(bb, 0, 32);
    *state = 0;
    int i = 0, row = 0, col = 0;
    for (; fen[i]; i++) {
        char c = fen[i];
        if (c == '/' || c == '\\') { row++; col = 0; }
        else if (c >= '1' && c <= '8') col += c - '0';
        else {
            const char *pcs = " pnbrqk";
            int pc = 0, black = 0;
            for (int j = 1; j <= 6; j++) {
                if (c == pcs[j]) { pc = j; black = 1; }
                if (c == pcs[j] - 32) { pc = j; }
            }
            if (pc) {
                int sq = (7 - row) * 8 + col;
                if (sq < 0 || sq > 63) return -1;
                if (black) bb[0] |= BIT(sq);
                if (pc & 1) bb[1] |= BIT(sq);
                if (pc & 2) bb[2] |= BIT(sq);
                if (pc & 4) bb[3] |= BIT(sq);
                col++;
            }
        }
        if (row >= 7 && col == 8) break;
    }
    if (!fen[i]) return -1;
    i++;
    while (fen[i] == ' ') i++;
    *color = (fen[i] == 'b' || fen[i] == 'B') ? BLACK : WHITE;
    if (fen[i]) i++;
    while (fen[i] == ' ') i++;
    while (fen[i] && fen[i] != ' ') {
        switch (fen[i]) {
        case 'K': *state |= 1; break;
        case 'Q': *state |= 2; break;
        case 'k': *state |= 4; break;
        case 'q': *state |= 8; break;
        }
        i++;
    }
    while (fen[i] == ' ') i++;
    if (fen[i] >= 'a' && fen[i] <= 'h') *state |= (u8)((fen[i] - 'a' + 1) << 4);
    return 0;
}

/* FEN writer.  The EP square is written only when a pawn of the side to move
 * stands next to the double-pushed pawn (the only case in which it can matter),
 * so FENs do not depend on which EP-flag convention produced the state. */
int oracle_write_fen(const uint64_t bb[4], uint8_t state, uint8_t color, char *out, int outLen)
{
    char buf[128];
    int n = 0;
    const char *pcs = ".PNBRQK";
    for (int r = 7; r >= 0; r--) {
        int emptyRun = 0;
        for (int f = 0; f < 8; f++) {
            int sq = r * 8 + f;
            int pc = piece_at(bb, sq);
            if (!pc) { emptyRun++; continue; }
            if (emptyRun) { buf[n++] = (char)('0' + emptyRun); emptyRun = 0; }
            char ch = pcs[pc];
            if (bb[0] & BIT(sq)) ch = (char)(ch + 32);
            buf[n++] = ch;
        }
        if (emptyRun) buf[n++] = (char)('0' + emptyRun);
        if (r) buf[n++] = '/';
    }
    buf[n++] = ' ';
    buf[n++] = color == WHITE ? 'w' : 'b';
    buf[n++] = ' ';
    if (!(state & 15)) buf[n++] = '-';
    if (state & 1) buf[n++] = 'K';
    if (state & 2) buf[n++] = 'Q';
    if (state & 4) buf[n++] = 'k';
    if (state & 8) buf[n++] = 'q';
    buf[n++] = ' ';
    int ep = state >> 4;
    int wrote = 0;
    if (ep) {
        int f = ep - 1;
        int capSq = color == WHITE ? 32 + f : 24 + f;
        u64 adj = (east(BIT(capSq)) | west(BIT(capSq)));
        Pieces p;
        derive(bb, &p);
        u64 mine = color == WHITE ? p.white : p.black;
        if (adj & p.pawns & mine) {
            buf[n++] = (char)('a' + f);
            buf[n++] = color == WHITE ? '6' : '3';
            wrote = 1;
        }
    }
    if (!wrote) buf[n++] = '-';
    buf[n++] = ' '; buf[n++] = '0'; buf[n++] = ' '; buf[n++] = '1';
    buf[n] = 0;
    if (n + 1 > outLen) return -1;
    memcpy(out, buf, (size_t)n + 1);
    return n;
}

/* -------- seeded random walk -------- */
static u64 splitmix64(u64 *s)
{
    u64 z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Lemire's nearly-divisionless unbiased bounded draw */
static u32 bounded(u64 *s, u32 n)
{
    u64 x = splitmix64(s);
    unsigned __int128 m = (unsigned __int128)x * n;
    u64 l = (u64)m;
    if (l < n) {
        u64 t = (0 - (u64)n) % n;
        while (l < t) {
            x = splitmix64(s);
            m = (unsigned __int128)x * n;
            l = (u64)m;
        }
    }
    return (u32)(m >> 64);
}

static int canon_key(u16 m)
{
    int from = m & 63, to = (m >> 6) & 63, fl = m >> 12;
    int promo = (fl & 8) ? (fl & 3) + 1 : 0;   /* 0 none, 1 N, 2 B, 3 R, 4 Q */
    return (from * 64 + to) * 8 + promo;
}

static void sort_canon(u16 *m, u32 n)
{
    for (u32 i = 1; i < n; i++) {
        u16 v = m[i];
        int kv = canon_key(v);
        int j = (int)i - 1;
        while (j >= 0 && canon_key(m[j]) > kv) { m[j + 1] = m[j]; j--; }
        m[j + 1] = v;
    }
}

int oracle_random_walk_stream(u64 *rng, int plies, int requireNonTerminal,
                              uint64_t bb[4], uint8_t *state, uint8_t *color)
{
    oracle_init();
    int redraws = 0;
    for (;;) {
        oracle_parse_fen("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", bb, state, color);
        int ok = 1;
        for (int ply = 0; ply < plies; ply++) {
            u16 moves[256];
            u32 n = oracle_gen_moves(bb, *state, *color, moves);
            if (!n) { ok = 0; break; }
            sort_canon(moves, n);
            u32 pick = bounded(rng, n);
            oracle_make_move(bb, state, *color, moves[pick]);
            *color ^= 1;
        }
        if (ok && requireNonTerminal && oracle_count_moves(bb, *state, *color) == 0) ok = 0;
        if (ok) return redraws;
        redraws++;
    }
}

int oracle_random_position(uint64_t seed, int plies, int requireNonTerminal,
                           uint64_t bb[4], uint8_t *state, uint8_t *color)
{
    u64 rng = seed;
    return oracle_random_walk_stream(&rng, plies, requireNonTerminal, bb, state, color);
}

/* draws used by the batched and serving workloads: plies and depth from the
 * same stream before the walk */
int oracle_random_request(uint64_t seed, int pliesLo, int pliesHi, int depthLo, int depthHi,
                          uint64_t bb[4], uint8_t *state, uint8_t *color, int *plies, int *depth)
{
    u64 rng = seed;
    *plies = pliesLo + (int)bounded(&rng, (u32)(pliesHi - pliesLo + 1));
    *depth = depthLo + (int)bounded(&rng, (u32)(depthHi - depthLo + 1));
    return oracle_random_walk_stream(&rng, *plies, 1, bb, state, color);
}
