/* oracle_perft <fen> <depth> [threads] [-divide] [-brute]
 * Command-line front end of the CPU oracle (test infrastructure only).
 * Prints "Perft(dd): <count>" for each depth 1..depth, the reference's -cpu
 * output form (launcher.cu:750-753) without the timing fields. */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: oracle_perft <fen> <depth> [threads] [-divide] [-brute]\n");
        return 2;
    }
    int threads = 1, divide = 0, brute = 0;
    for (int i = 3; i < argc; i++) {
        if (!strcmp(argv[i], "-divide")) divide = 1;
        else if (!strcmp(argv[i], "-brute")) brute = 1;
        else threads = atoi(argv[i]);
    }
    uint64_t bb[4];
    uint8_t st, col;
    if (oracle_parse_fen(argv[1], bb, &st, &col)) {
        fprintf(stderr, "bad fen\n");
        return 2;
    }
    int depth = atoi(argv[2]);
    if (divide) {
        uint16_t mv[256];
        uint32_t n = oracle_gen_moves(bb, st, col, mv);
        uint64_t total = 0;
        for (uint32_t i = 0; i < n; i++) {
            uint64_t c[4] = { bb[0], bb[1], bb[2], bb[3] };
            uint8_t s2 = st;
            oracle_make_move(c, &s2, col, mv[i]);
            uint64_t v = oracle_perft_mt(c, s2, !col, depth - 1, threads);
            total += v;
            printf("%c%d%c%d %d: %llu\n", 'a' + (mv[i] & 7), ((mv[i] >> 3) & 7) + 1,
                   'a' + ((mv[i] >> 6) & 7), ((mv[i] >> 9) & 7) + 1, mv[i] >> 12, (unsigned long long)v);
        }
        printf("total: %llu\n", (unsigned long long)total);
        return 0;
    }
    for (int d = 1; d <= depth; d++) {
        uint64_t v = brute ? oracle_perft_brute(bb, st, col, d) : oracle_perft_mt(bb, st, col, d, threads);
        printf("Perft(%02d): %llu\n", d, (unsigned long long)v);
        fflush(stdout);
    }
    return 0;
}
