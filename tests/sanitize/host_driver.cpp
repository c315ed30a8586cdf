// host-only exercise of the C ABI under ASan/UBSan (no GPU calls)
#include "../../include/perft_gpu.h"  // host-only exercise of the C ABI; built by test_sanitize.py
#include <cstdio>
#include <cstring>
int main() {
  const char* fens[] = {"rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1",
    "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -",
    "8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - 0 1", "3Q4/1Q4Q1/4Q3/2Q4R/Q4Q2/3Q4/1Q4Rp/1K1BBNNk w - - 0 1",
    "8/8/1k6/2b5/2pP4/8/5K2/8 b - d3 0 1", "bad fen", ""};
  uint64_t want[] = {197281, 4085603, 43238, 0, 13931};
  int bad = 0;
  for (int i = 0; i < 7; i++) {
    pg_position p; int s = pg_parse_fen(fens[i], &p);
    if (i >= 5) { if (s == 0) bad++; continue; }
    char buf[128]; pg_write_fen(&p, buf, sizeof buf);
    uint16_t mv[256]; int n; pg_gen_moves(&p, mv, &n);
    uint64_t c = 0; pg_perft_cpu(&p, i == 3 ? 2 : 4, 0, &c);
    if (i != 3 && c != want[i]) { printf("mismatch %d %llu\n", i, (unsigned long long)c); bad++; }
  }
  pg_tt_init(0, 16);
  pg_position p; pg_parse_fen(fens[1], &p);
  uint64_t c1 = 0, c2 = 0; pg_perft_cpu(&p, 4, 1, &c1); pg_perft_cpu(&p, 4, 1, &c2);
  if (c1 != 4085603 || c2 != 4085603) bad++;
  pg_tt_save("/tmp/asan_tt.bin"); pg_tt_free(); pg_tt_init(0, 16);
  if (pg_tt_load("/tmp/asan_tt.bin") != 0) bad++;
  pg_perft_cpu(&p, 4, 1, &c1); if (c1 != 4085603) bad++;
  pg_tt_free();
  for (uint64_t s = 0; s < 300; s++) { pg_position q; int pl, d; pg_random_request(s, 0, 80, 3, 6, &q, &pl, &d); char b[128]; pg_write_fen(&q, b, 128); }
  char b[64]; pg_u128_to_string(~0ULL, ~0ULL, b, 64); if (strcmp(b, "340282366920938463463374607431768211455")) bad++;
  pg_u128_to_string(5, 0, b, 2);  // truncation path
  uint64_t tries = 0; if (pg_magic_check(100, &tries) != 0) bad++;  // magic table search and lookups
  pg_position k; pg_parse_fen(fens[0], &k); k.bb[2] &= ~(1ULL << 4);  // no white king
  if (pg_perft_cpu(&k, 3, 0, &c1) != PG_ERR_INVALID) bad++;
  printf("asan driver: %s\n", bad ? "FAIL" : "ok"); return bad;
}
