"""Writes the committed workload fixtures and golden counts under tests/golden/.

  python tests/golden/make_fixtures.py [--threads N]

Positions come from the oracle's seeded random walk (SplitMix64 stream,
canonically sorted legal moves, Lemire pick; oracle_random_position /
oracle_random_request).  The engine's own generator must reproduce every FEN
(tests/test_host.py checks that), so the product never needs this script.

Golden counts come from the oracle (oracle/oracle.c), which is pinned against
published perft values in tests/test_oracle.py.  Where a full-size count is
beyond the oracle (config 4 at depth 8, the deep serving requests), the file
keeps the FEN and depth and a reduced-depth count instead.

Configs (sizes and seeds as the benchmark defines them):
  c1  seed 1, 4 plies, perft 1..5
  c2  seed 2, 8 plies (non-terminal), perft 1..7
  c3  seeds 0..4095: plies U[6,40] then depth U[3,6] from the seed's stream
  c4  seeds 100..163, 10 plies, depth 8 (golden at depth 5)
  c5  request ids 0..9999: plies U[0,60] then depth U[3,7]; golden for every
      request of depth <= 5 and the first 64 of depth 6
  ep  hand-set positions where a side in check by a double-pushed pawn has two
      en-passant captures (which the reference's perft(1) counts once), and
      positions one ply before them, perft 1..6
  hm  high-mobility positions: black to move, and after almost every black
      move white has 81-149 legal moves, so the fused two-ply leaf's children
      exceed the reference's 80-move leaf cap (switches.h:24,
      perft_kernels.cu:385) at plies 1 and 3; perft 1..5
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


EP_EVASION = [
    ("wk-e4 checked by d5, c5/e5 take", "4k3/8/8/2PpP3/4K3/8/8/8 w - d6 0 1"),
    ("bk-d5 checked by e4, d4/f4 take", "8/8/8/3k4/3pPp2/8/8/4K3 b - e3 0 1"),
    ("before d7-d5+", "4k3/3p4/8/2P1P3/4K3/8/8/8 b - - 0 1"),
    ("before e2-e4+", "8/8/8/3k4/3p1p2/8/4P3/4K3 w - - 0 1"),
    ("full board, before d7-d5+", "rnbqkbnr/ppp1pppp/8/2P1P3/4K3/8/PP1P1PPP/RNBQ1BNR b kq - 0 1"),
]

# found by a seeded search over queen placements (black king h8 behind pawns,
# rook a8, knight b8; white king e1, rooks a1/h1, 4-8 queens)
HIGH_MOBILITY = [
    ("eight queens, a", "rn5k/5ppp/1QQQQ3/8/6QQ/3Q4/6Q1/R3K2R b KQ - 0 1"),
    ("eight queens, b", "rn5k/5ppp/1Q2Q3/3Q4/5Q2/7Q/Q1QQ4/R3K2R b KQ - 0 1"),
]


def dump(name, rows):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        for r in rows:
            f.write(json.dumps(r, separators=(",", ":")) + "\n")
    print(f"wrote {path} ({len(rows)} rows)", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    T = a.threads
    only = set(a.only.split(",")) if a.only else {"c1", "c2", "c3", "c4", "c5", "ep", "hm"}
    t0 = time.time()

    if "c1" in only:
        p = oracle.random_position(1, 4, True)
        counts = [oracle.perft(p, d, T) for d in range(1, 6)]
        dump("c1.jsonl", [{"seed": 1, "plies": 4, "fen": oracle.write_fen(p), "depth": 5, "counts": counts}])

    if "c2" in only:
        p = oracle.random_position(2, 8, True)
        counts = []
        for d in range(1, 8):
            counts.append(oracle.perft(p, d, T))
            print(f"c2 depth {d}: {counts[-1]} ({time.time() - t0:.0f}s)", flush=True)
        dump("c2.jsonl", [{"seed": 2, "plies": 8, "fen": oracle.write_fen(p), "depth": 7, "counts": counts}])

    if "c3" in only:
        reqs = [oracle.random_request(s, (6, 40), (3, 6)) for s in range(4096)]
        counts = oracle.perft_batch([r[0] for r in reqs], [r[2] for r in reqs], T)
        dump("c3.jsonl", [{"seed": s, "plies": r[1], "fen": oracle.write_fen(r[0]), "depth": r[2], "count": c}
                          for s, (r, c) in enumerate(zip(reqs, counts))])
        print(f"c3 total nodes {sum(counts)} ({time.time() - t0:.0f}s)", flush=True)

    if "c4" in only:
        ps = [oracle.random_position(s, 10, True) for s in range(100, 164)]
        counts = oracle.perft_batch(ps, [5] * len(ps), T)
        dump("c4.jsonl", [{"seed": 100 + i, "plies": 10, "fen": oracle.write_fen(p), "depth": 8,
                           "golden_depth": 5, "golden_count": c} for i, (p, c) in enumerate(zip(ps, counts))])

    if "c5" in only:
        reqs = [oracle.random_request(s, (0, 60), (3, 7)) for s in range(10000)]
        pick = [i for i, r in enumerate(reqs) if r[2] <= 5]
        pick += [i for i, r in enumerate(reqs) if r[2] == 6][:64]
        counts = oracle.perft_batch([reqs[i][0] for i in pick], [reqs[i][2] for i in pick], T)
        gold = dict(zip(pick, counts))
        dump("c5.jsonl", [{"seed": s, "plies": r[1], "fen": oracle.write_fen(r[0]), "depth": r[2],
                           **({"count": gold[s]} if s in gold else {})} for s, r in enumerate(reqs)])
    if "ep" in only:
        rows = []
        for name, fen in EP_EVASION:
            p = oracle.parse_fen(fen)
            rows.append({"name": name, "fen": fen, "counts": [oracle.perft(p, d, T) for d in range(1, 7)]})
        path = os.path.join(HERE, "ep_evasion.json")
        with open(path, "w") as f:
            json.dump({"_source": "oracle/oracle.c perft, whose perft(1) counts a checked side's en-passant "
                                  "captures once (MoveGeneratorBitboard.h:1639-1649)", "positions": rows}, f, indent=1)
        print(f"wrote {path}", flush=True)
    if "hm" in only:
        rows = []
        for name, fen in HIGH_MOBILITY:
            p = oracle.parse_fen(fen)
            kids = [oracle.count_moves(oracle.make_move(p, m)) for m in oracle.gen_moves(p)]
            rows.append({"name": name, "fen": fen, "child_moves_max": max(kids),
                         "children_above_80": sum(1 for k in kids if k > 80),
                         "counts": [oracle.perft(p, d, T) for d in range(1, 6)]})
        path = os.path.join(HERE, "high_mobility.json")
        with open(path, "w") as f:
            json.dump({"_source": "oracle/oracle.c perft; positions whose leaf children have more than 80 legal "
                                  "moves", "positions": rows}, f, indent=1)
        print(f"wrote {path}", flush=True)
    print(f"done in {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
