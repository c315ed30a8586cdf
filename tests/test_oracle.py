"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Checks: published perft counts (known_positions.json), the fast legal
generator against the independent brute-force generator (pseudo-legal moves
filtered by king safety) on seeded random positions, the count against the
brute generator's count under the reference's perft(1) convention (a side in
check counts its en-passant captures once) on every node of full trees, FEN
round trips, and that a planted mistake in the oracle's inputs is caught by
these checks.
"""
import pytest

from conftest import load_json

NODE_BUDGET = 300_000_000


def _known():
    out = []
    for p in load_json("known_positions.json")["positions"]:
        if "counts" in p:
            for d, c in enumerate(p["counts"], 1):
                if c <= NODE_BUDGET:
                    out.append((p["name"], p["fen"], d, c))
        elif p["count"] <= NODE_BUDGET:
            out.append((p["name"], p["fen"], p["depth"], p["count"]))
    return out


@pytest.mark.parametrize("name,fen,depth,count", _known())
def test_published_counts(orc, name, fen, depth, count):
    assert orc.perft(orc.parse_fen(fen), depth, threads=8) == count


def _canon(moves):
    # drop the capture flag (bit 14 of the flags nibble): generators may differ there
    return sorted(m & ~(4 << 12) if (m >> 12) in (4, 12, 13, 14, 15) else m for m in moves)


def test_fast_equals_brute_on_random_positions(orc):
    for seed in range(3000):
        p = orc.random_position(seed, seed % 90, require_nonterminal=False)
        fast, brute = orc.gen_moves(p), orc.brute_gen_moves(p)
        assert _canon(fast) == _canon(brute), (seed, orc.write_fen(p))
        assert orc.count_moves(p) == orc.brute_count_moves(p)


def test_fast_equals_brute_perft3(orc):
    for seed in range(300):
        p = orc.random_position(1000 + seed, 10 + seed % 60, require_nonterminal=False)
        assert orc.perft(p, 3) == orc.perft_brute(p, 3), orc.write_fen(p)


@pytest.mark.parametrize("fen", [p["fen"] for p in load_json("known_positions.json")["positions"]][:8])
def test_count_equals_generate_on_tree(orc, fen):
    root = orc.parse_fen(fen)
    frontier = [root]
    for _ in range(3):
        nxt = []
        for p in frontier:
            mv = orc.gen_moves(p)
            assert orc.count_moves(p) == orc.brute_count_moves(p) <= len(mv)
            nxt.extend(orc.make_move(p, m) for m in mv)
        frontier = nxt[:4000]


# A side in check by a pawn that just made a double push, with two pawns that
# can take it en passant: both captures are legal moves, and the reference's
# countMovesOutOfCheck counts them as one (MoveGeneratorBitboard.h:1639-1649).
EP_DOUBLE_EVASION = [
    # white king e4 checked by d7-d5; c5xd6 and e5xd6 both legal
    ("4k3/8/8/2PpP3/4K3/8/8/8 w - d6 0 1", 9, 2),
    # black king d5 checked by e2-e4; d4xe3 and f4xe3 both legal
    ("8/8/8/3k4/3pPp2/8/8/4K3 b - e3 0 1", 9, 2),
]


@pytest.mark.parametrize("fen,legal,ep", EP_DOUBLE_EVASION)
def test_count_takes_ep_evasions_once(orc, fen, legal, ep):
    p = orc.parse_fen(fen)
    mv = orc.gen_moves(p)
    assert len(mv) == legal == len(orc.brute_gen_moves(p))
    assert sum((m >> 12) == 5 for m in mv) == ep
    assert orc.count_moves(p) == orc.brute_count_moves(p) == legal - ep + 1
    assert orc.perft(p, 1) == legal - ep + 1
    # one ply deeper every leaf is counted, both captures included
    assert orc.perft(p, 2) == sum(orc.count_moves(orc.make_move(p, m)) for m in mv)


def test_fen_round_trip(orc):
    for seed in range(500):
        p = orc.random_position(seed, seed % 70)
        q = orc.parse_fen(orc.write_fen(p))
        assert q.bb == p.bb and q.color == p.color and (q.state & 15) == (p.state & 15)


def test_walk_is_deterministic(orc):
    a = [orc.write_fen(orc.random_position(s, 20)) for s in range(50)]
    b = [orc.write_fen(orc.random_position(s, 20)) for s in range(50)]
    assert a == b and len(set(a)) > 40


def test_planted_mistake_is_caught(orc):
    # moving a white pawn one square changes Kiwipete's perft(3): the published
    # count distinguishes the two positions
    good = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    bad = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPP1/R3K2R w KQkq -"
    assert orc.perft(orc.parse_fen(good), 3) == 97862
    assert orc.perft(orc.parse_fen(bad), 3) != 97862
