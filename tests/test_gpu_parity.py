"""GPU parity: the HIP path, called through the C ABI, against the oracle's
golden counts (bit-exact; every count is an integer).

Covers the plain per-position BFS (both leaf forms), the launcher's host
recursion above the launch depth, forced arena OOM with fallback, the
kernel-level entry points one by one, and the perft_gpu command line."""
import ctypes

import numpy as np
import pytest

from conftest import load_json, load_jsonl

pytestmark = pytest.mark.gpu

GPU_NODE_BUDGET = 20_000_000_000


def _known(limit):
    out = []
    for p in load_json("known_positions.json")["positions"]:
        pairs = list(enumerate(p["counts"], 1)) if "counts" in p else [(p["depth"], p["count"])]
        out += [(p["name"], p["fen"], d, c) for d, c in pairs if c <= limit]
    return out


@pytest.mark.parametrize("leaf", [1, 2])
def test_known_positions_bfs(gpu, leaf):
    gpu.set_leaf_plies(leaf)
    try:
        for name, fen, d, c in _known(GPU_NODE_BUDGET if leaf == 1 else 2_000_000_000):
            print(f"[gpu] bfs leaf={leaf} {name} d={d}", flush=True)
            assert gpu.perft_bfs(gpu.parse_fen(fen), d) == c, (name, d)
    finally:
        gpu.set_leaf_plies(2)


@pytest.mark.parametrize("variant", [1])
def test_leaf2_variants(gpu, variant):
    # the 32-entry-chunk shape of the fused leaf (the default takes 64)
    gpu.lib().pg_set_leaf2_variant(variant)
    try:
        for name, fen, d, c in _known(2_000_000_000):
            print(f"[gpu] leaf2 variant={variant} {name} d={d}", flush=True)
            assert gpu.perft_bfs(gpu.parse_fen(fen), d) == c, (name, d)
        r = load_jsonl("c2.jsonl")[0]
        assert gpu.perft_bfs(gpu.parse_fen(r["fen"]), 7) == r["counts"][6]
    finally:
        gpu.lib().pg_set_leaf2_variant(0)


@pytest.mark.parametrize("leaf", [1, 2])
def test_ep_evasion_positions(gpu, leaf):
    # two en-passant captures out of a pawn check: levels are sized with both
    # moves, and the last ply counts them once, as the reference's perft(1) does
    gpu.set_leaf_plies(leaf)
    try:
        for p in load_json("ep_evasion.json")["positions"]:
            pos = gpu.parse_fen(p["fen"])
            print(f"[gpu] ep-evasion leaf={leaf} {p['name']}", flush=True)
            assert [gpu.perft_bfs(pos, d) for d in range(1, 7)] == p["counts"], p["name"]
            assert gpu.perft(pos, 6, launch_depth=3) == p["counts"][5], p["name"]
    finally:
        gpu.set_leaf_plies(2)


def test_ep_evasion_level_kernels(gpu):
    # the leaf entry point counts a checked side's en-passant captures once
    p = gpu.parse_fen("4k3/3p4/8/2P1P3/4K3/8/8/8 b - - 0 1")
    mv = [m for m in gpu.gen_moves(p) if (m & 63, (m >> 6) & 63) == (51, 35)]  # d7-d5+
    assert len(mv) == 1
    child = gpu.make_move(p, mv[0])
    assert len(gpu.gen_moves(child)) == 9 and gpu.count_moves(child) == 8
    pb = _d(gpu, np.array([list(p.bb)], dtype=np.uint64))
    ps = _d(gpu, np.array([p.state], dtype=np.uint8))
    di = _d(gpu, np.zeros(1, dtype=np.uint32))
    dm = _d(gpu, np.array(mv, dtype=np.uint16))
    want = {1: 8, 2: sum(gpu.count_moves(gpu.make_move(child, m)) for m in gpu.gen_moves(child))}
    for plies in (1, 2):
        s = gpu.DeviceArray(8).zero()
        assert gpu.lib().pg_leaf(pb.ptr, ps.ptr, di.ptr, dm.ptr, 1, 1, plies, s.ptr, None) == 0
        gpu.device_synchronize()
        assert int(s.to_numpy(np.uint64, 1)[0]) == want[plies]


def test_position3_perft10_reference_count(gpu):
    # The reference README's Position 3 perft 10 (README.md:74) is the one
    # published count that includes its once-only en-passant evasion count
    # (MoveGeneratorBitboard.h:1639-1649): 221,239 fewer leaves than the legal
    # moves.  A shared mistake of oracle and engine (both evasions counted)
    # would give 860,322,602,309 instead.
    k = {p["name"]: p for p in load_json("known_positions.json")["positions"]}["position3"]
    p = gpu.parse_fen(k["fen"])
    print("[gpu] position3 perft 10 (launcher, no tables)", flush=True)
    assert gpu.perft(p, 10) == 860_322_381_070
    print("[gpu] position3 perft 9 (one BFS call)", flush=True)
    assert gpu.perft_bfs(p, 9) == k["counts"][8]


@pytest.mark.parametrize("leaf", [1, 2])
def test_leaf_children_above_80_moves(gpu, leaf):
    # the reference's fused leaf keeps its child's moves in CMove[80]
    # (switches.h:24, perft_kernels.cu:385); these children have up to 149
    # moves, so a 64-entry chunk's pair list spans several LDS windows
    gpu.set_leaf_plies(leaf)
    try:
        for h in load_json("high_mobility.json")["positions"]:
            assert h["child_moves_max"] > 80
            pos = gpu.parse_fen(h["fen"])
            print(f"[gpu] high mobility leaf={leaf} {h['name']}", flush=True)
            assert [gpu.perft_bfs(pos, d) for d in range(1, 6)] == h["counts"], h["name"]
            assert gpu.perft(pos, 5, launch_depth=3) == h["counts"][4], h["name"]
    finally:
        gpu.set_leaf_plies(2)
    if leaf == 2:
        # the 32-entry-chunk shape and the leaf entry point on the same children
        gpu.lib().pg_set_leaf2_variant(1)
        try:
            for h in load_json("high_mobility.json")["positions"]:
                assert gpu.perft_bfs(gpu.parse_fen(h["fen"]), 5) == h["counts"][4]
        finally:
            gpu.lib().pg_set_leaf2_variant(0)
        h = load_json("high_mobility.json")["positions"][0]
        p = gpu.parse_fen(h["fen"])
        mv = gpu.gen_moves(p)
        pb = _d(gpu, np.array([list(p.bb)], dtype=np.uint64))
        ps = _d(gpu, np.array([p.state], dtype=np.uint8))
        di = _d(gpu, np.zeros(len(mv), dtype=np.uint32))
        dm = _d(gpu, np.array(mv, dtype=np.uint16))
        s = gpu.DeviceArray(8).zero()
        assert gpu.lib().pg_leaf(pb.ptr, ps.ptr, di.ptr, dm.ptr, 1, len(mv), 2, s.ptr, None) == 0
        gpu.device_synchronize()
        assert int(s.to_numpy(np.uint64, 1)[0]) == h["counts"][2]


def test_config1_all_depths(gpu):
    r = load_jsonl("c1.jsonl")[0]
    p = gpu.parse_fen(r["fen"])
    assert [gpu.perft_bfs(p, d) for d in range(1, 6)] == r["counts"]


def test_config2_full_depth(gpu):
    r = load_jsonl("c2.jsonl")[0]
    p = gpu.parse_fen(r["fen"])
    print("[gpu] config2 depth 1..7", flush=True)
    assert [gpu.perft_bfs(p, d) for d in range(1, 8)] == r["counts"]


@pytest.mark.parametrize("frontier", [1, 0])
def test_launcher_host_recursion(gpu, frontier):
    # frontier=1: host frontier merged and batched; frontier=0: the reference's
    # serial host recursion with one GPU call per node
    r = load_jsonl("c2.jsonl")[0]
    p = gpu.parse_fen(r["fen"])
    gpu.lib().pg_set_frontier_batch(frontier)
    try:
        for ld in (2, 3, 4):
            assert gpu.perft(p, 6, launch_depth=ld) == r["counts"][5]
        assert gpu.perft(gpu.parse_fen(gpu.START_FEN), 7, launch_depth=3) == 3195901860
    finally:
        gpu.lib().pg_set_frontier_batch(1)


def test_launcher_auto_frontier(gpu):
    # without the tables the automatic launch depth leaves a frontier of up to
    # four plies from depth 7 (identical positions merged, then the batched
    # engine); every count against the published ones
    cases = [(gpu.START_FEN, 7, 3195901860), (gpu.START_FEN, 8, 84998978956),
             ("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -", 7, 374190009323),
             ("8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - -", 8, 3009794393)]
    for fen, d, want in cases:
        print(f"[gpu] launcher auto frontier d={d} {fen}", flush=True)
        assert gpu.perft(gpu.parse_fen(fen), d) == want, (fen, d)
        assert gpu.lib().pg_get_effective_ld() <= max(4, d - 4), (fen, d)


def test_launcher_frontier_over_host_threads(gpu):
    # a five-ply frontier from the start position: its last ply (72,078
    # parents, 4,865,609 children, 822,518 unique) is expanded over the host
    # threads into hashed shards
    print("[gpu] five-ply frontier, startpos perft 8 at launch depth 3", flush=True)
    assert gpu.perft(gpu.parse_fen(gpu.START_FEN), 8, launch_depth=3) == 84998978956


def test_forced_oom_reports_and_falls_back(gpu):
    kiwi = gpu.parse_fen("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -")
    buf = gpu.DeviceArray(64 * 1024)
    s, _ = gpu.perft_bfs_status(kiwi, 5, buf.ptr.value, buf.nbytes)
    assert s == gpu.PG_ERR_ARENA_OOM
    # a buffer that fits gives the count
    big = gpu.DeviceArray(256 << 20)
    assert gpu.perft_bfs_status(kiwi, 4, big.ptr.value, big.nbytes) == (0, 4085603)
    before = gpu.lib().pg_get_oom_fallback_count()
    # perft 6 with the two-ply leaf holds ~4.1 M leaf entries (25 MB): an 8 MB
    # arena cannot take it, so the launcher splits one ply down
    gpu.lib().pg_set_arena_limit(8 << 20)
    try:
        assert gpu.perft(kiwi, 6, launch_depth=6) == 8031647685
    finally:
        gpu.lib().pg_set_arena_limit(0)
    assert gpu.lib().pg_get_oom_fallback_count() > before


def test_config3_sample_plain_batch(gpu):
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] <= 5][:300]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    got = gpu.perft_batch_plain(ps, [r["depth"] for r in rows])
    assert got == [r["count"] for r in rows]


# ---------------------------------------------------------------- kernel ABI
# Device buffers come from the library itself (pg_device_malloc/pg_memcpy):
# a second HIP runtime in this process (torch's) cannot share the device.
def _d(gpu, a):
    return gpu.DeviceArray.from_numpy(a)


@pytest.mark.parametrize("n", [1, 255, 256, 257, 16_385, 100_000, 1_048_576, 1_048_577, 3_000_001, 4_194_305,
                               9_000_000, 67_110_147])
def test_scan_matches_cumsum(gpu, n):
    # 1,048,576 entries (64 groups) is the largest level the one-block group
    # scan takes; one more entry takes the group pass and the top scan
    rng = np.random.default_rng(n)
    host = rng.integers(0, 219, size=n, dtype=np.uint32)
    d = _d(gpu, host)
    tiles = (n + 255) // 256
    scratch = gpu.DeviceArray(8 * (tiles + 4) + 4 * tiles)  # the documented size
    total = ctypes.c_uint64()
    assert gpu.lib().pg_scan_u32(d.ptr, n, None, scratch.ptr, ctypes.byref(total), None) == 0
    gpu.device_synchronize()
    assert np.array_equal(d.to_numpy(np.uint32, n), np.cumsum(host, dtype=np.uint64).astype(np.uint32))
    assert total.value == int(host.sum(dtype=np.uint64))


def test_scan_total_past_u32(gpu):
    # a total of 2^32 or more: the entries wrap like u32 (the BFS refuses such a
    # level as out of arena), the 64-bit total is exact
    n = 20_000_000
    host = np.full(n, 218, dtype=np.uint32)
    d = _d(gpu, host)
    tiles = (n + 255) // 256
    scratch = gpu.DeviceArray(8 * (tiles + 4) + 4 * tiles)
    total = ctypes.c_uint64()
    assert gpu.lib().pg_scan_u32(d.ptr, n, None, scratch.ptr, ctypes.byref(total), None) == 0
    gpu.device_synchronize()
    assert total.value == 218 * n > 2**32
    assert np.array_equal(d.to_numpy(np.uint32, n), np.cumsum(host, dtype=np.uint64).astype(np.uint32))


def test_bfs_levels_above_a_million_entries(gpu):
    # the BFS's large levels (group pass, top scan, apply with partitions): the
    # config-2 tree with the one-ply leaf scans its perft(5) level (15,143,580
    # entries), startpos perft 7 its 4,865,609-entry level, and the table BFS
    # the same levels through its merge pass
    r = load_jsonl("c2.jsonl")[0]
    pos = gpu.parse_fen(r["fen"])
    start = gpu.parse_fen("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1")
    gpu.set_leaf_plies(1)
    try:
        for d in (6, 7):
            print(f"[gpu] config-2 perft {d}, one-ply leaf", flush=True)
            assert gpu.perft_bfs(pos, d) == r["counts"][d - 1]
        assert gpu.perft_bfs(start, 7) == 3_195_901_860
        gpu.tt_init(1024, 0)
        try:
            print("[gpu] config-2 perft 7 with the table, one-ply leaf", flush=True)
            assert gpu.perft_bfs(pos, 7, use_tt=True) == r["counts"][6]
        finally:
            gpu.tt_free()
    finally:
        gpu.set_leaf_plies(2)


@pytest.mark.parametrize("n,shift", [(9_000_000, 0), (9_000_000, 1), (700_000, 1), (67_110_147, 1)])
def test_scan_with_given_tile_sums(gpu, n, shift):
    # caller-provided tile sums, 16-byte aligned (shift 0) or not (shift 1: 4 bytes
    # in); groups of 64 tile sums, and the top scan takes rows of 4,096 group
    # sums (67,108,864 entries), so 67,110,147 entries take two rows, the
    # second partial
    rng = np.random.default_rng(n + shift)
    host = rng.integers(0, 219, size=n, dtype=np.uint32)
    tiles = (n + 255) // 256
    pad = np.zeros(tiles * 256 - n, dtype=np.uint32)
    sums = np.concatenate([host, pad]).reshape(tiles, 256).sum(axis=1, dtype=np.uint64).astype(np.uint32)
    d = _d(gpu, host)
    ts = _d(gpu, np.concatenate([np.zeros(shift, dtype=np.uint32), sums]))
    scratch = gpu.DeviceArray(8 * (tiles + 4))  # the documented size
    total = ctypes.c_uint64()
    ptr = ctypes.c_void_p(ts.ptr.value + 4 * shift)
    assert gpu.lib().pg_scan_u32(d.ptr, n, ptr, scratch.ptr, ctypes.byref(total), None) == 0
    gpu.device_synchronize()
    assert np.array_equal(d.to_numpy(np.uint32, n), np.cumsum(host, dtype=np.uint64).astype(np.uint32))
    assert total.value == int(host.sum(dtype=np.uint64))


@pytest.mark.parametrize("n,zero_frac", [(1, 0.0), (1000, 0.0), (50_000, 0.3), (20_000, 0.97)])
def test_expand_matches_searchsorted(gpu, n, zero_frac):
    rng = np.random.default_rng(n)
    counts = rng.integers(1, 60, size=n).astype(np.uint32)
    counts[rng.random(n) < zero_frac] = 0
    counts[-1] = max(counts[-1], 1)
    incl = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    c = int(incl[-1])
    d_incl = _d(gpu, incl)
    out = gpu.DeviceArray(4 * c)
    scratch = gpu.DeviceArray(4 * (c // 2048 + 2))
    assert gpu.lib().pg_expand(d_incl.ptr, n, c, out.ptr, scratch.ptr, None) == 0
    gpu.device_synchronize()
    want = np.searchsorted(incl, np.arange(c), side="right")
    assert np.array_equal(out.to_numpy(np.uint32, c), want)


def test_small_level_boundary_against_oracle(gpu, orc):
    # A level of at most 1,024 entries runs in one block (k_level_small),
    # a larger one through make+count, scan, expand and genmoves.  The second
    # level holds perft(2) entries: positions with exactly 1,024 and 1,025,
    # and a few on either side, at perft 4 and 5 against the oracle.
    want = {1024: 2, 1025: 2, "below": 3, "above": 3}
    picked = []
    for seed in range(20_000):
        if not any(want.values()):
            break
        p = orc.random_position(seed, 4 + seed % 40, require_nonterminal=True)
        c2 = orc.perft(p, 2)
        key = c2 if c2 in (1024, 1025) else ("below" if 900 <= c2 < 1024 else "above" if 1025 < c2 <= 1200 else None)
        if key is not None and want.get(key, 0) > 0:
            want[key] -= 1
            picked.append((orc.write_fen(p), c2))
    assert not any(want.values()), want
    for fen, c2 in picked:
        for d in (4, 5):
            print(f"[gpu] small-level boundary perft2={c2} d={d} {fen}", flush=True)
            assert gpu.perft_bfs(gpu.parse_fen(fen), d) == orc.perft(orc.parse_fen(fen), d, threads=8), (fen, d)


def test_makecount_and_genmoves_match_host(gpu):
    # one colour per call: white-to-move parents
    roots = [p for p in (gpu.random_position(s, 2 * (s % 30)) for s in range(400)) if p.color == 0]
    pidx, moves, children = [], [], []
    for i, p in enumerate(roots):
        for m in gpu.gen_moves(p):
            pidx.append(i)
            moves.append(m)
            children.append(gpu.make_move(p, m))
    n = len(moves)
    L = gpu.lib()
    pb = _d(gpu, np.array([list(p.bb) for p in roots], dtype=np.uint64))
    ps = _d(gpu, np.array([p.state for p in roots], dtype=np.uint8))
    di = _d(gpu, np.array(pidx, dtype=np.uint32))
    dm = _d(gpu, np.array(moves, dtype=np.uint16))
    ob, os_ = gpu.DeviceArray(32 * n), gpu.DeviceArray(n)
    cnt, tiles = gpu.DeviceArray(4 * n), gpu.DeviceArray(4 * ((n + 255) // 256))
    assert L.pg_level_makecount(pb.ptr, ps.ptr, di.ptr, dm.ptr, 0, n, ob.ptr, os_.ptr, cnt.ptr, tiles.ptr, None) == 0
    gpu.device_synchronize()
    got_b = ob.to_numpy(np.uint64, 4 * n).reshape(n, 4)
    got_c = cnt.to_numpy(np.uint32, n)
    # the level count sizes the next level: every generated move, which is the
    # perft(1) count except for a checked side's double en-passant capture
    for k, ch in enumerate(children):
        assert tuple(int(x) for x in got_b[k]) == tuple(ch.bb)
        assert got_c[k] == len(gpu.gen_moves(ch))
    assert int(tiles.to_numpy(np.uint32, (n + 255) // 256).astype(np.int64).sum()) == int(got_c.sum())
    # genmoves over the children (black to move)
    incl = _d(gpu, np.cumsum(got_c).astype(np.uint32))
    c = int(got_c.sum())
    out = gpu.DeviceArray(2 * c)
    assert L.pg_genmoves(ob.ptr, os_.ptr, incl.ptr, 1, n, out.ptr, None) == 0
    gpu.device_synchronize()
    got_m = out.to_numpy(np.uint16, c)
    off = 0
    for k, ch in enumerate(children):
        seg = sorted(int(x) for x in got_m[off:off + got_c[k]])
        assert seg == sorted(gpu.gen_moves(ch))
        off += got_c[k]


def test_leaf_entry_point(gpu):
    p = gpu.parse_fen("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -")
    mv = gpu.gen_moves(p)
    pb = _d(gpu, np.array([list(p.bb)], dtype=np.uint64))
    ps = _d(gpu, np.array([p.state], dtype=np.uint8))
    di = _d(gpu, np.zeros(len(mv), dtype=np.uint32))
    dm = _d(gpu, np.array(mv, dtype=np.uint16))
    for plies, want in ((1, 2039), (2, 97862)):
        s = gpu.DeviceArray(8).zero()
        assert gpu.lib().pg_leaf(pb.ptr, ps.ptr, di.ptr, dm.ptr, 0, len(mv), plies, s.ptr, None) == 0
        gpu.device_synchronize()
        assert int(s.to_numpy(np.uint64, 1)[0]) == want


def test_cli_gpu(gpu):
    r = gpu.run_cli(["r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -", "5", "-nott",
                     "-arena", "4096"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert gpu.parse_cli_counts(r.stdout) == {1: 48, 2: 2039, 3: 97862, 4: 4085603, 5: 193690690}
    assert "Launch depth:" in r.stdout
    r = gpu.run_cli(["rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", "6", "-arena", "4096",
                     "-dtt", "512", "-htt", "256"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert gpu.parse_cli_counts(r.stdout)[6] == 119060324


def test_cli_gpu_divide_and_suite(gpu, tmp_path):
    # -divide on the GPU: one line per root move, and the total is perft(depth)
    kiwi = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    r = gpu.run_cli([kiwi, "5", "-divide", "-arena", "4096"])
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if len(l) > 4 and l[4] in ":qrbn"]
    assert len(lines) == 48
    assert sum(int(l.split(":")[1]) for l in lines) == 193690690
    assert "Total: 193690690" in r.stdout
    # an EPD suite on the GPU, one line failing on purpose
    f = tmp_path / "suite.epd"
    f.write_text(kiwi + " ;D1 48 ;D2 2039 ;D3 97862 ;D4 4085603\n"
                 "8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ;D1 14 ;D2 191 ;D3 2812 ;D4 43238 ;D5 674624\n")
    r = gpu.run_cli(["-suite", str(f), "5", "-arena", "4096"])
    assert r.returncode == 0 and "9 checks, 0 failures" in r.stdout, r.stdout + r.stderr
    f.write_text(kiwi + " ;D1 48 ;D2 2040\n")
    r = gpu.run_cli(["-suite", str(f), "5", "-arena", "4096"])
    assert r.returncode == 1 and "1 failures" in r.stdout, r.stdout


def test_cli_gpu_tt_snapshot(gpu, tmp_path):
    # the CLI writes both tiers after every depth and a second run resumes
    # from the file: same counts, and the run says it loaded the snapshot
    start = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1"
    snap = tmp_path / "tt.bin"
    args = [start, "6", "-arena", "4096", "-dtt", "64", "-htt", "64"]
    r = gpu.run_cli(args + ["-ttsave", str(snap)])
    assert r.returncode == 0, r.stdout + r.stderr
    assert gpu.parse_cli_counts(r.stdout)[6] == 119060324
    assert snap.stat().st_size > 32 << 20
    r = gpu.run_cli(args + ["-ttload", str(snap)])
    assert r.returncode == 0 and "TT snapshot loaded" in r.stdout, r.stdout + r.stderr
    assert gpu.parse_cli_counts(r.stdout) == {1: 20, 2: 400, 3: 8902, 4: 197281, 5: 4865609, 6: 119060324}


def test_genmoves_writes_nothing_for_zero_count_entries(gpu):
    # entries whose count is 0 (a table hit or a duplicate in TT mode) still
    # have legal moves; genmoves must not write them over their neighbours'
    ps = [gpu.random_position(s, 10 + s % 20) for s in range(600)]
    ps = [p for p in ps if p.color == 0]
    n = len(ps)
    counts = np.array([len(gpu.gen_moves(p)) for p in ps], dtype=np.uint32)
    counts[::3] = 0
    incl = np.cumsum(counts).astype(np.uint32)
    c = int(incl[-1])
    b = _d(gpu, np.array([list(p.bb) for p in ps], dtype=np.uint64))
    s = _d(gpu, np.array([p.state for p in ps], dtype=np.uint8))
    d_incl = _d(gpu, incl)
    out = _d(gpu, np.full(c + 64, 0xFFFF, dtype=np.uint16))
    assert gpu.lib().pg_genmoves(b.ptr, s.ptr, d_incl.ptr, 0, n, out.ptr, None) == 0
    gpu.device_synchronize()
    got = out.to_numpy(np.uint16, c + 64)
    off = 0
    for k, p in enumerate(ps):
        assert sorted(int(x) for x in got[off:off + counts[k]]) == (sorted(gpu.gen_moves(p)) if counts[k] else [])
        off += counts[k]
    assert (got[c:] == 0xFFFF).all()


@pytest.mark.parametrize("tables", [True, False])
def test_reference_launcher_interface(gpu, tables):
    # a C++ program written only against the reference's launcher names
    # (include/perft_launcher.h), linked with the library: the whole depth in
    # one GPU call, a forced OOM reported through g_lastBfsOom, the -cpu
    # dispatch, and perft.cu's per-depth loop with launch depth 4 (a device
    # table for the GPU calls, host recursion above them with the host tier,
    # when the tables are on)
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(gpu.LIB_PATH), "bin", "launcher_caller")
    kiwi = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    want = [48, 2039, 97862, 4085603, 193690690]
    args = [exe, kiwi, "5", "4"] + ([] if tables else ["-nott"])
    print(f"[gpu] launcher_caller tables={tables}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    kv = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if parts and parts[0] in ("bfs", "bfs_small", "cpu3", "estimate", "oom_fallbacks"):
            kv.setdefault(parts[0], []).append(parts[1:])
    assert kv["bfs"][0][:3] == ["193690690", "oom", "0"]
    assert kv["bfs_small"][0] == ["0", "oom", "1"]
    assert kv["cpu3"][0] == ["97862"]
    assert int(kv["estimate"][0][0]) >= 1
    assert all(v[0] == "0" for v in kv["oom_fallbacks"])
    counts = gpu.parse_cli_counts(r.stdout)  # the last Perft line of each depth
    perft_lines = [line for line in r.stdout.splitlines() if line.startswith("Perft(")]
    # five GPU depths, depth 5 again after clearDeviceTTs, then three -cpu depths
    assert [int(line.split()[1].rstrip(",")) for line in perft_lines] == want + [want[4]] + want[:3]
    assert counts[5] == want[4]
    assert ("Device TT probes" in r.stdout) == tables
    assert ("Host TT probes" in r.stdout) == tables
    assert r.stdout.rstrip().endswith("done")


@pytest.mark.parametrize("mode", [1, 2])
def test_magic_table_leaf(gpu, mode):
    # the fused leaf with fancy magic lookups (tables in device memory, or the
    # bishop table in LDS) gives the golden counts: the published positions,
    # the en-passant-evasion and high-mobility positions and config 2
    assert gpu.lib().pg_set_slider_attacks(mode) == 0, gpu.last_error()
    try:
        for name, fen, d, c in _known(2_000_000_000):
            print(f"[gpu] magic leaf mode={mode} {name} d={d}", flush=True)
            assert gpu.perft_bfs(gpu.parse_fen(fen), d) == c, (name, d)
        for p in load_json("ep_evasion.json")["positions"] + load_json("high_mobility.json")["positions"]:
            pos = gpu.parse_fen(p["fen"])
            assert [gpu.perft_bfs(pos, d) for d in range(1, len(p["counts"]) + 1)] == p["counts"], p["name"]
        r = load_jsonl("c2.jsonl")[0]
        assert gpu.perft_bfs(gpu.parse_fen(r["fen"]), 7) == r["counts"][6]
    finally:
        assert gpu.lib().pg_set_slider_attacks(0) == 0
