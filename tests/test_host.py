"""CPU-side checks of the engine library: it loads and exports the whole C ABI,
and its host code (FEN, generator, make, -cpu perft, workload walk) agrees
with the oracle.  No GPU is touched here."""
import os
import re

import pytest

from conftest import ROOT, load_json, load_jsonl


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "perft_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol(engine):
    L = engine.lib()
    names = _declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_python_mirror_covers_the_abi(engine):
    mirrored = {n for n, _, _ in engine._SIGS}
    assert set(_declared_symbols()) <= mirrored


def _to_engine(engine, orc_pos):
    p = engine.Position()
    for i in range(4):
        p.bb[i] = orc_pos.bb[i]
    p.state, p.color = orc_pos.state, orc_pos.color
    return p


def _canon(moves):
    return sorted(m & ~(4 << 12) if (m >> 12) in (4, 12, 13, 14, 15) else m for m in moves)


def _strip_flags(moves):
    # the engine sets the double-push flag only when an EP capture can follow
    return sorted((m & 0xFFF) | (((m >> 12) & 0xB) << 12 if (m >> 12) >= 8 else (((m >> 12) if (m >> 12) in (2, 3, 5) else 0) << 12))
                  for m in moves)


def test_fen_parse_matches_oracle(engine, orc):
    for p in load_json("known_positions.json")["positions"]:
        a, b = engine.parse_fen(p["fen"]), orc.parse_fen(p["fen"])
        assert tuple(a.bb) == b.bb and a.state == b.state and a.color == b.color


def test_invalid_fen_rejected(engine):
    with pytest.raises(engine.PerftError):
        engine.parse_fen("8/8/8/8/8/8/8/8 w - - 0 1")


def test_generator_matches_oracle_on_random_positions(engine, orc):
    for seed in range(4000):
        o = orc.random_position(seed, seed % 100, require_nonterminal=False)
        e = _to_engine(engine, o)
        assert engine.count_moves(e) == orc.count_moves(o), orc.write_fen(o)
        assert _strip_flags(engine.gen_moves(e)) == _strip_flags(orc.gen_moves(o)), orc.write_fen(o)


def test_make_move_matches_oracle(engine, orc):
    for seed in range(300):
        o = orc.random_position(seed, seed % 60)
        e = _to_engine(engine, o)
        def key(m):
            return (m & 0xFFF, ((m >> 12) & 3) if (m >> 12) >= 8 else -1)
        by_key = {key(m): m for m in orc.gen_moves(o)}
        for m in engine.gen_moves(e):
            om = by_key[key(m)]
            ec, oc = engine.make_move(e, m), orc.make_move(o, om)
            assert tuple(ec.bb) == oc.bb and ec.color == oc.color
            assert (ec.state & 15) == (oc.state & 15)
            # same legal continuations (EP state may differ only when no EP capture exists)
            assert engine.count_moves(ec) == orc.count_moves(oc)


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_cpu_perft_matches_known(engine, depth):
    for p in load_json("known_positions.json")["positions"]:
        if "counts" in p and len(p["counts"]) >= depth and p["counts"][depth - 1] < 5_000_000:
            assert engine.perft_cpu(engine.parse_fen(p["fen"]), depth) == p["counts"][depth - 1]


def test_cpu_perft_ep_evasions(engine, orc):
    # a side in check with two en-passant captures: perft(1) counts them once
    for p in load_json("ep_evasion.json")["positions"]:
        e = engine.parse_fen(p["fen"])
        assert [engine.perft_cpu(e, d) for d in range(1, 6)] == p["counts"][:5], p["name"]
        assert engine.count_moves(e) == orc.count_moves(orc.parse_fen(p["fen"]))


def test_cpu_perft_with_host_tt_matches(engine):
    fen = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    engine.tt_init(0, 64)
    try:
        assert engine.perft_cpu(engine.parse_fen(fen), 4, use_tt=True) == 4085603
        assert engine.perft_cpu(engine.parse_fen(fen), 4, use_tt=True) == 4085603
        st = engine.tt_stats()
        assert st[4] > 0 and st[5] > 0  # host hits and stores happened
    finally:
        engine.tt_free()


def test_host_tt_clear_takes_fresh_heads(engine):
    # a clear hands the bucket heads back and takes zeroed ones: nothing is
    # found right after it, counts stay exact, and a 256 GB tier is taken and
    # cleared without touching its pages (resident memory stays small)
    import os

    def resident_bytes():
        with open("/proc/self/statm") as f:
            return int(f.read().split()[1]) * os.sysconf("SC_PAGE_SIZE")

    fen = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    p = engine.parse_fen(fen)
    engine.tt_init(0, 64)
    try:
        assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
        assert engine.lib().pg_tt_clear() == 0
        before = engine.tt_stats()
        assert engine.perft_cpu(p, 2, use_tt=True) == 2039
        after = engine.tt_stats()
        assert after[4] == before[4]  # no host hit: the cleared tier holds nothing
        assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
    finally:
        engine.tt_free()
    rss0 = resident_bytes()
    engine.tt_init(0, 262144)
    try:
        assert engine.lib().pg_tt_clear() == 0
        assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
        st = engine.tt_stats()
        assert st[5] > 0  # the 256 GB tier was allocated and stored
        assert resident_bytes() - rss0 < 4 << 30  # under 4 GiB of the tier's pages touched
    finally:
        engine.tt_free()


def test_workload_walk_matches_oracle(engine, orc):
    for s in range(200):
        e, pl, d = engine.random_request(s, (6, 40), (3, 6))
        o, pl2, d2 = orc.random_request(s, (6, 40), (3, 6))
        assert (pl, d) == (pl2, d2)
        assert engine.write_fen(e) == orc.write_fen(o)
    for s in range(50):
        assert engine.write_fen(engine.random_position(s, 30)) == orc.write_fen(orc.random_position(s, 30))


@pytest.mark.parametrize("name", ["c1.jsonl", "c2.jsonl", "c3.jsonl", "c4.jsonl", "c5.jsonl"])
def test_fixture_fens_reproduced_by_engine(engine, name):
    rows = load_jsonl(name)
    for r in rows[:400]:
        if name in ("c1.jsonl", "c2.jsonl", "c4.jsonl"):
            p = engine.random_position(r["seed"], r["plies"])
        elif name == "c3.jsonl":
            p, pl, d = engine.random_request(r["seed"], (6, 40), (3, 6))
            assert (pl, d) == (r["plies"], r["depth"])
        else:
            p, pl, d = engine.random_request(r["seed"], (0, 60), (3, 7))
            assert (pl, d) == (r["plies"], r["depth"])
        assert engine.write_fen(p) == r["fen"]


def test_cli_cpu_mode(engine):
    r = engine.run_cli(["r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -", "3", "-cpu", "-nott"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "CPU mode" in r.stdout
    assert engine.parse_cli_counts(r.stdout) == {1: 48, 2: 2039, 3: 97862}
    assert re.search(r"Perft\(03\): 97862, time: [0-9.e+-]+ seconds", r.stdout)


def test_u128_printing(engine):
    import ctypes
    buf = ctypes.create_string_buffer(64)
    for v in [0, 1, 2**64 - 1, 2**64, 61885021521585529237, 2**128 - 1, 10**30]:
        engine.lib().pg_u128_to_string(v & (2**64 - 1), v >> 64, buf, 64)
        assert buf.value.decode() == str(v)


def test_cli_divide_cpu(engine, orc):
    fen = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    r = engine.run_cli([fen, "3", "-cpu", "-nott", "-divide"])
    assert r.returncode == 0, r.stdout
    lines = [l for l in r.stdout.splitlines() if len(l) > 4 and l[4] in ":qrbn"]
    assert len(lines) == 48
    assert "Total: 97862" in r.stdout
    # each root move's count equals the oracle's perft(2) of the child
    root = orc.parse_fen(fen)
    by_text = {}
    for m in orc.gen_moves(root):
        t = "abcdefgh"[m & 7] + str(((m >> 3) & 7) + 1) + "abcdefgh"[(m >> 6) & 7] + str(((m >> 9) & 7) + 1)
        if (m >> 12) >= 8:
            t += "nbrq"[(m >> 12) & 3]
        by_text[t] = orc.perft(orc.make_move(root, m), 2)
    for l in lines:
        mv, n = l.split(":")
        assert by_text[mv.strip()] == int(n)


def test_cli_suite_cpu(engine, tmp_path):
    f = tmp_path / "suite.epd"
    f.write_text("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1 ;D1 20 ;D2 400 ;D3 8902\n"
                 "8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ;D1 14 ;D2 191 ;D3 2812 ;D4 43238\n")
    r = engine.run_cli(["-suite", str(f), "4", "-cpu"])
    assert r.returncode == 0 and "7 checks, 0 failures" in r.stdout, r.stdout
    f.write_text("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1 ;D1 21\n")
    r = engine.run_cli(["-suite", str(f), "4", "-cpu"])
    assert r.returncode == 1 and "1 failures" in r.stdout


def test_cli_tt_snapshot_flags_cpu(engine, tmp_path):
    # -ttsave writes the tables after every depth; -ttload resumes from them.
    # Counts are the published Kiwipete ones either way; a truncated snapshot
    # is refused with a warning and the run goes on from empty tables.
    fen = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    want = {1: 48, 2: 2039, 3: 97862, 4: 4085603}
    snap = tmp_path / "tt.bin"
    r = engine.run_cli([fen, "4", "-cpu", "-htt", "16", "-ttsave", str(snap)])
    assert r.returncode == 0, r.stdout + r.stderr
    assert engine.parse_cli_counts(r.stdout) == want
    assert snap.stat().st_size > 0
    r = engine.run_cli([fen, "4", "-cpu", "-htt", "16", "-ttload", str(snap)])
    assert r.returncode == 0 and "TT snapshot loaded" in r.stdout, r.stdout
    assert engine.parse_cli_counts(r.stdout) == want
    cut = tmp_path / "cut.bin"
    cut.write_bytes(snap.read_bytes()[: snap.stat().st_size // 2])
    r = engine.run_cli([fen, "4", "-cpu", "-htt", "16", "-ttload", str(cut)])
    assert r.returncode == 0 and "warning: TT snapshot not loaded" in r.stdout, r.stdout
    assert engine.parse_cli_counts(r.stdout) == want


def test_host_tt_snapshot_resume(engine, tmp_path):
    # host tier only (no GPU here): save after a run, reload into fresh tables,
    # and the reloaded table answers the same queries exactly
    fen = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"
    p = engine.parse_fen(fen)
    engine.tt_init(0, 32)
    try:
        assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
        snap = str(tmp_path / "tt.bin")
        assert engine.lib().pg_tt_save(snap.encode()) == 0
    finally:
        engine.tt_free()
    engine.tt_init(0, 32)
    try:
        assert engine.lib().pg_tt_load(snap.encode()) == 0
        before = engine.tt_stats()
        assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
        after = engine.tt_stats()
        assert after[4] - before[4] >= 1  # the root itself is a hit after the reload
        engine.tt_free()
        engine.tt_init(0, 64)  # different geometry is refused
        assert engine.lib().pg_tt_load(snap.encode()) == engine.PG_ERR_INVALID
    finally:
        engine.tt_free()


def test_corrupt_snapshot_links_are_refused(engine, tmp_path):
    # a snapshot whose chain links point outside its entries is refused and
    # leaves the tables cleared; later counts stay exact
    import struct
    p = engine.parse_fen("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -")
    engine.tt_init(0, 32)
    try:
        assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
        snap = tmp_path / "tt.bin"
        assert engine.lib().pg_tt_save(str(snap).encode()) == 0
        raw = bytearray(snap.read_bytes())
        dev_slots, buckets, used = struct.unpack_from("<3Q", raw, 8)
        assert dev_slots == 0 and used > 2
        pool = 32 + 4 * buckets
        for label, off, value in (("head", 32 + 4 * next(i for i in range(buckets)
                                                           if struct.unpack_from("<I", raw, 32 + 4 * i)[0]), 0xFFFFFFF0),
                                  ("next", pool + 32 * 1 + 24, used + 5)):
            bad = bytearray(raw)
            struct.pack_into("<I", bad, off, value)
            path = tmp_path / f"bad_{label}.bin"
            path.write_bytes(bytes(bad))
            assert engine.lib().pg_tt_load(str(path).encode()) == engine.PG_ERR_INVALID, label
            assert "outside" in engine.last_error()
            assert engine.perft_cpu(p, 4, use_tt=True) == 4085603
        assert engine.lib().pg_tt_load(str(snap).encode()) == 0
    finally:
        engine.tt_free()


def test_invalid_positions_are_refused(engine):
    # every entry point that takes a caller-built position checks it before
    # the move generator sees it (a kingless board would index past the
    # per-square tables)
    import ctypes
    good = engine.parse_fen(engine.START_FEN)
    bad = []
    p = engine.Position(); p.bb[:] = list(good.bb); p.bb[2] &= ~(1 << 4); bad.append(("no white king", p))  # noqa: E702
    p = engine.Position(); p.bb[:] = list(good.bb); p.bb[0] |= 1 << 30; bad.append(("black empty square", p))  # noqa: E702
    p = engine.Position(); p.bb[:] = list(good.bb); p.bb[2] |= 1 << 3; bad.append(("code 7 on d1", p))  # noqa: E702
    p = engine.Position(); p.bb[:] = list(good.bb); p.color = 2; bad.append(("side 2", p))  # noqa: E702
    L = engine.lib()
    n = ctypes.c_int()
    moves = (ctypes.c_uint16 * 256)()
    cnt = ctypes.c_uint64()
    u = engine.U128()
    for why, p in bad:
        ref = ctypes.byref(p)
        assert L.pg_perft_cpu(ref, 3, 0, ctypes.byref(cnt)) == engine.PG_ERR_INVALID, why
        assert "invalid position" in engine.last_error()
        assert L.pg_gen_moves(ref, moves, ctypes.byref(n)) == engine.PG_ERR_INVALID, why
        assert L.pg_count_moves(ref, ctypes.byref(n)) == engine.PG_ERR_INVALID, why
        assert L.pg_make_move(ref, 0) == engine.PG_ERR_INVALID, why
        assert L.pg_perft_bfs(ref, 3, 0, None, 0, ctypes.byref(cnt)) == engine.PG_ERR_INVALID, why
        assert L.pg_perft(ref, 3, 0, 0, ctypes.byref(u)) == engine.PG_ERR_INVALID, why
        arr = (engine.Position * 2)(good, p)
        d = (ctypes.c_uint8 * 2)(3, 3)
        out = (ctypes.c_uint64 * 2)()
        assert L.pg_perft_batch(2, arr, d, None, out) == engine.PG_ERR_INVALID, why
        assert L.pg_perft_batch_plain(2, arr, d, 0, out) == engine.PG_ERR_INVALID, why
    assert engine.perft_cpu(good, 3) == 8902


def test_launcher_interface_is_exported(engine):
    # every function and global that include/perft_launcher.h declares (the
    # reference's launcher.h names, C++ linkage) is defined in the library,
    # with the reference's parameter types
    import subprocess
    so = os.path.join(ROOT, "perft_gpu_2026_amd", "libperft_gpu.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    funcs = {
        "initGPU(int)", "initMoveGen()",
        "estimateLaunchDepth(QuadBitBoard*, GameState*, unsigned char, float*)",
        "perftLauncher(QuadBitBoard*, GameState*, unsigned char, unsigned int, int)",
        "perft_gpu_host_bfs(QuadBitBoard*, GameState*, unsigned char, int, void*, unsigned long)",
        "perftCPU(QuadBitBoard*, GameState*, unsigned char, unsigned int)",
        "perft_cpu_dispatch(QuadBitBoard*, GameState*, unsigned char, unsigned int)",
        "initTT(int, int, int, float)", "clearDeviceTTs()", "freeTT()", "printTTStats()", "resetCallStats()",
        "getOomFallbackCount()", "getEffectiveLD()", "setEffectiveLD(int)",
        "readFENString(char const*, QuadBitBoard*, GameState*, unsigned char*)",
    }
    text = {line.split(" ", 2)[2] for line in out.splitlines() if line.count(" ") >= 2 and " T " in line}
    assert funcs <= text, funcs - text
    header = open(os.path.join(ROOT, "include", "perft_launcher.h")).read()
    for var in ("preAllocatedBufferHost", "g_preAllocatedMemorySize", "g_useTT", "g_lastBfsPeakMemory",
                "g_lastBfsOom", "g_deviceTTBudgetMB", "g_hostTTBudgetMB"):
        assert re.search(rf"extern [\w *]+\b{var};", header), var
        assert re.search(rf" [BD] {var}$", out, flags=re.M), var
    assert os.path.exists(os.path.join(ROOT, "perft_gpu_2026_amd", "bin", "launcher_caller"))


def test_magic_tables_match_ray_walk(engine):
    # the fancy magic tables of the table-lookup leaf: the builder checks
    # every blocker subset of every square as it fills them; here 2,000 random
    # full-board occupancies per square and slider go through the finished
    # tables against a ray walk
    import ctypes
    tries = ctypes.c_uint64()
    assert engine.lib().pg_magic_check(2000, ctypes.byref(tries)) == 0
    assert tries.value > 128
