"""GPU tests of the transposition table, the batched path and the serving loop,
each against the oracle's golden counts or the plain path (bit-exact)."""
import time

import pytest

from conftest import load_json, load_jsonl

pytestmark = pytest.mark.gpu

KIWI = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -"


@pytest.fixture()
def tt(gpu):
    gpu.tt_init(1024, 256)
    yield gpu
    gpu.tt_free()


@pytest.mark.parametrize("leaf", [1, 2])
def test_tt_bfs_equals_no_tt(tt, leaf):
    tt.set_leaf_plies(leaf)
    try:
        for p in load_json("known_positions.json")["positions"][:7]:
            for d, c in enumerate(p["counts"], 1):
                if c > 3_000_000_000 or d < 3:
                    continue
                pos = tt.parse_fen(p["fen"])
                print(f"[gpu] tt leaf={leaf} {p['name']} d={d}", flush=True)
                assert tt.perft_bfs(pos, d, use_tt=True) == c
                assert tt.perft_bfs(pos, d, use_tt=True) == c  # second call hits the table
        st = tt.tt_stats()
        assert st[0] > 0 and st[1] > 0 and st[2] > 0, st
    finally:
        tt.set_leaf_plies(2)


def test_tt_launcher_with_host_tier(tt):
    r = load_jsonl("c2.jsonl")[0]
    p = tt.parse_fen(r["fen"])
    # with the tables the launcher recurses serially above the launch depth;
    # launch depth 4 leaves 19,494 host nodes after 3 plies, where moves of
    # both sides transpose and the host tier hits
    for ld in (4, 5):
        print(f"[gpu] launcher tt ld={ld}", flush=True)
        assert tt.perft(p, 7, launch_depth=ld, use_tt=True) == r["counts"][6]
    st = tt.tt_stats()
    assert st[3] > 0 and st[4] > 0  # host tier probed and hit


def test_config4_with_64gb_table(gpu):
    # config 4's 64 positions (10 random plies, seeds 100..163) through the
    # launcher with a 64 GB device table and a host tier: the oracle's depth-5
    # counts, then TT on == TT off at depth 6 with two host plies above the
    # launch depth (host tier probed) and with one whole-depth BFS call each
    rows = load_jsonl("c4.jsonl")
    assert len(rows) == 64
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    gpu.tt_init(65536, 2048)
    try:
        for r, p in zip(rows, ps):
            print(f"[gpu] config4 seed {r['seed']} depth {r['golden_depth']} with the tables", flush=True)
            assert gpu.perft(p, r["golden_depth"], use_tt=True) == r["golden_count"], r["seed"]
        print("[gpu] config4 depth 6, tables on (launch depth 4) and off", flush=True)
        on = [gpu.perft(p, 6, launch_depth=4, use_tt=True) for p in ps]
        on_bfs = [gpu.perft_bfs(p, 6, use_tt=True) for p in ps]
        off = [gpu.perft_bfs(p, 6) for p in ps]
        assert on == off and on_bfs == off
        st = gpu.tt_stats()
        assert st[1] > 0 and st[3] > 0 and st[4] > 0, st  # device hits, host probes and hits
    finally:
        gpu.tt_free()


def test_high_mobility_leaf_children_with_tables_and_batch(gpu):
    # leaf children with 81-149 moves through the TT leaf and the batch leaves
    hm = load_json("high_mobility.json")["positions"]
    ps = [gpu.parse_fen(h["fen"]) for h in hm for _ in range(5)]
    ds = [d for _ in hm for d in range(1, 6)]
    want = [c for h in hm for c in h["counts"]]
    for leaf in (1, 2):
        assert gpu.perft_batch(ps, ds, leaf_plies=leaf) == want, leaf
    gpu.tt_init(1024, 0)
    try:
        for leaf in (1, 2):
            gpu.set_leaf_plies(leaf)
            for h in hm:
                p = gpu.parse_fen(h["fen"])
                assert [gpu.perft_bfs(p, d, use_tt=True) for d in range(1, 6)] == h["counts"], (leaf, h["name"])
            assert gpu.perft_batch(ps, ds, leaf_plies=leaf, use_tt=True) == want, leaf
    finally:
        gpu.set_leaf_plies(2)
        gpu.tt_free()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_tt_launcher_frontier_modes(gpu, mode):
    # the launcher with the tables: serial recursion (0), the unique frontier
    # over parallel host threads with their own streams and arena quarters
    # (1), and the unique frontier through the batched engine (2)
    assert gpu.lib().pg_set_tt_frontier(mode) == 0
    gpu.tt_init(4096, 512)
    try:
        r = load_jsonl("c2.jsonl")[0]
        p = gpu.parse_fen(r["fen"])
        for ld in (4, 5):
            print(f"[gpu] tt frontier mode={mode} c2 ld={ld}", flush=True)
            assert gpu.perft(p, 7, launch_depth=ld, use_tt=True) == r["counts"][6]
        rows = load_jsonl("c4.jsonl")[:8]
        ps = [gpu.parse_fen(x["fen"]) for x in rows]
        got = [gpu.perft(q, 6, launch_depth=3, use_tt=True) for q in ps]
        assert got == [gpu.perft_bfs(q, 6) for q in ps]
        # a frontier node that does not fit its slot's quarter of an 8 MB arena
        # is recomputed through the serial recursion
        gpu.lib().pg_tt_clear()
        gpu.lib().pg_set_arena_limit(8 << 20)
        try:
            assert gpu.perft(gpu.parse_fen(KIWI), 6, launch_depth=5, use_tt=True) == 8031647685
        finally:
            gpu.lib().pg_set_arena_limit(0)
    finally:
        gpu.tt_free()
        gpu.lib().pg_set_tt_frontier(0)


def test_tt_tiny_table_still_exact(gpu):
    gpu.tt_init(1, 0)  # 64 Ki slots: constant overwrites and collisions of slots
    try:
        p = gpu.parse_fen(KIWI)
        assert gpu.perft_bfs(p, 5, use_tt=True) == 193690690
        assert gpu.perft_bfs(p, 5, use_tt=True) == 193690690
    finally:
        gpu.tt_free()


def test_tt_oversized_request_is_halved(gpu, capfd):
    # 512 GB does not fit beside the arena: the table halves until it does
    gpu.tt_init(512 * 1024, 0)
    try:
        assert "does not fit" in capfd.readouterr().err
        p = gpu.parse_fen(KIWI)
        before = gpu.tt_stats()
        assert gpu.perft_bfs(p, 5, use_tt=True) == 193690690
        assert gpu.tt_stats()[0] > before[0]  # the device table exists and was probed
    finally:
        gpu.tt_free()


@pytest.mark.parametrize("leaf", [1, 2])
def test_batch_matches_golden_c3(gpu, leaf):
    rows = load_jsonl("c3.jsonl")
    rows = [r for r in rows if r["depth"] <= 5][:1500] + [r for r in rows if r["depth"] == 6][:40]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    got = gpu.perft_batch(ps, [r["depth"] for r in rows], leaf_plies=leaf)
    assert got == [r["count"] for r in rows]


@pytest.mark.parametrize("leaf", [1, 2])
def test_batch_ep_evasions(gpu, leaf):
    # every depth of the en-passant-evasion positions in one batch, TT off and on
    rows = load_json("ep_evasion.json")["positions"]
    ps = [gpu.parse_fen(r["fen"]) for r in rows for _ in range(6)]
    ds = [d for _ in rows for d in range(1, 7)]
    want = [c for r in rows for c in r["counts"]]
    assert gpu.perft_batch(ps, ds, leaf_plies=leaf) == want
    gpu.tt_init(256, 0)
    try:
        assert gpu.perft_batch(ps, ds, leaf_plies=leaf, use_tt=True) == want
    finally:
        gpu.tt_free()


def test_batch_autotuned_with_tt(gpu):
    # no forced leaf: the autotuner picks leaf plies and table use per depth
    # group (timed on samples of the group); any choice gives the golden counts
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] <= 5][:1200]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    gpu.tt_init(1024, 0)
    try:
        for rep in range(2):
            assert gpu.perft_batch(ps, [r["depth"] for r in rows], use_tt=True) == [r["count"] for r in rows]
    finally:
        gpu.tt_free()


def test_batch_deep_group_not_tuned(gpu):
    # depth >= 7: the tuner picks the two-ply leaf without trials (a one-ply
    # trial on such roots costs minutes); counts as golden
    r = load_jsonl("c2.jsonl")[0]
    p = gpu.parse_fen(r["fen"])
    t0 = time.perf_counter()
    assert gpu.perft_batch([p, p], [7, 7]) == [r["counts"][6]] * 2
    assert time.perf_counter() - t0 < 30.0


def test_batch_grid_choices(gpu):
    # the persistent grid of the level kernels and the table leaf held at
    # 4, 16 and 32 blocks per CU, and the tuner's own choice, give the
    # golden counts; the tuner reports a grid from its candidate set
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] <= 5][:400]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    ds = [r["depth"] for r in rows]
    want = [r["count"] for r in rows]
    for gpc in (4, 16, 32):
        assert gpu.perft_batch(ps, ds, grid_per_cu=gpc) == want, gpc
    gpu.tt_init(512, 0)
    try:
        assert gpu.perft_batch(ps, ds, grid_per_cu=4, use_tt=True) == want
        assert gpu.perft_batch(ps, ds, use_tt=True) == want
    finally:
        gpu.tt_free()
    assert gpu.perft_batch(ps, ds) == want
    tuned = gpu.batch_tuned()
    assert tuned and all(t["grid_per_cu"] in (4, 8, 16) and t["leaf_plies"] in (1, 2) for t in tuned), tuned


def test_batch_table_budget(gpu):
    # the part of the device table a group uses: 1 MB (65,536 entries, far
    # fewer than the stores, so most are evicted), 128 MB, and the tuner's
    # choice (the whole 512 MB table or its first 128 MiB) give the golden
    # counts, and so does a plain call afterwards on the table they filled
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] in (4, 5)][:300]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    ds = [r["depth"] for r in rows]
    want = [r["count"] for r in rows]
    gpu.tt_init(512, 0)
    try:
        for mb in (1, 128, 0):
            assert gpu.perft_batch(ps, ds, use_tt=True, tt_budget_mb=mb) == want, mb
        assert gpu.perft_batch(ps, ds, use_tt=True, tt_budget_mb=1) == want  # on a filled table
        tuned = [t for t in gpu.batch_tuned() if t["tt_requested"] and t["tt_used"]]
        assert all(t["tt_span_log2"] in (23, 25) for t in tuned), tuned
        for r, p in zip(rows[:8], ps[:8]):
            assert gpu.perft(p, r["depth"], launch_depth=r["depth"], use_tt=True) == r["count"]
    finally:
        gpu.tt_free()


def test_batch_graph_replay_and_eager_agree(gpu):
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] in (3, 4)][:600]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    ds = [r["depth"] for r in rows]
    want = [r["count"] for r in rows]
    eager = gpu.perft_batch(ps, ds, leaf_plies=2, use_graph=False)
    for rep in range(3):  # replays of the captured graphs, eager work in between
        assert gpu.perft_batch(ps, ds, leaf_plies=2, use_graph=True) == want
        assert gpu.perft_bfs(ps[rep], ds[rep]) == want[rep]
    assert eager == want


def test_batch_edge_cases(gpu):
    start = gpu.parse_fen(gpu.START_FEN)
    assert gpu.perft_batch([], []) == []
    assert gpu.perft_batch([start], [0]) == [1]
    assert gpu.perft_batch([start, start, start], [1, 2, 3]) == [20, 400, 8902]
    black = gpu.parse_fen("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R b KQkq -")
    # colour-flipped roots (black to move) in the same batch as white ones
    white = gpu.parse_fen(KIWI)
    assert gpu.perft_batch([white, black], [4, 4]) == [4085603, gpu.perft_bfs(black, 4)]


def test_batch_oom_split(gpu):
    # a batch whose levels exceed a small arena is split (halves, then the
    # host split of a single root) until every piece fits; counts stay exact
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] == 5][:64]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    gpu.lib().pg_set_arena_limit(64 << 20)
    try:
        got = gpu.perft_batch(ps, [5] * len(ps), leaf_plies=1)
    finally:
        gpu.lib().pg_set_arena_limit(0)
    assert got == [r["count"] for r in rows]


def _serve(gpu, reqs, budget=8 << 30, slots=4, interarrival=0.0, merge=1):
    import ctypes
    cfg = gpu.ServeConfig(budget, 512 << 20, slots, merge)
    h = ctypes.c_void_p()
    assert gpu.lib().pg_serve_create(ctypes.byref(cfg), ctypes.byref(h)) == 0, gpu.last_error()
    out = {}
    try:
        for rid, (p, d) in reqs.items():
            assert gpu.lib().pg_serve_submit(h, rid, ctypes.byref(p), d) == 0
            if interarrival:
                time.sleep(interarrival)
        res = (gpu.ServeResult * 64)()
        n = ctypes.c_int()
        deadline = time.time() + 300
        while len(out) < len(reqs) and time.time() < deadline:
            gpu.lib().pg_serve_poll(h, res, 64, 100, ctypes.byref(n))
            for i in range(n.value):
                assert res[i].status == 0
                out[res[i].id] = res[i].count
    finally:
        gpu.lib().pg_serve_destroy(h)
    return out


def test_serving_alone_equals_in_mix(gpu):
    rows = load_jsonl("c5.jsonl")
    sample = [r for r in rows if "count" in r][:120]
    reqs = {r["seed"]: (gpu.parse_fen(r["fen"]), r["depth"]) for r in sample}
    want = {r["seed"]: r["count"] for r in sample}
    mixed = _serve(gpu, reqs, slots=4)
    assert mixed == want
    # queued requests of one depth merged into multi-root passes on the slots'
    # batched engines (all submitted at once, so the queue holds many)
    assert _serve(gpu, reqs, slots=2, merge=16) == want
    for rid in list(reqs)[:10]:  # each alone, one slot, nothing else in flight
        assert _serve(gpu, {rid: reqs[rid]}, slots=1) == {rid: want[rid]}


@pytest.mark.parametrize("leaf", [1, 2])
def test_batch_with_tt_matches_golden(gpu, leaf):
    gpu.tt_init(1024, 0)
    try:
        rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] <= 5][:800]
        ps = [gpu.parse_fen(r["fen"]) for r in rows]
        ds = [r["depth"] for r in rows]
        want = [r["count"] for r in rows]
        for rep in range(2):  # the second pass is served mostly by the table
            assert gpu.perft_batch(ps, ds, leaf_plies=leaf, use_tt=True) == want
        st = gpu.tt_stats()
        assert st[1] > 0 and st[2] > 0, st
    finally:
        gpu.tt_free()


def test_tt_snapshot_resume_device(gpu, tmp_path):
    p = gpu.parse_fen(KIWI)
    gpu.tt_init(256, 16)
    try:
        assert gpu.perft(p, 5, launch_depth=4, use_tt=True) == 193690690
        snap = str(tmp_path / "tt.bin").encode()
        assert gpu.lib().pg_tt_save(snap) == 0
    finally:
        gpu.tt_free()
    gpu.tt_init(256, 16)
    try:
        assert gpu.lib().pg_tt_load(snap) == 0
        before = gpu.tt_stats()
        assert gpu.perft(p, 5, launch_depth=4, use_tt=True) == 193690690
        after = gpu.tt_stats()
        assert after[4] > before[4]  # host tier answers from the snapshot
    finally:
        gpu.tt_free()


def test_serving_refuses_what_would_share_its_memory(gpu):
    # while a server exists its slots own the arena and the table: the plain
    # entry points, table management and a second server are refused with
    # PG_ERR_BUSY, a full queue refuses new requests, and the served results
    # stay golden
    import ctypes
    L = gpu.lib()
    rows = [r for r in load_jsonl("c5.jsonl") if "count" in r and r["depth"] >= 5][:24]
    cfg = gpu.ServeConfig(8 << 30, 512 << 20, 2, 1, 8)  # two slots, at most 8 waiting
    h = ctypes.c_void_p()
    assert L.pg_serve_create(ctypes.byref(cfg), ctypes.byref(h)) == 0, gpu.last_error()
    got, busy = {}, 0
    try:
        p = gpu.parse_fen(KIWI)
        c = ctypes.c_uint64()
        u = gpu.U128()
        arr = (gpu.Position * 1)(p)
        d = (ctypes.c_uint8 * 1)(3)
        out = (ctypes.c_uint64 * 1)()
        h2 = ctypes.c_void_p()
        refused = {
            "pg_perft_bfs": L.pg_perft_bfs(ctypes.byref(p), 3, 0, None, 0, ctypes.byref(c)),
            "pg_perft": L.pg_perft(ctypes.byref(p), 3, 0, 0, ctypes.byref(u)),
            "pg_perft_batch": L.pg_perft_batch(1, arr, d, None, out),
            "pg_perft_batch_plain": L.pg_perft_batch_plain(1, arr, d, 0, out),
            "pg_tt_init": L.pg_tt_init(64, 0),
            "pg_tt_clear": L.pg_tt_clear(),
            "pg_tt_free": L.pg_tt_free(),
            "pg_shutdown": L.pg_shutdown(),
            "pg_serve_create": L.pg_serve_create(ctypes.byref(cfg), ctypes.byref(h2)),
        }
        assert all(v == gpu.PG_ERR_BUSY for v in refused.values()), refused
        want = {r["seed"]: r["count"] for r in rows}
        pending = list(rows)
        res = (gpu.ServeResult * 64)()
        n = ctypes.c_int()
        deadline = time.time() + 120
        while len(got) < len(rows) and time.time() < deadline:
            while pending:
                r = pending[0]
                s = L.pg_serve_submit(h, r["seed"], ctypes.byref(gpu.parse_fen(r["fen"])), r["depth"])
                if s == gpu.PG_ERR_BUSY:
                    busy += 1
                    break
                assert s == 0
                pending.pop(0)
            L.pg_serve_poll(h, res, 64, 20, ctypes.byref(n))
            for i in range(n.value):
                assert res[i].status == 0
                got[res[i].id] = res[i].count
    finally:
        assert L.pg_serve_destroy(h) == 0
    assert busy > 0  # 24 requests against 2 slots and 8 places in the queue
    assert got == want
    # after the server is gone the plain path runs again
    assert gpu.perft_bfs(gpu.parse_fen(KIWI), 4) == 4085603
