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
        gpu.set_leaf_plies(1)


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


def test_forced_oom_reports_and_falls_back(gpu):
    kiwi = gpu.parse_fen("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -")
    import torch
    buf = torch.empty(64 * 1024, dtype=torch.uint8, device="cuda")
    s, _ = gpu.perft_bfs_status(kiwi, 5, buf.data_ptr(), buf.numel())
    assert s == gpu.PG_ERR_ARENA_OOM
    before = gpu.lib().pg_get_oom_fallback_count()
    gpu.lib().pg_set_arena_limit(8 << 20)
    try:
        assert gpu.perft(kiwi, 5, launch_depth=5) == 193690690
    finally:
        gpu.lib().pg_set_arena_limit(0)
    assert gpu.lib().pg_get_oom_fallback_count() > before


def test_config3_sample_plain_batch(gpu):
    rows = [r for r in load_jsonl("c3.jsonl") if r["depth"] <= 5][:300]
    ps = [gpu.parse_fen(r["fen"]) for r in rows]
    got = gpu.perft_batch_plain(ps, [r["depth"] for r in rows])
    assert got == [r["count"] for r in rows]


# ---------------------------------------------------------------- kernel ABI
def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("n", [1, 255, 256, 257, 100_000, 3_000_001])
def test_scan_matches_cumsum(gpu, n):
    import torch
    rng = np.random.default_rng(n)
    host = rng.integers(0, 219, size=n, dtype=np.uint32)
    d = torch.from_numpy(host.astype(np.int32)).cuda()
    tiles = (n + 255) // 256
    scratch = torch.zeros(8 * (tiles + 1) + 4 * tiles + 256, dtype=torch.uint8, device="cuda")
    total = ctypes.c_uint64()
    assert gpu.lib().pg_scan_u32(_p(d), n, None, _p(scratch), ctypes.byref(total), _stream()) == 0
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint32), np.cumsum(host, dtype=np.uint64).astype(np.uint32))
    assert total.value == int(host.sum(dtype=np.uint64))


@pytest.mark.parametrize("n,zero_frac", [(1, 0.0), (1000, 0.0), (50_000, 0.3), (20_000, 0.97)])
def test_expand_matches_searchsorted(gpu, n, zero_frac):
    import torch
    rng = np.random.default_rng(n)
    counts = rng.integers(1, 60, size=n).astype(np.uint32)
    counts[rng.random(n) < zero_frac] = 0
    counts[-1] = max(counts[-1], 1)
    incl = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    c = int(incl[-1])
    d_incl = torch.from_numpy(incl.view(np.int32)).cuda()
    out = torch.empty(c, dtype=torch.int32, device="cuda")
    scratch = torch.empty(4 * (c // 2048 + 2), dtype=torch.uint8, device="cuda")
    assert gpu.lib().pg_expand(_p(d_incl), n, c, _p(out), _p(scratch), _stream()) == 0
    torch.cuda.synchronize()
    want = np.searchsorted(incl, np.arange(c), side="right")
    assert np.array_equal(out.cpu().numpy(), want)


def _level_inputs(gpu, orc, nroots):
    roots = [gpu.random_position(s, 5 + s % 50) for s in range(nroots)]
    boards = np.array([list(p.bb) for p in roots], dtype=np.uint64)
    states = np.array([p.state for p in roots], dtype=np.uint8)
    colors = [p.color for p in roots]
    return roots, boards, states, colors


def test_makecount_and_genmoves_match_host(gpu, orc):
    import torch
    # one colour per call: white-to-move parents
    roots = [p for p in (gpu.random_position(s, 2 * (s % 30)) for s in range(400)) if p.color == 0]
    pidx, moves, children = [], [], []
    for i, p in enumerate(roots):
        for m in gpu.gen_moves(p):
            pidx.append(i)
            moves.append(m)
            children.append(gpu.make_move(p, m))
    n = len(moves)
    pb = torch.from_numpy(np.array([list(p.bb) for p in roots], dtype=np.uint64).view(np.int64)).cuda()
    ps = torch.from_numpy(np.array([p.state for p in roots], dtype=np.uint8)).cuda()
    di = torch.from_numpy(np.array(pidx, dtype=np.int32)).cuda()
    dm = torch.from_numpy(np.array(moves, dtype=np.uint16).view(np.int16)).cuda()
    ob = torch.empty(n * 4, dtype=torch.int64, device="cuda")
    os_ = torch.empty(n, dtype=torch.uint8, device="cuda")
    cnt = torch.empty(n, dtype=torch.int32, device="cuda")
    tiles = torch.empty((n + 255) // 256, dtype=torch.int32, device="cuda")
    L = gpu.lib()
    assert L.pg_level_makecount(_p(pb), _p(ps), _p(di), _p(dm), 0, n, _p(ob), _p(os_), _p(cnt), _p(tiles),
                                _stream()) == 0
    torch.cuda.synchronize()
    got_b = ob.cpu().numpy().view(np.uint64).reshape(n, 4)
    got_c = cnt.cpu().numpy()
    for k, ch in enumerate(children):
        assert tuple(int(x) for x in got_b[k]) == tuple(ch.bb)
        assert got_c[k] == gpu.count_moves(ch)
    assert int(tiles.cpu().numpy().astype(np.int64).sum()) == int(got_c.sum())
    # genmoves over the children (black to move)
    incl = torch.from_numpy(np.cumsum(got_c).astype(np.int32)).cuda()
    c = int(got_c.sum())
    out = torch.empty(c, dtype=torch.int16, device="cuda")
    assert L.pg_genmoves(_p(ob), _p(os_), _p(incl), 1, n, _p(out), _stream()) == 0
    torch.cuda.synchronize()
    got_m = out.cpu().numpy().view(np.uint16)
    off = 0
    for k, ch in enumerate(children):
        seg = sorted(int(x) for x in got_m[off:off + got_c[k]])
        assert seg == sorted(gpu.gen_moves(ch))
        off += got_c[k]


def test_leaf_entry_point(gpu):
    import torch
    p = gpu.parse_fen("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -")
    mv = gpu.gen_moves(p)
    pb = torch.from_numpy(np.array([list(p.bb)], dtype=np.uint64).view(np.int64)).cuda()
    ps = torch.tensor([p.state], dtype=torch.uint8, device="cuda")
    di = torch.zeros(len(mv), dtype=torch.int32, device="cuda")
    dm = torch.from_numpy(np.array(mv, dtype=np.uint16).view(np.int16)).cuda()
    for plies, want in ((1, 2039), (2, 97862)):
        s = torch.zeros(1, dtype=torch.int64, device="cuda")
        assert gpu.lib().pg_leaf(_p(pb), _p(ps), _p(di), _p(dm), 0, len(mv), plies, _p(s), _stream()) == 0
        torch.cuda.synchronize()
        assert int(s.item()) == want


def test_cli_gpu(gpu):
    r = gpu.run_cli(["r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -", "5", "-nott",
                     "-arena", "4096"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert gpu.parse_cli_counts(r.stdout) == {1: 48, 2: 2039, 3: 97862, 4: 4085603, 5: 193690690}
    assert "Launch depth:" in r.stdout
