"""Benchmark: perft nodes/s on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--leaf 1|2]
                  [--no-levels] [--no-batch] [--no-tt] [--no-ref-positions]
                  [--serve N] [--serve-rate R] [--copy-probe | --valu-probe | --tt-leg-only]

Workload (one step): perft(7) of the config-2 position -- the start position
after 8 seeded random plies (seed 2, SplitMix64 walk over canonically sorted
legal moves, re-drawn if terminal), no transposition table, one plain BFS
call of the whole depth on the GPU (BFS levels + leaf kernel).  13,069,312,759
leaf nodes per step; every step's count is checked against the oracle's golden
count.  Inputs live on the device for the whole timed region (the root position
is 33 bytes; every BFS level is built in HBM).

The JSON line carries:
  roofline      the dominant kernel (the fused leaf) against its VALU issue
                slots (bound "valu": a SIMD issues one VALU instruction per
                four-cycle slot, 1,024 SIMDs x 2.4 GHz / 4 = 614.4 G slots/s):
                its PMC instruction count per launch, less the second of each
                dual-issued pair, over the live HIP-event launch time on the
                library stream.  The same object carries the leaf's HBM
                figures (algorithmic bytes 6K + 33P per launch, K leaf
                entries, P parent boards; PMC traffic) and the share of slots
                the calibration kernel fills.  `roofline_levels` gives the
                HBM-bound BFS level kernels (scan, expand) at config 2's own
                level sizes (small: launch-latency bound), and
                `roofline_levels_full_size` gives scan, expand, make+count and
                genmoves on the largest level of the tree (config 2 run with
                the one-ply leaf).
  valu_issue    the leaf's VALU issue rate per SIMD and cycle: its measured
                instruction count per launch over the live launch time, and
                its lane instructions per counted position.
  cpu_baseline  the oracle (a C restatement of the reference's single-threaded
                -cpu path) on this host, one core, on a bounded sample of the
                same tree: perft(6) under the first root moves.
After the timed config-2 steps (the `value`) it runs more legs, each checked
against golden counts: the level kernels at full size (above); the
4096-position config-3 batch, batched vs plain; config 4 (64 positions,
perft 8) with the transposition tables, reporting table probes/s; the
reference README's published positions (startpos perft 9, Kiwipete perft 7,
Position 3 perft 10 without the tables, each at the launcher's automatic
launch depth, which merges repeated positions of a host-expanded frontier,
and at the memory-based launch depth, which computes every node as the
reference's -nott run does; startpos perft 9-11 with the tables) with the
published times beside them; and the config-5 serving loop, replaying N requests
(default all 10,000) with Poisson arrivals (rate R requests/s; default 0.75 x
the capacity measured by a closed-loop pass), reporting p50/p99 latency and
nodes/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNEL_TABLE_STEPS = 3  # untimed steps with every kernel timed, after the timed ones


def load_fixture(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return [json.loads(line) for line in f if line.strip()]


def cpu_baseline(fen: str, depth: int, budget_s: float = 15.0) -> dict:
    """Time the oracle single-threaded on perft(depth-1) of successive root children."""
    import oracle

    root = oracle.parse_fen(fen)
    moves = oracle.gen_moves(root)
    nodes, t0, used = 0, time.perf_counter(), 0
    for m in moves:
        nodes += oracle.perft(oracle.make_move(root, m), depth - 1, threads=1)
        used += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": nodes / dt, "unit": "nodes/s", "cores": 1, "kind": "port",
            "sample": f"oracle perft({depth - 1}) under the first {used} of {len(moves)} root moves of the "
                      f"config-2 position: {nodes} nodes in {dt:.1f} s, single-threaded like the reference -cpu path"}


def cpu_configs(pg, threads: int = 16, c3_sample: int = 128) -> dict:
    """The CPU side of configs 1 and 3, each count checked against golden:
    config 1 in full (perft 5) on one core, by the oracle and by this engine's
    own drop-in -cpu path (pg_perft_cpu, the reference's perft_cpu_dispatch);
    and the first `c3_sample` config-3 positions as independent single-threaded
    perfts on `threads` cores at once (one position per worker, the way the
    reference's -cpu path runs a batch: one process per FEN), reported as
    sum of counts / makespan."""
    import concurrent.futures as cf

    import oracle

    out = {}
    c1 = load_fixture("c1.jsonl")[0]
    want1 = c1["counts"][-1]
    po, pe = oracle.parse_fen(c1["fen"]), pg.parse_fen(c1["fen"])
    for kind, fn in (("oracle", lambda: oracle.perft(po, 5, threads=1)), ("engine_cpu", lambda: pg.perft_cpu(pe, 5))):
        fn()  # warm
        t0 = time.perf_counter()
        c = fn()
        dt = time.perf_counter() - t0
        assert c == want1, (kind, c, want1)
        out[f"config1_{kind}"] = {"value": c / dt, "unit": "nodes/s", "cores": 1, "nodes": c, "seconds": dt}
    # deepest first, so that no worker starts a depth-6 perft at the end
    rows = sorted(load_fixture("c3.jsonl")[:c3_sample], key=lambda r: -r["depth"])
    want = [r["count"] for r in rows]
    eng = [pg.parse_fen(r["fen"]) for r in rows]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:  # ctypes releases the GIL inside each call
        got = list(ex.map(lambda a: pg.perft_cpu(a[0], a[1]), zip(eng, [r["depth"] for r in rows])))
    dt_e = time.perf_counter() - t0
    assert got == want
    orc = [oracle.parse_fen(r["fen"]) for r in rows]
    t0 = time.perf_counter()
    got = oracle.perft_batch(orc, [r["depth"] for r in rows], threads)
    dt_o = time.perf_counter() - t0
    assert list(got) == want
    nodes = sum(want)
    sample = f"the first {len(rows)} config-3 positions (depth 3-6, {nodes} nodes), one single-threaded perft per worker"
    out["config3_engine_cpu"] = {"value": nodes / dt_e, "unit": "nodes/s", "cores": threads, "seconds": dt_e,
                                 "sample": sample}
    out["config3_oracle"] = {"value": nodes / dt_o, "unit": "nodes/s", "cores": threads, "seconds": dt_o,
                             "sample": sample}
    return out


# HBM bytes per launch of the default config-2 leaf, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE in KiB per dispatch) of this bench, each scaled by
# what the counter reads for a plain 256 MiB-in / 256 MiB-out elementwise
# kernel (FETCH_SIZE reads half the bytes: x2.0; WRITE_SIZE exact: x1.0);
# the writes are mostly the chunk queue's atomics.
# Commands and figures: profiles/leaf2_config2_pmc.md.
PMC_TRAFFIC = {"leaf2": 66200.6 * 1024 * 2.0 + 7650.3 * 1024 * 1.0}

# VALU work per launch of the same leaf: SQ_INSTS_VALU (wave64 instructions)
# and SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64), the share of lanes
# active per VALU instruction, from a third rocprofv3 --pmc run of this bench
# (profiles/leaf2_config2_pmc.md).  Divided by the live launch time, the
# instruction count gives the issue rate per SIMD and cycle (1,024 SIMDs at
# the 2.4 GHz peak clock).
PMC_VALU = {"leaf2": (4.2418e9, 0.914, 0.2193)}  # + the share of issue slots that issued two (VALU2 / VALU)
SIMDS, CLOCK_HZ = 1024, 2.4e9


def valu_issue(stat: dict, counts: int):
    """counts: the leaf's make + countMoves calls per launch (perft(depth - 1))."""
    if stat["name"] not in PMC_VALU:
        return None
    insts, lanes, _ = PMC_VALU[stat["name"]]
    avg_s = stat["ms"] / stat["launches"] * 1e-3
    return {"kernel": stat["name"], "wave_insts_per_launch": insts, "lanes_active": lanes,
            "wave_insts_per_simd_cycle": round(insts / (SIMDS * CLOCK_HZ * avg_s), 4),
            "lane_insts_per_count": round(insts * 64 / counts)}


PROBE_KINDS = ["v_xor_b32", "v_bitop3_b32", "v_lshlrev_b64", "v_bcnt_u32_b32", "v_lshl_add_u64", "leaf_mix",
               "v_lshlrev_b32", "v_alignbit_b32", "v_cndmask_b32", "v_add_u32", "v_and_or_b32", "v_sub_co_u32",
               "v_and_b32", "v_or_b32", "v_mov_b32", "v_bfrev_b32", "v_or3_b32", "v_ffbl_b32", "v_add3_u32",
               "v_cmp_eq_u32", "v_xor_b32_e64"]


def valu_calibration(pg, iters: int = 100_000) -> dict:
    """The VALU issue ceiling: the calibration kernel (exactly 32 VALU
    instructions per loop iteration in 8 independent chains) per instruction
    form and occupancy, as wave instructions per second over the whole chip and
    per SIMD and cycle at the 2.4 GHz peak clock.  Each launch runs once warm,
    then the best of three is kept."""
    out = {}
    for kind, name in enumerate(PROBE_KINDS):
        for waves in (2, 4, 8):
            pg.valu_probe(kind, waves, iters // 10)
            best = min(pg.valu_probe(kind, waves, iters)[0] for _ in range(3))
            insts = pg.valu_probe(kind, waves, 1)[1] * iters
            rate = insts / (best * 1e-3)
            out[f"{name}_w{waves}"] = {"wave_insts_per_s": rate, "ms": round(best, 4),
                                       "wave_insts_per_simd_cycle": round(rate / (SIMDS * CLOCK_HZ), 4)}
    return out


def calibrated_issue_ceiling(calib: dict) -> float:
    """The largest share of VALU issue cycles the calibration kernel fills
    with instructions that never issue two per cycle (64-bit shift, popcount,
    64-bit add), at eight waves per SIMD: what a saturating instruction stream
    reaches with this clock and launch shape."""
    return max(calib[f"{PROBE_KINDS[k]}_w8"]["wave_insts_per_s"] for k in (2, 3, 4)) / (SIMDS * CLOCK_HZ / 4)


def roofline_valu(stat: dict, hbm: dict, ceiling: float) -> dict:
    """The leaf against its VALU ceiling.  A SIMD issues one VALU instruction
    per four-cycle issue slot (1,024 SIMDs x 2.4 GHz / 4 = 614.4 G slots/s),
    and some 32-bit forms of two waves issue together in one slot
    (SQ_ACTIVE_INST_VALU2).  achieved = the issue slots the leaf fills per
    second: its PMC VALU instructions per launch, less the second
    instructions of dual-issued slots, over the live launch time."""
    insts, lanes, dual = PMC_VALU[stat["name"]]
    avg_s = stat["ms"] / stat["launches"] * 1e-3
    peak = SIMDS * CLOCK_HZ / 4 / 1e9
    achieved = insts * (1 - dual) / avg_s / 1e9
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": round(peak, 2), "unit": "G issue slots/s",
            "frac": round(achieved / peak, 4), "traffic": hbm["traffic"], "kernel": stat["name"],
            "avg_launch_ms": round(avg_s * 1e3, 4), "wave_insts_per_launch": insts, "dual_issue_share": dual,
            "lanes_active": lanes, "wave_insts_per_simd_cycle": round(insts / (SIMDS * CLOCK_HZ * avg_s), 4),
            "calibrated_ceiling_frac": round(ceiling, 4),
            "op": "one four-cycle VALU issue slot of one SIMD",
            "hbm": {k: hbm[k] for k in ("achieved", "peak", "unit", "frac")},
            "algorithmic_bytes_per_launch": hbm["algorithmic_bytes_per_launch"]}


def roofline(stat: dict, traffic=None) -> dict:
    avg_ms = stat["ms"] / stat["launches"]
    bytes_per_launch = stat["bytes"] / stat["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "kernel": stat["name"], "avg_launch_ms": round(avg_ms, 4),
            "algorithmic_bytes_per_launch": round(bytes_per_launch)}


def levels_leg(pg, pos, depth: int, want: int, leaf_plies: int, steps: int = 3) -> dict:
    """The BFS level kernels at full size: with the one-ply leaf the config-2
    BFS builds its perft(6) level itself (15,143,580 boards, 414,873,802
    children), so each level kernel's largest launch runs on the largest level
    this tree has.  Per kernel: the mean HIP-event time of those launches (one
    per step) and their algorithmic bytes.  scan (8N) and expand (4N + 4C) are
    HBM-bound; makecount and genmoves are VALU-bound and listed for scale."""
    pg.set_leaf_plies(1)
    try:
        assert pg.perft_bfs(pos, depth) == want
        pg.profile_reset()
        pg.profile_enable(True)
        for _ in range(steps):
            c = pg.perft_bfs(pos, depth)
            if c != want:
                raise SystemExit(f"count mismatch in the level leg: {c} != {want}")
        pg.device_synchronize()
        pg.profile_enable(False)
        stats = pg.profile_get()
    finally:
        pg.set_leaf_plies(leaf_plies)
    out = {"workload": "config-2 perft 7 with the one-ply leaf: largest level 15,143,580 boards -> 414,873,802 "
                       "children; the largest launch of each kernel, one per step",
           "steps": steps}
    for k in ("scan", "expand", "makecount", "genmoves"):
        v = stats.get(k)
        if not v or not v["big_ms"]:
            continue
        gbs = v["big_bytes"] / (v["big_ms"] * 1e-3) / 1e9
        out[k] = {"bound": "hbm" if k in ("scan", "expand") else "valu", "achieved": round(gbs, 1),
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                  "avg_launch_ms": round(v["big_ms"], 4), "launches": v["big_launches"],
                  "algorithmic_bytes_per_launch": round(v["big_bytes"])}
    return out


def batch_leg(pg) -> dict:
    rows = load_fixture("c3.jsonl")
    ps = [pg.parse_fen(r["fen"]) for r in rows]
    ds = [r["depth"] for r in rows]
    want = [r["count"] for r in rows]
    total = sum(want)
    t0 = time.perf_counter()
    got = pg.perft_batch(ps, ds)  # cold: autotune + graph capture + first launches
    t_cold = time.perf_counter() - t0
    assert got == want
    t0 = time.perf_counter()
    got = pg.perft_batch(ps, ds)
    t_b = time.perf_counter() - t0
    assert got == want
    tuned = pg.batch_tuned()
    # the same batch with the grid held at 4, 8 and 16 blocks per CU (leaf
    # plies and table use as tuned), warm, for the tuner's grid choice
    sweep = {}
    for gpc in (4, 8, 16):
        assert pg.perft_batch(ps, ds, grid_per_cu=gpc) == want
        t0 = time.perf_counter()
        assert pg.perft_batch(ps, ds, grid_per_cu=gpc) == want
        sweep[gpc] = time.perf_counter() - t0
    pg.perft_batch_plain(ps[:64], ds[:64])
    t0 = time.perf_counter()
    got = pg.perft_batch_plain(ps, ds)
    t_plain = time.perf_counter() - t0
    assert got == want
    # beside the -nott workload: the same batch with a 64 GB device table,
    # emptied before the timed pass so that it finds only what the pass stores
    pg.tt_init(65536, 0)
    try:
        assert pg.perft_batch(ps, ds, use_tt=True) == want  # tunes the table groups
        pg.lib().pg_tt_clear()
        t0 = time.perf_counter()
        assert pg.perft_batch(ps, ds, use_tt=True) == want
        t_tt = time.perf_counter() - t0
        tuned_tt = [t for t in pg.batch_tuned() if t["tt_requested"]]
    finally:
        pg.tt_free()
    return {"workload": "config-3: 4096 positions, 6-40 random plies, depth 3-6, no TT", "nodes": total,
            "batched_s": t_b, "batched_nodes_per_s": total / t_b, "batched_cold_s": t_cold, "plain_s": t_plain,
            "plain_nodes_per_s": total / t_plain, "batched_over_plain": t_plain / t_b,
            "batched_cold_over_plain": t_plain / t_cold, "tuned": tuned, "grid_sweep_s": sweep,
            "with_table": {"device_table_gb": 64, "batched_s": t_tt, "batched_nodes_per_s": total / t_tt,
                           "tuned": tuned_tt}}


HOST_TT_MB = 262144  # config 4's 256 GB host tier (its untouched pages cost no memory)


def tt_leg(pg) -> dict:
    """Config 4: 64 positions (10 random plies), perft 8 with the transposition
    tables (64 GB device table + 256 GB host tier), through the launcher (host
    recursion above the launch depth, frontier batched on the device)."""
    rows = load_fixture("c4.jsonl")
    ps = [pg.parse_fen(r["fen"]) for r in rows]
    pg.tt_init(65536, HOST_TT_MB)
    # exactness: depth-5 golden counts (oracle), then TT on == TT off at depth 7
    for r, p in zip(rows, ps):
        assert pg.perft(p, r["golden_depth"], use_tt=True) == r["golden_count"]
    d7_tt = [pg.perft(p, 7, use_tt=True) for p in ps]
    d7 = [pg.perft(p, 7) for p in ps]
    assert d7_tt == d7
    pg.lib().pg_tt_clear()
    stats0 = pg.tt_stats()
    t0 = time.perf_counter()
    counts = [pg.perft(p, 8, use_tt=True) for p in ps]
    dt = time.perf_counter() - t0
    stats = [b - a for a, b in zip(stats0, pg.tt_stats())]
    assert stats[3] > 0, "the host tier took no probe: it was not allocated"
    # TT off on the first 4 positions at full depth as a spot check
    spot = [pg.perft(p, 8) for p in ps[:4]]
    assert spot == counts[:4], (spot, counts[:4])
    pg.tt_free()
    nodes = sum(counts)
    return {"workload": "config-4: 64 positions (10 plies, seeds 100-163), perft 8, device TT 64 GB + host TT 256 GB, "
                        "launcher (auto launch depth), one position after another, tables shared",
            "nodes": nodes, "seconds": dt, "nodes_per_s": nodes / dt,
            "device_tt_probes": stats[0], "device_tt_hits": stats[1], "device_tt_stores": stats[2],
            "host_tt_probes": stats[3], "host_tt_hits": stats[4], "host_tt_stores": stats[5],
            "tt_probes_per_s": (stats[0] + stats[3]) / dt,
            "checked": "depth-5 oracle counts; TT on == TT off at depth 7 (all 64) and depth 8 (first 4)"}


def tt_leg_timed(pg) -> dict:
    """Config 4's timed leg alone, as tt_leg runs it (for counter runs): the
    tables cleared, then perft 8 of the 64 positions with them, and the table
    tallies of the leg."""
    ps = [pg.parse_fen(r["fen"]) for r in load_fixture("c4.jsonl")]
    pg.tt_init(65536, HOST_TT_MB)
    pg.lib().pg_tt_clear()
    stats0 = pg.tt_stats()
    t0 = time.perf_counter()
    nodes = sum(pg.perft(p, 8, use_tt=True) for p in ps)
    dt = time.perf_counter() - t0
    stats = [b - a for a, b in zip(stats0, pg.tt_stats())]
    pg.tt_free()
    return {"workload": "config-4 timed leg", "nodes": nodes, "seconds": dt, "device_tt_probes": stats[0],
            "device_tt_hits": stats[1], "device_tt_stores": stats[2], "host_tt_probes": stats[3],
            "host_tt_hits": stats[4], "host_tt_stores": stats[5]}


# The reference's own published single-position runs, for context:
# README.md:51-52 and :58-60 (RTX PRO 6000 Blackwell) and :74 (RTX 4090).
REF_RUNS = [("startpos", 9, "1.867 s on RTX PRO 6000"), ("kiwipete", 7, "0.177 s on RTX PRO 6000"),
            ("position3", 10, "1.55 s on RTX 4090")]
REF_RUNS_TT = [("startpos", 9, "0.174 s on RTX PRO 6000"), ("startpos", 10, "2.040 s on RTX PRO 6000"),
               ("startpos", 11, "24.33 s on RTX PRO 6000")]


def ref_positions_leg(pg, tt_max_depth: int, tt_only: bool = False) -> dict:
    """The reference README's published positions through the launcher, each
    count checked against the published value: without the tables, then
    startpos with them (64 GB device table, 256 GB host tier, cleared before
    each run).

    Without the tables the launcher's default expands up to four plies above
    its launch depth on the host, merges identical positions (counted once,
    times their multiplicity) and runs the rest through the batched engine:
    `seconds` is that (second call; `seconds_first_call` includes the batched
    engine's graph captures and tuning trials).  `every_node_seconds` runs the
    launcher at the memory-based launch depth instead, where the frontier is
    at most one ply and no position repeats, so every node of the tree is
    computed as the reference's -nott run computes it."""
    import ctypes
    known = {p["name"]: p for p in json.load(open(os.path.join(ROOT, "tests", "golden", "known_positions.json")))["positions"]}
    out = {}
    for name, d, theirs in ([] if tt_only else REF_RUNS):
        k = known[name]
        want = k["counts"][d - 1]
        p = pg.parse_fen(k["fen"])
        pg.perft(p, d - 2)  # warm (launch depth estimate, first launches)
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            c = pg.perft(p, d)
            times.append(time.perf_counter() - t0)
            assert c == want, (name, d, c, want)
        bf, ld = ctypes.c_float(), ctypes.c_int()
        pg.lib().pg_estimate_launch_depth(ctypes.byref(p), ctypes.byref(bf), ctypes.byref(ld))
        pg.perft(p, d - 2, launch_depth=ld.value)
        t0 = time.perf_counter()
        c = pg.perft(p, d, launch_depth=ld.value)
        dt_all = time.perf_counter() - t0
        assert c == want, (name, d, c, want)
        dt = times[1]
        out[f"{name}_perft{d}"] = {"nodes": c, "seconds": dt, "nodes_per_s": c / dt, "seconds_first_call": times[0],
                                   "every_node_seconds": dt_all, "every_node_launch_depth": ld.value,
                                   "reference_published": theirs}
    pg.tt_init(65536, HOST_TT_MB)
    try:
        for name, d, theirs in REF_RUNS_TT:
            if d > tt_max_depth:
                continue
            k = known[name]
            want = k["counts"][d - 1]
            p = pg.parse_fen(k["fen"])
            pg.lib().pg_tt_clear()
            print(f"[bench] {name} perft {d} with the tables", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            c = pg.perft(p, d, use_tt=True)
            dt = time.perf_counter() - t0
            assert c == want, (name, d, c, want)
            out[f"{name}_perft{d}_tt"] = {"nodes": c, "seconds": dt, "nodes_per_s": c / dt,
                                          "reference_published": theirs + ", TT"}
    finally:
        pg.tt_free()
    return out


def serve_leg(pg, n_req: int, rate: float, tt_gb: int, slots: int, merge: int = 1) -> dict:
    import ctypes
    import random
    import threading

    rows = load_fixture("c5.jsonl")[:n_req]
    reqs = [(r["seed"], pg.parse_fen(r["fen"]), r["depth"]) for r in rows]
    gold = {r["seed"]: r["count"] for r in rows if "count" in r}
    L = pg.lib()

    def run(arrivals):
        cfg = pg.ServeConfig(200 << 30, tt_gb << 30, slots, merge)
        h = ctypes.c_void_p()
        if L.pg_serve_create(ctypes.byref(cfg), ctypes.byref(h)) != 0:
            raise RuntimeError(pg.last_error())
        done = {}
        stop = threading.Event()

        def poller():
            res = (pg.ServeResult * 256)()
            n = ctypes.c_int()
            while not stop.is_set() or len(done) < len(reqs):
                L.pg_serve_poll(h, res, 256, 50, ctypes.byref(n))
                for i in range(n.value):
                    r = res[i]
                    done[r.id] = (r.count, r.submitTime, r.startTime, r.endTime, r.status)
                    if len(done) % 50 == 0:
                        print(f"[bench] serving: {len(done)}/{len(reqs)} done", file=sys.stderr, flush=True)
                if stop.is_set() and len(done) >= len(reqs):
                    break

        th = threading.Thread(target=poller)
        th.start()
        t0 = L.pg_clock()
        for (rid, p, d), at in zip(reqs, arrivals):
            delay = t0 + at - L.pg_clock()
            if delay > 0:
                time.sleep(delay)
            L.pg_serve_submit(h, rid, ctypes.byref(p), d)
        stop.set()
        th.join()
        L.pg_serve_destroy(h)
        return done

    # closed loop (everything at t=0) measures capacity
    done = run([0.0] * len(reqs))
    first = min(v[1] for v in done.values())
    last = max(v[3] for v in done.values())
    capacity = len(reqs) / (last - first)
    print(f"[bench] serving closed loop: {len(reqs)} requests in {last - first:.2f} s", file=sys.stderr, flush=True)
    lam = rate or 0.75 * capacity
    rnd = random.Random(5)
    t, arrivals = 0.0, []
    for _ in reqs:
        t += rnd.expovariate(lam)
        arrivals.append(t)
    done = run(arrivals)
    lat = sorted(v[3] - v[1] for v in done.values())
    nodes = sum(v[0] for v in done.values())
    span = max(v[3] for v in done.values()) - min(v[1] for v in done.values())
    # every served count is checked: against the oracle's golden count where
    # the fixture has one (depth <= 5 and 64 of depth 6), the rest against the
    # batched engine without the table, after the server is gone
    rest = [(rid, p, d) for rid, p, d in reqs if rid not in gold]
    print(f"[bench] serving: checking {len(rest)} results against the batched engine without the table",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    plain = dict(zip([r[0] for r in rest], pg.perft_batch([r[1] for r in rest], [r[2] for r in rest])))
    t_check = time.perf_counter() - t0
    want = {**gold, **plain}
    bad = [k for k, v in done.items() if v[4] != 0 or v[0] != want[k]]
    merging = (f"up to {merge} queued requests of one depth <= 5 merged per pass" if merge > 1
               else "one request per pass")
    return {"workload": f"config-5: first {len(reqs)} requests (0-60 plies, depth 3-7), TT shared "
                        f"({tt_gb} GB), 200 GB budget, {slots} slots, {merging}, Poisson arrivals",
            "rate_req_per_s": lam, "capacity_req_per_s": capacity, "p50_ms": 1e3 * lat[len(lat) // 2],
            "p99_ms": 1e3 * lat[min(len(lat) - 1, int(0.99 * len(lat)))], "nodes_per_s": nodes / span,
            "golden_checked": sum(1 for k in done if k in gold), "checked_without_table": len(plain),
            "check_s": t_check, "checked": len(done), "mismatches": len(bad) + (len(reqs) - len(done))}


def copy_probe(pg) -> None:
    """Calibration kernel of known size for HBM counters and the achievable
    copy rate: three device-to-device copies of 256 MiB (the runtime's copy
    kernel reads 256 MiB and writes 256 MiB each)."""
    n = 256 << 20
    a, b = pg.DeviceArray(n), pg.DeviceArray(n)
    a.zero()
    for _ in range(3):
        rc = pg.lib().pg_memcpy(b.ptr, a.ptr, n, 2)
        if rc:
            raise RuntimeError(pg.last_error())
    pg.device_synchronize()
    print(f"copy probe: 3 device copies of {n} bytes (read {n} + write {n} each)", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--leaf", type=int, default=0)
    ap.add_argument("--leaf2-variant", type=int, default=0)
    ap.add_argument("--leaf-grid", type=int, default=0, help="leaf blocks per CU (0: library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-levels", action="store_true", help="skip the full-size level-kernel leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the config-3 batched-vs-plain leg")
    ap.add_argument("--no-tt", action="store_true", help="skip the config-4 transposition-table leg")
    ap.add_argument("--no-ref-positions", action="store_true", help="skip the reference README positions")
    ap.add_argument("--ref-tt-only", action="store_true",
                    help="reference-positions leg: only the startpos runs with the tables")
    ap.add_argument("--ref-tt-depth", type=int, default=11,
                    help="deepest startpos perft run with the tables in the reference-positions leg (8 = none)")
    ap.add_argument("--serve", type=int, default=10000, help="config-5 requests to replay (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=0.0)
    ap.add_argument("--serve-slots", type=int, default=8, help="requests in flight in the serving leg")
    ap.add_argument("--serve-batch", type=int, default=1,
                    help="queued requests of one depth (at most 5) a free slot merges into one multi-root pass")
    ap.add_argument("--serve-tt-gb", type=int, default=16,
                    help="shared transposition table of the serving leg, inside its 200 GB budget")
    ap.add_argument("--copy-probe", action="store_true", help="only run the HBM calibration copies and exit")
    ap.add_argument("--valu-probe", action="store_true", help="only run the VALU calibration kernel and exit")
    ap.add_argument("--tt-leg-only", action="store_true",
                    help="only run config 4's timed leg once (tables cleared first; for counter runs) and exit")
    ap.add_argument("--tt-frontier", type=int, default=-1,
                    help="launcher with the tables: 0 serial recursion, 1 parallel frontier, 2 batched frontier "
                         "(-1: the library default)")
    ap.add_argument("--slider", type=int, default=0,
                    help="slider attacks of the fused leaf: 0 fills (default), 1 magic tables, 2 magic with LDS bishops")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the engine runs on one GPU and has no multi-GPU mode "
              "(one would split a position's tree at the root across GPUs, with one 8-byte sum per GPU).",
              file=sys.stderr)
        sys.exit(2)

    import perft_gpu_2026_amd as pg

    if a.copy_probe:
        pg.init(arena_bytes=1 << 30)
        copy_probe(pg)
        return
    if a.valu_probe:
        pg.init(arena_bytes=1 << 30)
        print(json.dumps(valu_calibration(pg)), flush=True)
        return
    if a.tt_leg_only:
        pg.init(0)
        print(json.dumps(tt_leg_timed(pg)), flush=True)
        return
    fx = load_fixture("c2.jsonl")[0]
    depth, want = fx["depth"], fx["counts"][-1]
    print(f"[bench] init; config-2 position {fx['fen']} depth {depth}", file=sys.stderr, flush=True)
    pg.init(0)
    if a.leaf:
        pg.set_leaf_plies(a.leaf)
    pg.lib().pg_set_leaf2_variant(a.leaf2_variant)
    if a.tt_frontier >= 0 and pg.lib().pg_set_tt_frontier(a.tt_frontier) != 0:
        raise SystemExit(f"pg_set_tt_frontier({a.tt_frontier}): {pg.last_error()}")
    if a.slider and pg.lib().pg_set_slider_attacks(a.slider) != 0:
        raise SystemExit(f"pg_set_slider_attacks({a.slider}): {pg.last_error()}")
    pmc_known = a.leaf2_variant == 0 and a.slider == 0  # the PMC figures are of the default leaf
    if a.leaf_grid:
        pg.lib().pg_set_leaf_grid(a.leaf_grid)
    pos = pg.parse_fen(fx["fen"])
    for i in range(a.warmup):
        c = pg.perft_bfs(pos, depth)
        assert c == want, (c, want)
        print(f"[bench] warmup {i} ok", file=sys.stderr, flush=True)
    # the timed steps time only the leaf (the dominant kernel) with HIP
    # events: every timed launch adds two stream events, about 2 % of a step
    # when all of a step's 21 launches carry them
    pg.profile_reset()
    pg.profile_kernels(["leaf1", "leaf2"])
    pg.profile_enable(True)
    pg.device_synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        c = pg.perft_bfs(pos, depth)
        if c != want:
            raise SystemExit(f"count mismatch at step {i}: {c} != {want}")
    pg.device_synchronize()
    t1 = time.perf_counter()
    pg.profile_enable(False)
    leaf_stats = pg.profile_get()
    # then every kernel, on untimed steps, for the per-kernel table and the
    # level kernels' roofline at this workload's sizes
    pg.profile_reset()
    pg.profile_kernels(None)
    pg.profile_enable(True)
    for i in range(KERNEL_TABLE_STEPS):
        assert pg.perft_bfs(pos, depth) == want
    pg.profile_enable(False)
    stats = pg.profile_get()
    for k in ("leaf1", "leaf2"):
        stats.pop(k, None)
    stats.update(leaf_stats)
    for k, v in stats.items():
        v["name"] = k
    dt = t1 - t0
    value = want * a.steps / dt
    leaf = stats.get("leaf2") or stats.get("leaf1")
    # the HBM-bound level kernels (make+count and genmoves are VALU-bound)
    levels = [v for k, v in stats.items() if k in ("scan", "expand")]
    top_level = max(levels, key=lambda v: v["ms"]) if levels else None
    print("[bench] VALU calibration", file=sys.stderr, flush=True)
    calib = valu_calibration(pg)
    leaf_roofline = None
    if leaf:
        hbm = roofline(leaf, PMC_TRAFFIC.get(leaf["name"]) if pmc_known else None)
        if leaf["name"] in PMC_VALU and pmc_known:
            leaf_roofline = roofline_valu(leaf, hbm, calibrated_issue_ceiling(calib))
        else:
            leaf_roofline = dict(hbm, note="no PMC instruction count for this leaf form: HBM bytes only")
    out = {
        "metric": "perft nodes/s on 1 MI355X, batched vs plain BFS; p50/p99 query latency",
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "config-2: startpos + 8 random plies (seed 2), perft 7, no TT, one plain BFS call",
                   "fen": fx["fen"], "depth": depth, "nodes_per_step": want, "leaf_plies": a.leaf or 2,
                   "slider_attacks": ["fills", "magic", "magic_lds_bishops"][a.slider]},
        "roofline": leaf_roofline,
        "roofline_levels": roofline(top_level) if top_level else None,
        "valu_issue": valu_issue(leaf, fx["counts"][depth - 2]) if leaf and pmc_known else None,
        "valu_calibration": calib,
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                        "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] else None}
                    for k, v in stats.items()},
    }
    if not a.no_levels:
        print("[bench] level kernels at full size", file=sys.stderr, flush=True)
        out["roofline_levels_full_size"] = levels_leg(pg, pos, depth, want, a.leaf or 2)
    if not a.no_batch:
        print("[bench] config-3 batch", file=sys.stderr, flush=True)
        out["batch_config3"] = batch_leg(pg)
    if not a.no_tt:
        print("[bench] config-4 transposition tables", file=sys.stderr, flush=True)
        out["tt_config4"] = tt_leg(pg)
    if not a.no_ref_positions:
        print("[bench] reference README positions", file=sys.stderr, flush=True)
        out["reference_positions"] = ref_positions_leg(pg, a.ref_tt_depth, a.ref_tt_only)
    if a.serve:
        print("[bench] config-5 serving", file=sys.stderr, flush=True)
        out["serve_config5"] = serve_leg(pg, a.serve, a.serve_rate, a.serve_tt_gb, a.serve_slots, a.serve_batch)
    if not a.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(fx["fen"], depth)
        out["cpu_baseline"].update(cpu_configs(pg, threads=min(16, os.cpu_count() or 1)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
