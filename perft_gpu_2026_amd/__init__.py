"""perft_gpu_2026_amd -- MI355X-native chess perft engine.

The engine is native: C++ host code and hand-written HIP kernels for gfx950 in
``csrc/``, built into ``libperft_gpu.so`` (C ABI declared in
``include/perft_gpu.h``) and the ``bin/perft_gpu`` command line, a drop-in for
the reference's ``perft_gpu <fen> <depth> [-nott] [-dtt] [-htt] [-ld] [-cpu]``
(perft.cu:13-186).

This module is a thin ctypes mirror of that C ABI for Python callers (tests,
bench).  It never computes a count itself: every call goes to the library, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_uint8, c_uint16, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PERFT_GPU_LIB") or os.path.join(_HERE, "libperft_gpu.so")
CLI_PATH = os.path.join(_HERE, "bin", "perft_gpu")

PG_OK = 0
PG_ERR_ARENA_OOM = 1
PG_ERR_INVALID = 2
PG_ERR_HIP = 3
PG_ERR_NOT_INIT = 4
PG_ERR_NO_DEVICE = 5
PG_ERR_BUSY = 6

START_FEN = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1"


class Position(ctypes.Structure):
    """pg_position: quad-bitboard (chess.h:66-76) + state byte + side to move."""

    _fields_ = [("bb", c_uint64 * 4), ("state", c_uint8), ("color", c_uint8), ("_pad", c_uint8 * 6)]

    def key(self):
        return (tuple(self.bb), self.state, self.color)


class U128(ctypes.Structure):
    _fields_ = [("lo", c_uint64), ("hi", c_uint64)]

    def value(self) -> int:
        return (self.hi << 64) | self.lo


class BatchOptions(ctypes.Structure):
    _fields_ = [("gridPerCU", c_int), ("leafPlies", c_int), ("useGraph", c_int), ("useTT", c_int),
                ("ttBudgetMB", c_int), ("reserved", c_int * 3)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbmBudgetBytes", c_uint64), ("ttBytes", c_uint64), ("maxInflight", c_int), ("maxBatch", c_int),
                ("maxQueued", c_int), ("reserved", c_int * 3)]


class ServeResult(ctypes.Structure):
    _fields_ = [("id", c_uint64), ("count", c_uint64), ("submitTime", c_double), ("startTime", c_double),
                ("endTime", c_double), ("status", c_int), ("_pad", c_int)]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", c_uint64), ("ms", c_double), ("bytes", c_double),
                ("items", c_double), ("big_bytes", c_double), ("big_ms", c_double), ("big_launches", c_uint64)]


class PerftError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: status {status}: {last_error()}")
        self.status = status


_lib = None

# (name, restype, argtypes) of every C-ABI entry point
_SIGS = [
    ("pg_init", c_int, [c_int, c_uint64]),
    ("pg_shutdown", c_int, []),
    ("pg_last_error", c_char_p, []),
    ("pg_arena_bytes", c_uint64, []),
    ("pg_set_arena_limit", None, [c_uint64]),
    ("pg_set_frontier_batch", None, [c_int]),
    ("pg_set_tt_frontier", c_int, [c_int]),
    ("pg_parse_fen", c_int, [c_char_p, POINTER(Position)]),
    ("pg_write_fen", c_int, [POINTER(Position), c_char_p, c_int]),
    ("pg_gen_moves", c_int, [POINTER(Position), POINTER(c_uint16), POINTER(c_int)]),
    ("pg_count_moves", c_int, [POINTER(Position), POINTER(c_int)]),
    ("pg_make_move", c_int, [POINTER(Position), c_uint16]),
    ("pg_random_position", c_int, [c_uint64, c_int, c_int, POINTER(Position)]),
    ("pg_random_request", c_int, [c_uint64, c_int, c_int, c_int, c_int, POINTER(Position), POINTER(c_int),
                                  POINTER(c_int)]),
    ("pg_perft_bfs", c_int, [POINTER(Position), c_int, c_int, c_void_p, c_uint64, POINTER(c_uint64)]),
    ("pg_perft", c_int, [POINTER(Position), c_int, c_int, c_int, POINTER(U128)]),
    ("pg_perft_cpu", c_int, [POINTER(Position), c_int, c_int, POINTER(c_uint64)]),
    ("pg_estimate_launch_depth", c_int, [POINTER(Position), POINTER(c_float), POINTER(c_int)]),
    ("pg_get_effective_ld", c_int, []),
    ("pg_set_effective_ld", None, [c_int]),
    ("pg_get_oom_fallback_count", c_uint64, []),
    ("pg_last_bfs_peak_bytes", c_uint64, []),
    ("pg_set_leaf_plies", None, [c_int]),
    ("pg_set_leaf2_variant", None, [c_int]),
    ("pg_set_leaf_grid", None, [c_int]),
    ("pg_set_slider_attacks", c_int, [c_int]),
    ("pg_magic_check", c_int, [c_uint64, POINTER(c_uint64)]),
    ("pg_u128_to_string", c_int, [c_uint64, c_uint64, c_char_p, c_int]),
    ("pg_tt_init", c_int, [c_uint64, c_uint64]),
    ("pg_tt_clear", c_int, []),
    ("pg_tt_free", c_int, []),
    ("pg_tt_stats", c_int, [POINTER(c_uint64)]),
    ("pg_tt_save", c_int, [c_char_p]),
    ("pg_tt_load", c_int, [c_char_p]),
    ("pg_level_makecount", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_uint64, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    ("pg_scan_u32", c_int, [c_void_p, c_uint64, c_void_p, c_void_p, POINTER(c_uint64), c_void_p]),
    ("pg_scan_u32_static", c_int, [c_void_p, c_uint64, c_void_p, POINTER(c_uint64), c_void_p]),
    ("pg_expand", c_int, [c_void_p, c_uint64, c_uint64, c_void_p, c_void_p, c_void_p]),
    ("pg_genmoves", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_uint64, c_void_p, c_void_p]),
    ("pg_leaf", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_uint64, c_int, c_void_p, c_void_p]),
    ("pg_perft_batch", c_int, [c_int, POINTER(Position), POINTER(c_uint8), POINTER(BatchOptions), POINTER(c_uint64)]),
    ("pg_perft_batch_plain", c_int, [c_int, POINTER(Position), POINTER(c_uint8), c_int, POINTER(c_uint64)]),
    ("pg_batch_tuned", c_int, [POINTER(c_int), c_int, POINTER(c_int)]),
    ("pg_serve_create", c_int, [POINTER(ServeConfig), POINTER(c_void_p)]),
    ("pg_serve_submit", c_int, [c_void_p, c_uint64, POINTER(Position), c_int]),
    ("pg_serve_poll", c_int, [c_void_p, POINTER(ServeResult), c_int, c_int, POINTER(c_int)]),
    ("pg_serve_destroy", c_int, [c_void_p]),
    ("pg_clock", c_double, []),
    ("pg_valu_probe", c_int, [c_int, c_int, c_uint64, POINTER(c_double), POINTER(c_uint64)]),
    ("pg_profile_enable", c_int, [c_int]),
    ("pg_profile_kernels", c_int, [c_uint32]),
    ("pg_profile_reset", c_int, []),
    ("pg_profile_get", c_int, [POINTER(KernelStat), c_int, POINTER(c_int)]),
    ("pg_device_synchronize", c_int, []),
    ("pg_device_malloc", c_int, [c_uint64, POINTER(c_void_p)]),
    ("pg_device_free", c_int, [c_void_p]),
    ("pg_memcpy", c_int, [c_void_p, c_void_p, c_uint64, c_int]),
]


def lib() -> ctypes.CDLL:
    """Load libperft_gpu.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               "g.build()' or make -C perft_gpu_2026_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    try:
        return (lib().pg_last_error() or b"").decode()
    except Exception:  # noqa: BLE001 - error text is best effort
        return ""


def _check(status: int, what: str) -> None:
    if status != PG_OK:
        raise PerftError(status, what)


# ------------------------------------------------------------------ host side
def parse_fen(fen: str) -> Position:
    p = Position()
    _check(lib().pg_parse_fen(fen.encode(), byref(p)), "pg_parse_fen")
    return p


def write_fen(p: Position) -> str:
    buf = ctypes.create_string_buffer(128)
    _check(lib().pg_write_fen(byref(p), buf, 128), "pg_write_fen")
    return buf.value.decode()


def gen_moves(p: Position) -> list[int]:
    mv = (c_uint16 * 256)()
    n = c_int()
    _check(lib().pg_gen_moves(byref(p), mv, byref(n)), "pg_gen_moves")
    return list(mv[: n.value])


def count_moves(p: Position) -> int:
    n = c_int()
    _check(lib().pg_count_moves(byref(p), byref(n)), "pg_count_moves")
    return n.value


def make_move(p: Position, move: int) -> Position:
    c = Position.from_buffer_copy(p)
    _check(lib().pg_make_move(byref(c), move), "pg_make_move")
    return c


def random_position(seed: int, plies: int, require_nonterminal: bool = True) -> Position:
    p = Position()
    _check(lib().pg_random_position(seed, plies, int(require_nonterminal), byref(p)), "pg_random_position")
    return p


def random_request(seed: int, plies: tuple[int, int], depth: tuple[int, int]) -> tuple[Position, int, int]:
    p = Position()
    pl, d = c_int(), c_int()
    _check(lib().pg_random_request(seed, plies[0], plies[1], depth[0], depth[1], byref(p), byref(pl), byref(d)),
           "pg_random_request")
    return p, pl.value, d.value


def perft_cpu(p: Position, depth: int, use_tt: bool = False) -> int:
    c = c_uint64()
    _check(lib().pg_perft_cpu(byref(p), depth, int(use_tt), byref(c)), "pg_perft_cpu")
    return c.value


# ------------------------------------------------------------------ GPU side
_initialised = False


def init(arena_bytes: int = 0, device: int = 0) -> None:
    """pg_init: select the device and allocate the BFS arena (0 = automatic)."""
    global _initialised
    if not _initialised:
        _check(lib().pg_init(device, arena_bytes), "pg_init")
        _initialised = True


def shutdown() -> None:
    global _initialised
    if _initialised:
        lib().pg_shutdown()
        _initialised = False


def set_leaf_plies(plies: int) -> None:
    lib().pg_set_leaf_plies(plies)


def perft_bfs(p: Position, depth: int, use_tt: bool = False, buffer_ptr: int = 0, buffer_bytes: int = 0) -> int:
    """One GPU BFS call for the whole depth (perft_gpu_host_bfs)."""
    c = c_uint64()
    _check(lib().pg_perft_bfs(byref(p), depth, int(use_tt), c_void_p(buffer_ptr or None), buffer_bytes, byref(c)),
           "pg_perft_bfs")
    return c.value


def perft_bfs_status(p: Position, depth: int, buffer_ptr: int = 0, buffer_bytes: int = 0) -> tuple[int, int]:
    c = c_uint64()
    s = lib().pg_perft_bfs(byref(p), depth, 0, c_void_p(buffer_ptr or None), buffer_bytes, byref(c))
    return s, c.value


def perft(p: Position, depth: int, launch_depth: int = 0, use_tt: bool = False) -> int:
    """perftLauncher: host recursion above the launch depth, GPU BFS below."""
    r = U128()
    _check(lib().pg_perft(byref(p), depth, launch_depth, int(use_tt), byref(r)), "pg_perft")
    return r.value()


def tt_init(device_mb: int, host_mb: int) -> None:
    _check(lib().pg_tt_init(device_mb, host_mb), "pg_tt_init")


def tt_free() -> None:
    lib().pg_tt_free()


def tt_stats() -> list[int]:
    out = (c_uint64 * 6)()
    _check(lib().pg_tt_stats(out), "pg_tt_stats")
    return list(out)


def _pack(positions, depths):
    n = len(positions)
    arr = (Position * n)(*positions)
    d = (c_uint8 * n)(*depths)
    return n, arr, d


def perft_batch(positions, depths, leaf_plies: int = 0, grid_per_cu: int = 0, use_graph: bool = True,
                use_tt: bool = False, tt_budget_mb: int = 0) -> list[int]:
    """Batched perft; 0 for leaf_plies, grid_per_cu and tt_budget_mb means autotuned."""
    n, arr, d = _pack(positions, depths)
    out = (c_uint64 * max(n, 1))()
    opt = BatchOptions(grid_per_cu, leaf_plies, int(use_graph), int(use_tt), tt_budget_mb)
    _check(lib().pg_perft_batch(n, arr, d, byref(opt), out), "pg_perft_batch")
    return list(out[:n])


def batch_tuned() -> list[dict]:
    """The batch autotuner's cached choices."""
    rows = (c_int * (7 * 256))()
    n = c_int()
    _check(lib().pg_batch_tuned(rows, 256, ctypes.byref(n)), "pg_batch_tuned")
    keys = ("depth", "size_log2", "tt_requested", "leaf_plies", "tt_used", "grid_per_cu", "tt_span_log2")
    return [dict(zip(keys, rows[7 * i:7 * i + 7])) for i in range(n.value)]


def perft_batch_plain(positions, depths, use_tt: bool = False) -> list[int]:
    n, arr, d = _pack(positions, depths)
    out = (c_uint64 * max(n, 1))()
    _check(lib().pg_perft_batch_plain(n, arr, d, int(use_tt), out), "pg_perft_batch_plain")
    return list(out[:n])


def run_cli(args: list[str], timeout: float = 600) -> subprocess.CompletedProcess:
    """Run the perft_gpu command line (the drop-in binary)."""
    if not os.path.exists(CLI_PATH):
        raise RuntimeError(f"{CLI_PATH} is missing: build first")
    return subprocess.run([CLI_PATH] + args, capture_output=True, text=True, timeout=timeout)


def parse_cli_counts(stdout: str) -> dict[int, int]:
    out = {}
    for line in stdout.splitlines():
        if line.startswith("Perft("):
            d = int(line[6:8])
            out[d] = int(line.split(":")[1].split(",")[0])
    return out


def valu_probe(kind: int, waves_per_simd: int, iters: int) -> tuple[float, int]:
    """(milliseconds, wave instructions) of one launch of the VALU calibration kernel."""
    ms, insts = ctypes.c_double(), c_uint64()
    _check(lib().pg_valu_probe(kind, waves_per_simd, iters, ctypes.byref(ms), ctypes.byref(insts)), "pg_valu_probe")
    return ms.value, insts.value


def profile_enable(on: bool = True) -> None:
    lib().pg_profile_enable(int(on))


PROFILED_KERNELS = ("makecount", "scan", "expand", "genmoves", "leaf1", "leaf2", "batch_group", "tt")


def profile_kernels(names=None) -> None:
    """Time only the named kernels when profiling is on (all when None)."""
    mask = 0xFFFFFFFF if names is None else sum(1 << PROFILED_KERNELS.index(n) for n in names)
    _check(lib().pg_profile_kernels(mask), "pg_profile_kernels")


def profile_reset() -> None:
    lib().pg_profile_reset()


def profile_get() -> dict[str, dict]:
    arr = (KernelStat * 16)()
    n = c_int()
    _check(lib().pg_profile_get(arr, 16, byref(n)), "pg_profile_get")
    return {arr[i].name.decode(): {"launches": arr[i].launches, "ms": arr[i].ms, "bytes": arr[i].bytes,
                                   "items": arr[i].items, "big_bytes": arr[i].big_bytes, "big_ms": arr[i].big_ms,
                                   "big_launches": arr[i].big_launches} for i in range(n.value)}


def device_synchronize() -> None:
    _check(lib().pg_device_synchronize(), "pg_device_synchronize")


class DeviceArray:
    """A device buffer owned through the library (numpy in, numpy out)."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 1)
        self.ptr = c_void_p()
        _check(lib().pg_device_malloc(self.nbytes, byref(self.ptr)), "pg_device_malloc")

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            _check(lib().pg_memcpy(d.ptr, a.ctypes.data_as(c_void_p), a.nbytes, 0), "pg_memcpy")
        return d

    def to_numpy(self, dtype, count):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(lib().pg_memcpy(out.ctypes.data_as(c_void_p), self.ptr, out.nbytes, 1), "pg_memcpy")
        return out

    def zero(self):
        import numpy as np
        z = np.zeros(self.nbytes, dtype=np.uint8)
        _check(lib().pg_memcpy(self.ptr, z.ctypes.data_as(c_void_p), self.nbytes, 0), "pg_memcpy")
        return self

    def free(self):
        if self.ptr:
            lib().pg_device_free(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
