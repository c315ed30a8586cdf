"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/_build/liboracle.so (plain-C restatement of the
reference's -cpu perft path, see oracle.h).  Only tests/, the smoke entry and
bench.py's cpu_baseline leg load this module, and only as the checker / the
timed CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_char_p, c_int, c_uint8, c_uint16, c_uint32, c_uint64

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64x4 = POINTER(c_uint64)
        L.oracle_init.restype = None
        L.oracle_parse_fen.argtypes = [c_char_p, u64x4, POINTER(c_uint8), POINTER(c_uint8)]
        L.oracle_write_fen.argtypes = [u64x4, c_uint8, c_uint8, c_char_p, c_int]
        L.oracle_count_moves.argtypes = [u64x4, c_uint8, c_uint8]
        L.oracle_count_moves.restype = c_uint32
        L.oracle_gen_moves.argtypes = [u64x4, c_uint8, c_uint8, POINTER(c_uint16)]
        L.oracle_gen_moves.restype = c_uint32
        L.oracle_brute_gen_moves.argtypes = [u64x4, c_uint8, c_uint8, POINTER(c_uint16)]
        L.oracle_brute_gen_moves.restype = c_uint32
        L.oracle_brute_count_moves.argtypes = [u64x4, c_uint8, c_uint8]
        L.oracle_brute_count_moves.restype = c_uint32
        L.oracle_make_move.argtypes = [u64x4, POINTER(c_uint8), c_uint8, c_uint16]
        L.oracle_make_move.restype = None
        for f in ("oracle_perft", "oracle_perft_brute"):
            getattr(L, f).argtypes = [u64x4, c_uint8, c_uint8, c_uint32]
            getattr(L, f).restype = c_uint64
        L.oracle_perft_mt.argtypes = [u64x4, c_uint8, c_uint8, c_uint32, c_int]
        L.oracle_perft_mt.restype = c_uint64
        L.oracle_perft_batch.argtypes = [c_int, u64x4, POINTER(c_uint8), POINTER(c_uint8), POINTER(c_uint32),
                                         POINTER(c_uint64), c_int]
        L.oracle_perft_batch.restype = None
        L.oracle_random_position.argtypes = [c_uint64, c_int, c_int, u64x4, POINTER(c_uint8), POINTER(c_uint8)]
        L.oracle_random_request.argtypes = [c_uint64, c_int, c_int, c_int, c_int, u64x4, POINTER(c_uint8),
                                            POINTER(c_uint8), POINTER(c_int), POINTER(c_int)]
        L.oracle_init()
        _lib = L
    return _lib


class Pos:
    """(bb[4], state, color) triple understood by the oracle."""

    __slots__ = ("bb", "state", "color")

    def __init__(self, bb, state, color):
        self.bb = tuple(int(x) for x in bb)
        self.state = int(state)
        self.color = int(color)

    def arr(self):
        return (c_uint64 * 4)(*self.bb)

    def key(self):
        return (self.bb, self.state, self.color)


def parse_fen(fen: str) -> Pos:
    bb = (c_uint64 * 4)()
    st, col = c_uint8(), c_uint8()
    if lib().oracle_parse_fen(fen.encode(), bb, byref(st), byref(col)) != 0:
        raise ValueError(f"bad FEN {fen!r}")
    return Pos(bb, st.value, col.value)


def write_fen(p: Pos) -> str:
    buf = ctypes.create_string_buffer(128)
    lib().oracle_write_fen(p.arr(), p.state, p.color, buf, 128)
    return buf.value.decode()


def count_moves(p: Pos) -> int:
    return lib().oracle_count_moves(p.arr(), p.state, p.color)


def gen_moves(p: Pos) -> list[int]:
    mv = (c_uint16 * 256)()
    n = lib().oracle_gen_moves(p.arr(), p.state, p.color, mv)
    return list(mv[:n])


def brute_gen_moves(p: Pos) -> list[int]:
    mv = (c_uint16 * 512)()
    n = lib().oracle_brute_gen_moves(p.arr(), p.state, p.color, mv)
    return list(mv[:n])


def brute_count_moves(p: Pos) -> int:
    return lib().oracle_brute_count_moves(p.arr(), p.state, p.color)


def make_move(p: Pos, m: int) -> Pos:
    bb = p.arr()
    st = c_uint8(p.state)
    lib().oracle_make_move(bb, byref(st), p.color, m)
    return Pos(bb, st.value, p.color ^ 1)


def perft(p: Pos, depth: int, threads: int = 1) -> int:
    if threads > 1:
        return lib().oracle_perft_mt(p.arr(), p.state, p.color, depth, threads)
    return lib().oracle_perft(p.arr(), p.state, p.color, depth)


def perft_brute(p: Pos, depth: int) -> int:
    return lib().oracle_perft_brute(p.arr(), p.state, p.color, depth)


def perft_batch(positions, depths, threads: int = 1) -> list[int]:
    n = len(positions)
    bbs = (c_uint64 * (4 * n))()
    sts, cols = (c_uint8 * n)(), (c_uint8 * n)()
    ds, out = (c_uint32 * n)(*depths), (c_uint64 * n)()
    for i, p in enumerate(positions):
        bbs[4 * i:4 * i + 4] = list(p.bb)
        sts[i], cols[i] = p.state, p.color
    lib().oracle_perft_batch(n, bbs, sts, cols, ds, out, threads)
    return list(out)


def random_position(seed: int, plies: int, require_nonterminal: bool = True) -> Pos:
    bb = (c_uint64 * 4)()
    st, col = c_uint8(), c_uint8()
    lib().oracle_random_position(seed, plies, int(require_nonterminal), bb, byref(st), byref(col))
    return Pos(bb, st.value, col.value)


def random_request(seed: int, plies: tuple[int, int], depth: tuple[int, int]):
    bb = (c_uint64 * 4)()
    st, col = c_uint8(), c_uint8()
    pl, d = c_int(), c_int()
    lib().oracle_random_request(seed, plies[0], plies[1], depth[0], depth[1], bb, byref(st), byref(col), byref(pl),
                                byref(d))
    return Pos(bb, st.value, col.value), pl.value, d.value
