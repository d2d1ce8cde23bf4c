"""ctypes view of the CPU oracle (oracle/bin/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg, as the checker. The product never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "bin", "liborc.so")

KIND_NAMES = ["fold", "check", "call", "bet", "raise", "allin"]


class OrcSpot(ctypes.Structure):
    _fields_ = [
        ("n_board", ctypes.c_int),
        ("board", ctypes.c_int * 5),
        ("w0", ctypes.c_double * 1326),
        ("w1", ctypes.c_double * 1326),
        ("pot", ctypes.c_int64),
        ("stack", ctypes.c_int64),
        ("n_bets", ctypes.c_int),
        ("bets", ctypes.c_double * 16),
        ("n_raises", ctypes.c_int),
        ("raises", ctypes.c_double * 16),
        ("bet_allin", ctypes.c_int),
        ("bet_geo", ctypes.c_int),
        ("max_raises", ctypes.c_int),
        ("add_allin", ctypes.c_double),
        ("force_allin", ctypes.c_double),
        ("merging", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_hand_key.restype = ctypes.c_uint64
        L.orc_hand_key.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.orc_parse_range.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_char_p, ctypes.c_int]
        L.orc_bet_actions.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(OrcSpot),
                                      ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
        L.orc_schedule.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcSpot)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        L.orc_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        L.orc_continue.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_half_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_char_p, ctypes.c_int]
        L.orc_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.orc_root_value.restype = ctypes.c_double
        L.orc_root_value.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]
        L.orc_root_strategy.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_int]
        L.orc_node_strategy.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.orc_rows.restype = ctypes.c_int64
        L.orc_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                               ctypes.c_int64]
        L.orc_set_rows.restype = ctypes.c_int64
        L.orc_set_rows.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_float), ctypes.c_int64]
        L.orc_decide_meta.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int)]
        L.orc_seed_strategy.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_count5.argtypes = [ctypes.POINTER(ctypes.c_int64)]
        L.orc_alg_bytes.restype = ctypes.c_double
        L.orc_alg_bytes.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def make_spot(board, w0, w1, pot, stack, bets=(), raises=(), bet_allin=False,
              bet_geo=False, max_raises=0, add_allin=1.5, force_allin=0.15,
              merging=0.1):
    s = OrcSpot()
    s.n_board = len(board)
    for i, c in enumerate(board):
        s.board[i] = int(c)
    for i in range(1326):
        s.w0[i] = float(w0[i])
        s.w1[i] = float(w1[i])
    s.pot = int(pot)
    s.stack = int(stack)
    s.n_bets = len(bets)
    for i, b in enumerate(bets):
        s.bets[i] = float(b)
    s.n_raises = len(raises)
    for i, r in enumerate(raises):
        s.raises[i] = float(r)
    s.bet_allin = int(bool(bet_allin))
    s.bet_geo = int(bool(bet_geo))
    s.max_raises = int(max_raises)
    s.add_allin = add_allin
    s.force_allin = force_allin
    s.merging = merging
    return s


def hand_key(cards):
    arr = (ctypes.c_int * len(cards))(*[int(c) for c in cards])
    return int(lib().orc_hand_key(arr, len(cards)))


def parse_range(text):
    w = (ctypes.c_double * 1326)()
    err = ctypes.create_string_buffer(256)
    rc = lib().orc_parse_range(text.encode(), w, err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return np.frombuffer(w, dtype=np.float64).copy()


def bet_actions(pot_base, stack, spot, sc0, sc1, actor, after_allin, num_bets, num_streets):
    out = (ctypes.c_int64 * 64)()
    n = lib().orc_bet_actions(pot_base, stack, ctypes.byref(spot), sc0, sc1, actor,
                              int(after_allin), num_bets, num_streets, out, 64)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]


def count5():
    out = (ctypes.c_int64 * 10)()
    lib().orc_count5(out)
    return [int(x) for x in out]


def schedule(algo, t, iters):
    out = (ctypes.c_double * 3)()
    lib().orc_schedule(algo.encode(), t, iters, out)
    return tuple(out)


class Oracle:
    """One compiled spot on the CPU oracle."""

    def __init__(self, spot, threads=1):
        self._spot = spot
        lib().orc_set_threads(int(threads))
        self.h = lib().orc_create(ctypes.byref(spot))
        if not self.h:
            raise ValueError("empty ranges")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def counts(self):
        out = (ctypes.c_int64 * 8)()
        lib().orc_counts(self.h, out)
        keys = ["nodes", "max_depth", "decide", "reg_total", "showdown", "fold", "chance",
                "river_boards"]
        return dict(zip(keys, [int(x) for x in out]))

    def solve(self, iters, algo="dcfr"):
        if lib().orc_solve(self.h, int(iters), algo.encode()) != 0:
            raise ValueError(f"bad algo {algo}")

    def continue_solve(self, iters):
        if lib().orc_continue(self.h, int(iters)) != 0:
            raise ValueError("continuation needs a dcfr solve")

    def half_step(self, tr, t, algo="dcfr", planned=None):
        if lib().orc_half_step(self.h, tr, t, algo.encode(), planned or t + 1) != 0:
            raise ValueError(f"bad algo {algo}")

    def stats(self):
        out = (ctypes.c_double * 4)()
        if lib().orc_stats(self.h, out) != 0:
            raise ValueError("no disjoint pair mass")
        return {"ev_oop": out[0], "ev_ip": out[1], "exploitability": out[2],
                "pair_mass": out[3]}

    def root_value(self, tr, mode):
        n = self.counts()  # noqa: F841
        buf = (ctypes.c_float * 1326)()
        v = lib().orc_root_value(self.h, tr, mode, buf)
        return v, np.frombuffer(buf, dtype=np.float32).copy()

    def root_strategy(self):
        out = (ctypes.c_double * 64)()
        n = lib().orc_root_strategy(self.h, out, 64)
        return [out[i] for i in range(n)]

    def node_values(self, node, tr, mode):
        """One walk (mode 1 = EV, 2 = BR) recording node's traverser values and
        the opponent reach arriving there; with the cards of both lists."""
        L = lib()
        L.orc_node_values.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_node_cards.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        cfv, reach = (ctypes.c_float * 1326)(), (ctypes.c_float * 1326)()
        nt, no = ctypes.c_int(), ctypes.c_int()
        if L.orc_node_values(self.h, node, tr, mode, cfv, reach, 1326, ctypes.byref(nt),
                             ctypes.byref(no)) != 0:
            raise IndexError(node)
        cards = []
        for p in (tr, 1 - tr):
            buf = (ctypes.c_int * 2652)()
            n = L.orc_node_cards(self.h, node, p, buf, 2652)
            cards.append([(buf[2 * i], buf[2 * i + 1]) for i in range(n)])
        return (np.frombuffer(cfv, dtype=np.float32)[:nt.value].copy(),
                np.frombuffer(reach, dtype=np.float32)[:no.value].copy(), cards[0], cards[1])

    def node_strategy(self, idx):
        out = (ctypes.c_double * 64)()
        n = lib().orc_node_strategy(self.h, idx, out, 64)
        if n < 0:
            raise IndexError(idx)
        return [out[i] for i in range(n)]

    def rows(self, which):
        total = self.counts()["reg_total"]
        buf = np.zeros(total, dtype=np.float32)
        lib().orc_rows(self.h, which, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), total)
        return buf

    def set_rows(self, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        r = lib().orc_set_rows(self.h, which, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               arr.size)
        if r < 0:
            raise ValueError("row size mismatch")

    def decide_meta(self, idx):
        out = (ctypes.c_int * 6)()
        if lib().orc_decide_meta(self.h, idx, out) != 0:
            raise IndexError(idx)
        keys = ["node_id", "player", "depth", "n_board", "na", "n_combos"]
        return dict(zip(keys, list(out)))

    def seed_strategy(self, mode):
        lib().orc_seed_strategy(self.h, int(mode))

    def alg_bytes(self):
        return float(lib().orc_alg_bytes(self.h))
