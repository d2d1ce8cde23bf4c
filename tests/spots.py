"""Spots shared by the tests."""
import numpy as np

import orc
from conftest import board_of, card

_IDX = {}
_k = 0
for _x in range(52):
    for _y in range(_x + 1, 52):
        _IDX[(_x, _y)] = _k
        _k += 1


def weights(combos, seed):
    """Weights in [0.2, 1) for the named two-card combos ("AhKh"), else 0."""
    rng = np.random.default_rng(seed)
    w = np.zeros(1326)
    for s in combos:
        a, b = card(s[:2]), card(s[2:])
        w[_IDX[(min(a, b), max(a, b))]] = rng.uniform(0.2, 1.0)
    return w


# Tiny spots: few combos per side, with an identical combo on both sides
# (AhKh) to exercise the same-hand correction.
TINY_SPOTS = {
    "river": dict(board="Qs9h2d7c3s", oop=["AhKh", "QhQd", "9c9d", "JhTc", "7h7d"],
                  ip=["AhKh", "KcQc", "8h8d", "AsAd", "7h6h"], pot=100, stack=200,
                  bets=[0.5, 1.0], raises=[2.5], allin=True),
    "turn": dict(board="Qs9h2d7c", oop=["AhKh", "QhQd", "JhTc"], ip=["AhKh", "KcQc", "8h8d"],
                 pot=60, stack=100, bets=[0.75], raises=[], allin=True),
    # Every OOP combo holds the As: after an As river, OOP's list is empty
    # and the showdown / fold groups of that class have no opponent combos.
    "blocked": dict(board="Qs9h2d7c", oop=["AsKh", "AsQd", "AsJc"],
                    ip=["KcQc", "8h8d", "AhKh", "AsTd"], pot=60, stack=100, bets=[0.75],
                    raises=[], allin=True),
    # Two chance levels (turn deal, then river deal): 31,124 nodes over
    # 2,352 river boards. Check / all-in on every street keeps the brute
    # force affordable while every betting line still reaches both deals.
    "flop": dict(board="Qs9h2d", oop=["AhKh", "QhQd", "JhTc", "9c8c"],
                 ip=["AhKh", "KcQc", "8h8d", "Ts9s"], pot=60, stack=30, bets=[], raises=[],
                 allin=True),
}


def tiny_dict(name):
    """A tiny spot as a generator-style dict (board ids, weights, config)."""
    s = TINY_SPOTS[name]
    return {"board": board_of(s["board"]), "w0": weights(s["oop"], 1),
            "w1": weights(s["ip"], 2), "pot": s["pot"], "stack": s["stack"],
            "bets": s["bets"], "raises": s["raises"], "allin": s["allin"], "max_raises": 0,
            "algo": "dcfr", "name": name}


def orc_spot(s):
    return orc.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"], s["bets"],
                         s["raises"], s["allin"], s.get("geo", False), s["max_raises"])


def tiny(name):
    return orc_spot(tiny_dict(name))
