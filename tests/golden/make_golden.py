"""Writes the golden fixtures under tests/golden/ from the reference's own
CPU C++, compiled (checker only) into oracle/_ref/ by `make -C oracle ref`.

Run in a container where the reference tree is mounted:
    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (data only — inputs and the reference's outputs):
  hand_keys.json    hand keys of seeded 5/6/7-card hands + the evaluator
                    edge cases of tests/test_main.cpp:24-64
  ranges.json       parseRange weight vectors / errors for a list of strings
                    (src/range.cpp:176-254, tests/test_main.cpp:910-927)
  bet_actions.json  pf::betActions outputs over a seeded grid of betting
                    states and bet configs (src/postflop_cfr.cpp:18-152)
  tiny_truth.json   tools/tiny_check.cpp's TINY / M7 / M9 root EVs
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
TINY = os.path.join(ROOT, "oracle", "_ref", "tiny_check")


def run_dump(lines):
    p = subprocess.run([REF_DUMP], input="\n".join(lines) + "\n", capture_output=True,
                       text=True, check=True)
    out = p.stdout.splitlines()
    assert len(out) == len(lines), (len(out), len(lines))
    return out


def C(rank, suit):
    return suit * 13 + rank


EDGE_HANDS = [
    [C(12, 0), C(11, 0), C(10, 0), C(9, 0), C(8, 0), C(5, 1), C(3, 2)],
    [C(7, 0), C(7, 1), C(7, 2), C(7, 3), C(12, 0), C(11, 1), C(5, 2)],
    [C(7, 0), C(7, 1), C(7, 2), C(3, 3), C(3, 1), C(12, 0), C(11, 1)],
    [C(12, 2), C(10, 2), C(8, 2), C(5, 2), C(3, 2), C(7, 1), C(6, 0)],
    [C(4, 0), C(5, 1), C(6, 2), C(7, 3), C(8, 0), C(12, 1), C(2, 2)],
    [C(12, 0), C(0, 1), C(1, 2), C(2, 3), C(3, 0), C(7, 1), C(9, 2)],
    [C(9, 0), C(9, 1), C(9, 2), C(4, 3), C(6, 1), C(11, 0), C(12, 1)],
    [C(5, 0), C(5, 1), C(2, 2), C(2, 3), C(12, 0), C(11, 1), C(10, 2)],
    [C(11, 0), C(11, 1), C(2, 2), C(4, 3), C(6, 1), C(9, 0), C(7, 1)],
    [C(12, 0), C(10, 1), C(8, 2), C(6, 3), C(4, 0), C(2, 1), C(5, 2)],
    [C(3, 0), C(3, 1), C(3, 2), C(3, 3), C(2, 0), C(11, 1), C(12, 2)],
    [C(0, 0), C(0, 1), C(0, 2), C(2, 0), C(2, 1), C(2, 2), C(5, 3)],
    [C(4, 0), C(4, 1), C(4, 2), C(4, 3), C(2, 0), C(2, 1), C(2, 2)],
    [C(5, 0), C(5, 1), C(5, 2), C(3, 0), C(3, 1), C(6, 2), C(7, 3)],
    # three pairs, steel wheel, broadway straight flush with a pair
    [C(9, 0), C(9, 1), C(5, 2), C(5, 3), C(2, 0), C(2, 1), C(12, 2)],
    [C(12, 1), C(0, 1), C(1, 1), C(2, 1), C(3, 1), C(3, 2), C(3, 3)],
    [C(8, 3), C(9, 3), C(10, 3), C(11, 3), C(12, 3), C(12, 0), C(12, 1)],
]

RANGE_STRINGS = [
    "AA", "AKs", "AKo", "AK", "TT+", "A9s+", "T9s+", "A5s-A2s", "98s-65s", "55-22",
    "AA:0.5,KK", "AK,AA", "A5s-A2s:0.5", "A5s:0.5-A2s", "TT+:0.5", "22+,A9s+,KTs+,QJs,JTs,T9s",
    "22+,A9s+,KTs+,QJs,JTs,T9s,AQo+", "XX", "AKx", "KQo-JTo", "K9s-K6s", "QQ-JJ", " AA , KK ",
    "", "AA,", "JJ-QQ", "AKs-AQo", "T9+", "KTo+", "A2+", "AA:0.25,AA:0.25", "AT+:0.5",
    "AKs:0.3,QQ:0.75", "TT+,AKo", "AQ+,KQs", "99+", "KK-QQ:0.4", "A", "AKs+", "22-AA",
    "K2s+", "65s+", "54o+", "J9+", "QJ:0.125", "AKs,AKs", "T9s-54s", "J8o-J5o", "Q7-Q4",
    "AA-22", "KQ-JT", "KdKc", "A5s-A2o", "77+,A2s+,K9s+,Q9s+,J9s+,T8s+,97s+,86s+,75s+,64s+,ATo+,KJo+",
]


def gen_hand_keys(rng):
    hands = [list(h) for h in EDGE_HANDS]
    for n, count in ((5, 1500), (6, 1500), (7, 6000)):
        for _ in range(count):
            hands.append([int(x) for x in rng.choice(52, size=n, replace=False)])
    lines = ["K %d %s" % (len(h), " ".join(str(c) for c in h)) for h in hands]
    keys = [int(x) for x in run_dump(lines)]
    return {"hands": hands, "keys": keys}


def gen_ranges():
    out = []
    for s in RANGE_STRINGS:
        line = run_dump(["R " + s])[0]
        if line.startswith("OK"):
            w = {}
            for tok in line.split()[1:]:
                i, v = tok.split(":")
                w[int(i)] = float(v)
            out.append({"text": s, "ok": True, "weights": w})
        else:
            out.append({"text": s, "ok": False, "error": line[4:]})
    return out


CONFIGS = [
    {"bets": [0.5, 0.75], "raises": [2.5, 3.0], "allin": 0, "geo": 0, "max_raises": 0},
    {"bets": [0.75], "raises": [2.5, 3.0], "allin": 1, "geo": 0, "max_raises": 0},
    {"bets": [0.5, 1.0], "raises": [2.5], "allin": 1, "geo": 0, "max_raises": 2},
    {"bets": [0.33, 0.75, 1.5], "raises": [2.5, 4.0], "allin": 1, "geo": 0, "max_raises": 3},
    {"bets": [0.25, 0.5, 0.75, 1.25], "raises": [2.5, 3.0], "allin": 0, "geo": 0, "max_raises": 0},
    {"bets": [0.33], "raises": [], "allin": 0, "geo": 1, "max_raises": 1},
    {"bets": [], "raises": [3.0], "allin": 1, "geo": 1, "max_raises": 0},
    {"bets": [0.5], "raises": [], "allin": 0, "geo": 0, "max_raises": 0},
    {"bets": [0.1, 0.15, 0.2, 0.6, 0.65], "raises": [2.0, 2.2], "allin": 0, "geo": 0,
     "max_raises": 0},
]


def gen_bets(rng):
    cases = []
    for k in range(3000):
        cfg = CONFIGS[k % len(CONFIGS)]
        pot_base = int(rng.choice([20, 60, 100, 137, 200, 300, 500, 1000]))
        stack = int(rng.choice([50, 100, 150, 240, 300, 500, 900, 970, 1500, 3000]))
        mode = int(rng.integers(0, 4))
        if mode == 0:
            sc0 = sc1 = 0
        elif mode == 1:
            sc0, sc1 = 0, int(rng.integers(1, stack + 1))
        elif mode == 2:
            sc0, sc1 = int(rng.integers(1, stack + 1)), 0
        else:
            a = int(rng.integers(1, stack + 1))
            b = int(rng.integers(a, stack + 1))
            sc0, sc1 = (a, b) if rng.integers(0, 2) else (b, a)
        actor = int(rng.integers(0, 2))
        after = int(rng.integers(0, 2)) if sc0 != sc1 else 0
        num_bets = int(rng.integers(0, 4))
        streets = int(rng.integers(1, 4))
        add_t, force_t, merge_t = 1.5, 0.15, 0.1
        if k % 7 == 0:
            add_t, force_t, merge_t = 3.0, 0.2, 0.0
        cases.append({"pot_base": pot_base, "stack": stack, "sc0": sc0, "sc1": sc1,
                      "actor": actor, "after_allin": after, "num_bets": num_bets,
                      "num_streets": streets, "cfg": cfg, "add": add_t, "force": force_t,
                      "merge": merge_t})
    lines = []
    for c in cases:
        cfg = c["cfg"]
        lines.append("B %d %d %d %d %d %d %d %d %d %s %d %s %d %d %d %r %r %r" % (
            c["pot_base"], c["stack"], c["sc0"], c["sc1"], c["actor"], c["after_allin"],
            c["num_bets"], c["num_streets"], len(cfg["bets"]),
            " ".join(repr(x) for x in cfg["bets"]), len(cfg["raises"]),
            " ".join(repr(x) for x in cfg["raises"]), cfg["allin"], cfg["geo"],
            cfg["max_raises"], c["add"], c["force"], c["merge"]))
    outs = run_dump(lines)
    for c, o in zip(cases, outs):
        c["actions"] = [[int(t.split(":")[0]), int(t.split(":")[1])] for t in o.split()]
    return cases


def gen_tiny():
    out = subprocess.run([TINY], capture_output=True, text=True, check=True).stdout
    res = {"spot": {"board": "Qs9h2d7c", "oop": "TT+", "ip": "99+", "pot": 200, "stack": 3000,
                    "bets": [0.5], "raises": []}}
    for line in out.splitlines():
        for tag in ("TINY", "M7", "M9"):
            if line.startswith(tag + ":"):
                parts = line.split()
                res[tag] = {"ev0": float(parts[2]), "ev1": float(parts[4])}
    return res


def main():
    if not os.path.exists(REF_DUMP):
        sys.exit("build the reference checker first: make -C oracle ref")
    rng = np.random.default_rng(20261203)
    fixtures = {
        "hand_keys.json": gen_hand_keys(rng),
        "ranges.json": gen_ranges(),
        "bet_actions.json": gen_bets(rng),
        "tiny_truth.json": gen_tiny(),
    }
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()
