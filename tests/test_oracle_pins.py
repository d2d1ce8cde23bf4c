"""Pins of the CPU oracle (the checker) before anything is compared with it.

Each check is chosen so that a plausible slip in the oracle fails it:
golden vectors from the reference's own C++ (hand keys, ranges, bet lists),
the reference's tiny turn ground truth (tools/tiny_check.cpp), textbook
5-card enumeration counts, closed forms of the discount schedule, the
postflop-solver degenerate known answers (game/tests.rs) and an independent
per-pair f64 brute-force DCFR compared half-step by half-step.
"""
import json
import math
import os

import numpy as np
import pytest

import orc
from bruteforce import BR, EV, BruteForce
from conftest import board_of
import spots


def load(golden_dir, name):
    with open(os.path.join(golden_dir, name)) as f:
        return json.load(f)


def test_hand_keys_match_reference_golden(golden_dir):
    g = load(golden_dir, "hand_keys.json")
    bad = [(h, k) for h, k in zip(g["hands"], g["keys"]) if orc.hand_key(h) != k]
    assert not bad, bad[:5]
    assert len(g["keys"]) > 9000


def test_five_card_enumeration_textbook_counts():
    c = orc.count5()
    # high card .. straight flush, all 2,598,960 hands; 7,462 distinct ranks
    assert c[:9] == [1302540, 1098240, 123552, 54912, 10200, 5108, 3744, 624, 40]
    assert c[9] == 7462


def test_ranges_match_reference_golden(golden_dir):
    for e in load(golden_dir, "ranges.json"):
        if e["ok"]:
            w = orc.parse_range(e["text"])
            want = np.zeros(1326)
            for i, v in e["weights"].items():
                want[int(i)] = v
            assert np.array_equal(w, want), e["text"]
        else:
            with pytest.raises(ValueError) as ei:
                orc.parse_range(e["text"])
            assert str(ei.value) == e["error"], e["text"]


def test_range_totals_of_reference_unit_test():
    # tests/test_main.cpp:910-927
    for text, total in [("AA", 6), ("AKs", 4), ("AKo", 12), ("AK", 16), ("TT+", 30),
                        ("A9s+", 20), ("T9s+", 20), ("A5s-A2s", 16), ("98s-65s", 16),
                        ("55-22", 24), ("AA:0.5,KK", 9), ("AK,AA", 22)]:
        assert orc.parse_range(text).sum() == pytest.approx(total, abs=1e-9)


def test_bet_actions_match_reference_golden(golden_dir):
    cases = load(golden_dir, "bet_actions.json")
    for c in cases:
        cfg = c["cfg"]
        sp = orc.make_spot([0, 1, 2], np.zeros(1326), np.zeros(1326), 1, 1, bets=cfg["bets"],
                           raises=cfg["raises"], bet_allin=cfg["allin"], bet_geo=cfg["geo"],
                           max_raises=cfg["max_raises"], add_allin=c["add"],
                           force_allin=c["force"], merging=c["merge"])
        got = orc.bet_actions(c["pot_base"], c["stack"], sp, c["sc0"], c["sc1"], c["actor"],
                              c["after_allin"], c["num_bets"], c["num_streets"])
        assert [list(a) for a in got] == c["actions"], c
    assert len(cases) >= 3000


def _tiny_spot():
    return orc.make_spot(board_of("Qs9h2d7c"), orc.parse_range("TT+"), orc.parse_range("99+"),
                         200, 3000, bets=[0.5], raises=[])


def test_tiny_turn_ground_truth(golden_dir):
    g = load(golden_dir, "tiny_truth.json")
    o = orc.Oracle(_tiny_spot())
    st = o.stats()  # untrained: uniform average strategy
    assert abs(st["ev_oop"] - g["TINY"]["ev0"]) < 1e-4
    assert abs(st["ev_ip"] - g["TINY"]["ev1"]) < 1e-4
    for mode, tag in ((7, "M7"), (9, "M9")):
        o.seed_strategy(mode)
        st = o.stats()
        assert abs(st["ev_oop"] - g[tag]["ev0"]) < 1e-4, tag
        assert abs(st["ev_ip"] - g[tag]["ev1"]) < 1e-4, tag


def test_schedule_closed_forms():
    assert orc.schedule("dcfr", 0, 10) == (0.0, 0.5, 0.0)
    assert orc.schedule("dcfr", 1, 10) == (0.0, 0.5, 0.0)
    assert orc.schedule("dcfr", 2, 10) == (0.5, 0.5, 0.125)
    p3 = 3 ** 1.5
    assert orc.schedule("dcfr", 4, 10) == pytest.approx((p3 / (p3 + 1), 0.5, 0.0))
    # t=20: restart at 16, (4/5)^3
    p = 19 ** 1.5
    assert orc.schedule("dcfr", 20, 10) == pytest.approx((p / (p + 1), 0.5, 0.8 ** 3))
    assert orc.schedule("dcfr-g2", 3, 10) == pytest.approx((2 ** 1.5 / (2 ** 1.5 + 1), 0.5,
                                                            (3 / 4) ** 2))
    n = 100.0
    tt = 10.0
    a, b, g = 1 + 3 * tt / n, -1 - 2 * tt / n, 30 - 5 * tt / n
    assert orc.schedule("hs30", 9, 100) == pytest.approx(
        (tt ** a / (tt ** a + 1), tt ** b / (tt ** b + 1), (tt / (tt + 1)) ** g))


def _check_only(board, oop, ip):
    # postflop-solver game/tests.rs spots: pot 60, stack 970, no bet sizes
    return orc.make_spot(board_of(board), orc.parse_range(oop), orc.parse_range(ip), 60, 970)


def test_degenerate_known_answers():
    low = "KK-22,K9-K2,Q8-Q2,J8-J2,T8-T2,92+,82+,72+,62+"
    st = orc.Oracle(_check_only("AcAdKh", "AA", low)).stats()  # tests.rs:439-496
    assert st["ev_oop"] == pytest.approx(60.0, abs=1e-4)
    assert st["ev_ip"] == pytest.approx(0.0, abs=1e-4)
    st = orc.Oracle(_check_only("AcAdKh", low, "AA")).stats()  # tests.rs:551-585
    assert st["ev_oop"] == pytest.approx(0.0, abs=1e-4)
    assert st["ev_ip"] == pytest.approx(60.0, abs=1e-4)
    st = orc.Oracle(_check_only("2c6dTh", "AA", "AA")).stats()  # tests.rs:621-653
    assert st["ev_oop"] == pytest.approx(30.0, abs=1e-4)
    assert st["ev_ip"] == pytest.approx(30.0, abs=1e-4)
    # tests.rs:687-703: no disjoint pair at all must be refused
    with pytest.raises(ValueError):
        orc.Oracle(_check_only("Td9d6h", "TT", "TT")).stats()


@pytest.mark.slow
def test_all_check_full_ranges_is_symmetric():
    # tests.rs:8-97: identical full ranges, check-down flop => 30/30
    w = np.ones(1326)
    sp = orc.make_spot(board_of("Td9d6h"), w, w, 60, 970)
    st = orc.Oracle(sp, threads=8).stats()
    assert st["ev_oop"] == pytest.approx(30.0, abs=1e-3)
    assert st["ev_ip"] == pytest.approx(30.0, abs=1e-3)


def test_check_down_equals_brute_force_equity():
    # check-only turn tree: EV = pot x pair-weighted equity over the river
    sp = orc.make_spot(board_of("Qs9h2d7c"), orc.parse_range("AK,QJs"),
                       orc.parse_range("99,T8s,A2s"), 100, 1000)
    st = orc.Oracle(sp).stats()
    board = board_of("Qs9h2d7c")
    num = den = 0.0
    combos = [(a, b) for a in range(52) for b in range(a + 1, 52)]
    for i, (a, b) in enumerate(combos):
        if sp.w0[i] <= 0 or a in board or b in board:
            continue
        for j, (c, d) in enumerate(combos):
            if sp.w1[j] <= 0 or c in board or d in board or {a, b} & {c, d}:
                continue
            pw = sp.w0[i] * sp.w1[j]
            eq = 0.0
            rivers = [r for r in range(52) if r not in board and r not in (a, b, c, d)]
            for r in rivers:
                k0 = orc.hand_key([a, b] + board + [r])
                k1 = orc.hand_key([c, d] + board + [r])
                eq += 1.0 if k0 > k1 else 0.5 if k0 == k1 else 0.0
            num += pw * eq / len(rivers)
            den += pw
    assert st["ev_oop"] == pytest.approx(100 * num / den, abs=1e-4)
    assert st["ev_oop"] + st["ev_ip"] == pytest.approx(100, abs=1e-4)


def test_reference_root_labels():
    # python/test_pps.py:55-56: turn pot 200, stack 500, bets "0.75,a"
    sp = orc.make_spot(board_of("Qs9h2d7c"), orc.parse_range("TT+,AKo"),
                       orc.parse_range("AQ+,KQs"), 200, 500, bets=[0.75], raises=[2.5, 3],
                       bet_allin=True)
    o = orc.Oracle(sp)
    m = o.decide_meta(0)
    assert m["player"] == 0 and m["depth"] == 0
    acts = orc.bet_actions(200, 500, sp, 0, 0, 0, False, 0, 2)
    assert acts == [(1, 0), (3, 150), (5, 500)]
    assert o.decide_meta(1)["player"] == 1


def _tiny(name):
    return spots.tiny(name)


@pytest.mark.parametrize("name,steps", [("river", [0, 1, 2, 7, 30]), ("turn", [0, 1, 4, 15]),
                                        ("blocked", [0, 1, 6]),
                                        ("flop", [0, 3])])
def test_half_steps_match_per_pair_brute_force(name, steps):
    sp = _tiny(name)
    o = orc.Oracle(sp)
    bf = BruteForce(sp)
    assert o.counts()["decide"] == len(bf.decide)
    for t in steps:
        o.solve(t, "dcfr")
        for tr in (0, 1):
            bf.set_rows(0, o.rows(0))
            bf.set_rows(1, o.rows(1))
            o.half_step(tr, t)
            bf.half(tr, t)
            for which in (0, 1):
                a, c = o.rows(which), bf.rows(which)
                err = np.max(np.abs(a - c) / np.maximum(1.0, np.abs(c)))
                assert err < 2e-4, (name, t, tr, which, err)


@pytest.mark.parametrize("name", ["river", "turn", "blocked", "flop"])
def test_walks_match_per_pair_brute_force(name):
    sp = _tiny(name)
    o = orc.Oracle(sp)
    o.solve(12, "dcfr")
    bf = BruteForce(sp)
    bf.set_rows(0, o.rows(0))
    bf.set_rows(1, o.rows(1))
    for tr in (0, 1):
        for mode in (EV, BR):
            v, cfv = o.root_value(tr, mode)
            vb, cfvb = bf.root_value(tr, mode)
            assert v == pytest.approx(vb, abs=1e-4)
            assert np.allclose(cfv[:len(cfvb)], cfvb, atol=1e-4, rtol=1e-5)
    st = o.stats()
    sb = bf.stats()
    for k in sb:
        assert st[k] == pytest.approx(sb[k], abs=1e-4)
    assert st["exploitability"] >= 0


def test_zero_sum_and_exploitability_sign():
    sp = _tiny("turn")
    o = orc.Oracle(sp)
    for iters in (0, 5, 40):
        o.solve(iters, "dcfr")
        st = o.stats()
        assert st["ev_oop"] + st["ev_ip"] == pytest.approx(60.0, abs=1e-3)
        assert st["exploitability"] >= -1e-4


def test_exploitability_decreases_with_iterations():
    sp = _tiny("turn")
    o = orc.Oracle(sp)
    o.solve(10, "dcfr")
    e10 = o.stats()["exploitability"]
    o.solve(300, "dcfr")
    e300 = o.stats()["exploitability"]
    assert e300 < 0.2 * e10
    assert math.isfinite(e300)


def _aggregate(bf, u, strat):
    """Reach-free strategy aggregate of decide node u: each combo's strategy
    sums normalised (uniform when they are all zero), averaged with the
    acting player's range weights (cuda/gpu_cfr.cu:2109-2150)."""
    n = bf.nodes[u]
    p = n["player"]
    w = np.array([bf.base[p][k][2] for k in n["lists"][p]])
    tot = strat.sum(axis=0)
    na = strat.shape[0]
    sigma = np.where(tot > 0, strat / np.where(tot > 0, tot, 1.0), 1.0 / na)
    return (sigma * w).sum(axis=1) / w.sum()


@pytest.mark.parametrize("name", ["river", "turn"])
def test_strategy_aggregates_equal_numpy_restatement(name):
    """orc_root_strategy / orc_node_strategy against an aggregate computed
    here from the brute force's node lists and base weights, on strategy
    rows with all-zero combos (the uniform case) and uneven magnitudes."""
    sp = _tiny(name)
    o = orc.Oracle(sp)
    bf = BruteForce(sp)
    rng = np.random.default_rng(7)
    flat = rng.uniform(0.0, 5.0, o.counts()["reg_total"]).astype(np.float32)
    flat[rng.uniform(size=flat.shape) < 0.2] = 0.0
    o.set_rows(1, flat)
    bf.set_rows(1, flat)
    for idx in (0, 1):
        u = bf.decide[idx]
        strat = np.array(bf.nodes[u]["strat"], dtype=np.float64)
        strat[:, ::3] = 0.0  # every third combo: no strategy mass at all
        # write the zeroed columns back through the flat rows
        off = sum(bf.nodes[v]["strat"].size for v in bf.decide[:idx])
        flat[off:off + strat.size] = strat.astype(np.float32).ravel()
        o.set_rows(1, flat)
        want = _aggregate(bf, u, strat)
        got = o.root_strategy() if idx == 0 else o.node_strategy(idx)
        assert np.allclose(got, want, rtol=0, atol=1e-12), (idx, got, want)
        assert abs(sum(got) - 1.0) < 1e-12
    # decide nodes past the first street have no aggregate (chance-compacted
    # lists), as the reference returns none for them
    later = [i for i, u in enumerate(bf.decide) if len(bf.nodes[u]["board"]) > sp.n_board]
    if later:
        assert o.node_strategy(later[0]) == []
