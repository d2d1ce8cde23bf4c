"""GPU parity: the HIP engine against the CPU oracle, through the C ABI.

Bars (floating point; integer work is pinned bit-exact in test_host_cpu):
  - one CFR half-step from identical rows: regrets within 1e-3 relative
    (denominator max(1, |x|)), strategy sums within 1e-5;
  - EV / best-response root walks: values within 1e-4 chips, root cfv rows
    within 1e-4 + 1e-5 relative;
  - full solves at 500 iterations (the reference's gates): root EV within
    0.01 chips, exploitability ratio in [0.5, 2], root and node-1 strategy
    within 0.10;
  - determinism: two solves give bitwise-identical rows.
"""
import json
import os

import numpy as np
import pytest

import orc
import spots
from pps import capi, configs

pytestmark = pytest.mark.gpu


def gpu_solver(s):
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    cf = capi.make_config(s["bets"], s["raises"], s["allin"], s.get("geo", False),
                          s["max_raises"])
    return capi.Solver(sp, cf)


def rel_err(a, c):
    return float(np.max(np.abs(a - c) / np.maximum(1.0, np.abs(c)))) if len(c) else 0.0


@pytest.mark.parametrize("name,steps", [("river", [0, 1, 2, 7, 30]), ("turn", [0, 1, 4, 15]),
                                        ("blocked", [0, 1, 6]),
                                        ("flop", [0, 1, 4, 15])])
def test_half_steps_match_oracle(name, steps):
    s = spots.tiny_dict(name)
    g = gpu_solver(s)
    o = orc.Oracle(spots.orc_spot(s))
    assert g.num_nodes == o.counts()["nodes"]
    for t in steps:
        o.solve(t, "dcfr")
        for tr in (0, 1):
            R, S = o.rows(0), o.rows(1)
            g.set_rows(0, R)
            g.set_rows(1, S)
            o.half_step(tr, t)
            g.half_step(tr, t)
            assert rel_err(g.rows(0), o.rows(0)) < 1e-3, (name, t, tr)
            assert rel_err(g.rows(1), o.rows(1)) < 1e-5, (name, t, tr)


@pytest.mark.parametrize("name", ["river", "turn", "blocked", "flop"])
def test_walks_match_oracle(name):
    s = spots.tiny_dict(name)
    o = orc.Oracle(spots.orc_spot(s))
    o.solve(12, "dcfr")
    g = gpu_solver(s)
    g.set_rows(0, o.rows(0))
    g.set_rows(1, o.rows(1))
    for tr in (0, 1):
        for mode in (1, 2):
            v, cfv = o.root_value(tr, mode)
            vg, cfvg = g.root_value(tr, mode)
            n = len(np.nonzero(s["w0" if tr == 0 else "w1"])[0])
            assert vg == pytest.approx(v, abs=1e-4)
            assert np.allclose(cfvg[:n], cfv[:n], atol=1e-4, rtol=1e-5)


def test_plain_turn_half_step_at_full_size():
    s = configs.plain_turn()
    g = gpu_solver(s)
    o = orc.Oracle(spots.orc_spot(s), threads=8)
    o.solve(3, s["algo"])
    for tr in (0, 1):
        g.set_rows(0, o.rows(0))
        g.set_rows(1, o.rows(1))
        o.half_step(tr, 3, s["algo"], 500)
        g.half_step(tr, 3, s["algo"], 500)
        assert rel_err(g.rows(0), o.rows(0)) < 1e-3
        assert rel_err(g.rows(1), o.rows(1)) < 1e-5


def full_range_flop():
    """Every combo on both sides (1,176 each, weights in (0, 1]): each card's
    opponent segment holds 48 combos on the flop and 47 after the turn, the
    largest per-card segments any tree has; check / all-in over two deals."""
    s = configs.random_spot(11, 3, 1.0, 60, 30, [], [], True, 0)
    s.update(name="full-range-flop", algo="dcfr")
    return s


def test_flop_half_step_at_full_range():
    s = full_range_flop()
    g = gpu_solver(s)
    o = orc.Oracle(spots.orc_spot(s), threads=8)
    assert g.num_nodes == o.counts()["nodes"] == 31124
    o.solve(3, s["algo"])
    for tr in (0, 1):
        g.set_rows(0, o.rows(0))
        g.set_rows(1, o.rows(1))
        o.half_step(tr, 3, s["algo"], 500)
        g.half_step(tr, 3, s["algo"], 500)
        assert rel_err(g.rows(0), o.rows(0)) < 1e-3, tr
        assert rel_err(g.rows(1), o.rows(1)) < 1e-5, tr
    for tr in (0, 1):
        for mode in (1, 2):
            v, _ = o.root_value(tr, mode)
            vg, _ = g.root_value(tr, mode)
            assert vg == pytest.approx(v, abs=1e-4), (tr, mode)


def _gates(g, ref, pot):
    st = g.stats()
    assert abs(st["ev_oop"] - ref["ev_oop"]) < 0.01, (st, ref["ev_oop"])
    assert abs(st["ev_oop"] + st["ev_ip"] - pot) < 1e-2
    ratio = st["exploitability"] / ref["exploitability"]
    assert 0.5 <= ratio <= 2.0, (st["exploitability"], ref["exploitability"])
    assert np.max(np.abs(np.array(g.root_strategy()) - ref["root_strategy"])) < 0.10
    ns = g.node_strategy(1)
    if ref["node1_strategy"]:
        assert np.max(np.abs(np.array(ns) - ref["node1_strategy"])) < 0.10
    return st


@pytest.mark.parametrize("maker", [configs.plain_turn, configs.large_turn])
def test_full_size_gates_at_500_iterations(golden_dir, maker):
    s = maker()
    with open(os.path.join(golden_dir, "oracle_runs.json")) as f:
        ref = json.load(f)[s["name"]]
    g = gpu_solver(s)
    assert g.num_nodes == ref["counts"]["nodes"]
    g.solve(500, s["algo"])
    assert g.iterations_run == 500
    _gates(g, ref, s["pot"])


def test_solves_are_bitwise_deterministic():
    s = configs.plain_turn()
    g = gpu_solver(s)
    g.solve(40, "dcfr")
    a0, a1 = g.rows(0), g.rows(1)
    e0 = g.stats()
    h = gpu_solver(s)
    h.solve(40, "dcfr")
    assert np.array_equal(a0, h.rows(0)) and np.array_equal(a1, h.rows(1))
    assert h.stats() == e0
    g.solve(40, "dcfr")  # fresh solve on the same handle
    assert np.array_equal(a0, g.rows(0))


def test_warm_start_equals_long_solve_bitwise():
    s = configs.plain_turn()
    a = gpu_solver(s)
    a.solve(60, "dcfr")
    b = gpu_solver(s)
    b.solve(25, "dcfr")
    b.continue_solve(35)
    assert b.total_iterations == 60
    assert np.array_equal(a.rows(0), b.rows(0)) and np.array_equal(a.rows(1), b.rows(1))


def _text_solver(board, oop, ip, pot, stack, bets, raises, allin, geo=False):
    sp = capi.make_spot([configs.card_id(board[i:i + 2]) for i in range(0, len(board), 2)],
                        orc.parse_range(oop), orc.parse_range(ip), pot, stack)
    return capi.Solver(sp, capi.make_config(bets, raises, allin, geo, 0))


def test_shortstack_flop_gates_at_500_iterations(golden_dir):
    """The reference's shortstack flop geometry (tools/quality/bench_flop.sh:47:
    Qs9h2d, pot 500, stack 150, "0.4,a", raises 2.5x) with its gate ranges,
    against the oracle's 500-iteration solve."""
    with open(os.path.join(golden_dir, "oracle_flop.json")) as f:
        ref = json.load(f)["shortstack-flop"]
    sp = ref["spot"]
    g = _text_solver(ref["board"], ref["oop"], ref["ip"], sp["pot"], sp["stack"], sp["bets"],
                     sp["raises"], sp["allin"])
    assert g.num_nodes == ref["counts"]["nodes"]
    g.solve(500, "dcfr")
    _gates(g, ref, sp["pot"])


def test_k83_scenario_reaches_the_reference_answer():
    """tools/scenarios/k83_utg_bb.sh: flop Kc8h3s, BB (~786 combos) vs UTG,
    pot 500, stack 3750, geometric + all-in bets, 2.5x raises, solved to
    exploitability 0.5 chips. The reference records (lines 45-48) EV
    169.90 / 330.10 at exploitability 0.364 after 256 iterations, with the
    postflop-solver oracle at 169.91 / 330.09, and BB checking 99.9995 % at
    the root. EVs agree within the exploitability band."""
    bb = ("22+,A2s+,K2s+,Q2s+,J2s+,T2s+,98s-92s,87s-82s,76s-72s,65s-62s,54s-52s,43s-42s,32s,"
          "A2o+,K2o+,Q8o+,J9o+,T9o,98o,97o,87o")
    utg = "55+,A5s+,K9s+,Q9s+,J9s+,T9s,98s,AJo+,KQo"
    g = _text_solver("Kc8h3s", bb, utg, 500, 3750, [], [2.5], True, geo=True)
    assert g.num_nodes == 579660
    g.solve(100000, "dcfr", target=0.5)
    st = g.stats()
    assert st["exploitability"] <= 0.5
    assert g.iterations_run <= 512, g.iterations_run
    assert abs(st["ev_oop"] + st["ev_ip"] - 500) < 1e-2
    band = 0.5
    assert abs(st["ev_oop"] - 169.90) <= band, st
    assert abs(st["ev_ip"] - 330.10) <= band, st
    assert g.root_strategy()[0] >= 0.9999, g.root_strategy()


def test_large_flop_properties_at_full_size():
    """The large flop workload (3.57 M nodes, 30.6 GB of rows) at its full
    size: zero-sum EVs, exploitability falling from 200 to 1000 iterations,
    and a second solve of 200 iterations bitwise equal to the first."""
    s = configs.large_flop()
    g = gpu_solver(s)
    assert g.num_nodes == 3572482
    g.solve(200, s["algo"])
    a = g.stats()
    ra = g.root_strategy()
    g.continue_solve(800)
    b = g.stats()
    assert g.total_iterations == 1000
    for st in (a, b):
        assert abs(st["ev_oop"] + st["ev_ip"] - s["pot"]) < 1e-2, st
        assert st["exploitability"] >= 0
    assert b["exploitability"] < 0.5 * a["exploitability"], (a, b)
    g.close()
    h = gpu_solver(s)
    h.solve(200, s["algo"])
    assert h.stats() == a
    assert h.root_strategy() == ra


def test_detailed_profile_leaves_the_rows_as_they_were():
    """profile(detail=True) re-runs the backward launches (their update would
    apply twice); the rows and the iteration count are restored, so a
    continued solve equals an uninterrupted one bit for bit."""
    s = configs.plain_turn()
    a = gpu_solver(s)
    a.solve(30, "dcfr")
    r0, r1 = a.rows(0), a.rows(1)
    prof = a.profile(3, detail=True)
    assert prof["iters"] == 3 and prof["bwd_launches"] > 0
    assert a.total_iterations == 30
    assert np.array_equal(a.rows(0), r0) and np.array_equal(a.rows(1), r1)
    a.continue_solve(20)
    b = gpu_solver(s)
    b.solve(50, "dcfr")
    assert np.array_equal(a.rows(0), b.rows(0)) and np.array_equal(a.rows(1), b.rows(1))


def _pps_from_oracle_rows(s, o, tmp_path):
    """The product solver loaded from a PPS1 file holding the oracle's
    current rows (the file's row layout is the oracle's: decide nodes in
    creation order, action-major)."""
    import pps
    from test_solution_file_cpu import pps1
    reg, strat = o.rows(0), o.rows(1)
    rows = np.concatenate([reg, strat]).astype("<f4")
    data = pps1(board=tuple(s["board"]), pot=s["pot"], stack=s["stack"], w0=s["w0"], w1=s["w1"],
                bets=tuple(s["bets"]), raises=tuple(s["raises"]), allin=1 if s["allin"] else 0,
                cap=s["max_raises"], nodes=o.counts()["nodes"], iterations=25, algo=b"dcfr",
                rows=rows, n_rows=len(reg))
    path = tmp_path / "oracle_rows.sol"
    path.write_bytes(data)
    return pps.Solver.load(str(path))


@pytest.mark.parametrize("name", ["turn", "river", "blocked", "flop"])
def test_node_ev_matches_oracle_node_values(name, tmp_path):
    """node_ev (walks restricted to the node's subtree) against the oracle's
    own walk on the same rows: per traverser combo, the oracle's value at the
    node divided by the opponent reach mass arriving there without shared
    cards, on the root, node 1 and the first decide node past a deal."""
    s = spots.tiny_dict(name)
    o = orc.Oracle(spots.orc_spot(s))
    o.solve(25, "dcfr")
    g = _pps_from_oracle_rows(s, o, tmp_path)
    assert g.num_nodes == o.counts()["nodes"] and g.total_iterations == 25
    nodes = [0, 1]
    deeper = [i for i in range(o.counts()["decide"])
              if o.decide_meta(i)["n_board"] > len(s["board"])]
    nodes += deeper[:1]
    for idx in nodes:
        ev = g.node_ev(idx)
        node = o.decide_meta(idx)["node_id"]
        for tr, side in ((0, "oop"), (1, "ip")):
            cfv, reach, tc, oc = o.node_values(node, tr, 1)
            mass = np.array([sum(r for r, (x, y) in zip(reach, oc) if not {x, y} & {a, b})
                             for a, b in tc])
            got = np.asarray(ev["sides"][side]["ev"])
            assert len(got) == len(cfv), (name, idx, side)
            live = mass > 0
            want = cfv[live] / mass[live]
            err = np.abs(got[live] - want) / np.maximum(1.0, np.abs(want))
            assert np.all(err <= 1e-4), (name, idx, side, err.max())
            assert np.allclose(ev["sides"][side]["mass"], mass, rtol=1e-5, atol=1e-6)


def test_terminal_split_changes_no_rows_and_no_results():
    """PPSG_TUNE_TERM_SPLIT k puts the fold and showdown nodes at depth <= k
    on a second stream (forked after the forward launch of depth k - 1,
    joined before the backward launch of depth k - 1), and
    ppsg_autotune_param times the candidates over CFR half-steps with dry
    backward updates: the rows stay bitwise the same, and solves at every
    split (none, shallow, mid-tree, all terminals) equal the one-stream
    solve."""
    import ctypes
    L = capi.lib()
    L.pps_solver_engine.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(ctypes.c_int)]
    L.ppsg_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.ppsg_autotune_param.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_double)]
    TERM_SPLIT = 1
    s = configs.large_turn()
    a = gpu_solver(s)
    a.solve(30, "dcfr")
    r0, r1 = a.rows(0), a.rows(1)
    eng, slot = ctypes.c_void_p(), ctypes.c_int()
    assert L.pps_solver_engine(a.h, ctypes.byref(eng), ctypes.byref(slot)) == 0
    cands = [0, 2, 3, 5, 7, 40]
    arr = (ctypes.c_int * len(cands))(*cands)
    ms = (ctypes.c_double * len(cands))()
    chosen = ctypes.c_int(-1)
    assert L.ppsg_autotune_param(eng, TERM_SPLIT, arr, len(cands), 2, ctypes.byref(chosen), ms) == 0
    assert chosen.value in cands and all(x > 0.0 for x in ms)
    assert np.array_equal(a.rows(0), r0) and np.array_equal(a.rows(1), r1)
    a.continue_solve(20)
    b = gpu_solver(s)
    b.solve(50, "dcfr")
    assert np.array_equal(a.rows(0), b.rows(0)) and np.array_equal(a.rows(1), b.rows(1))
    for k in (1, 3, 6, 100):
        assert L.ppsg_tune(eng, TERM_SPLIT, k) == 0
        a.solve(50, "dcfr")
        assert np.array_equal(a.rows(0), b.rows(0)), k
        assert np.array_equal(a.rows(1), b.rows(1)), k
        assert a.stats() == b.stats(), k
    assert L.ppsg_tune(eng, TERM_SPLIT, -1) != 0


def test_shape_autotuner_changes_no_rows_and_no_results():
    """ppsg_autotune_shapes times every depth's launches at each candidate
    shape (forward launches rewrite reach only, backward ones run a dry
    update): the rows stay bitwise the same, and a solve continued at the
    tuned shapes equals one at the default shapes (the shape only decides
    which lane takes which combo)."""
    import ctypes
    L = capi.lib()
    L.pps_solver_engine.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(ctypes.c_int)]
    L.ppsg_autotune_shapes.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.ppsg_set_shapes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    s = configs.large_turn()
    a = gpu_solver(s)
    a.solve(30, "dcfr")
    r0, r1 = a.rows(0), a.rows(1)
    eng, slot = ctypes.c_void_p(), ctypes.c_int()
    assert L.pps_solver_engine(a.h, ctypes.byref(eng), ctypes.byref(slot)) == 0
    table = (ctypes.c_int * 1024)()
    assert L.ppsg_autotune_shapes(eng, 2, table, 1024) == 0
    depths = table[0]
    assert depths >= 8
    shapes = [table[1 + k] for k in range(6 * depths)]
    assert set(shapes[0::3]) <= {1, 2, 4} and set(shapes[1::3]) <= {2, 4, 8}
    assert set(shapes[2::3]) <= {1, 2}
    assert np.array_equal(a.rows(0), r0) and np.array_equal(a.rows(1), r1)
    a.continue_solve(20)
    b = gpu_solver(s)
    b.solve(50, "dcfr")
    assert np.array_equal(a.rows(0), b.rows(0)) and np.array_equal(a.rows(1), b.rows(1))
    # every candidate shape at every depth: still the same rows
    for rd, rc, rb in ((1, 2, 2), (4, 8, 1)):
        t = [depths] + [rd, rc, rb] * (2 * depths)
        arr = (ctypes.c_int * len(t))(*t)
        assert L.ppsg_set_shapes(eng, arr, len(t)) == 0
        c = gpu_solver(s)
        c.solve(50, "dcfr")
        a.solve(50, "dcfr")
        assert np.array_equal(a.rows(0), c.rows(0)) and np.array_equal(a.rows(1), c.rows(1))
