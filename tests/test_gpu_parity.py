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


@pytest.mark.parametrize("name,steps", [("river", [0, 1, 2, 7, 30]), ("turn", [0, 1, 4, 15])])
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


@pytest.mark.parametrize("name", ["river", "turn"])
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
