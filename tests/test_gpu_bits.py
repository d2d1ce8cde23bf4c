"""Bit-for-bit pins of the engine's results on fixed solves.

The engine's sums have a fixed order, so a solve is a pure function of its
inputs: these tests record the SHA-256 of both node-order tables (regret,
then strategy sums, f32) after fixed solves of the plain and large turn
spots, and the exact exploitability and EV of a served batch. A change that
only moves work around (launch shapes, tiles, memory layout, load order)
must leave them as they are; a change to the arithmetic shows up here
first and must pass the oracle parity tests (test_gpu_parity.py) before the
recorded values are updated.
"""
import hashlib

import numpy as np
import pytest

from pps import capi, configs

pytestmark = pytest.mark.gpu


def spot_cfg(s):
    return (capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"]),
            capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"]))


@pytest.mark.parametrize("name,iters,digest,ev_oop,expl", [
    ("plain-turn", 100, "cc053de370756b001519e4519520f7ba39ddf1e6bde4cc15caf237a787b228ab",
     47.81108402089765, 0.08781354679636877),
    ("large-turn", 60, "5e709bc7ce5aa73f9315be6b257c3ff8bcb4066639aba12724dc15b1b1964006",
     26.22252828607608, 0.8788157595749126),
])
def test_solve_tables_are_pinned(name, iters, digest, ev_oop, expl):
    spot = {"plain-turn": configs.plain_turn, "large-turn": configs.large_turn}[name]()
    s = capi.Solver(*spot_cfg(spot))
    try:
        s.solve(iters, spot["algo"])
        h = hashlib.sha256()
        for which in (0, 1):
            h.update(np.asarray(s.rows(which), dtype=np.float32).tobytes())
        st = s.stats()
    finally:
        s.close()
    assert h.hexdigest() == digest
    assert st["ev_oop"] == ev_oop and st["exploitability"] == expl


@pytest.mark.parametrize("term_split", [0, 5])
def test_served_batch_is_pinned(term_split):
    """The same served batch with every launch on one stream and with the
    terminals at depth <= 5 on a second stream beside the deeper launches."""
    spots = configs.batch_stream(4)
    pairs = [spot_cfg(s) for s in spots]
    res, _ = capi.serve([p[0] for p in pairs], [p[1] for p in pairs], "dcfr", None,
                        target_frac=0.005, check_every=32, max_iters=4096, real_time=0,
                        prefill=1, term_split=term_split)
    got = [(r["iterations"], r["expl"], r["ev0"]) for r in res]
    assert got == [(128, 0.5087439806226541, 71.6570791858915),
                   (64, 0.28165960617813823, 54.26376348232555),
                   (64, 0.5142293226740549, 58.04267755971645),
                   (64, 0.5652470131910121, 56.25281028438216)]


def test_stream_group_sizes_leave_results_unchanged():
    """A stream gives each spot the terminal group size its live count calls
    for at admission (PPSG_TUNE_ADD_GROUP: 4 up to 31 live spots, 8 from
    32); the spots admitted with different sizes solve exactly as they do
    one at a time."""
    spots = configs.batch_stream(40)
    pairs = [spot_cfg(s) for s in spots]
    S, C = [p[0] for p in pairs], [p[1] for p in pairs]
    opt = dict(target_frac=0.005, check_every=32, max_iters=4096, real_time=0)
    streamed, rep = capi.serve(S, C, "dcfr", None, prefill=0, **opt)
    alone, _ = capi.serve(S, C, "dcfr", None, prefill=1, max_live=1, **opt)
    assert rep["peak_live"] > 32
    got = [(r["iterations"], r["expl"], r["ev0"], r["ev1"]) for r in streamed]
    want = [(r["iterations"], r["expl"], r["ev0"], r["ev1"]) for r in alone]
    assert got == want


def test_solvers_in_two_threads():
    """Two solvers, each with its own engine, solved at the same time from
    two threads: one thread's graph capture must not break the other's
    copies (the engines keep every copy and fill on their own streams), and
    both give the pinned tables."""
    import threading
    cases = {"plain-turn": (configs.plain_turn, 100, "cc053de37075"),
             "large-turn": (configs.large_turn, 60, "5e709bc7ce5a")}
    got, errs = {}, []

    def run(name):
        make, iters, _ = cases[name]
        spot = make()
        try:
            for _ in range(2):  # twice: captures and copies interleave
                s = capi.Solver(*spot_cfg(spot))
                try:
                    s.solve(iters, spot["algo"])
                    h = hashlib.sha256()
                    for which in (0, 1):
                        h.update(np.asarray(s.rows(which), dtype=np.float32).tobytes())
                    got.setdefault(name, []).append(h.hexdigest()[:12])
                finally:
                    s.close()
        except Exception as e:  # surfaced below
            errs.append(f"{name}: {e}")

    th = [threading.Thread(target=run, args=(n,)) for n in cases]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for name, (_, _, digest) in cases.items():
        assert got[name] == [digest, digest], (name, got[name])


def test_multi_iteration_replays_match_single_iteration_replays():
    """A one-spot solver replays 8 iterations per graph once its lists are
    stable (here in the second and third of its 128-iteration chunks); a
    served spot, on a batch engine, replays one iteration per graph. The
    same spot for 300 iterations both ways: the same EVs, exploitability
    and root strategy."""
    spot = configs.batch_turn(1001)
    sp, cf = spot_cfg(spot)
    r = capi.serve([sp], [cf], "dcfr", None, target_frac=0.0, check_every=300, max_iters=300,
                   real_time=0, prefill=1)[0][0]
    assert r["ok"] and r["iterations"] == 300
    g = capi.Solver(sp, cf)
    try:
        g.solve(300, "dcfr")
        st = g.stats()
        # (the served figures come from the device root sums, the solver's
        # from its own walks: the same f64 operations, held to 1e-12 as the
        # batch tests hold them)
        assert abs(st["ev_oop"] - r["ev0"]) <= 1e-12 and abs(st["ev_ip"] - r["ev1"]) <= 1e-12
        assert abs(st["exploitability"] - r["expl"]) <= 1e-12
        assert max(abs(a - b) for a, b in zip(g.root_strategy(), r["root_strategy"])) <= 1e-12
    finally:
        g.close()
