"""pps.Server: spots resident in one arena, solves batched by the driver
thread. Every served result must equal, bit for bit, the same solve on a
solver of its own (pps.Solver), whatever else runs at the same time."""
import gc
import os
import tempfile
import threading

import numpy as np
import pytest

import pps

pytestmark = pytest.mark.gpu

ARENA = 4 << 30

# (board, oop, ip, pot, stack, bets, raises): turn and river spots of
# different sizes and bet trees
SPOTS = [
    ("Qs9h2d7c", "TT+,AKo", "AQ+,KQs", 200, 500, "0.75,a", "2.5,3"),
    ("Kd8s4h2c", "22+,A2s+,KTo+", "55+,ATs+,KQo", 100, 300, "0.5,1", "3"),
    ("Ah7c3d9s", "QQ+,AK", "JJ-22,AQ", 60, 200, "0.33,a", "2.5"),
    ("Qs9h2d7c3s", "TT+,AKo,AQs", "AQ+,KQs,99", 120, 240, "0.5,1,a", "3"),
    ("Jc8d5h", "AA,KK,AKs", "QQ,JJ,AQs", 80, 160, "0.5", ""),
]
# (spot index, iterations, algo, target)
JOBS = [(0, 300, "dcfr", -1.0), (1, 400, "dcfr-g2", -1.0), (2, 250, "hs30", -1.0),
        (3, 2000, "dcfr", 0.6), (4, 300, "dcfr", -1.0)]


def spec(i):
    b, o, p, pot, st, bets, raises = SPOTS[i]
    return dict(board=b, oop=o, ip=p, pot=pot, stack=st, bets=bets, raises=raises)


def result(s):
    """Everything a solve determines: counters, stats, root strategy."""
    st = s.stats()
    strat = s.strategy(0)
    return (s.iterations_run, s.total_iterations, st["ev_oop"], st["ev_ip"], st["exploitability"],
            np.asarray(strat["freqs"]).tobytes())


def own_results():
    out = []
    for i, it, algo, target in JOBS:
        s = pps.Solver(**spec(i))
        s.solve(it, algo, target)
        out.append(result(s))
        del s
        gc.collect()
    return out


@pytest.fixture(scope="module")
def expected():
    return own_results()


def test_jobs_started_together_match_own_solvers(expected):
    srv = pps.Server(arena_bytes=ARENA)
    served = [srv.add(**spec(i)) for i, _, _, _ in JOBS]
    # a long companion job keeps the batch wide while the others start
    companion = srv.add(**spec(0))
    companion.start_solve(2000000)
    for s, (i, it, algo, target) in zip(served, JOBS):
        j = s.start_solve(it, algo, target)
        assert j["kind"] == "solve" and j["total"] == it
    for s in served:
        j = s.wait()
        assert j["status"] == "done" and not j["running"] and j["error"] is None
    j = companion.job()
    assert j["running"] and j["status"] == "running", j
    assert companion.cancel()
    assert companion.wait()["status"] == "cancelled"
    got = [result(s) for s in served]
    for k, (a, b) in enumerate(zip(got, expected)):
        assert a == b, f"job {JOBS[k]}: served {a[:5]} vs own {b[:5]}"
    st = srv.stats()
    assert st["peak_running"] >= 2, st
    assert st["resident"] == len(served) + 1
    # the target job stopped where its own solver stopped, before the cap
    assert got[3][0] == expected[3][0] and got[3][0] < 2000


def test_blocking_solves_from_threads_match(expected):
    srv = pps.Server(arena_bytes=ARENA)
    served = [srv.add(**spec(i)) for i, _, _, _ in JOBS]
    errors = []

    def go(s, job):
        try:
            s.solve(job[1], job[2], job[3])
        except Exception as ex:  # reported below
            errors.append(repr(ex))

    threads = [threading.Thread(target=go, args=(s, j)) for s, j in zip(served, JOBS)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert [result(s) for s in served] == expected


def test_continue_reset_cancel_and_errors():
    own = pps.Solver(**spec(0))
    own.solve(200)
    own.continue_solve(300)
    want = result(own)
    srv = pps.Server(arena_bytes=ARENA)
    s = srv.add(**spec(0))
    s.solve(200)
    s.continue_solve(300)
    assert result(s) == want
    assert s.total_iterations == 500
    # a second job while one runs is refused
    s.start_continue(100000)
    with pytest.raises(RuntimeError):
        s.start_continue(10)
    assert s.cancel()
    j = s.wait()
    assert j["status"] == "cancelled" and 500 < s.total_iterations < 100500
    s.reset()
    assert s.total_iterations == 0
    with pytest.raises(RuntimeError):
        s.continue_solve(10)  # nothing to continue after a reset
    with pytest.raises(RuntimeError):
        s.solve(10, "nope")
    # hs30 cannot be continued, as with a solver of its own
    s.solve(100, "hs30")
    with pytest.raises(RuntimeError):
        s.continue_solve(10)
    # a job that asks for no iterations ends at once
    s.solve(0)
    assert s.job()["status"] == "done" and s.total_iterations == 0


def test_queries_between_steps_and_save_load():
    srv = pps.Server(arena_bytes=ARENA)
    s = srv.add(**spec(1))
    s.start_solve(3000)
    # queries while the job runs see a consistent step boundary
    for _ in range(5):
        st = s.stats()
        assert abs(st["ev_oop"] + st["ev_ip"] - SPOTS[1][3]) < 0.05
        ev = s.node_ev(0)
        assert len(ev["action_ev"]) == len(ev["actions"])
    s.wait()
    own = pps.Solver(**spec(1))
    own.solve(3000)
    assert result(s) == result(own)
    # node EV and runout summaries match the solver of its own exactly
    a, b = s.node_ev(1), own.node_ev(1)
    assert a["sides"]["oop"]["agg_ev"] == b["sides"]["oop"]["agg_ev"]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "x.sol")
        s.save(path)
        t = srv.load(path)
        assert t.total_iterations == 3000
        assert t.stats() == s.stats()
        t.continue_solve(200)
        own.continue_solve(200)
        assert result(t) == result(own)


def test_arena_full_and_removal():
    srv = pps.Server(arena_bytes=64 << 20, max_spots=64)
    held = []
    with pytest.raises(pps.ArenaFull):
        for _ in range(64):
            held.append(srv.add(**spec(1)))
    n = len(held)
    assert n >= 1
    assert srv.stats()["resident"] == n
    held.pop()  # dropping a served solver frees its slot
    gc.collect()
    assert srv.stats()["resident"] == n - 1
    held.append(srv.add(**spec(1)))
    assert srv.stats()["resident"] == n


def test_c_abi_server_matches_own_solvers():
    """The same server through the C ABI (pps_server_*): served handles
    answer every pps_solver_* query, and their solves equal solvers of
    their own."""
    from pps import capi, configs
    spots = [configs.batch_turn(sd) for sd in (1001, 1013, 1002)] + [configs.river_spot(3003)]

    def pair(s):
        return (capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"]),
                capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"]))

    want = []
    for s in spots:
        own = capi.Solver(*pair(s))
        own.solve(256, "dcfr", -1.0)
        want.append((own.total_iterations, own.stats(), own.root_strategy()))
        own.close()
    srv = capi.Server(arena_bytes=ARENA)
    served = [srv.add(*pair(s)) for s in spots]
    for h in served:
        h.start_solve(256)
        with pytest.raises(capi.CapiError):
            h.start_solve(10)  # one job at a time
    for h in served:
        j = h.job(wait=True)
        assert j["status"] == "done" and j["done"] == 256 and not j["running"], j
    got = [(h.total_iterations, h.stats(), h.root_strategy()) for h in served]
    assert got == want
    assert srv.stats()["resident"] == len(spots) and srv.stats()["peak_running"] >= 1
    served[0].solve(100)  # blocking solve through the same call as a solver of its own
    assert served[0].total_iterations == 100
    for h in served:
        h.close()
    assert srv.stats()["resident"] == 0
    srv.close()
