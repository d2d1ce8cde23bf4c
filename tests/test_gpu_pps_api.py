"""The reference's pps API tests (python/test_pps.py), run against this
package's `pps` module on the GPU, plus the tiny-truth gate of
tools/quality/gpu_postflop_parity.sh:124-148 through the gpu_pfflop CLI."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pps
from conftest import PKG

pytestmark = pytest.mark.gpu

BOARD, OOP, IP, POT, STACK, BETS, RAISES = "Qs9h2d7c", "TT+,AKo", "AQ+,KQs", 200, 500, "0.75,a", "2.5,3"


def fresh():
    return pps.Solver(board=BOARD, oop=OOP, ip=IP, pot=POT, stack=STACK, bets=BETS, raises=RAISES)


def close(a, b, tol, what):
    d = np.max(np.abs(np.asarray(a, dtype=float) - np.asarray(b, dtype=float)))
    assert d <= tol, f"{what}: max diff {d:g} > {tol:g}"


@pytest.fixture(scope="module")
def solved():
    s = fresh()
    s.solve(300)
    return s


def test_construct_and_solve(solved):
    assert solved.num_nodes > 0 and solved.num_decide_nodes > 0
    assert solved.iterations_run == 300 and solved.total_iterations == 300
    st = solved.stats()
    close(st["ev_oop"] + st["ev_ip"], POT, 0.05, "EV zero-sum")
    assert 0.0 <= st["exploitability"] < 5.0


def test_labeled_tree():
    s = fresh()
    nodes = s.decide_nodes()
    assert len(nodes) == s.num_decide_nodes
    root = nodes[0]
    assert root["player"] == 0 and root["depth"] == 0
    assert [a["kind"] for a in root["actions"]] == ["check", "bet", "allin"]
    assert [a["amount"] for a in root["actions"]] == [0, 150, 500]
    assert len(root["children"]) == len(root["actions"])
    assert all(0 <= c < s.num_nodes for c in root["children"])


def test_combo_strategy(solved):
    for idx in (0, 1, 2):
        r = solved.strategy(idx)
        f = np.asarray(r["freqs"])
        assert f.ndim == 2 and f.shape[0] == len(r["actions"])
        assert f.shape[1] == len(r["cards"]) == len(r["combos"])
        close(f.sum(axis=0), 1.0, 1e-9, f"strategy({idx}) row sums")
        for c in r["cards"]:
            assert len(c) == 4 and c[0] in "23456789TJQKA" and c[1] in "cdhs"
    assert "aggregate" in solved.strategy(0)
    assert "aggregate" in solved.strategy(1)
    idx_deep = next(n["index"] for n in solved.decide_nodes() if n["n_board"] == 5)
    assert "aggregate" not in solved.strategy(idx_deep)


def test_root_ev_consistency(solved):
    ev = solved.root_ev()
    st = solved.stats()
    for p, agg, pl in (("oop", st["ev_oop"], 0), ("ip", st["ev_ip"], 1)):
        w = np.asarray(solved.player_weights(pl))
        m = np.asarray(ev[p]["mass"])
        v = np.asarray(ev[p]["ev"])
        ok = m > 0.0
        close((w[ok] * m[ok] * v[ok]).sum() / (w[ok] * m[ok]).sum(), agg, 2e-3, f"root EV ({p})")


def test_warm_start(solved):
    b = fresh()
    b.solve(150)
    b.continue_solve(150)
    assert b.total_iterations == 300
    sa, sb = solved.stats(), b.stats()
    # this engine is deterministic: the warm start is the same trajectory
    close(sa["ev_oop"], sb["ev_oop"], 1e-12, "warm-start EV oop")
    close(np.asarray(solved.strategy(0)["freqs"]), np.asarray(b.strategy(0)["freqs"]), 0.0,
          "warm-start root freqs")
    b.reset()
    assert b.total_iterations == 0
    b.solve(300)
    close(sa["ev_oop"], b.stats()["ev_oop"], 1e-12, "post-reset EV")


def test_save_load(solved):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "spot.sol")
        solved.save(path)
        t = pps.Solver.load(path)
        assert t.num_nodes == solved.num_nodes
        assert t.total_iterations == 300
        close(solved.stats()["ev_oop"], t.stats()["ev_oop"], 1e-9, "loaded EV")
        close(np.asarray(solved.strategy(1)["freqs"]), np.asarray(t.strategy(1)["freqs"]), 1e-12,
              "loaded freqs")
        u = fresh()
        u.solve(300)
        u.continue_solve(200)
        t.continue_solve(200)
        assert t.total_iterations == 500 == u.total_iterations
        close(u.stats()["ev_oop"], t.stats()["ev_oop"], 1e-12, "post-continue EV")
        with pytest.raises((FileNotFoundError, RuntimeError)):
            pps.Solver.load(os.path.join(d, "nonexistent.sol"))
        with open(path, "r+b") as f:  # wrong magic
            f.write(b"XXXX")
        with pytest.raises(RuntimeError, match="magic.*not a PPS1 solution file"):
            pps.Solver.load(path)


def read_solution_rows(path):
    """The regret and strategy-sum tables of a PPS1 v1 file (layout of
    cuda/gpu_cfr.cu:2686-2853): header fields, then regTotal f32 each."""
    import struct
    with open(path, "rb") as f:
        b = f.read()
    at = 4 + 4 + 4 + 5 + 8 + 8 + 2 * 1326 * 8
    for _ in range(2):  # bet fractions, raise multiples
        (n,) = struct.unpack_from("<i", b, at)
        at += 4 + 8 * n
    at += 1 + 1 + 4 + 3 * 8 + 8 + 8
    (alen,) = struct.unpack_from("<i", b, at)
    at += 4 + alen
    (total,) = struct.unpack_from("<q", b, at)
    at += 8
    reg = np.frombuffer(b, dtype="<f4", count=total, offset=at)
    strat = np.frombuffer(b, dtype="<f4", count=total, offset=at + 4 * total)
    assert at + 8 * total == len(b)
    return reg, strat


def test_solution_file_rows_in_node_order(solved):
    """Solution files hold each decide node's action-major block in node
    order (the reference's table layout), whatever the device layout: the
    strategy sums of decide node i, found at the sum of the earlier nodes'
    block sizes, normalise to strategy(i)."""
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "spot.sol")
        solved.save(path)
        reg, strat = read_solution_rows(path)
    assert len(reg) == len(strat)
    at = 0
    for i, info in enumerate(solved.decide_nodes()):
        na, nc = len(info["actions"]), info["n_combos"]
        blk = strat[at:at + na * nc].astype(float).reshape(na, nc)
        at += na * nc
        s = blk.sum(axis=0)
        want = np.where(s > 0, blk / np.where(s > 0, s, 1), 1.0 / na)
        close(want, np.asarray(solved.strategy(i)["freqs"]), 1e-6, f"decide node {i}")
    assert at == len(strat)


def test_errors(solved):
    with pytest.raises((ValueError, RuntimeError)):
        pps.Solver(board="Qs9h", oop=OOP, ip=IP, pot=POT, stack=STACK)
    with pytest.raises((ValueError, RuntimeError)):
        pps.Solver(board=BOARD, oop="XX", ip=IP, pot=POT, stack=STACK)
    with pytest.raises((IndexError, RuntimeError)):
        solved.strategy(solved.num_decide_nodes + 5)
    with pytest.raises(RuntimeError, match="unknown algo"):
        fresh().solve(10, algo="nope")
    with pytest.raises(RuntimeError, match="no prior solve"):
        fresh().continue_solve(10)
    h = fresh()
    h.solve(10, algo="hs30")
    with pytest.raises(RuntimeError, match="DCFR only"):
        h.continue_solve(10)


def test_target_solve():
    s = fresh()
    s.solve(max_iters=100000, target=1.0)
    assert s.iterations_run < 100000
    assert s.stats()["exploitability"] <= 1.0
    it0 = s.total_iterations
    s.continue_solve(max_iters=100000, target=0.5)
    assert s.total_iterations > it0


def test_node_ev():
    s = fresh()
    s.solve(400)
    st = s.stats()
    ev = s.node_ev(0)
    close(ev["sides"]["oop"]["agg_ev"], st["ev_oop"], 1e-4, "node_ev root agg oop")
    close(ev["sides"]["ip"]["agg_ev"], st["ev_ip"], 1e-4, "node_ev root agg ip")
    side = ev["sides"]["oop"]
    freqs = np.asarray(s.strategy(0)["freqs"])
    mix = (freqs * np.asarray([a["per_combo"] for a in ev["action_ev"]])).sum(axis=0)
    live = np.asarray(side["mass"]) > 0
    close(mix[live], np.asarray(side["ev"])[live], 1e-4, "node_ev action mix identity")
    for a in ev["action_ev"]:
        assert np.all(np.isfinite(np.asarray(a["per_combo"])))
    ev1 = s.node_ev(1)
    assert ev1["decider"] == 1
    assert len(ev1["action_ev"]) == len(ev1["actions"])
    with pytest.raises((IndexError, RuntimeError)):
        s.node_ev(s.num_decide_nodes + 5)


def test_runout_summary():
    s = fresh()
    s.solve(200)
    root = s.decide_nodes()[0]
    # OOP checks, IP checks behind -> deal
    ip_node = next(n for n in s.decide_nodes() if n["node_id"] == root["children"][0])
    rs = s.runout_summary(ip_node["index"], 0)
    # one branch per card off the 4-card board
    assert len(rs["branches"]) == 48
    cards = [b["card"] for b in rs["branches"]]
    assert len(set(cards)) == 48 and not set(cards) & {"Qs", "9h", "2d", "7c"}
    dn = s.decide_nodes()
    for b in rs["branches"]:
        assert 0 <= b["next_decide"] < s.num_decide_nodes
        assert dn[b["next_decide"]]["n_board"] == 5 and dn[b["next_decide"]]["player"] == 0
        close(np.sum(b["freq"]), 1.0, 1e-6, "branch freq sums")
        assert np.isfinite(b["agg_ev"]["oop"]) and np.isfinite(b["agg_ev"]["ip"])
    with pytest.raises(RuntimeError, match="does not lead to a deal"):
        s.runout_summary(0, 1)


def _cli(*args, env=None):
    exe = os.path.join(PKG, "bin", "gpu_pfflop")
    e = dict(os.environ)
    e.update(env or {})
    out = subprocess.run([exe, *args], capture_output=True, text=True, env=e, timeout=120)
    assert out.returncode == 0, out.stderr
    return out


def test_cli_tiny_truth(golden_dir):
    with open(os.path.join(golden_dir, "tiny_truth.json")) as f:
        g = json.load(f)
    for env, tag in (({}, "TINY"), ({"GPU_CFR_SEED_STRAT": "7"}, "M7"),
                     ({"GPU_CFR_SEED_STRAT": "9"}, "M9")):
        out = _cli("postflop", "Qs9h2d7c", "TT+", "99+", "200", "3000", "0.5", "", "0", env=env)
        ev0 = float(next(l for l in out.stdout.splitlines() if l.startswith("EV 0")).split()[2])
        assert abs(ev0 - g[tag]["ev0"]) < 1e-4, (tag, ev0)


def test_cli_protocol():
    out = _cli("postflop", BOARD, OOP, IP, "200", "500", "0.75,a", "2.5,3", "200", "--root",
               "--strat", "1")
    lines = out.stdout.splitlines()
    assert lines[0].startswith("ROOTSTRAT ") and len(lines[0].split()) == 4
    assert lines[1].startswith("STRAT 1 ")
    assert lines[2].startswith("EV 0 ") and lines[3].startswith("EV 1 ")
    assert lines[4].startswith("EXPLOITABILITY ")
    err = out.stderr
    assert "gpu_cfr: nodes=" in err and "gpu_cfr: compile" in err and "iters in" in err


def test_file_written_from_the_layout_loads(tmp_path):
    """A file built here from the PPS1 v1 layout (test_solution_file_cpu.pps1)
    loads into the product with its rows and iteration count, and the
    product's own file reads back through pps_solution_inspect."""
    from pps import capi
    import spots
    from test_solution_file_cpu import pps1
    s = spots.tiny_dict("turn")
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    cf = capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"])
    ci = capi.compile_info(sp, cf)
    rng = np.random.default_rng(3)
    rows = rng.uniform(-5, 5, 2 * ci["reg_total"]).astype("<f4")
    rows[ci["reg_total"]:] = np.abs(rows[ci["reg_total"]:])
    data = pps1(board=tuple(s["board"]), pot=s["pot"], stack=s["stack"], w0=s["w0"], w1=s["w1"],
                bets=tuple(s["bets"]), raises=tuple(s["raises"]), allin=1, cap=s["max_raises"],
                nodes=ci["nodes"], iterations=77, algo=b"dcfr", rows=rows,
                n_rows=ci["reg_total"])
    path = tmp_path / "layout.sol"
    path.write_bytes(data)
    g = capi.Solver(path=str(path))
    assert g.total_iterations == 77
    assert np.array_equal(g.rows(0), rows[:ci["reg_total"]])
    assert np.array_equal(g.rows(1), rows[ci["reg_total"]:])
    g.continue_solve(5)
    assert g.total_iterations == 82
    own = tmp_path / "own.sol"
    g.save(str(own))
    got = capi.solution_info(own)
    assert got["iterations"] == 82 and got["nodes"] == ci["nodes"]
    assert got["rows"] == ci["reg_total"] and got["algo"] == "dcfr"
    assert got["board"] == list(s["board"]) and got["pot"] == s["pot"]


def test_cli_profile_mode():
    """GPU_CFR_PROFILE (cuda/gpu_cfr.cu:1895-1987): at most 64 iterations run
    launch by launch, with per-depth timings and the serialized per-iteration
    total on stderr; only those iterations count."""
    out = _cli("postflop", BOARD, OOP, IP, "200", "500", "0.75,a", "2.5,3", "200",
               env={"GPU_CFR_PROFILE": "1"})
    err = out.stderr.splitlines()
    depth_lines = [l for l in err if l.startswith("gpu_cfr: d")]
    assert len(depth_lines) >= 7 and depth_lines[0].startswith("gpu_cfr: d0 ")
    assert "decide=1 " in depth_lines[0] and " fwd " in depth_lines[0] and " bwd " in depth_lines[0]
    assert any(l.startswith("gpu_cfr: terminals fold") for l in err)
    per = [l for l in err if "per-iter (serialized)" in l]
    assert per and per[0].endswith("over 64 iters")
    assert "gpu_cfr: 64 iters in" in out.stderr
    ev = [float(l.split()[2]) for l in out.stdout.splitlines() if l.startswith("EV ")]
    assert abs(sum(ev) - 200) < 1e-2
