"""Product host layer on the CPU, through the C ABI (no GPU needed):
integer work bit-exact against the reference's golden vectors and the
oracle, compiled trees identical in shape to the oracle's, and the library
exporting every symbol its headers declare."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import orc
from conftest import ROOT, board_of
from pps import capi, configs


def load(golden_dir, name):
    with open(os.path.join(golden_dir, name)) as f:
        return json.load(f)


def test_library_exports_every_declared_symbol():
    lib = capi.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = set()
    for hdr in ("pps_capi.h", "ppsg.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        declared |= set(re.findall(r"\b((?:pps|ppsg)_[a-z_0-9]+)\s*\(", text))
    missing = declared - exported
    assert not missing, missing
    assert set(capi.CAPI_SYMBOLS + capi.PPSG_SYMBOLS) == declared
    capi.lib()  # loads without a GPU


def test_hand_keys_match_reference_golden(golden_dir):
    g = load(golden_dir, "hand_keys.json")
    for h, k in zip(g["hands"], g["keys"]):
        assert capi.hand_key(h) == k, h


def test_ranges_match_reference_golden(golden_dir):
    for e in load(golden_dir, "ranges.json"):
        if e["ok"]:
            want = np.zeros(1326)
            for i, v in e["weights"].items():
                want[int(i)] = v
            assert np.array_equal(capi.parse_range(e["text"]), want), e["text"]
        else:
            with pytest.raises(capi.CapiError) as ei:
                capi.parse_range(e["text"])
            assert str(ei.value) == e["error"], e["text"]


def test_bet_actions_match_reference_golden(golden_dir):
    for c in load(golden_dir, "bet_actions.json"):
        cfg = c["cfg"]
        bc = capi.make_config(cfg["bets"], cfg["raises"], cfg["allin"], cfg["geo"],
                              cfg["max_raises"], c["add"], c["force"], c["merge"])
        got = capi.bet_actions(c["pot_base"], c["stack"], bc, c["sc0"], c["sc1"], c["actor"],
                               c["after_allin"], c["num_bets"], c["num_streets"])
        assert [list(a) for a in got] == c["actions"], c


@pytest.mark.parametrize("algo", ["dcfr", "dcfr-g2", "hs30"])
def test_schedule_equals_oracle(algo):
    for t in list(range(0, 70)) + [255, 256, 1023, 1024, 4097]:
        assert capi.schedule(algo, t, 2000) == orc.schedule(algo, t, 2000), (algo, t)


def _both(s):
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    cf = capi.make_config(s["bets"], s["raises"], s["allin"], s.get("geo", False),
                          s["max_raises"])
    osp = orc.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"], s["bets"],
                        s["raises"], s["allin"], s.get("geo", False), s["max_raises"])
    return capi.compile_info(sp, cf), orc.Oracle(osp)


@pytest.mark.parametrize("maker", [configs.plain_turn, configs.large_turn,
                                   lambda: configs.batch_turn(1003),
                                   lambda: configs.river_spot(7)])
def test_compiled_tree_matches_oracle(maker):
    s = maker()
    info, o = _both(s)
    c = o.counts()
    for k in ("nodes", "max_depth", "decide", "reg_total", "showdown", "fold", "chance"):
        assert info[k] == c[k], k
    assert info["alg_bytes"] == pytest.approx(o.alg_bytes(), rel=1e-12)


def test_shortstack_dead_street_tree():
    # all-in on the turn: the river is a pure runout (dead-street rule)
    s = configs.random_spot(11, 4, 0.2, 500, 150, [0.5, 0.75], [2.5, 3.0], False, 0)
    info, o = _both(s)
    assert info["nodes"] == o.counts()["nodes"]
    assert info["chance"] > 0


def test_flop_tree_compiles_with_geometric_size():
    s = configs.random_spot(5, 3, 0.05, 60, 300, [0.75], [], True, 1)
    s["geo"] = True
    info, o = _both(s)
    assert info["nodes"] == o.counts()["nodes"]
    assert info["reg_total"] == o.counts()["reg_total"]


def test_errors_are_reported_not_raised_across_the_abi():
    with pytest.raises(capi.CapiError, match="bad range token"):
        capi.parse_range("XX")
    sp = capi.make_spot(board_of("Qs9h2d7c"), np.zeros(1326), np.ones(1326), 100, 100)
    with pytest.raises(capi.CapiError, match="empty"):
        capi.compile_info(sp, capi.make_config([0.5]))


def test_arena_footprint_covers_the_rows():
    # host-only planning of a spot's arena block: at least the f32 rows
    # (regret + strategy, two reach layouts, cfv) and grows with the tree
    small, large = configs.batch_turn(1001), configs.batch_turn(1006)
    fp = []
    for s in (small, large):
        sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
        cf = capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"])
        info = capi.compile_info(sp, cf)
        b = capi.compile_footprint(sp, cf)
        assert b >= 2 * 4 * info["reg_total"]
        assert b % 256 == 0
        fp.append(b)
    assert fp[1] > fp[0]


def test_serve_rejects_bad_options_before_touching_the_device():
    s = configs.river_spot(3000)
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    cf = capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"])
    with pytest.raises(capi.CapiError, match="algo"):
        capi.serve([sp], [cf], "hs30")
    with pytest.raises(capi.CapiError, match="bad options"):
        capi.serve([sp], [cf], "dcfr", check_every=0)
    with pytest.raises(ValueError):
        capi.serve([sp], [cf], "dcfr", no_such_option=1)
    assert capi.serve([], [], "dcfr")[0] == []


def test_workload_streams_are_seeded_and_size_filtered():
    a = configs.batch_stream(12, first_seed=1000)
    b = configs.batch_stream(12, first_seed=1000)
    assert [x["nodes"] for x in a] == [x["nodes"] for x in b]
    assert all(2000 <= x["nodes"] <= 60000 for x in a)
    # seed 1003 compiles to 537 nodes: resampled at 1003 + RESAMPLE_STEP
    assert a[3]["name"] == f"batch-turn-{1003 + configs.RESAMPLE_STEP}"
    s = configs.serving_stream(40, qps=10.0, seed=7)
    t = [x["arrival_ms"] for x in s]
    assert t == sorted(t) and t[0] > 0
    rivers = [x for x in s if len(x["board"]) == 5]
    assert 0 < len(rivers) < len(s)
    assert all(2000 <= x["nodes"] <= 60000 for x in s if len(x["board"]) == 4)


def test_suit_isomorphism_shrinks_only_symmetric_trees():
    """Compile only (no device): with suit-symmetric ranges, monotone and
    two-tone boards deal fewer subtrees; a rainbow turn board, or ranges
    with random weights, compile to the same tree."""
    from conftest import board_of

    def info(board, oop, ip, iso):
        w0 = capi.parse_range(oop) if isinstance(oop, str) else oop
        sp = capi.make_spot(board_of(board), w0, capi.parse_range(ip), 60, 200)
        cf = capi.make_config([0.5], [2.5], False, False, 0)
        return capi.compile_info(sp, cf, isomorphism=iso)

    sym = ("TT+,AKo,AQs,KQs", "99+,AJs+,KQo")
    for board, least in (("Ks8s5s2s", 2.0), ("Ks8s5s", 3.0), ("Qs9s2d", 1.5)):
        full, iso = info(board, *sym, False), info(board, *sym, True)
        assert full["nodes"] >= least * iso["nodes"], (board, full["nodes"], iso["nodes"])
        assert iso["alg_bytes"] < full["alg_bytes"]
    assert info("Qs9h2d7c", *sym, False) == info("Qs9h2d7c", *sym, True)
    # a heart combo weighted apart from its club / diamond images: hearts no
    # longer swap, clubs and diamonds still do
    w = capi.parse_range(sym[0])
    ah, kh = board_of("AhKh")
    k = [i for i, (x, y) in enumerate((x, y) for x in range(52) for y in range(x + 1, 52))
         if {x, y} == {ah, kh}][0]
    w[k] = 0.5
    a, b, c = info("Ks8s5s2s", w, sym[1], True), info("Ks8s5s2s", w, sym[1], False), \
        info("Ks8s5s2s", *sym, True)
    assert c["nodes"] < a["nodes"] < b["nodes"]
    s = configs.plain_turn()
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    cf = capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"])
    assert capi.compile_info(sp, cf) == capi.compile_info(sp, cf, isomorphism=True)
