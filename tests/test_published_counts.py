"""Tree sizes the reference publishes, asserted for both tree builders.

The reference's performance tables list the node count (and, for flops, the
depth) of the trees its gpu_pfflop builds for named geometries. Node counts
depend only on the geometry (pot, stack, bet and raise menus, street), not
on the ranges, so they pin the bet abstraction and the tree rules
(cuda/gpu_cfr.cu:1116-1326) of the product compiler and of the oracle with
numbers neither of them produced. The ranges are the ones the reference's
bench scripts pass (tools/quality/bench_turn.sh:14-15).
"""
import pytest

import orc
from conftest import board_of
from pps import capi

OOP = "22+,A9s+,KTs+,QJs,JTs,T9s"
IP = "22+,A9s+,KTs+,QJs,JTs,T9s,AQo+"

# label, board, pot, stack, bets, raises, nodes, depth (None: not published)
PUBLISHED = [
    # docs/STATUS.md:67-69, geometries of tools/quality/bench_turn.sh:44-47
    ("turn-shortstack", "Qs9h2d7c", 500, 150, "0.5,0.75", "2.5,3.0", 537, None),
    ("turn-standard", "Qs9h2d7c", 200, 500, "0.5,0.75", "2.5,3.0", 5943, None),
    ("turn-wide", "Qs9h2d7c", 300, 1500, "0.25,0.5,0.75,1.25", "2.5,3.0", 51903, None),
    ("turn-deep", "AhTd9s6h", 300, 1500, "0.25,0.5,0.75,1.25", "2.5,3.0", 51903, None),
    # docs/STATUS.md:80-82: the practical abstraction "0.75,a", raises 2.5x/3x
    ("turn-practical", "Qs9h2d7c", 200, 500, "0.75,a", "2.5,3", 2085, None),
    ("turn-practical-1500", "Qs9h2d7c", 300, 1500, "0.75,a", "2.5,3", 5943, None),
    # docs/STATUS.md:75: flop "standard (200 pot, 500 stack, 2 sizes)", depth 11
    ("flop-standard", "Qs9h2d", 200, 500, "0.5,0.75", "2.5,3.0", 615234, 11),
]
# docs/STATUS.md:76: flop "wide (1500 stack, 4 sizes)", depth 14 (product
# compiler only: the oracle keeps per-node rows and is not sized for it)
WIDE_FLOP = ("flop-wide", "Qs9h2d", 300, 1500, "0.25,0.5,0.75,1.25", "2.5,3.0", 11716550, 14)


def _menu(bets, raises):
    sizes = [float(t) for t in bets.split(",") if t not in ("a", "A", "allin")]
    allin = any(t in ("a", "A", "allin") for t in bets.split(","))
    return sizes, [float(t) for t in raises.split(",")], allin


def _product(board, pot, stack, bets, raises):
    sizes, rais, allin = _menu(bets, raises)
    sp = capi.make_spot(board_of(board), orc.parse_range(OOP), orc.parse_range(IP), pot, stack)
    return capi.compile_info(sp, capi.make_config(sizes, rais, allin, False, 0))


@pytest.mark.parametrize("case", PUBLISHED + [WIDE_FLOP], ids=lambda c: c[0])
def test_product_compiler_builds_the_published_tree_sizes(case):
    _, board, pot, stack, bets, raises, nodes, depth = case
    info = _product(board, pot, stack, bets, raises)
    assert info["nodes"] == nodes
    if depth is not None:
        assert info["max_depth"] == depth


@pytest.mark.parametrize("case", PUBLISHED, ids=lambda c: c[0])
def test_oracle_builds_the_published_tree_sizes(case):
    _, board, pot, stack, bets, raises, nodes, depth = case
    sizes, rais, allin = _menu(bets, raises)
    sp = orc.make_spot(board_of(board), orc.parse_range(OOP), orc.parse_range(IP), pot, stack,
                       bets=sizes, raises=rais, bet_allin=allin)
    c = orc.Oracle(sp).counts()
    assert c["nodes"] == nodes
    if depth is not None:
        assert c["max_depth"] == depth
