"""Solution files on the host: the reader of the product (pps_solution_inspect,
no device) against files written here from the PPS1 v1 layout of the
reference (cuda/gpu_cfr.cu:2686-2853): magic "PPS1", version 1, board count +
5 card bytes, pot, stack, both ranges (1326 f64 each), bet and raise size
lists (i32 count + f64 entries), all-in and geometric flags (i8), raise cap
(i32), three f64 thresholds, node count, iterations (i64), algorithm name
(i32 length + bytes), row count (i64), then regret and strategy-sum rows
(f32). Every corruption is refused with the field it is in.
"""
import struct

import numpy as np
import pytest

from pps import capi


def pps1(board=(50, 33, 13, 3), pot=200, stack=500, w0=None, w1=None, bets=(0.75,),
         raises=(2.5, 3.0), allin=1, geo=0, cap=0, thresholds=(1.5, 0.15, 0.1), nodes=2085,
         iterations=300, algo=b"dcfr", rows=None, n_rows=12, magic=b"PPS1", version=1,
         n_board=None):
    w0 = np.linspace(0.0, 1.0, 1326) if w0 is None else np.asarray(w0, dtype=float)
    w1 = np.full(1326, 0.5) if w1 is None else np.asarray(w1, dtype=float)
    cards = list(board) + [0] * (5 - len(board))
    out = magic + struct.pack("<I", version)
    out += struct.pack("<i", len(board) if n_board is None else n_board) + bytes(cards)
    out += struct.pack("<qq", pot, stack) + w0.astype("<f8").tobytes() + w1.astype("<f8").tobytes()
    for lst in (bets, raises):
        out += struct.pack("<i", len(lst)) + struct.pack(f"<{len(lst)}d", *lst)
    out += struct.pack("<bbi", allin, geo, cap) + struct.pack("<3d", *thresholds)
    out += struct.pack("<qq", nodes, iterations) + struct.pack("<i", len(algo)) + algo
    out += struct.pack("<q", n_rows)
    r = np.arange(2 * n_rows, dtype="<f4") if rows is None else rows
    return out + r.tobytes()


def info(tmp_path, data):
    p = tmp_path / "spot.sol"
    p.write_bytes(data)
    return capi.solution_info(p)


def refused(tmp_path, data, *words):
    with pytest.raises(capi.CapiError) as e:
        info(tmp_path, data)
    for w in words:
        assert w in str(e.value), str(e.value)
    return str(e.value)


def test_reads_every_header_field(tmp_path):
    got = info(tmp_path, pps1())
    assert got == {"n_board": 4, "board": [50, 33, 13, 3], "pot": 200, "stack": 500, "n_bets": 1,
                   "n_raises": 2, "nodes": 2085, "iterations": 300, "rows": 12, "algo": "dcfr"}
    got = info(tmp_path, pps1(board=(1, 2, 3), bets=(), raises=(), algo=b"dcfr-g2", n_rows=0))
    assert got["n_board"] == 3 and got["n_bets"] == 0 and got["rows"] == 0
    assert got["algo"] == "dcfr-g2"


def test_refuses_each_corruption_by_field(tmp_path):
    refused(tmp_path, pps1(magic=b"XXXX"), "magic", "not a PPS1")
    refused(tmp_path, pps1(version=2), "version", "version 2")
    refused(tmp_path, pps1(n_board=6), "board size")
    refused(tmp_path, pps1(board=(50, 33, 13, 50)), "board", "appears twice")
    refused(tmp_path, pps1(board=(50, 33, 13, 60)), "board", "card id 60")
    refused(tmp_path, pps1(pot=0), "pot")
    refused(tmp_path, pps1(stack=-5), "stack")
    bad = np.linspace(0.0, 1.0, 1326)
    bad[77] = -1.0
    refused(tmp_path, pps1(w0=bad), "OOP range", "combo 77")
    bad[77] = float("nan")
    refused(tmp_path, pps1(w1=bad), "IP range", "combo 77")
    refused(tmp_path, pps1(bets=tuple([0.5] * 40)), "bet sizes", "40 entries")
    refused(tmp_path, pps1(raises=(2.5, -1.0)), "raise sizes")
    refused(tmp_path, pps1(allin=7), "all-in bet flag")
    refused(tmp_path, pps1(cap=-1), "raise cap")
    refused(tmp_path, pps1(thresholds=(1.5, float("inf"), 0.1)), "thresholds")
    refused(tmp_path, pps1(nodes=0), "node count")
    refused(tmp_path, pps1(iterations=-3), "iterations")
    refused(tmp_path, pps1(algo=b"cfr+"), "algorithm", "cfr+")
    refused(tmp_path, pps1(algo=b""), "algorithm", "length 0")


def test_refuses_truncated_or_padded_files(tmp_path):
    good = pps1(n_rows=12)
    refused(tmp_path, good[:-4], "strategy rows", "needs 48 bytes")
    refused(tmp_path, good[:-60], "regret rows")
    refused(tmp_path, good[:100], "OOP range")
    refused(tmp_path, good[:2], "magic", "needs 4 bytes")
    refused(tmp_path, good + b"\0\0\0", "end of data", "3 more bytes")
    msg = refused(tmp_path, pps1(n_rows=1 << 40, rows=np.zeros(4, dtype="<f4")), "regret rows")
    assert "at byte" in msg
    with pytest.raises(capi.CapiError):
        capi.solution_info(tmp_path / "missing.sol")
