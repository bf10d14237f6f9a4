"""The checkpoint Header of incflo::WriteHeader (src/utilities/io.cpp:16-68)
written and read back through the C ABI (host only, no GPU)."""
import numpy as np

from incflo_amd import boxes as B
from incflo_amd import mgx


def test_checkpoint_header_round_trip(tmp_path):
    n = (32, 16, 24)
    bxs = B.chop_domain(n, 16)
    lev = mgx.make_level(n, bxs, prob_lo=(0.0, -1.0, 0.5), prob_hi=(2.0, 1.0, 3.5))
    path = str(tmp_path / "chk00007")
    mgx.write_checkpoint_header(path, [lev], 7, 0.125, 0.01, 0.011, 0.0105)
    text = open(path + "/Header").read().split("\n")
    assert text[0] == "Checkpoint version: 1" and text[1] == "0" and text[2] == "7"
    h = mgx.read_checkpoint_header(path)
    assert h["finest_level"] == 0 and h["nstep"] == 7
    assert (h["cur_time"], h["dt"], h["prev_dt"], h["prev_prev_dt"]) == (0.125, 0.01, 0.011, 0.0105)
    assert h["prob_lo"] == [0.0, -1.0, 0.5] and h["prob_hi"] == [2.0, 1.0, 3.5]
    assert h["boxes"][0] == [(tuple(lo), tuple(hi)) for lo, hi in bxs]
