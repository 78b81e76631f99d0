"""The tile-shape table and the autotuner's shape rules, on the CPU: the
autotuner admits exactly the shapes the library admits (same limits as
csrc/bssn_common.h), every committed table entry is such a shape, and the
library refuses a forced shape outside them before touching a GPU."""
import os
import re

import pytest

from sympygr_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_caps():
    text = open(os.path.join(ROOT, "sympygr_amd/csrc/bssn_common.h")).read()
    caps = {k: int(re.search(rf"constexpr int {k} = (\d+);", text).group(1))
            for k in ("GEN_U_CAP", "GEN_DX_CAP", "GEN_DY_CAP", "T")}
    assert "constexpr int TILE_PTS = T * T * T;" in text
    caps["TILE_PTS"] = caps["T"] ** 3
    return caps


def test_autotuner_limits_match_the_library():
    c = _header_caps()
    assert (autotune.U_CAP, autotune.DX_CAP, autotune.DY_CAP, autotune.TILE_PTS) == \
        (c["GEN_U_CAP"], c["GEN_DX_CAP"], c["GEN_DY_CAP"], c["TILE_PTS"])


def test_table_entries_are_admissible_shapes():
    text = open(autotune.TABLE).read()
    rows = re.findall(r"\{\{(\d+), (\d+), (\d+)\}, \{(\d+), (\d+), (\d+)\}\}", text)
    seen = set()
    for r in rows:
        n = tuple(int(x) for x in r[:3])
        s = tuple(int(x) for x in r[3:])
        assert autotune.shape_ok(*s), (n, s)
        assert n not in seen
        seen.add(n)


def test_candidates_cover_8_cubed_and_fit():
    for n in (9, 13, 17, 25, 32):
        cand = autotune.candidates(n, 4)
        assert (8, 8, 8) in cand
        assert all(autotune.shape_ok(*s) for s in cand)
    # a shape that fits a 13-point block exactly beats 8^3 in the model
    assert autotune.cost(13, (13, 13, 3)) < autotune.cost(13, (8, 8, 8))


@pytest.mark.parametrize("bad", ["1,8,8", "16,16,4", "30,8,2", "8,8", "a,b,c"])
def test_library_refuses_bad_forced_shapes(bad, monkeypatch):
    import ctypes
    import sympygr_amd as sgr
    monkeypatch.setenv("SGR_TILE_SHAPE", bad)
    blk = sgr.SgrBlock()
    blk.offset = 0
    for k in range(3):
        blk.sz[k] = 19
        blk.pmin[k] = 0.0
        blk.pmax[k] = 1.8
    blk.bflag = 0
    plan = ctypes.c_void_p()
    prm = sgr.params()
    rc = sgr.lib().sgr_plan_create(ctypes.byref(plan), ctypes.byref(blk), 1, ctypes.byref(prm), 0)
    assert rc == -1  # SGR_E_INVALID
    assert "SGR_TILE_SHAPE" in sgr.last_error()
