"""Host code under the address and undefined-behaviour sanitizers (CPU only).

* The library's ghost-table builder (sympygr_amd/csrc/ghost_table.cpp, plain
  C++) is compiled with g++ -fsanitize=address,undefined together with a
  small driver (tests/ghost_table_sanitize.cpp) and run over the benchmark
  octree, the small mixed-size octree, a uniform tiling and two refused sets;
  its statistics must equal the library's own build of the same code.
* The CPU oracle (oracle/bssn_oracle.cpp, oracle/ghost_oracle.cpp) is
  compiled with the same sanitizers and loaded into a child Python with the
  sanitizer runtime preloaded; its RHS, constraints and ghost refresh must
  equal the regular build's bit for bit.
"""
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

import sympygr_amd as sgr
from sympygr_amd import data

from helpers import small_octree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sympygr_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def _gxx(args, cwd):
    r = subprocess.run(["g++"] + args, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.fixture(scope="module")
def ghost_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("asan_ghost")
    exe = str(d / "ghost_table_sanitize")
    _gxx(["-std=c++17"] + SAN + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", CSRC,
                                 os.path.join(ROOT, "tests", "ghost_table_sanitize.cpp"),
                                 os.path.join(CSRC, "ghost_table.cpp"), "-o", exe], cwd=str(d))
    return exe, d


def _write_blocks(path, bl):
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(bl)))
        for b in range(len(bl)):
            f.write(struct.pack("<Q3I I3d3d", bl.offsets[b], *bl.sz[b], bl.bflag[b], *bl.pmin[b], *bl.pmax[b]))


def _run(ghost_exe, bl, name):
    exe, d = ghost_exe
    path = str(d / f"{name}.bin")
    _write_blocks(path, bl)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, path], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (name, r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
    lines = r.stdout.strip().splitlines()
    complete, patches, nodes, interp, boxes = (int(v) for v in lines[0].split())
    return complete, patches, nodes, interp, boxes, (lines[1] if len(lines) > 1 else "")


def _overlap_set():
    """Two sources whose owned boxes overlap in a sliver that holds a ghost
    node of a third block, but neither reaches the other's interior."""
    bl = data.BlockList()
    n = 13
    for pmin, flags in (((0.0, 0.0, 0.0), 1 | 4 | 8 | 16 | 32),
                        ((6.75, 0.0, 0.0), 2 | 4 | 8 | 16 | 32),
                        ((3.25, 7.0, 0.0), 1 | 2 | 8 | 16 | 32)):
        bl.add((n, n, n), pmin, tuple(p + (n - 1) for p in pmin), flags)
    return bl


@pytest.mark.parametrize("case", ["bench_octree", "small_octree", "uniform17", "overlap", "gap"])
def test_ghost_table_builder_under_sanitizers(ghost_exe, case):
    if case == "bench_octree":
        bl = data.octree_blocks()
    elif case == "small_octree":
        bl = small_octree()
    elif case == "uniform17":
        bl = data.uniform_tiling(nb=3, n=17, lo=-1.0, hi=1.0, stagger=0.5, faces=True)
    elif case == "overlap":
        bl = _overlap_set()
    else:  # no flagged faces: the outer ghost nodes have no owner
        bl = data.uniform_tiling(nb=2, n=16, lo=-1.0, hi=1.0, stagger=0.5, faces=False)
    complete, patches, nodes, interp, boxes, msg = _run(ghost_exe, bl, case)
    if case == "overlap":
        assert not complete and "two blocks" in msg, msg
    elif case == "gap":
        assert not complete and "no owning block" in msg, msg
    else:
        assert complete, msg
        st = sgr.ghost_table_stats(bl)
        assert (patches, nodes, interp) == (st["patches"], st["nodes"], st["interpolated"])
        assert boxes > 0


def test_library_refuses_a_doubly_owned_ghost_node():
    with pytest.raises(sgr.SgrError, match="two blocks"):
        sgr.ghost_table_stats(_overlap_set())


CHILD = r"""
import os, sys
import numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
import pyoracle as po
from sympygr_amd import data
sys.path.insert(0, os.path.join(sys.argv[1], "tests"))
from helpers import small_octree
u = data.perturbed_flat_block(seed=2, n=22, amp=1e-2)
sz, pmin, pmax = (22, 22, 22), (-1.05,) * 3, (1.05,) * 3
out = [po.rhs_block(u, sz, pmin, pmax, bf, po.make_params(upwind=up)) for bf in (0, 63) for up in (True, False)]
out.append(po.constraints_block(u, sz, pmin, pmax, 21))
bl = small_octree()
v = data.fill_blocks(bl, kind="perturbed", seed=4, amp=1e-2, lo=-2.0, L=4.0)
out.append(po.ghost_fill(v, bl))
np.savez(sys.argv[2], *out)
"""


def test_oracle_under_sanitizers(tmp_path):
    lib = str(tmp_path / "liboracle_asan.so")
    _gxx(["-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off"] + SAN +
         [os.path.join(ROOT, "oracle", "bssn_oracle.cpp"), os.path.join(ROOT, "oracle", "ghost_oracle.cpp"),
          "-o", lib], cwd=str(tmp_path))
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["g++", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    res = {}
    for tag, extra in (("asan", {"SGR_ORACLE_LIB": lib, "LD_PRELOAD": f"{asan}:{ubsan}",
                                 "ASAN_OPTIONS": "detect_leaks=0:verify_asan_link_order=0"}),
                       ("plain", {})):
        env = dict(os.environ, OMP_NUM_THREADS="2", **extra)
        env.pop("SGR_ORACLE_LIB", None) if tag == "plain" else None
        r = subprocess.run([sys.executable, str(script), ROOT, str(tmp_path / f"{tag}.npz")], env=env,
                           capture_output=True, text=True, timeout=900)
        assert r.returncode == 0, (tag, r.stderr[-4000:])
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
        res[tag] = np.load(tmp_path / f"{tag}.npz")
    for k in res["plain"].files:
        np.testing.assert_array_equal(res["asan"][k], res["plain"][k], err_msg=k)
