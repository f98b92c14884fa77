"""The deblocking kernels (maroontree_amd/csrc/mt_deblock.hip) run on the CPU through the lane
emulation of tests/emu/hip/hip_runtime.h (a thread per lane, the workgroup barriers, LDS as
static storage) and equal the oracle's sequential deblocking (oracle/oracle_lf.cpp):
  - mt_deblock_frame on random quadtree partitions at 4:2:0 and 4:4:4, 8 / 10 / 12-bit,
    frames of several 64 x 64 tiles with partial tiles, visible sizes below the coded size,
    chroma levels 0, luma level 0 (a copy), with a zero status word;
  - mt_deblock_plane with the alignment rule (no edge flags), as the chroma planes use it;
  - maps that place two edges within each other's reach raise MT_DEBLOCK_ORDER.
No GPU is involved; the GPU parity tests run the same kernels."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import lfgen
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
FLAGS = ["-std=c++20", "-O1", "-fwrapv", "-ffp-contract=off", "-w", "-I", EMU, "-I", os.path.join(ROOT, "include")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("emu_lf")
    out = str(d / "emu_deblock")
    subprocess.run(["g++", *FLAGS, "-o", out, os.path.join(EMU, "emu_deblock.cpp"), "-lpthread"], check=True)
    return out


def _run(exe, d, params, planes, maps):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write(params)
    for k, p in planes.items():
        np.ascontiguousarray(p, np.uint16).tofile(os.path.join(d, k + ".bin"))
    np.concatenate([np.ascontiguousarray(m, np.uint8).reshape(-1) for m in maps]).tofile(os.path.join(d, "maps.bin"))
    subprocess.run([exe, d], check=True, timeout=600)
    rc, st = open(os.path.join(d, "status.txt")).read().split()
    outs = {k: np.fromfile(os.path.join(d, "o" + k + ".bin"), np.uint16).reshape(planes[k].shape)
            for k in planes}
    return int(rc), int(st), outs


@pytest.mark.parametrize("seed,w8,h8,sub,bd,lv", [
    (11, 136, 96, 1, 8, (12, 4, 6)), (12, 128, 136, 0, 10, (30, 20, 0)), (13, 72, 200, 1, 12, (63, 32, 32)),
    (14, 200, 72, 0, 8, (0, 5, 5)), (15, 64, 64, 1, 8, (20, 0, 0))])
def test_emulated_frame_matches_oracle(exe, tmp_path, seed, w8, h8, sub, bd, lv):
    rng = np.random.default_rng(seed)
    leaves = lfgen.random_leaves(rng, w8, h8)
    m = lfgen.maps_from_leaves(leaves, w8, h8)
    y = lfgen.blocky_plane(rng, leaves, w8, h8, bd)
    u = lfgen.blocky_plane(rng, leaves, w8 >> sub, h8 >> sub, bd, sub)
    v = lfgen.blocky_plane(rng, leaves, w8 >> sub, h8 >> sub, bd, sub)
    dw, dh = w8 - int(rng.integers(0, 8)), h8 - int(rng.integers(0, 8))
    sharp = int(rng.integers(0, 8))
    maps = [m[k] for k in ("blk4", "blk4h", "blk4v", "blk4t", "pblk4", "pblk4h")]
    params = f"frame {w8} {h8} {dw} {dh} {sub} {sub} 0 {sharp} {bd} {lv[0]} {lv[1]} {lv[2]}\n"
    rc, st, got = _run(exe, str(tmp_path / "f"), params, {"y": y, "u": u, "v": v}, maps)
    assert rc == 0 and st == 0
    want = [y.copy(), u.copy(), v.copy()]
    O.frame_deblock(want[0], want[1], want[2], dw, dh, m, sub, sub, sharp, bd, *lv)
    for k, w in zip("yuv", want):
        assert (got[k] == w).all(), k


def test_emulated_plane_alignment_rule(exe, tmp_path):
    rng = np.random.default_rng(21)
    w, h = 96, 72
    leaves = lfgen.random_leaves(rng, w, h)
    m = lfgen.maps_from_leaves(leaves, w, h)
    y = lfgen.blocky_plane(rng, leaves, w, h, 10)
    for luma in (0, 1):
        params = f"plane {w} {h} {w - 3} {h - 5} 0 25 3 {luma} 16 10\n"
        rc, st, got = _run(exe, str(tmp_path / f"p{luma}"), params, {"y": y}, [m["blk4"], m["blk4h"]])
        assert rc == 0 and st == 0
        want = y.copy()
        O.filter_plane(want, w - 3, h - 5, m["blk4"], m["blk4h"], None, None, 25, 3, luma, 16, 10)
        assert (got["y"] == want).all(), luma


def test_emulated_overlapping_edges_are_reported(exe, tmp_path):
    # 32-wide transforms claimed in the maps, but an edge flagged at every 4-unit column:
    # 16-wide filters 4 samples apart
    w, h = 64, 64
    bw = np.full((h // 4, w // 4), 8, np.uint8)
    flags = np.ones_like(bw)
    y = np.full((h, w), 512, np.uint16)
    rc, st, _ = _run(exe, str(tmp_path / "c"), f"plane {w} {h} {w} {h} 1 20 0 1 16 10\n", {"y": y},
                     [bw, bw, flags, flags])
    assert rc == 0 and st & 1
