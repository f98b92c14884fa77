"""The library's trellis (mt_rate.hip: the class sort and the lane-group kernels over
trellis_group) compiled with g++ against the CPU emulation of the HIP built-ins
(tests/emu/hip/hip_runtime.h), one thread per lane, equals the oracle's sequential trellis
(trellis_optimize and trellis_optimize_ctx) level for level on random blocks of every
reference size, type, bit depth, quantizer, context and lambda, with the variants mixed in
one wave and alone; an invalid descriptor gets MT_ERR_ARG and keeps its levels.  No GPU is
involved."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import trellisgen as TG

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("emu") / "emu_trellis")
    subprocess.run(["g++", "-std=c++20", "-O1", "-fwrapv", "-ffp-contract=off", "-w", "-I", EMU, "-I",
                    os.path.join(ROOT, "include"), "-o", path, os.path.join(EMU, "emu_trellis.cpp"), "-lpthread"],
                   check=True)
    return path


# mixed variants in one wave (the per-element running sums), only the context variant (the
# four-per-access running sums), only the plain variant
@pytest.mark.parametrize("seed,n,ctx_frac", [(2024, 700, 0.6), (7, 700, 1.0), (8, 400, 0.0)])
def test_emulated_trellis_kernel_matches_oracle(tmp_path, exe, seed, n, ctx_frac):
    desc, levels, targets, want = TG.make(seed, n, ctx_frac)
    desc[5]["tx_size"] = 11  # 32x64: no reference transform
    desc[9]["flags"] |= 4    # an unknown flag
    paths = [str(tmp_path / f) for f in ("b", "l", "t", "ol", "os")]
    desc.tofile(paths[0])
    levels.tofile(paths[1])
    targets.tofile(paths[2])
    subprocess.run([exe] + paths, check=True, capture_output=True)
    got = np.fromfile(paths[3], np.int32)
    st = np.fromfile(paths[4], np.int32)
    bad = np.nonzero(st)[0]
    assert list(bad) == [5, 9]
    for i in (5, 9):  # untouched
        o = int(desc[i]["offset"])
        nxt = int(desc[i + 1]["offset"])
        assert (got[o:nxt] == levels[o:nxt]).all()
        want[o:nxt] = levels[o:nxt]
    diff = np.nonzero(got != want)[0]
    assert len(diff) == 0, f"{len(diff)} levels differ, first at {diff[:5]}"
    assert (want != levels).sum() > 1000  # the trellis changed many levels
