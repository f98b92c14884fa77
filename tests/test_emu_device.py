"""The library's device sources run on the CPU: tests/emu/hip/hip_runtime.h emulates the HIP
built-ins the kernels use (a thread per lane, barriers, shuffles, ballots; __mul24 checks
its 24-bit operand range), g++ compiles mtgpu.hip and mt_rdo.hip against it, and
  - mt_tx_chain (bucketing, every tile size, the f64 64x64 kernel) on blocks with extreme
    12-bit residuals at qindex 0..6, with and without QM, equals the oracle's chain (SSE and
    checksum of every block) without a 24-bit range violation;
  - mt_intra_rdo (plan, candidate generation, argmin, decisions, the superblock
    reconstruction kernel, final accounting; candidate evaluation through the oracle) equals
    the oracle's RDO: reconstruction, every node record, the stats; also with the trellis
    (MT_RDO_TRELLIS) in the candidates and in the reconstruction;
  - mt_tx_chain_trellis (the chain kernels' split halves and the trellis) equals the
    oracle's chain with the trellis.
No GPU is involved; the same kernels have their GPU parity tests."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import _lib
from maroontree_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
FLAGS = ["-std=c++20", "-O1", "-fwrapv", "-ffp-contract=off", "-w", "-I", EMU, "-I", os.path.join(ROOT, "include")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("emu")
    O.lib()  # the oracle library is built
    gpu_o = str(d / "emu_gpu.o")
    subprocess.run(["g++", *FLAGS, "-c", os.path.join(EMU, "emu_gpu.cpp"), "-o", gpu_o], check=True)
    rate_o = str(d / "emu_rate.o")
    subprocess.run(["g++", *FLAGS, "-c", os.path.join(EMU, "emu_rate.cpp"), "-o", rate_o], check=True)
    chain = str(d / "emu_chain")
    subprocess.run(["g++", *FLAGS, "-o", chain, os.path.join(EMU, "emu_chain.cpp"), gpu_o, rate_o, "-lpthread"],
                   check=True)
    chain_tr = str(d / "emu_chain_trellis")
    subprocess.run(["g++", *FLAGS, "-o", chain_tr, os.path.join(EMU, "emu_chain_trellis.cpp"), gpu_o, rate_o,
                    "-lpthread"], check=True)
    rdo = str(d / "emu_rdo")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", *FLAGS, "-o", rdo, os.path.join(EMU, "emu_rdo.cpp"), gpu_o, rate_o, "-L" + odir,
                    "-loracle", "-lpthread", "-Wl,-rpath," + odir], check=True)
    return d, chain, rdo, chain_tr


def test_emulated_chain_matches_oracle_on_extreme_inputs(bins):
    d, chain, _, _ = bins
    n = 1400
    blocks, pool = mt.gen_tx_blocks(77, n, all_sizes=True)
    pool = pool.copy()
    rng = np.random.default_rng(5)
    for i in range(n):
        s = int(blocks["tx_size"][i])
        w, h = mt.tx_width(s), mt.tx_height(s)
        o = int(blocks["offset"][i])
        kind = i % 5
        if kind == 0:
            continue  # the generator's residual and quantizer
        blocks["bd_qm"][i] = (2 << 4) | int(rng.choice([15, 0, 4, 8]))
        blocks["qidx"][i] = int(rng.integers(0, 7))
        y, x = np.mgrid[0:h, 0:w]
        if kind == 1:
            r = np.where((x + y) % 2 == 0, 4095, -4095)
        elif kind == 2:
            r = np.full((h, w), 4095 if i % 2 else -4095)
        elif kind == 3:
            r = np.where(x % 2 == 0, 4095, -4095) * np.where(y < h // 2, 1, -1)
        else:
            r = rng.choice([-4095, 4095], (h, w))
        pool[o:o + w * h] = r.reshape(-1).astype(np.int16)
    blocks.tofile(d / "blocks.bin")
    pool.tofile(d / "pool.bin")
    r = subprocess.run([chain, str(d / "blocks.bin"), str(d / "pool.bin"), str(d / "res.bin")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.fromfile(d / "res.bin", _lib.RES_DTYPE)
    bdi = (blocks["bd_qm"] >> 4) & 3
    bd = np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8)).astype(np.uint8)
    sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], bd, blocks["bd_qm"] & 15,
                           blocks["offset"].astype(np.uint64), pool)
    bad = np.nonzero((got["sse"] != sse) | (got["hash"] != h))[0]
    assert len(bad) == 0, bad[:10]
    assert (blocks["tx_size"] == 4).any()  # the 64x64 kernel ran


@pytest.mark.parametrize("bd,qidx,w,h,flags", [(10, 77, 128, 72, 7), (8, 160, 72, 64, 3), (10, 30, 128, 72, 7 | 32),
                                               (8, 120, 136, 64, 1 | 4 | 32)])
def test_emulated_rdo_matches_oracle(bins, bd, qidx, w, h, flags):
    d, _, rdo, _ = bins
    img = W.synth_image(100 + qidx, w, h, bd)
    path = d / f"img_{qidx}.raw"
    img.astype("<u2").tofile(path)
    qm = mt.qm_level_law(qidx, 0)
    r = subprocess.run([rdo, str(path), str(w), str(h), str(bd), str(qidx), str(qm), str(flags)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    emu = lines[0].split("bits ")[1]
    orc = lines[1].split("bits ")[1]
    assert emu == orc, (emu, orc)
    assert lines[-1] == "recon diffs 0 node diffs 0", lines[-1]


def test_emulated_trellis_chain_matches_oracle(bins):
    """mt_tx_chain_trellis: bucketing, the forward halves of every chain kernel, the trellis,
    the inverse halves with their measurement; SSE and checksum of every block equal the
    oracle's chain with the trellis."""
    import trellisgen as TG
    d, _, _, chain_tr = bins
    blocks, pool = mt.gen_tx_blocks(91, 900, all_sizes=True)
    tr = TG.chain_descriptors(blocks, 5)
    paths = [str(d / f) for f in ("tb", "tt", "tp", "to")]
    blocks.tofile(paths[0])
    tr.tofile(paths[1])
    pool.tofile(paths[2])
    subprocess.run([chain_tr] + paths, check=True, capture_output=True)
    got = np.fromfile(paths[3], _lib.RES_DTYPE)
    sse, h = TG.oracle_chain(blocks, pool, tr)
    bad = np.nonzero((got["sse"] != sse) | (got["hash"] != h))[0]
    assert len(bad) == 0, (len(bad), bad[:5], got[bad[:3]], sse[bad[:3]])
    plain, ph = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"],
                              np.where(((blocks["bd_qm"] >> 4) & 3) == 1, 10,
                                       np.where(((blocks["bd_qm"] >> 4) & 3) == 2, 12, 8)),
                              blocks["bd_qm"] & 0x4F, blocks["offset"].astype(np.uint64), pool)
    assert (ph != h).sum() > 100  # the trellis changed many blocks
