"""mt_trellis on the GPU against the oracle's trellis (trellis_optimize and
trellis_optimize_ctx of the reference's src/trellis.rs), level for level: random blocks of
every reference size, transform type, bit depth, luma and chroma quantizers, QM, CDF
snapshot, dc-sign context, lambda and band tilt, with the forward transform's own levels and
targets; invalid descriptors; the reference's own random blocks at 16x16."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
import trellisgen as TG

pytestmark = pytest.mark.gpu


def run(dev, desc, levels, targets):
    b = torch.from_numpy(desc.view(np.uint8).copy()).to(dev)
    lv = torch.from_numpy(levels.copy()).to(dev)
    tg = torch.from_numpy(targets).to(dev)
    wk = torch.full_like(tg, float("nan"))
    st = torch.full((len(desc),), 7, dtype=torch.int32, device=dev)
    mt.trellis(b, lv, tg, wk, st)
    torch.cuda.synchronize()
    return lv.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("seed,n,ctx_frac", [(1, 4000, 0.6), (2, 4000, 0.6), (3, 20000, 0.6), (4, 8000, 1.0),
                                              (5, 4000, 0.0)])
def test_trellis_matches_oracle(dev, seed, n, ctx_frac):
    """Variants mixed in each wave, the context variant alone, the plain variant alone."""
    desc, levels, targets, want = TG.make(seed, n, ctx_frac)
    print(f"trellis seed {seed}: {n} blocks, context variant {ctx_frac}", flush=True)
    got, st = run(dev, desc, levels, targets)
    assert not st.any()
    diff = np.nonzero(got != want)[0]
    assert len(diff) == 0, f"{len(diff)} levels differ, first at {diff[:5]}"
    assert (want != levels).sum() > 0


def test_trellis_invalid_descriptors(dev):
    desc, levels, targets, want = TG.make(9, 50)
    desc[3]["tx_size"] = 12  # 64x32: no reference transform
    desc[7]["flags"] = 0x81
    desc[11]["bd_qm"] = 3 << 4  # no such bit depth
    i = int(np.nonzero(desc["flags"] & mt.TRELLIS_CTX)[0][-1])
    desc[i]["eob_bin"] = 16 if desc[i]["eob_bin"] != 16 else 0  # not the block's coefficient count
    got, st = run(dev, desc, levels, targets)
    bad = sorted({3, 7, 11, i})
    assert list(np.nonzero(st)[0]) == bad and (st[bad] == -1).all()
    for k in bad:
        o = int(desc[k]["offset"])
        e = int(desc[k + 1]["offset"]) if k + 1 < len(desc) else len(levels)
        assert (got[o:e] == levels[o:e]).all()


@pytest.mark.gpu
def test_trellis_lambda0_law():
    """trellis_lambda() * hbd_trellis_scale: 0.05 at 8 bits and at 4:4:4; above 8 bits at
    4:2:0 the scale ramps from 1 at qindex 36 to 0.174 at 20 and below."""
    assert mt.trellis_lambda0(10, 8, 2) == np.float32(0.05)
    assert mt.trellis_lambda0(10, 10, 0) == np.float32(0.05)
    assert mt.trellis_lambda0(40, 10, 2) == np.float32(0.05)
    assert abs(mt.trellis_lambda0(10, 10, 2) - 0.05 * 0.174) < 1e-7
    assert abs(mt.trellis_lambda0(28, 12, 2) - 0.05 * (1 + (0.174 - 1) * 0.5)) < 1e-7


@pytest.mark.parametrize("seed,n,all_sizes", [(0x6D61726F6F6E, 1 << 16, False), (77, 20000, True)])
def test_chain_with_trellis_matches_oracle(dev, seed, n, all_sizes):
    """mt_tx_chain_trellis (forward + quantization, trellis, dequantization + inverse, SSE,
    checksum of the final levels) against the oracle's chain with the same trellis, block
    for block; unsupported sizes get sse = -1."""
    blocks, pool = mt.gen_tx_blocks(seed, n, all_sizes)
    tr = TG.chain_descriptors(blocks, seed & 0xFFFF)
    print(f"chain+trellis seed {seed:#x}: {n} blocks", flush=True)
    t = lambda a: torch.from_numpy(a.view(np.uint8).copy()).to(dev)  # noqa: E731
    pt = torch.from_numpy(pool).to(dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    lv = torch.zeros(len(pool), dtype=torch.int32, device=dev)
    tg = torch.zeros(len(pool), dtype=torch.float32, device=dev)
    rec = torch.zeros(len(pool), dtype=torch.int32, device=dev)
    mt.tx_chain_trellis(t(blocks), t(tr), pt, res, lv, tg, rec)
    torch.cuda.synchronize()
    got = res.cpu().numpy().view(mt.RES_DTYPE)
    sse, h = TG.oracle_chain(blocks, pool, tr)
    bad = np.nonzero((got["sse"] != sse) | (got["hash"] != h))[0]
    assert len(bad) == 0, (len(bad), bad[:5])


def test_slow_chain_descriptors_match_oracle(dev):
    """The Slow-mode descriptors of the chain benchmark (workload.chain_trellis_descriptors:
    the context trellis on every block with its qindex context and frame lambda) at a slice
    of the benchmark's seeded batch: block for block against the oracle."""
    from maroontree_amd import workload as W
    blocks, pool = mt.gen_tx_blocks(0x6D61726F6F6E, 1 << 16)
    tr = W.chain_trellis_descriptors(blocks, sub=2)
    print(f"slow chain descriptors: {len(blocks)} blocks", flush=True)
    t = lambda a: torch.from_numpy(a.view(np.uint8).copy()).to(dev)  # noqa: E731
    pt = torch.from_numpy(pool).to(dev)
    n = len(blocks)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    lv = torch.zeros(len(pool), dtype=torch.int32, device=dev)
    tg = torch.zeros(len(pool), dtype=torch.float32, device=dev)
    rec = torch.zeros(len(pool), dtype=torch.int32, device=dev)
    mt.tx_chain_trellis(t(blocks), t(tr), pt, res, lv, tg, rec)
    torch.cuda.synchronize()
    got = res.cpu().numpy().view(mt.RES_DTYPE)
    sse, h = TG.oracle_chain(blocks, pool, tr)
    assert (sse >= 0).all()
    bad = np.nonzero((got["sse"] != sse) | (got["hash"] != h))[0]
    assert len(bad) == 0, (len(bad), bad[:5])
