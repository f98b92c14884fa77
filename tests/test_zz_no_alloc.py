"""mt_tx_chain_noalloc (the allocation-free chain on the library's fixed device workspace)
against mt_tx_chain through the C ABI: every result, level, target and reconstructed
residual equal bit for bit, on two calls in a row (the second smaller, so stale counters or
order entries of the first would show), plus the size limit."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _both(dev, seed, n):
    blocks, pool = mt.gen_tx_blocks(seed, n)
    print(f"no-alloc chain: seed={seed:#x} n={n} pool={len(pool)}", flush=True)
    bt = torch.from_numpy(blocks.view(np.uint8).copy()).to(dev)
    pt = torch.from_numpy(pool).to(dev)
    outs = []
    for fn in (mt.tx_chain, mt.tx_chain_noalloc):
        res = torch.full((n * 16,), 0xA5, dtype=torch.uint8, device=dev)
        lev = torch.full((len(pool),), 0x7F7F7F7F, dtype=torch.int32, device=dev)
        tgt = torch.full((len(pool),), float("nan"), dtype=torch.float32, device=dev)
        rec = torch.full((len(pool),), -12345, dtype=torch.int32, device=dev)
        fn(bt, pt, res, lev, tgt, rec)
        torch.cuda.synchronize()
        outs.append((res.cpu().numpy(), lev.cpu().numpy(), tgt.cpu().numpy().view(np.uint32),
                     rec.cpu().numpy()))
    return blocks, pool, outs


def test_noalloc_chain_equals_chain_on_two_calls(dev):
    for seed, n in ((0x6D61726F6F6E, 60000), (0x1234, 7001)):
        blocks, pool, (a, b) = _both(dev, seed, n)
        for name, x, y in zip(("res", "levels", "targets", "recon"), a, b):
            bad = np.nonzero(x != y)[0]
            assert len(bad) == 0, f"seed {seed:#x}: {name} differs at {len(bad)} places, first {bad[:4]}"
        got = b[0].view(mt.RES_DTYPE)
        bdi = (blocks["bd_qm"] >> 4) & 3
        sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"],
                               np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8)),
                               blocks["bd_qm"] & 15, blocks["offset"].astype(np.uint64), pool)
        assert (got["sse"] == sse).all() and (got["hash"] == h).all()


def test_noalloc_chain_rejects_more_than_its_maximum(dev):
    n = mt.NOALLOC_MAX_BLOCKS + 1
    blocks = torch.zeros(n * 8, dtype=torch.uint8, device=dev)
    pool = torch.zeros(16, dtype=torch.int16, device=dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    with pytest.raises(mt.MtError, match="MT_NOALLOC_MAX_BLOCKS"):
        mt.tx_chain_noalloc(blocks, pool, res)
