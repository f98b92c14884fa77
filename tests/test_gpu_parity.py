"""Bit-exact parity of the HIP path (through the C ABI) with the CPU oracle.

Forward and inverse entry points per (size, type) on the reference's own test inputs
(src/dct.rs:3381-3429, src/idct.rs:734-827) plus random and extreme inputs; the batched chain
on the microbench generator, including every block of the full 2^20-block workload.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
from refgen import QUANT_PAIRS, forward_parity_inputs, itx_levels

pytestmark = pytest.mark.gpu

REF_SIZES = mt.TX_SIZES_REF
TYPES = mt.TX_TYPES_REF
SEED = 0x6D61726F6F6E


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def orc_q(q):
    """MtQuant -> oracle quant with the same fields."""
    o = O.OrcQuant()
    for n, _ in mt.MtQuant._fields_:
        setattr(o, n, getattr(q, n))
    return o


def quant_set():
    qs = [mt.quant_new(0, 8), mt.quant_new(60, 8), mt.quant_new(128, 8, 0), mt.quant_new(255, 8),
          mt.quant_new(40, 8, 0), mt.quant_new(96, 10, 7), mt.quant_new(180, 12, 14),
          mt.quant_new(91, 10, mt.qm_level_law(91, 0)), mt.quant_new(17, 10, 12),
          mt.quant_new_chroma(128, -7, 11, 12, 3), mt.quant_new_chroma(40, 0, 0, 10, 15)]
    # the raw multiplier pairs of the reference's selected-vs-scalar test
    for dc, ac in QUANT_PAIRS:
        q = mt.quant_new(100, 8)
        q.q_mult_dc, q.q_mult_ac = dc, ac
        qs.append(q)
    return qs


def fwd_inputs(n, rng):
    ins = [v for _, v in forward_parity_inputs(n)]
    for lim in (255, 1023, 4095):
        for _ in range(4):
            g1 = np.floor(-12.0 * np.log(1.0 - rng.random(n)))
            g2 = np.floor(-12.0 * np.log(1.0 - rng.random(n)))
            ins.append(np.clip(g1 - g2, -lim, lim).astype(np.int32))
        ins.append(rng.integers(-lim, lim + 1, n).astype(np.int32))
    return ins


@pytest.mark.parametrize("s", REF_SIZES, ids=[mt.TX_NAMES[s] for s in REF_SIZES])
def test_fwd_txfm_parity(dev, s):
    w, h, nc = mt.tx_width(s), mt.tx_height(s), mt.tx_ncoef(s)
    rng = np.random.default_rng(1000 + s)
    ins = fwd_inputs(w * h, rng)
    qs = quant_set()
    for t in TYPES:
        if not mt.tx_supported(s, t):
            continue
        say(f"fwd parity {mt.TX_NAMES[s]} {mt.TXTP_NAMES[t]}: {len(ins) * len(qs)} blocks")
        batch = np.stack([x for x in ins for _ in qs])
        bq = [q for _ in ins for q in qs]
        lv, tg = mt.fwd_txfm(s, t, torch.from_numpy(batch).to(dev), bq)
        lv, tg = lv.cpu().numpy(), tg.cpu().numpy()
        for i in range(len(batch)):
            el, et = O.fwd(s, t, batch[i], orc_q(bq[i]))
            if not ((lv[i] == el).all() and (tg[i].view(np.uint32) == et.view(np.uint32)).all()):
                j = int(np.nonzero((lv[i] != el) | (tg[i].view(np.uint32) != et.view(np.uint32)))[0][0])
                pytest.fail(f"{mt.TX_NAMES[s]} {mt.TXTP_NAMES[t]} block {i}: coef {j} "
                            f"gpu ({lv[i][j]}, {tg[i][j]!r}) oracle ({el[j]}, {et[j]!r})")
        assert lv.shape == (len(batch), nc)


@pytest.mark.parametrize("s", REF_SIZES, ids=[mt.TX_NAMES[s] for s in REF_SIZES])
def test_inv_txfm_parity(dev, s):
    nc, npx = mt.tx_ncoef(s), mt.tx_width(s) * mt.tx_height(s)
    rng = np.random.default_rng(2000 + s)
    ref_quants = [mt.quant_new(8, 8), mt.quant_new(96, 8), mt.quant_new_chroma(40, 0, 0, 10, 15),
                  mt.quant_new(40, 8, 0), mt.quant_new(96, 10, 7), mt.quant_new(180, 12, 14),
                  mt.quant_new_chroma(128, -7, 11, 12, 3)]
    levels, quants = [], []
    for qi, q in enumerate(ref_quants):
        for seed in (0x12345678, 0x87654321, 0xA5A55A5A, 0x01234567, 0x89ABCDEF, 0x55AA33CC):
            levels.append(itx_levels(nc, seed ^ qi))
            quants.append(q)
        # sparse large levels: exercise the 24-bit mask, cf_max and the stage clips
        for lim in (64, 5000, 1 << 20):
            lv = rng.integers(-lim, lim + 1, nc).astype(np.int32)
            lv[rng.random(nc) < 0.7] = 0
            levels.append(lv)
            quants.append(q)
    for t in TYPES:
        if not mt.tx_supported(s, t):
            continue
        say(f"inv parity {mt.TX_NAMES[s]} {mt.TXTP_NAMES[t]}: {len(levels)} blocks")
        out = mt.inv_txfm(s, t, torch.from_numpy(np.stack(levels)).to(dev), quants).cpu().numpy()
        for i in range(len(levels)):
            exp = O.inv(s, t, levels[i], orc_q(quants[i]))
            if not (out[i] == exp).all():
                j = int(np.nonzero(out[i] != exp)[0][0])
                pytest.fail(f"{mt.TX_NAMES[s]} {mt.TXTP_NAMES[t]} block {i}: px {j} gpu {out[i][j]} "
                            f"oracle {exp[j]}")
        assert out.shape == (len(levels), npx)


def _upload(blocks, pool, dev):
    bt = torch.from_numpy(blocks.view(np.uint8).copy()).to(dev)
    pt = torch.from_numpy(pool).to(dev)
    return bt, pt


def _bd(blocks):
    bdi = (blocks["bd_qm"] >> 4) & 3
    return np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8)).astype(np.uint8)


def test_chain_full_outputs_parity(dev):
    """Every output of the chain (levels, targets, reconstruction, SSE, nnz, checksum) on
    30k generated blocks of every reference class."""
    n = 30000
    blocks, pool = mt.gen_tx_blocks(SEED, n)
    say(f"chain full outputs: {n} blocks, {len(pool)} px")
    bt, pt = _upload(blocks, pool, dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    lv = torch.zeros(len(pool), dtype=torch.int32, device=dev)
    tg = torch.zeros(len(pool), dtype=torch.float32, device=dev)
    rc = torch.zeros(len(pool), dtype=torch.int32, device=dev)
    mt.tx_chain(bt, pt, res, lv, tg, rc)
    torch.cuda.synchronize()
    got = res.cpu().numpy().view(mt._lib.RES_DTYPE)
    lv, tg, rc = lv.cpu().numpy(), tg.cpu().numpy(), rc.cpu().numpy()
    bd = _bd(blocks)
    for i in range(n):
        s, t = int(blocks["tx_size"][i]), int(blocks["tx_type"][i])
        q = O.quant(int(blocks["qidx"][i]), int(bd[i]), int(blocks["bd_qm"][i] & 15))
        off = int(blocks["offset"][i])
        npx, nc = mt.tx_width(s) * mt.tx_height(s), mt.tx_ncoef(s)
        sse, el, et, er = O.chain_block(s, t, pool[off:off + npx], q)
        ok = (got["sse"][i] == sse and got["nnz"][i] == np.count_nonzero(el)
              and (lv[off:off + nc] == el).all()
              and (tg[off:off + nc].view(np.uint32) == et.view(np.uint32)).all()
              and (rc[off:off + npx] == er).all())
        if not ok:
            pytest.fail(f"block {i} ({mt.TX_NAMES[s]} {mt.TXTP_NAMES[t]} q{blocks['qidx'][i]} "
                        f"bd{bd[i]} qm{blocks['bd_qm'][i] & 15}): gpu sse {got['sse'][i]} oracle {sse}")
        h = O.lib().orc_block_hash(nc, O.ptr(el, ctypes.c_int32), O.ptr(et, ctypes.c_float), npx,
                                   O.ptr(er, ctypes.c_int32))
        assert got["hash"][i] == h


def test_chain_unsupported_sizes_flagged(dev):
    """All 19 AV1 sizes: the 6 without a reference forward transform come back sse = -1,
    the rest match the oracle."""
    n = 4000
    blocks, pool = mt.gen_tx_blocks(SEED + 1, n, all_sizes=True)
    say(f"chain 19 sizes: {n} blocks")
    bt, pt = _upload(blocks, pool, dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    mt.tx_chain(bt, pt, res)
    got = res.cpu().numpy().view(mt._lib.RES_DTYPE)
    sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], _bd(blocks),
                           blocks["bd_qm"] & 15, blocks["offset"].astype(np.uint64), pool)
    unsup = np.array([not mt.tx_supported(int(s), int(t)) for s, t in zip(blocks["tx_size"], blocks["tx_type"])])
    assert unsup.any()
    assert (got["sse"][unsup] == -1).all()
    assert (got["sse"] == sse).all()
    assert (got["hash"][~unsup] == h[~unsup]).all()


def test_chain_edge_batches(dev):
    """Empty and single-block batches, and a batch of one class only."""
    blocks, pool = mt.gen_tx_blocks(SEED + 2, 1)
    bt, pt = _upload(blocks, pool, dev)
    res = torch.zeros(16, dtype=torch.uint8, device=dev)
    mt.tx_chain(bt, pt, res)
    got = res.cpu().numpy().view(mt._lib.RES_DTYPE)
    sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], _bd(blocks),
                           blocks["bd_qm"] & 15, blocks["offset"].astype(np.uint64), pool)
    assert got["sse"][0] == sse[0] and got["hash"][0] == h[0]
    empty = torch.zeros(0, dtype=torch.uint8, device=dev)
    mt.tx_chain(empty, pt, res)  # n = 0 is a no-op
    # one class: 3000 blocks of 32x32 DCT_DCT
    b2, p2 = mt.gen_tx_blocks(SEED + 3, 3000)
    keep = b2["tx_size"] == 3
    b2 = b2[keep].copy()
    bt, pt = _upload(b2, p2, dev)
    res = torch.zeros(len(b2) * 16, dtype=torch.uint8, device=dev)
    mt.tx_chain(bt, pt, res)
    got = res.cpu().numpy().view(mt._lib.RES_DTYPE)
    sse, h = O.chain_batch(b2["tx_size"], b2["tx_type"], b2["qidx"], _bd(b2), b2["bd_qm"] & 15,
                           b2["offset"].astype(np.uint64), p2)
    assert (got["sse"] == sse).all() and (got["hash"] == h).all()


def test_chain_chroma_quantizer(dev):
    """Blocks with the chroma bit (bd_qm bit 6) take new_chroma_with_delta_qm (zero deltas:
    the chroma QM planes); the others the luma quantizer, in one mixed batch."""
    n = 20000
    blocks, pool = mt.gen_tx_blocks(SEED + 4, n)
    rng = np.random.default_rng(SEED + 4)
    chroma = rng.random(n) < 0.5
    blocks["bd_qm"][chroma] |= mt.TXBLK_CHROMA
    say(f"chain chroma quantizer: {n} blocks, {int(chroma.sum())} chroma")
    bt, pt = _upload(blocks, pool, dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    mt.tx_chain(bt, pt, res)
    got = res.cpu().numpy().view(mt._lib.RES_DTYPE)
    sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], _bd(blocks),
                           blocks["bd_qm"] & 0x4f, blocks["offset"].astype(np.uint64), pool)
    bad = np.nonzero((got["sse"] != sse) | (got["hash"] != h))[0]
    assert len(bad) == 0, f"{len(bad)} blocks differ, first {bad[:10]}"
    # the chroma planes of the QM change the result of some blocks
    sse_l, h_l = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], _bd(blocks),
                               blocks["bd_qm"] & 15, blocks["offset"].astype(np.uint64), pool)
    assert (h_l[chroma] != h[chroma]).any()


def test_chain_full_workload_every_block(dev):
    """The full 2^20-block microbench workload: SSE and checksum of every block equal the
    oracle's (the checksum covers every level, target and reconstructed sample)."""
    n = 1 << 20
    blocks, pool = mt.gen_tx_blocks(SEED, n)
    say(f"chain full workload: {n} blocks, {len(pool) / 1e6:.1f} Mpx")
    bt, pt = _upload(blocks, pool, dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    mt.tx_chain(bt, pt, res)
    torch.cuda.synchronize()
    got = res.cpu().numpy().view(mt._lib.RES_DTYPE)
    sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], _bd(blocks),
                           blocks["bd_qm"] & 15, blocks["offset"].astype(np.uint64), pool,
                           nthreads=16)
    bad = np.nonzero((got["sse"] != sse) | (got["hash"] != h))[0]
    assert len(bad) == 0, f"{len(bad)} blocks differ, first {bad[:10]}"
    # determinism: a second run gives the same bytes
    res2 = torch.zeros_like(res)
    mt.tx_chain(bt, pt, res2)
    assert torch.equal(res, res2)


def test_error_paths(dev):
    """Unsupported pairs and bad arguments return error codes, never abort."""
    x = torch.zeros((1, 64), dtype=torch.int32, device=dev)
    with pytest.raises(mt.MtError):
        mt.fwd_txfm(15, 0, x, [mt.quant_new(10)])   # 8x32: no reference forward
    with pytest.raises(mt.MtError):
        mt.fwd_txfm(13, 3, x, [mt.quant_new(10)])   # 4x16 ADST: not in the reference
    assert mt.lib().mt_tx_chain(None, 5, None, None, None, None, None, None) == -1
