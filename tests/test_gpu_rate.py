"""Bit-exact parity of the GPU coefficient rate (mt_block_rate) with the oracle's
real_block_bits_bounded (src/rate.rs:149) on every reference size, plane and qctx."""
import math
import sys

import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O

pytestmark = pytest.mark.gpu


def cls_of(s):
    w, h = mt.tx_width(s), mt.tx_height(s)
    return (int(math.log2(w)) - 2 + int(math.log2(h)) - 2 + 1) >> 1


def make_blocks(rng, n, qctxs=(0, 1, 2, 3)):
    sizes = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
    recs = np.zeros(n, mt.RATEBLK_DTYPE)
    pool = []
    off = 0
    for i in range(n):
        s = sizes[i % len(sizes)]
        nc = mt.tx_ncoef(s)
        kind = rng.integers(0, 6)
        lv = np.zeros(nc, np.int32)
        if kind == 0:
            pass  # all zero
        elif kind == 1:
            lv[0] = rng.integers(-40, 41)  # DC only (maybe zero)
        else:
            dens = [0.05, 0.2, 0.5, 0.9][rng.integers(0, 4)]
            mag = [3, 8, 40, 4000][rng.integers(0, 4)]
            lv = rng.integers(-mag, mag + 1, nc).astype(np.int32)
            lv[rng.random(nc) > dens] = 0
            if kind == 5:
                lv[O.scan(s)[nc - 1]] = rng.integers(1, 30)  # eob at the last scan position
        pool.append(lv)
        r = recs[i]
        r["offset"] = off
        r["tx_size"] = s
        r["plane"] = 0 if s == 4 else rng.integers(0, 2)
        bins = {16: [0, 1, 2], 32: [3, 4], 64: [5, 6, 7], 128: [8, 9, 10], 256: [11, 12],
                512: [13, 14], 1024: [15, 16]}[nc]
        r["eob_bin"] = bins[rng.integers(0, len(bins))]
        r["skip_ctx"] = rng.integers(0, 13)
        r["dcs_ctx"] = rng.integers(0, 3)
        if r["plane"] == 0 and rng.random() < 0.5:
            r["txtp_set"] = rng.integers(1, 4)
            r["txtp_mode"] = rng.integers(0, 13)
            r["txtp"] = rng.integers(0, 5 if r["txtp_set"] == 3 else 7)
        r["qctx"] = qctxs[rng.integers(0, len(qctxs))]
        if rng.random() < 0.2:
            r["flags"] = 1
            r["skip_one"] = rng.uniform(0.01, 9.0)
            r["skip_zero"] = rng.uniform(0.01, 9.0)
        r["bound"] = np.inf if rng.random() < 0.7 else rng.uniform(0.0, 400.0)
        off += nc
    return recs, np.concatenate(pool)


def oracle_bits(recs, pool):
    out = np.zeros(len(recs), np.float32)
    for i, r in enumerate(recs):
        s = int(r["tx_size"])
        nc = mt.tx_ncoef(s)
        w = min(mt.tx_width(s), 32)
        h = min(mt.tx_height(s), 32)
        sb = (float(r["skip_one"]), float(r["skip_zero"])) if r["flags"] & 1 else None
        c = O.rate_ctx(qctx=int(r["qctx"]), cls=cls_of(s), plane=int(r["plane"]), w=w, h=h,
                       eob_bin=int(r["eob_bin"]), skip_ctx=int(r["skip_ctx"]), dcs_ctx=int(r["dcs_ctx"]),
                       txtp_set=int(r["txtp_set"]), txtp_mode=int(r["txtp_mode"]), txtp=int(r["txtp"]),
                       skip_bits=sb)
        cf = pool[int(r["offset"]):int(r["offset"]) + nc]
        out[i] = O.block_rate(cf, O.scan(s), c, float(r["bound"]))
    return out


def test_block_rate_parity(dev):
    rng = np.random.default_rng(77)
    recs, pool = make_blocks(rng, 6000)
    print(f"rate parity: {len(recs)} blocks", flush=True)
    got = mt.block_rate(torch.from_numpy(recs.view(np.uint8).copy()).to(dev),
                        torch.from_numpy(pool).to(dev)).cpu().numpy()
    exp = oracle_bits(recs, pool)
    same = (got.view(np.uint32) == exp.view(np.uint32))
    assert same.all(), (f"{(~same).sum()} blocks differ; first {np.nonzero(~same)[0][:5]}: "
                        f"gpu {got[~same][:5]} oracle {exp[~same][:5]}")
    assert np.isinf(got).any() and (got > 0).all()


def test_block_rate_reference_goldens(dev):
    """The reference's golden rates (src/rate.rs:393-399) through the GPU."""
    s48, s84 = O.scan(5), O.scan(6)
    a = np.zeros(32, np.int32)
    b = np.zeros(32, np.int32)
    for i, lvl in ((0, 3), (1, -2), (4, 1), (17, 4)):
        a[s48[i]] = lvl
        b[s84[i]] = lvl
    recs = np.zeros(2, mt.RATEBLK_DTYPE)
    for k, (s, off) in enumerate(((5, 0), (6, 32))):
        recs[k]["offset"] = off
        recs[k]["tx_size"] = s
        recs[k]["plane"] = 1
        recs[k]["eob_bin"] = 4
        recs[k]["skip_ctx"] = 7
        recs[k]["bound"] = np.inf
    got = mt.block_rate(torch.from_numpy(recs.view(np.uint8).copy()).to(dev),
                        torch.from_numpy(np.concatenate([a, b])).to(dev)).cpu().numpy()
    assert abs(got[0] - 38.532284) < 1e-5 and abs(got[1] - 38.2208) < 1e-5, got


def test_block_rate_invalid_descriptor_is_nan(dev):
    recs = np.zeros(1, mt.RATEBLK_DTYPE)
    recs[0]["tx_size"] = 15  # 8x32: no reference scan
    recs = np.concatenate([recs, recs, recs])
    recs[1]["tx_size"] = 1
    recs[1]["eob_bin"] = 0  # a 16-coefficient eob table for a 64-coefficient block
    recs[2]["tx_size"] = 1
    recs[2]["eob_bin"] = 5
    recs[2]["bound"] = np.inf
    got = mt.block_rate(torch.from_numpy(recs.view(np.uint8).copy()).to(dev),
                        torch.zeros(1024, dtype=torch.int32, device=dev)).cpu().numpy()
    assert np.isnan(got[0]) and np.isnan(got[1]) and not np.isnan(got[2])


def test_block_rate_with_a_loaded_snapshot(dev):
    """Random decision CDF snapshots in slots 4 and 7: the device rate equals the oracle's
    against the same snapshot, bit for bit."""
    import oracle_lib as O2
    rng = np.random.default_rng(77)
    for slot in (4, 7):
        snap = O2.random_cdfs(rng)
        mt.rate_load_cdfs(slot, snap)
        O2.rate_load_cdfs(slot, snap)
    recs, pool = make_blocks(rng, 3000, qctxs=(4, 7))
    print("rate snapshots loaded", flush=True)
    got = mt.block_rate(torch.from_numpy(recs.view(np.uint8).copy()).to(dev),
                        torch.from_numpy(pool).to(dev)).cpu().numpy()
    exp = oracle_bits(recs, pool)
    same = (got.view(np.uint32) == exp.view(np.uint32))
    assert same.all(), f"{(~same).sum()} blocks differ; first {np.nonzero(~same)[0][:5]}"
