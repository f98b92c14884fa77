"""Random trellis workloads for the parity tests: descriptors (the library's mt_trellisblk),
the forward transform's levels and targets from random residuals (the oracle's forward),
and the oracle's trellis result of every block."""
import numpy as np

import maroontree_amd as mt
import oracle_lib as O

SIZES = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
# eob_bin CDF index by coefficient count: luma, chroma
EOB_BIN = {16: (0, 1), 32: (3, 4), 64: (5, 6), 128: (8, 9), 256: (11, 12), 512: (13, 14), 1024: (15, 16)}


def cls_of(s):
    w, h = O.tx_w(s), O.tx_h(s)
    return ((w.bit_length() - 3) + (h.bit_length() - 3) + 1) >> 1


def make(seed, n, ctx_frac=0.6):
    """n random blocks: (descriptors, levels, targets, want) with want the oracle's levels;
    a fraction ctx_frac of them with the context variant."""
    rng = np.random.default_rng(seed)
    desc = np.zeros(n, mt.TRELLISBLK_DTYPE)
    lv_parts, tg_parts = [], []
    off = 0
    for i in range(n):
        s = int(rng.choice(SIZES))
        w, h = O.tx_w(s), O.tx_h(s)
        t = 0
        types = [tt for tt in (0, 1, 2, 3, 9, 10, 11) if O.lib().orc_tx_supported(s, tt)]
        t = int(rng.choice(types))
        bd = int(rng.choice([8, 10, 12]))
        qidx = int(rng.integers(0, 256))
        qm = int(rng.choice([15, 15, 0, 4, 5, 9]))
        chroma = s != 4 and rng.random() < 0.3
        ctx = rng.random() < ctx_frac
        dcd, acd = (int(rng.integers(-12, 13)), int(rng.integers(-12, 13))) if chroma else (0, 0)
        q = O.quant_chroma(qidx, dcd, acd, bd, qm) if chroma else O.quant(qidx, bd, qm)
        scale = int(rng.choice([1, 2, 8, 32]))
        r = rng.integers(-(1 << bd) + 1, 1 << bd, w * h) // scale
        if rng.random() < 0.15:  # a few very large levels (the scalar round-down chunks)
            r = (rng.integers(-(1 << bd) + 1, 1 << bd, w * h)).astype(np.int64)
        lv, tg = O.fwd(s, t, r, q)
        ncoef = len(lv)
        nc = min(w, 32) * min(h, 32)
        pl = 1 if chroma else 0
        d = desc[i]
        d["offset"] = off
        d["tx_size"] = s
        d["qidx"] = qidx
        d["bd_qm"] = ({8: 0, 10: 1, 12: 2}[bd] << 4) | qm | (mt.TXBLK_CHROMA if chroma else 0)
        d["flags"] = mt.TRELLIS_CTX if ctx else 0
        d["dc_delta"], d["ac_delta"] = dcd, acd
        d["qctx"] = int(rng.integers(0, 4))
        d["eob_bin"] = EOB_BIN[nc][pl]
        d["dcs_ctx"] = int(rng.integers(0, 3))
        d["plane"] = pl
        d["lambda0"] = float(rng.choice([0.05, 0.05 * 0.4, 0.2, 0.05, 0.0]))
        d["band_tilt"] = float(rng.choice([0.0, 0.0, 1.9, 3.5]))
        lv_parts.append(lv)
        tg_parts.append(tg)
        off += ncoef
    levels = np.concatenate(lv_parts).astype(np.int32)
    targets = np.concatenate(tg_parts).astype(np.float32)
    want = levels.copy()
    for i in range(n):
        d = desc[i]
        s = int(d["tx_size"])
        w, h = O.tx_w(s), O.tx_h(s)
        nc = min(w, 32) * min(h, 32)
        o = int(d["offset"])
        bd = {0: 8, 1: 10, 2: 12}[(int(d["bd_qm"]) >> 4) & 3]
        qm = int(d["bd_qm"]) & 15
        chroma = bool(int(d["bd_qm"]) & mt.TXBLK_CHROMA)
        q = O.quant_chroma(int(d["qidx"]), int(d["dc_delta"]), int(d["ac_delta"]), bd, qm) if chroma \
            else O.quant(int(d["qidx"]), bd, qm)
        cf, tf = want[o:o + nc], targets[o:o + nc]
        if d["flags"] & mt.TRELLIS_CTX:
            c = O.rate_ctx(qctx=int(d["qctx"]), cls=cls_of(s), plane=int(d["plane"]), w=min(w, 32), h=min(h, 32),
                           eob_bin=int(d["eob_bin"]), dcs_ctx=int(d["dcs_ctx"]))
            want[o:o + nc] = O.trellis_ctx(cf, tf, float(q.dc), float(q.ac), O.scan(s), float(d["lambda0"]), c, qm,
                                           float(d["band_tilt"]))
        else:
            want[o:o + nc] = O.trellis(cf, tf, float(q.dc), float(q.ac), O.scan(s), float(d["lambda0"]))
    return desc, levels, targets, want


def chain_descriptors(blocks, seed):
    """Trellis descriptors for the chain's blocks (mt_txblk records): the same offset, size
    and quantizer, random variant, context, lambda and tilt."""
    rng = np.random.default_rng(seed)
    n = len(blocks)
    tr = np.zeros(n, mt.TRELLISBLK_DTYPE)
    tr["offset"] = blocks["offset"]
    tr["tx_size"] = blocks["tx_size"]
    tr["qidx"] = blocks["qidx"]
    tr["bd_qm"] = blocks["bd_qm"]
    tr["flags"] = (rng.random(n) < 0.6).astype(np.uint8) * mt.TRELLIS_CTX
    tr["qctx"] = rng.integers(0, 4, n)
    tr["dcs_ctx"] = rng.integers(0, 3, n)
    chroma = (blocks["bd_qm"] & mt.TXBLK_CHROMA) != 0
    tr["plane"] = chroma.astype(np.uint8)
    for i in range(n):
        s = int(blocks["tx_size"][i])
        if s not in SIZES:
            continue
        nc = min(O.tx_w(s), 32) * min(O.tx_h(s), 32)
        tr["eob_bin"][i] = EOB_BIN[nc][1 if chroma[i] else 0]
    tr["lambda0"] = rng.choice([0.05, 0.02, 0.2, 0.0], n).astype(np.float32)
    tr["band_tilt"] = rng.choice([0.0, 0.0, 1.9], n).astype(np.float32)
    return tr


def oracle_chain(blocks, pool, tr, nthreads=8):
    bdi = (blocks["bd_qm"] >> 4) & 3
    bd = np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8))
    return O.chain_batch_trellis(blocks["tx_size"], blocks["tx_type"], blocks["qidx"], bd,
                                 blocks["bd_qm"] & 0x4F, blocks["offset"].astype(np.uint64), pool, tr, nthreads)
