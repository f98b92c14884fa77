"""Synthetic planes and intra candidates for the candidate-evaluation tests, and the
stage-by-stage oracle composition of one candidate (the reference's candidate loop body,
src/coder/partition_search.rs:471-520).  Test infrastructure."""
import ctypes
import math

import numpy as np

import maroontree_amd as mt
import oracle_lib as O

PW, PH = 320, 256
REF_SIZES = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
EOB_BINS = {16: [0, 1, 2], 32: [3, 4], 64: [5, 6, 7], 128: [8, 9, 10], 256: [11, 12],
            512: [13, 14], 1024: [15, 16]}


def cls_of(s):
    w, h = mt.tx_width(s), mt.tx_height(s)
    return (int(math.log2(w)) - 2 + int(math.log2(h)) - 2 + 1) >> 1


def make_planes(rng, bd):
    mx = (1 << bd) - 1
    walk = np.cumsum(np.cumsum(rng.integers(-2, 3, (PH, PW)), 0), 1)
    src = np.clip(mx // 2 + walk // 4 + rng.integers(-6, 7, (PH, PW)), 0, mx).astype(np.uint16)
    recon = np.clip(src.astype(np.int64) + rng.integers(-9, 10, (PH, PW)), 0, mx).astype(np.uint16)
    return src, recon


def make_cands(rng, n, bd):
    bdi = {8: 0, 10: 1, 12: 2}[bd]
    c = np.zeros(n, mt.CAND_DTYPE)
    off = 0
    for i in range(n):
        s = REF_SIZES[i % len(REF_SIZES)]
        w, h = mt.tx_width(s), mt.tx_height(s)
        types = [t for t in mt.TX_TYPES_REF if mt.tx_supported(s, t)]
        r = c[i]
        r["offset"] = off
        if i % 11 == 0:
            r["x"], r["y"] = 0, int(rng.integers(0, PH - h + 1))
        elif i % 11 == 1:
            r["x"], r["y"] = int(rng.integers(0, PW - w + 1)), 0
        else:
            r["x"] = 4 * int(rng.integers(0, (PW - w) // 4 + 1))
            r["y"] = 4 * int(rng.integers(0, (PH - h) // 4 + 1))
        r["tx_size"] = s
        r["tx_type"] = types[int(rng.integers(0, len(types)))]
        modes = list(range(13)) + ([16, 17, 18, 19, 20] if max(w, h) <= 32 else [])
        r["mode"] = modes[int(rng.integers(0, len(modes)))]
        r["angle_delta"] = int(rng.integers(-3, 4)) if 1 <= r["mode"] <= 8 else 0
        r["pflags"] = int(rng.integers(0, 8))
        r["planes"] = (1 << 2) | (0 if i % 5 else 2)  # recon plane 1; source plane 0 or 2
        r["qidx"] = int(rng.integers(0, 256))
        qm = 15 if rng.random() < 0.4 else int(rng.integers(0, 15))
        r["bd_qm"] = qm | (bdi << 4)
        if i % 3 == 2:  # chroma quantizer with qindex deltas
            r["bd_qm"] |= mt.TXBLK_CHROMA
            r["dc_delta"] = int(rng.integers(-20, 21))
            r["ac_delta"] = int(rng.integers(-20, 21))
        nc = mt.tx_ncoef(s)
        r["rplane"] = 0 if s == 4 else int(rng.integers(0, 2))
        r["eob_bin"] = EOB_BINS[nc][int(rng.integers(0, len(EOB_BINS[nc])))]
        r["skip_ctx"] = int(rng.integers(0, 13))
        r["dcs_ctx"] = int(rng.integers(0, 3))
        if r["rplane"] == 0 and rng.random() < 0.5:
            r["txtp_set"] = int(rng.integers(1, 4))
            r["txtp_mode"] = int(rng.integers(0, 13))
            r["txtp"] = int(rng.integers(0, 5 if r["txtp_set"] == 3 else 7))
        r["qctx"] = int(rng.integers(0, 4))
        if rng.random() < 0.2:
            r["rflags"] = 1
            r["skip_one"] = rng.uniform(0.1, 9.0)
            r["skip_zero"] = rng.uniform(0.1, 9.0)
        r["wsse_scale"] = [1.0, 0.0, 1.0, 0.8][i % 4]
        r["lambda"] = rng.uniform(0.5, 400.0)
        r["extra_bits"] = rng.uniform(0.0, 12.0)
        off += w * h
    return c, off


def oracle_candidate(r, planes_np, views):
    s, t = int(r["tx_size"]), int(r["tx_type"])
    w, h = mt.tx_width(s), mt.tx_height(s)
    x, y = int(r["x"]), int(r["y"])
    bd = [8, 10, 12][(int(r["bd_qm"]) >> 4) & 3]
    qm = int(r["bd_qm"]) & 15
    si, ri = int(r["planes"]) & 3, (int(r["planes"]) >> 2) & 3
    src, recon = planes_np[si], planes_np[ri]
    _, rfw, rfh, _, _ = views[ri]
    _, _, _, svw, svh = views[si]
    pf = int(r["pflags"])
    pred = O.intra_pred(int(r["mode"]), int(r["angle_delta"]), recon, x, y, w, h, pf & 1, pf & 2,
                        rfw, rfh, pf & 4, bd).ravel().astype(np.int32)
    sblk = np.ascontiguousarray(src[y:y + h, x:x + w])
    resid = sblk.astype(np.int32).ravel() - pred
    if int(r["bd_qm"]) & mt.TXBLK_CHROMA:
        q = O.quant_chroma(int(r["qidx"]), int(r["dc_delta"]), int(r["ac_delta"]), bd, qm)
    else:
        q = O.quant(int(r["qidx"]), bd, qm)
    levels, targets = O.fwd(s, t, resid, q)
    if int(r["rflags"]) & 2:  # MT_CAND_TRELLIS: trellis_optimize_ctx in the rate context
        tctx = O.rate_ctx(qctx=int(r["qctx"]), cls=cls_of(s), plane=int(r["rplane"]), w=min(w, 32),
                          h=min(h, 32), eob_bin=int(r["eob_bin"]), dcs_ctx=int(r["dcs_ctx"]))
        lam0 = O.trellis_lambda0(int(r["qidx"]), bd, (int(r["rflags"]) >> 2) & 3)
        levels = O.trellis_ctx(levels, targets, float(q.dc), float(q.ac), O.scan(s), lam0, tctx, q.qm_level)
    inv = O.inv(s, t, levels, q)
    L = O.plib()
    vw, vh = min(max(svw - x, 0), w), min(max(svh - y, 0), h)
    sp, pp, ip = O.ptr(sblk, ctypes.c_uint16), O.ptr(pred, ctypes.c_int32), O.ptr(inv, ctypes.c_int32)
    sse = L.orc_sse_recon_vis(pp, ip, sp, w, w, vw, vh, bd)
    scale = float(r["wsse_scale"])
    if scale > 0:
        rec = np.clip(pred + inv, 0, (1 << bd) - 1).reshape(h, w)
        err = np.ascontiguousarray((sblk.astype(np.int64) - rec)[:vh, :vw].astype(np.int32))
        dist = L.orc_weighted_dist(O.ptr(err, ctypes.c_int32), vw, vh, scale, 1)
    else:
        dist = sse
    satd = L.orc_luma_satd(sp, w, w, h, vw, vh, bd, pp, ip)
    sb = (float(r["skip_one"]), float(r["skip_zero"])) if int(r["rflags"]) & 1 else None
    ctx = O.rate_ctx(qctx=int(r["qctx"]), cls=cls_of(s), plane=int(r["rplane"]), w=min(w, 32),
                     h=min(h, 32), eob_bin=int(r["eob_bin"]), skip_ctx=int(r["skip_ctx"]),
                     dcs_ctx=int(r["dcs_ctx"]), txtp_set=int(r["txtp_set"]),
                     txtp_mode=int(r["txtp_mode"]), txtp=int(r["txtp"]), skip_bits=sb)
    bits = np.float32(O.block_rate(levels, O.scan(s), ctx))
    bits = np.float32(bits + np.float32(r["extra_bits"]))
    cost = np.float32(np.float32(dist) + np.float32(np.float32(r["lambda"]) * bits))
    return sse, dist, satd, bits, cost, int(np.count_nonzero(levels))
