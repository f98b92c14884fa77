"""Pins of the rectangular partitions (PARTITION_HORZ / VERT) of the oracle's whole-frame
RDO (oracle/oracle_rdo.cpp):

- availability of the halves against a brute-force decode-time map of an SB whose nodes
  are all coded as HORZ or VERT pairs;
- reconstruction replayed leaf by leaf (rectangular leaves included) in coding order from
  the node records alone, on the pinned per-kernel oracle functions, with the rectangular
  transform sizes;
- decisions: every rectangular half of one SB rebuilt in Python on the decision-time plane,
  its best mode equal to the node's recorded half modes, and every node's cost the minimum
  of NONE, HORZ, VERT and SPLIT with their partition symbols (first wins on ties).
"""
import ctypes

import numpy as np
import pytest

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import workload as W
from test_oracle_rdo import FILTER_DIR, KF_CTX, base_of, brute_avail, geo, mode_of, nid

ANGLE, FILTER, MIN4, CFL, RECT = 1, 2, 4, 8, 16
TX = {(4, 4): 0, (8, 8): 1, (16, 16): 2, (32, 32): 3, (64, 64): 4, (4, 8): 5, (8, 4): 6, (8, 16): 7,
      (16, 8): 8, (16, 32): 9, (32, 16): 10}
EOB_LUMA = {16: 0, 32: 3, 64: 5, 128: 8, 256: 11, 512: 13, 1024: 15}
BSIZE = {(4, 4): 0, (4, 8): 1, (8, 4): 2, (8, 8): 3, (8, 16): 4, (16, 8): 5, (16, 16): 6, (16, 32): 7,
         (32, 16): 8, (32, 32): 9, (64, 64): 12}
PART_BYTE = {64: 0x10, 32: 0x18, 16: 0x1C, 8: 0x1E, 4: 0x1F}


def halves(x, y, s, k):
    h = s // 2
    return [(x, y, s, h), (x, y + h, s, h), (x, y, h, s), (x + h, y, h, s)][k]


def rect_decode_times(s, vert):
    """Coding time of every 4x4 unit around an SB whose s x s nodes are each coded as two
    halves (HORZ or VERT): the row above and the left SB before, the right SB after."""
    t = np.full((32, 48), 10 ** 9, np.int64)
    t[:16, :] = -2
    t[16:, :16] = -1
    order = []

    def rec(x, y, size):
        if size == s:
            order.append((x, y))
            return
        hh = size // 2
        for dy in (0, hh):
            for dx in (0, hh):
                rec(x + dx, y + dy, hh)
    rec(0, 0, 64)
    when = {}
    for k, (x, y) in enumerate(order):
        for j in range(2):
            bx, by, bw, bh = halves(x, y, s, (2 if vert else 0) + j)
            t[16 + by // 4:16 + (by + bh) // 4, 16 + bx // 4:16 + (bx + bw) // 4] = 2 * k + j
            when[(bx, by, bw, bh)] = 2 * k + j
    return t, when


def test_rect_availability_matches_coding_order():
    L = O.lib()
    L.orc_rdo_rect_avail.argtypes = [ctypes.c_int] * 4
    for vert in (False, True):
        for s in (8, 16, 32):
            t, when = rect_decode_times(s, vert)
            for y in range(0, 64, s):
                for x in range(0, 64, s):
                    for j in range(2):
                        k = (2 if vert else 0) + j
                        bx, by, bw, bh = halves(x, y, s, k)
                        now = when[(bx, by, bw, bh)]
                        tr = all(t[16 + (by - 1) // 4, 16 + xx // 4] < now for xx in range(bx + bw, bx + 2 * bw))
                        bl = by + 2 * bh <= 64 and all(t[16 + yy // 4, 16 + (bx - 1) // 4] < now
                                                       for yy in range(by + bh, by + 2 * bh))
                        got = L.orc_rdo_rect_avail(x, y, s, k)
                        assert bool(got & 1) == tr, ("tr", vert, s, x, y, j)
                        assert bool(got & 2) == bl, ("bl", vert, s, x, y, j)


def rect_avail(x, y, s, k):
    L = O.lib()
    L.orc_rdo_rect_avail.argtypes = [ctypes.c_int] * 4
    v = L.orc_rdo_rect_avail(x, y, s, k)
    return bool(v & 1), bool(v & 2)


def leaves(nodes_sb, x0, y0):
    """(x, y, w, h, mode id, availability) of an SB's leaves in coding order."""
    out, stack = [], [0]
    while stack:
        n = stack.pop()
        d = int(nodes_sb[n]["split"])
        if d == 2:
            continue
        L, ix, iy, s = geo(n)
        if d == 1:
            for q in (3, 2, 1, 0):
                stack.append(nid(L + 1, 2 * ix + (q & 1), 2 * iy + (q >> 1)))
            continue
        x, y = ix * s, iy * s
        if d == 0:
            out.append((x0 + x, y0 + y, s, s, int(nodes_sb[n]["mode"]), brute_avail(x, y, s)))
        else:
            k0 = 0 if d == 3 else 2
            for j in range(2):
                bx, by, bw, bh = halves(x, y, s, k0 + j)
                m = int(nodes_sb[n]["rmode0" if j == 0 else "rmode1"])
                out.append((x0 + bx, y0 + by, bw, bh, m, rect_avail(x, y, s, k0 + j)))
    return out


@pytest.mark.parametrize("bd,qidx,w,h,flags", [(10, 91, 136, 72, ANGLE | FILTER | MIN4 | RECT),
                                              (8, 60, 128, 64, ANGLE | RECT), (12, 150, 72, 80, MIN4 | RECT)])
def test_rect_reconstruction_replays_from_the_decisions(bd, qidx, w, h, flags):
    img = W.synth_image(21 + qidx, w, h, bd)
    qm = mt.qm_level_law(qidx, 0)
    lam = O.rdo_lambda(qidx, bd, 2)
    rec, nodes, modes, st = O.intra_rdo(img, bd, qidx, qm, flags, lam)
    assert np.isin(nodes["split"], (3, 4)).any()  # rectangular partitions win somewhere
    H, Wd = img.shape
    C = (Wd + 63) // 64
    rrec = img.copy()
    mmap = np.zeros((H // 4, Wd // 4), np.int64)
    q = O.quant(qidx, bd, qm)
    mx = (1 << bd) - 1
    nleaves = 0
    for sb in range(nodes.shape[0]):
        r, c = divmod(sb, C)
        for (x, y, bw, bh, m, (tr, bl)) in leaves(nodes[sb], 64 * c, 64 * r):
            mode, delta = mode_of(m)
            am = mmap[(y - 1) // 4, x // 4] if y > 0 else 0
            lm = mmap[y // 4, (x - 1) // 4] if x > 0 else 0
            smooth = am in (9, 10, 11) or lm in (9, 10, 11)
            pred = O.intra_pred(mode, delta, rrec, x, y, bw, bh, tr, bl, Wd, H, smooth, bd)
            resid = img[y:y + bh, x:x + bw].astype(np.int32) - pred
            lv, _ = O.fwd(TX[(bw, bh)], 0, resid.ravel(), q)
            inv = O.inv(TX[(bw, bh)], 0, lv, q).reshape(bh, bw)
            rrec[y:y + bh, x:x + bw] = np.clip(pred + inv, 0, mx)
            mmap[y // 4:(y + bh) // 4, x // 4:(x + bw) // 4] = base_of(m)
            nleaves += 1
    assert np.array_equal(rec, rrec)
    assert np.array_equal(modes, mmap)
    assert st["leaves"] == nleaves
    assert st["sse"] == int(((rec.astype(np.int64) - img) ** 2).sum())


def test_rect_decisions_are_the_argmin():
    """SB 3 of a 2 x 2 SB frame: every HORZ / VERT half of every eligible node rebuilt on the
    decision-time plane; the recorded half modes of rectangular decisions are their argmins,
    and every node's cost is the first minimum of NONE, HORZ, VERT and SPLIT."""
    bd, qidx, flags = 10, 60, ANGLE | FILTER | MIN4 | RECT
    img = W.synth_image(321, 128, 128, bd)
    qm = mt.qm_level_law(qidx, 0)
    lam = np.float32(O.rdo_lambda(qidx, bd, 2))
    rec, nodes, modes, _ = O.intra_rdo(img, bd, qidx, qm, flags, float(lam))
    L = O.lib()
    L.orc_cdf_cost_raw.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.orc_cdf_cost_raw.restype = ctypes.c_float
    tabs = {k: np.ascontiguousarray(v, dtype=np.uint16) for k, v in O.cdf_tables().items()
            if k in ("PART_SPLIT_CDF", "PART_BL8_CDF", "KF_Y_MODE_CDF", "ANGLE_DELTA_CDF",
                     "FILTER_INTRA_CDF", "FILTER_INTRA_MODE_CDF")}
    cost = lambda row, n, s: np.float32(L.orc_cdf_cost_raw(row.ctypes.data, n, s))  # noqa: E731

    def part(Lv, ctx, sym):
        if Lv == 3:
            return cost(tabs["PART_BL8_CDF"][ctx], 4, sym)
        return cost(tabs["PART_SPLIT_CDF"][Lv][ctx], 10, sym)

    # final partition context maps of the other SBs (above: width class, left: height class)
    pa = np.zeros((16, 16), np.int64)
    pl = np.zeros((16, 16), np.int64)
    for s2 in range(3):
        r, c = divmod(s2, 2)
        for (x, y, bw, bh, m, _) in leaves(nodes[s2], 64 * c, 64 * r):
            for yy in range(y, y + bh, 4):
                for xx in range(x, x + bw, 4):
                    pa[yy // 8, xx // 8] = PART_BYTE[bw]
                    pl[yy // 8, xx // 8] = PART_BYTE[bh]
    sb, x0, y0 = 3, 64, 64
    plane = rec.copy()
    plane[64:, 64:] = img[64:, 64:]
    checked_rect = 0
    for n in range(1, 85):
        Lv, ix, iy, s = geo(n)
        nd = nodes[sb][n]
        if nd["split"] == 2:
            continue
        x, y = ix * s, iy * s
        a = pa[(y0 + y - 1) // 8, (x0 + x) // 8] if y == 0 else 0
        lft = pl[(y0 + y) // 8, (x0 + x - 1) // 8] if x == 0 else 0
        sh = 4 - (Lv + 1)
        ctx = ((a >> sh) & 1) + (((lft >> sh) & 1) << 1)
        best = []
        for k in range(4):
            bx, by, bw, bh = halves(x, y, s, k)
            tr, bl = rect_avail(x, y, s, k)
            cands, ids = [], []
            for i in range(66):
                if 13 <= i <= 60 and (bw < 8 or bh < 8):
                    continue
                if i > 60 and (bw > 32 or bh > 32):
                    continue
                mode, delta = mode_of(i)
                b = base_of(i)
                am = modes[(y0 + by - 1) // 4, (x0 + bx) // 4] if by == 0 else 0
                lm = modes[(y0 + by) // 4, (x0 + bx - 1) // 4] if bx == 0 else 0
                smooth = am in (9, 10, 11) or lm in (9, 10, 11)
                bits = cost(tabs["KF_Y_MODE_CDF"][KF_CTX[am]][KF_CTX[lm]], 13, b)
                if 1 <= b <= 8 and bw >= 8 and bh >= 8:
                    bits = np.float32(bits + cost(tabs["ANGLE_DELTA_CDF"][b - 1], 7, delta + 3))
                elif b == 0:
                    use = int(i > 60)
                    bits = np.float32(bits + cost(tabs["FILTER_INTRA_CDF"][BSIZE[(bw, bh)]:], 2, use))
                    if use:
                        bits = np.float32(bits + cost(tabs["FILTER_INTRA_MODE_CDF"], 5, i - 61))
                cd = np.zeros(1, mt.CAND_DTYPE)
                cd["x"], cd["y"] = x0 + bx, y0 + by
                cd["tx_size"] = TX[(bw, bh)]
                cd["mode"], cd["angle_delta"] = mode, delta
                cd["pflags"] = int(tr) | (int(bl) << 1) | (int(smooth) << 2)
                cd["planes"] = 1 << 2
                cd["qidx"] = qidx
                cd["bd_qm"] = qm | (1 << 4)
                cd["eob_bin"] = EOB_LUMA[min(bw, 32) * min(bh, 32)]
                tset = 0 if (bw >= 32 or bh >= 32) else 3 if (bw == 16 and bh == 16) else (1 if min(bw, bh) == 4 else 2)
                cd["txtp_set"] = tset
                cd["txtp_mode"] = FILTER_DIR[i - 61] if i > 60 else b
                cd["txtp"] = 1 if tset else 0
                cd["qctx"] = 1
                cd["wsse_scale"] = 1.0
                cd["lambda"] = lam
                cd["extra_bits"] = bits
                cands.append(cd)
                ids.append(i)
            res = O.eval_batch(np.concatenate(cands), [(img, 128, 128, 128, 128), (plane, 128, 128, 128, 128)])
            j = int(np.argmin(res["cost"]))
            best.append((res["cost"][j], ids[j]))
        horz = np.float32(np.float32(best[0][0] + best[1][0]) + np.float32(lam * part(Lv, ctx, 1)))
        vert = np.float32(np.float32(best[2][0] + best[3][0]) + np.float32(lam * part(Lv, ctx, 2)))
        opts = [(np.float32(nd["none_cost"]), 0), (horz, 3), (vert, 4)]
        if nd["split"] == 1 or s > 4:
            ssum = np.float32(0)
            for q in range(4):
                ssum = np.float32(ssum + nodes[sb][nid(Lv + 1, 2 * ix + (q & 1), 2 * iy + (q >> 1))]["cost"])
            opts.append((np.float32(ssum + np.float32(lam * part(Lv, ctx, 3))), 1))
        bc, bd_ = opts[0]
        for cst, d in opts[1:]:
            if cst < bc:
                bc, bd_ = cst, d
        assert int(nd["split"]) == bd_, (n, int(nd["split"]), bd_)
        assert np.float32(nd["cost"]) == bc, n
        if bd_ in (3, 4):
            k0 = 0 if bd_ == 3 else 2
            assert (int(nd["rmode0"]), int(nd["rmode1"])) == (best[k0][1], best[k0 + 1][1])
            checked_rect += 1
    assert checked_rect >= 2
