"""Pins of the chroma part of the oracle's whole-frame RDO (oracle/oracle_rdo.cpp), the
checker of mt_intra_rdo with chroma.

- the frame chroma quantizer laws against the reference's own assertions
  (src/quant.rs:481-482) and their piecewise definition at chosen points;
- reconstruction: U and V replayed block by block in coding order from the decisions alone
  (uv mode, CfL alphas) by Python code on the per-kernel oracle functions; the luma part is
  unchanged by chroma;
- decisions: every chroma block's uv mode (and CfL alphas) is the argmin over its candidates
  rebuilt in Python on the decision-time planes, priced with uv_mode_bits
  (src/coder/lossy_state.rs:1302) from the CDF tables.
"""
import ctypes

import numpy as np
import pytest

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import workload as W
from test_oracle_rdo import brute_avail, geo, nid, mode_of, base_of

ANGLE, FILTER, MIN4, CFL, RECT = 1, 2, 4, 8, 16
TX = {4: 0, 8: 1, 16: 2, 32: 3}
TXR = {(4, 4): 0, (8, 8): 1, (16, 16): 2, (32, 32): 3, (4, 8): 5, (8, 4): 6, (8, 16): 7, (16, 8): 8,
       (16, 32): 9, (32, 16): 10}
# chroma_tx_for_mode (src/coder/chroma_tx.rs:46) in the AV1 TxType numbering
UV_TXTP = {0: 0, 1: 1, 2: 2, 3: 0, 4: 3, 5: 1, 6: 2, 7: 2, 8: 1, 9: 3, 10: 1, 11: 2, 12: 3}


def test_chroma_laws_pinned():
    # the reference's assertions (src/quant.rs:481-482)
    assert O.chroma_laws(12, 0)[2] == 2
    assert O.chroma_laws(7, 0)[2] == 15
    # 4:2:0 shares the luma QM law (level 8 at the top band, src/quant.rs:471-476)
    for q in (1, 7, 12, 20):
        assert O.chroma_laws(q, 2)[2] == 8
    # the piecewise deltas at chosen points (src/quant.rs:170-240, tuning 28 / 28 / 6)
    want = {  # q: (4:2:0 dc, 4:2:0 ac, 4:4:4 dc, 4:4:4 ac)
        10: (0, 0, 0, 6), 20: (0, 0, 0, 6), 42: (-14, -14, 0, 6), 58: (-24, -24, 0, 12), 64: (-28, -28, 0, 13),
        96: (-28, -28, 0, 24), 150: (-28, -28, 0, 24), 161: (-28, -28, 0, 24),
        163: (-29, -28, -1, 24), 200: (-41, -28, -13, 24), 255: (-59, -28, -31, 24)}
    for q, (d2, a2, d0, a0) in want.items():
        assert O.chroma_laws(q, 2)[:2] == (d2, a2), q
        assert O.chroma_laws(q, 0)[:2] == (d0, a0), q
    from maroontree_amd import rdo as R
    for q in range(256):
        for sub in (0, 2):
            assert R.chroma_laws(q, sub) == O.chroma_laws(q, sub), (q, sub)


def chroma_frame(seed, w, h, bd, fmt):
    """Seeded Y, Cb, Cr planes of the RGB image recipe (4:2:0 for fmt 1, 4:4:4 for fmt 2)."""
    ss = 1 if fmt == 1 else 0
    y, u, v = W.synth_yuv(seed, w, h, bd, 2 if ss else 0)
    return y, u, v, ss


def uv_blocks(nodes_sb, x0, y0, ss):
    """Chroma blocks of one SB in coding order: (node, half, cx, cy, cw, ch, nt, luma mode,
    cfl_ok, angle_ok, (tr, bl))."""
    from test_oracle_rdo_rect import halves, rect_avail
    out, stack = [], [0]
    while stack:
        n = stack.pop()
        sp = int(nodes_sb[n]["split"])
        if sp == 2:
            continue
        L, ix, iy, s = geo(n)
        group = ss and s == 8 and sp != 0
        if sp == 1 and not group:
            for q in (3, 2, 1, 0):
                stack.append(nid(L + 1, 2 * ix + (q & 1), 2 * iy + (q >> 1)))
            continue
        x, y = ix * s, iy * s
        if sp == 0 or group:
            c = 4 if group else s >> ss
            if not group:
                ym = base_of(int(nodes_sb[n]["mode"]))
            elif sp == 1:
                ym = base_of(int(nodes_sb[nid(4, 2 * ix + 1, 2 * iy + 1)]["mode"]))
            else:
                ym = base_of(int(nodes_sb[n]["rmode1"]))
            out.append((n, 0, (x0 + x) >> ss, (y0 + y) >> ss, c, c, 4 if c == 64 else 1, ym, group or s <= 32,
                        not group and s >= 8, brute_avail(x, y, s)))
            continue
        k0 = 0 if sp == 3 else 2
        for j in range(2):
            bx, by, bw, bh = halves(x, y, s, k0 + j)
            ym = base_of(int(nodes_sb[n]["rmode0" if j == 0 else "rmode1"]))
            out.append((n, j, (x0 + bx) >> ss, (y0 + by) >> ss, bw >> ss, bh >> ss, 1, ym, True, bw >= 8 and bh >= 8,
                        rect_avail(x, y, s, k0 + j)))
    return out


def cfl_pred(dc, luma, lx, ly, cw, ch, ss, alpha, bd):
    L = O.plib()
    lum = np.ascontiguousarray(luma[ly:ly + (ch << ss), lx:lx + (cw << ss)], dtype=np.uint16)
    ac = np.zeros(cw * ch, np.int32)
    L.orc_cfl_ac_sub(lum.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), lum.shape[1], cw, ch, ss, ss,
                     ac.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return np.array([L.orc_cfl_pred_pixel(int(dc), int(a), int(alpha), bd) for a in ac], np.int32).reshape(ch, cw)


def replay_chroma(planes_src, luma_rec, nodes, W_, H_, ss, bd, qidx, uvq, min4):
    """U / V reconstructions from the decisions alone."""
    dc_d, ac_d, qm = uvq
    q = O.quant_chroma(qidx, dc_d, ac_d, bd, qm)
    Wc, Hc = W_ >> ss, H_ >> ss
    recs = [p.copy() for p in planes_src]
    uvmap = np.zeros((Hc // 4, Wc // 4), np.int64)
    C = (W_ + 63) // 64
    mx = (1 << bd) - 1
    for sb in range(nodes.shape[0]):
        r, c = divmod(sb, C)
        x0, y0 = 64 * c, 64 * r
        blocks = uv_blocks(nodes[sb], x0, y0, ss)
        for (n, half, cx, cy, cw, ch, nt, ym, cfl_ok, ang_ok, av) in blocks:
            m = int(nodes[sb][n]["uv_mode1" if half else "uv_mode"])
            assert m <= 13 and (m < 13 or cfl_ok)
            uvmap[cy // 4:(cy + ch) // 4, cx // 4:(cx + cw) // 4] = m
        for (n, half, cx, cy, cw, ch, nt, ym, cfl_ok, ang_ok, av) in blocks:
            nd = nodes[sb][n]
            m = int(nd["uv_mode1" if half else "uv_mode"])
            alphas = ((int(nd["cfl_alpha_u1"]), int(nd["cfl_alpha_v1"])) if half else
                      (int(nd["cfl_alpha_u"]), int(nd["cfl_alpha_v"])))
            tw, th = (32, 32) if nt == 4 else (cw, ch)
            for p in range(2):
                rec, src = recs[p], planes_src[p]
                for t in range(nt):
                    tx, ty = cx + (t & 1) * tw, cy + (t >> 1) * th
                    tr, bl = brute_avail((t & 1) * 32, (t >> 1) * 32, 32) if nt == 4 else av
                    am = uvmap[(ty - 1) // 4, tx // 4] if ty > 0 else 0
                    lm = uvmap[ty // 4, (tx - 1) // 4] if tx > 0 else 0
                    smooth = am in (9, 10, 11) or lm in (9, 10, 11)
                    if m == 13:
                        dc = O.intra_pred(0, 0, rec, tx, ty, tw, th, False, False, Wc, Hc, False, bd)[0, 0]
                        pred = cfl_pred(dc, luma_rec, tx << ss, ty << ss, tw, th, ss, alphas[p], bd)
                        typ = 0
                    else:
                        pred = O.intra_pred(m, 0, rec, tx, ty, tw, th, tr, bl, Wc, Hc, smooth, bd)
                        typ = UV_TXTP[m] if max(tw, th) < 32 else 0
                    resid = src[ty:ty + th, tx:tx + tw].astype(np.int32) - pred
                    lv, _ = O.fwd(TXR[(tw, th)], typ, resid.ravel(), q)
                    inv = O.inv(TXR[(tw, th)], typ, lv, q).reshape(th, tw)
                    rec[ty:ty + th, tx:tx + tw] = np.clip(pred + inv, 0, mx)
    return recs


@pytest.mark.parametrize("fmt,bd,qidx,w,h,flags", [
    (1, 10, 91, 136, 72, ANGLE | FILTER | MIN4 | CFL), (1, 8, 180, 128, 64, ANGLE | CFL),
    (2, 8, 53, 128, 72, MIN4 | CFL), (2, 12, 30, 72, 64, ANGLE),
    (1, 10, 60, 136, 72, ANGLE | FILTER | MIN4 | CFL | RECT), (2, 8, 91, 128, 72, ANGLE | MIN4 | CFL | RECT)])
def test_chroma_reconstruction_replays_from_the_decisions(fmt, bd, qidx, w, h, flags):
    y, u, v, ss = chroma_frame(31 + qidx, w, h, bd, fmt)
    qm = mt.qm_level_law(qidx, 2 if ss else 0)
    lam = O.rdo_lambda(qidx, bd, 2 if ss else 0)
    uvq = O.chroma_laws(qidx, 2 if ss else 0)
    rec, ru, rv, nodes, st = O.intra_rdo_yuv(y, u, v, bd, qidx, qm, flags, lam, fmt, *uvq)
    rec_l, nodes_l, _, st_l = O.intra_rdo(y, bd, qidx, qm, flags & ~CFL, lam)
    if flags & RECT:
        assert np.isin(nodes["split"], (3, 4)).any()
    assert np.array_equal(rec, rec_l)  # chroma does not change the luma decisions
    assert np.array_equal(nodes["mode"], nodes_l["mode"]) and np.array_equal(nodes["split"], nodes_l["split"])
    pu, pv = replay_chroma([u, v], rec, nodes, w, h, ss, bd, qidx, uvq, flags & MIN4)
    assert np.array_equal(ru, pu) and np.array_equal(rv, pv)
    assert st["sse_u"] == int(((ru.astype(np.int64) - u) ** 2).sum())
    assert st["sse_v"] == int(((rv.astype(np.int64) - v) ** 2).sum())
    assert st["bits_uv"] > 0 and st["bits"] > st["bits_uv"]
    coded = nodes["uv_mode"] != 255
    assert coded.any()
    if flags & RECT:
        assert (nodes["uv_mode1"] != 255).any()  # chroma of second rectangular halves
    if flags & CFL:
        assert (nodes["uv_mode"] == 13).any()  # CfL wins somewhere on these images
    else:
        assert not (nodes["uv_mode"] == 13).any()


def ref_walk(a0, c0, steps, eval_cost):
    """cfl_alpha_rd_refine (src/cost.rs:332) as the reference writes it, over a cost callback."""
    if steps == 0 or a0 == 0:
        return a0
    best_a, best_c = a0, c0
    d = 1 if a0 > 0 else -1
    for s in range(1, steps + 1):
        a = a0 - d * s
        if a * d < 0:
            break
        c = eval_cost(a)
        if c < best_c:
            best_a, best_c = a, c
        else:
            break
    if best_a == a0:
        for s in range(1, steps + 1):
            a = a0 + d * s
            if abs(a) > 16:
                break
            c = eval_cost(a)
            if c < best_c:
                best_a, best_c = a, c
            else:
                break
    return best_a


@pytest.mark.parametrize("steps", [0, 3])
def test_chroma_decisions_are_the_argmin_over_candidates(steps):
    """Rebuild every chroma candidate of the last SB of a 2 x 2 SB frame (4:2:0, CfL) on the
    decision-time planes (final reconstruction outside the SB, source inside) and check each
    block's uv mode and CfL alphas; with CfL walk steps, each plane's alpha walked from its
    energy optimum over candidates rebuilt at given alphas (U against V's optimum, then V
    against U's result; walk cost dist + lambda * (bits + uv_mode_bits of the pair))."""
    bd, qidx, flags, fmt = 10, 91, ANGLE | MIN4 | CFL | (steps << 8), 1
    y, u, v, ss = chroma_frame(77, 128, 128, bd, fmt)
    qm = mt.qm_level_law(qidx, 2)
    lam = np.float32(O.rdo_lambda(qidx, bd, 2))
    uvq = O.chroma_laws(qidx, 2)
    rec, ru, rv, nodes, _ = O.intra_rdo_yuv(y, u, v, bd, qidx, qm, flags, float(lam), fmt, *uvq)
    tabs = {k: np.ascontiguousarray(t, dtype=np.uint16) for k, t in O.cdf_tables().items()
            if k in ("UV_MODE_CFL_CDF", "UV_MODE_NOCFL_CDF", "CFL_SIGN_CDF", "CFL_ALPHA_CDF", "ANGLE_DELTA_CDF")}
    L = O.lib()
    L.orc_cdf_cost_raw.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.orc_cdf_cost_raw.restype = ctypes.c_float
    cost = lambda row, n, s: np.float32(L.orc_cdf_cost_raw(row.ctypes.data, n, s))  # noqa: E731

    def uv_bits(ym, m, au, av, cfl_ok, ang_ok):
        if m == 13:
            su = 0 if au == 0 else 1 if au < 0 else 2
            sv = 0 if av == 0 else 1 if av < 0 else 2
            if su == 0 and sv == 0:  # both zero: priced as DC_PRED (src/coder/lossy_state.rs:1333)
                return cost(tabs["UV_MODE_CFL_CDF"][ym], 14, 0)
            b = np.float32(cost(tabs["UV_MODE_CFL_CDF"][ym], 14, 13) + cost(tabs["CFL_SIGN_CDF"], 8, su * 3 + sv - 1))
            if su:
                b = np.float32(b + cost(tabs["CFL_ALPHA_CDF"][(su == 2) * 3 + sv], 16, abs(au) - 1))
            if sv:
                b = np.float32(b + cost(tabs["CFL_ALPHA_CDF"][(sv == 2) * 3 + su], 16, abs(av) - 1))
            return b
        b = cost(tabs["UV_MODE_CFL_CDF"][ym], 14, m) if cfl_ok else cost(tabs["UV_MODE_NOCFL_CDF"][ym], 13, m)
        if ang_ok and 1 <= m <= 8:
            b = np.float32(b + cost(tabs["ANGLE_DELTA_CDF"][m - 1], 7, 3))
        return b

    sb, x0, y0 = 3, 64, 64
    pu, pv = ru.copy(), rv.copy()
    pu[32:64, 32:64] = u[32:64, 32:64]
    pv[32:64, 32:64] = v[32:64, 32:64]
    planes = [(y, 128, 128, 128, 128), (rec, 128, 128, 128, 128), (u, 64, 64, 64, 64), (pu, 64, 64, 64, 64),
              (v, 64, 64, 64, 64), (pv, 64, 64, 64, 64)]
    uvmap = np.zeros((16, 16), np.int64)  # final uv modes of the other SBs
    for s2 in range(3):
        r, c = divmod(s2, 2)
        for (n, half, cx, cy, cw, ch, *_rest) in uv_blocks(nodes[s2], 64 * c, 64 * r, ss):
            uvmap[cy // 4:(cy + ch) // 4, cx // 4:(cx + cw) // 4] = int(nodes[s2][n]["uv_mode1" if half else "uv_mode"])
    checked = moved = 0
    for (n, half, cx, cy, cw, ch, nt, ym, cfl_ok, ang_ok, (tr, bl)) in uv_blocks(nodes[sb], x0, y0, ss):
        cs = cw
        am = uvmap[(cy - 1) // 4, cx // 4] if cy == 32 else 0
        lm = uvmap[cy // 4, (cx - 1) // 4] if cx == 32 else 0
        smooth = am in (9, 10, 11) or lm in (9, 10, 11)
        best, bm, ba = np.float32(np.inf), None, (0, 0)

        def cands(m, given=None):
            cd = np.zeros(2, mt.CAND_DTYPE)
            for p in range(2):
                c = cd[p]
                c["x"], c["y"], c["tx_size"] = cx, cy, TX[cs]
                c["mode"] = m
                if m == 13:
                    c["pflags"] = 8 | 16 | (32 if given is not None else 0)
                    c["angle_delta"] = 0 if given is None else given[p]
                    c["txtp_mode"] = 1
                else:
                    c["tx_type"] = UV_TXTP[m] if cs < 32 else 0
                    c["pflags"] = int(tr) | (int(bl) << 1) | (int(smooth) << 2)
                sp, rp = 2 + 2 * p, 3 + 2 * p
                c["planes"] = (sp & 3) | ((rp & 3) << 2)
                c["planes_hi"] = (sp >> 2) | ((rp >> 2) << 4)
                c["qidx"] = qidx
                c["bd_qm"] = uvq[2] | (1 << 4) | 0x40
                c["dc_delta"], c["ac_delta"] = uvq[0], uvq[1]
                c["rplane"] = 1
                c["eob_bin"] = {4: 1, 8: 6, 16: 12, 32: 16}[cs]  # the chroma eob_bin CDFs
                c["skip_ctx"] = 7
                c["qctx"] = 2
                c["lambda"] = lam
            return O.eval_batch(cd, planes)

        for m in range(14 if cfl_ok else 13):
            res = cands(m)
            body = [res["cost"][0], res["cost"][1]]
            au, av = (int(res["alpha"][0]), int(res["alpha"][1])) if m == 13 else (0, 0)
            if m == 13 and steps:
                al = [au, av]
                for p in range(2):
                    def walk_cost(a, p=p):
                        r = cands(13, (a, a))[p]
                        pair = (a, al[1]) if p == 0 else (al[0], a)
                        return np.float32(np.float32(r["dist"]) +
                                          np.float32(lam * np.float32(r["bits"] + uv_bits(ym, 13, *pair, True, ang_ok))))
                    a1 = ref_walk(al[p], walk_cost(al[p]), steps, walk_cost)
                    if a1 != al[p]:
                        body[p] = cands(13, (a1, a1))["cost"][p]
                        moved += 1
                    al[p] = a1
                au, av = al
            s = np.float32(np.float32(0) + body[0])
            s = np.float32(s + body[1])
            if m == 13 and au == 0 and av == 0:
                continue
            tot = np.float32(s + np.float32(lam * uv_bits(ym, m, au, av, cfl_ok, ang_ok)))
            if tot < best:
                best, bm, ba = tot, m, (au, av)
        nd = nodes[sb][n]
        assert int(nd["uv_mode"]) == bm, (n, int(nd["uv_mode"]), bm)
        assert (int(nd["cfl_alpha_u"]), int(nd["cfl_alpha_v"])) == ba
        checked += 1
    assert checked >= 4
    print(f"{checked} blocks, {moved} walked alphas")
    assert moved > 0 or not steps
