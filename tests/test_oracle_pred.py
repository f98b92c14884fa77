"""Pins of the intra-prediction, CfL and distortion oracle (oracle/oracle_pred.cpp).

Intra prediction is AV1-normative.  The oracle (a restatement of the reference's
src/intrapred.rs, row-structured like the reference) is checked against an independent,
per-pixel restatement of the AV1 specification's intra prediction process (section 7.11.2:
edge preparation 7.11.2, Paeth 7.11.2.2, filter intra 7.11.2.3, directional 7.11.2.4, DC
7.11.2.5, smooth 7.11.2.6, edge filter strength / upsample selection 7.11.2.9-10, upsample
7.11.2.11, edge filter 7.11.2.12), vectorised in numpy below.  The zone-3 read is clamped at
maxBaseY as dav1d and libaom do.  The distortion kernels are pinned by brute force (SSE,
residual), an independent Hadamard-matrix form (SATD), an f64 transform (weighted SSE) and
the reference's known answers (src/rd_sse.rs:1982-2013, src/partition_rd.rs:149).
"""
import ctypes
import os
import re
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [(4, 4), (4, 8), (8, 4), (4, 16), (16, 4), (8, 8), (8, 16), (16, 8), (16, 16),
         (16, 32), (32, 16), (32, 32), (64, 64), (8, 32), (32, 8), (16, 64), (64, 16),
         (32, 64), (64, 32)]


def _hdr_table(name):
    hdr = open(os.path.join(ROOT, "maroontree_amd", "csrc", "av1_cdf_tables.h")).read()
    s = hdr.index(f"MT_{name}[")
    body = hdr[hdr.index("{", s) + 1:hdr.index("};", s)]
    return np.array([int(x) for x in re.findall(r"-?\d+", body)])


DERIV = _hdr_table("DR_INTRA_DERIVATIVE")
SMW = _hdr_table("SM_WEIGHTS")
TAPS = _hdr_table("INTRA_FILTER_TAPS").reshape(5, 8, 7)
NOMINAL = {1: 90, 2: 180, 3: 45, 4: 135, 5: 113, 6: 157, 7: 203, 8: 67}
KERNEL = [[0, 4, 8, 4, 0], [0, 5, 6, 5, 0], [2, 4, 4, 4, 2]]

# ---- the specification's process, per pixel ---------------------------------------------

OFF = 16  # edge arrays: logical index i lives at [i + OFF]


def spec_strength(w, h, ftype, delta):
    """7.11.2.9, as the spec's threshold table: (max w+h, [(min d, strength), ...])."""
    d, wh = abs(delta), w + h
    table0 = [(8, [(56, 1)]), (12, [(40, 1)]), (16, [(40, 1)]),
              (24, [(8, 1), (16, 2), (32, 3)]), (32, [(1, 1), (4, 2), (32, 3)]),
              (10 ** 9, [(1, 3)])]
    table1 = [(8, [(40, 1), (64, 2)]), (16, [(20, 1), (48, 2)]), (24, [(4, 3)]),
              (10 ** 9, [(1, 3)])]
    for lim, steps in (table1 if ftype else table0):
        if wh <= lim:
            s = 0
            for dmin, val in steps:
                if d >= dmin:
                    s = val
            return s


def spec_upsample_sel(w, h, ftype, delta):
    """7.11.2.10"""
    d = abs(delta)
    if d <= 0 or d >= 40:
        return 0
    return int(w + h <= (8 if ftype else 16))


def spec_filter_edge(buf, sz, strength):
    """7.11.2.12 on buf (logical index -1 .. sz-2 rewritten)."""
    if strength == 0:
        return
    edge = [int(buf[i - 1 + OFF]) for i in range(sz)]
    for i in range(1, sz):
        s = 0
        for j in range(5):
            k = min(max(i - 2 + j, 0), sz - 1)
            s += KERNEL[strength - 1][j] * edge[k]
        buf[i - 1 + OFF] = (s + 8) >> 4


def spec_upsample(buf, num_px, bd):
    """7.11.2.11"""
    dup = [0] * (num_px + 3)
    dup[0] = int(buf[-1 + OFF])
    for i in range(-1, num_px):
        dup[i + 2] = int(buf[i + OFF])
    dup[num_px + 2] = int(buf[num_px - 1 + OFF])
    buf[-2 + OFF] = dup[0]
    for i in range(num_px):
        s = -dup[i] + 9 * dup[i + 1] + 9 * dup[i + 2] - dup[i + 3]
        buf[2 * i - 1 + OFF] = min(max((s + 8) >> 4, 0), (1 << bd) - 1)
        buf[2 * i + OFF] = dup[i + 2]


def spec_edges(recon, x, y, w, h, have_tr, have_bl, fw, fh, bd):
    """7.11.2 steps building AboveRow[-1 .. w+h-1] and LeftCol[-1 .. w+h-1]."""
    ha, hl = y > 0, x > 0
    max_x, max_y = fw - 1, fh - 1
    n = w + h
    above = np.zeros(n * 2 + 2 * OFF, np.int64)
    left = np.zeros(n * 2 + 2 * OFF, np.int64)
    for i in range(n):
        if not ha and hl:
            above[i + OFF] = recon[y, x - 1]
        elif not ha:
            above[i + OFF] = (1 << (bd - 1)) - 1
        else:
            lim = min(max_x, x + (2 * w if have_tr else w) - 1)
            above[i + OFF] = recon[y - 1, min(lim, x + i)]
        if not hl and ha:
            left[i + OFF] = recon[y - 1, x]
        elif not hl:
            left[i + OFF] = (1 << (bd - 1)) + 1
        else:
            lim = min(max_y, y + (2 * h if have_bl else h) - 1)
            left[i + OFF] = recon[min(lim, y + i), x - 1]
    if ha and hl:
        c = recon[y - 1, x - 1]
    elif ha:
        c = recon[y - 1, x]
    elif hl:
        c = recon[y, x - 1]
    else:
        c = 1 << (bd - 1)
    above[-1 + OFF] = left[-1 + OFF] = c
    return above, left


def spec_predict(mode, delta, recon, x, y, w, h, have_tr, have_bl, fw, fh, ftype, bd):
    above, left = spec_edges(recon, x, y, w, h, have_tr, have_bl, fw, fh, bd)
    ha, hl = y > 0, x > 0
    i = np.arange(h)[:, None]  # row
    j = np.arange(w)[None, :]  # column
    A = lambda k: above[np.asarray(k) + OFF]  # noqa: E731
    L = lambda k: left[np.asarray(k) + OFF]  # noqa: E731
    if mode == 0:  # 7.11.2.5
        if ha and hl:
            s = above[OFF:OFF + w].sum() + left[OFF:OFF + h].sum()
            v = (s + ((w + h) >> 1)) // (w + h)
        elif ha:
            v = (above[OFF:OFF + w].sum() + (w >> 1)) >> (w.bit_length() - 1)
        elif hl:
            v = (left[OFF:OFF + h].sum() + (h >> 1)) >> (h.bit_length() - 1)
        else:
            v = 1 << (bd - 1)
        return np.full((h, w), v, np.int64)
    if mode >= 16:  # 7.11.2.3
        pred = np.zeros((h, w), np.int64)
        for i2 in range(h >> 1):
            for j4 in range(w >> 2):
                p = []
                for k in range(7):
                    if k < 5:
                        if i2 == 0:
                            p.append(A((j4 << 2) + k - 1))
                        elif j4 == 0 and k == 0:
                            p.append(L((i2 << 1) - 1))
                        else:
                            p.append(pred[(i2 << 1) - 1, (j4 << 2) + k - 1])
                    else:
                        if j4 == 0:
                            p.append(L((i2 << 1) + k - 5))
                        else:
                            p.append(pred[(i2 << 1) + k - 5, (j4 << 2) - 1])
                for k in range(8):
                    pr = int(sum(int(TAPS[mode - 16, k, t]) * int(p[t]) for t in range(7)))
                    r = (pr + 8) >> 4 if pr >= 0 else -((-pr + 8) >> 4)
                    pred[(i2 << 1) + (k >> 2), (j4 << 2) + (k & 3)] = min(max(r, 0), (1 << bd) - 1)
        return pred
    if mode == 12:  # 7.11.2.2
        base = A(j) + L(i) - A(-1)
        pl, pt, ptl = abs(base - L(i)), abs(base - A(j)), abs(base - A(-1))
        return np.where((pl <= pt) & (pl <= ptl), L(i) + 0 * j,
                        np.where(pt <= ptl, A(j) + 0 * i, A(-1) + 0 * i * j))
    if mode in (9, 10, 11):  # 7.11.2.6
        wy, wx = SMW[h + i], SMW[w + j]
        if mode == 9:
            s = wy * A(j) + (256 - wy) * L(h - 1) + wx * L(i) + (256 - wx) * A(w - 1)
            return (s + 256) >> 9
        if mode == 10:
            return (wy * A(j) + (256 - wy) * L(h - 1) + 128 + 0 * j) >> 8
        return (wx * L(i) + (256 - wx) * A(w - 1) + 128 + 0 * i) >> 8
    # 7.11.2.4 directional
    p = NOMINAL[mode] + 3 * delta
    upa = upl = 0
    if p != 90 and p != 180:
        if 90 < p < 180 and w + h >= 24:
            c = (L(0) * 5 + A(-1) * 6 + A(0) * 5 + 8) >> 4
            above[-1 + OFF] = left[-1 + OFF] = c
        if ha:
            num = min(w, fw - x) + (h if p < 90 else 0) + 1
            spec_filter_edge(above, num, spec_strength(w, h, ftype, p - 90))
        if hl:
            num = min(h, fh - y) + (w if p > 180 else 0) + 1
            spec_filter_edge(left, num, spec_strength(h, w, ftype, p - 180))
        upa = spec_upsample_sel(w, h, ftype, p - 90)
        if upa:
            spec_upsample(above, w + (h if p < 90 else 0), bd)
        upl = spec_upsample_sel(h, w, ftype, p - 180)
        if upl:
            spec_upsample(left, h + (w if p > 180 else 0), bd)
    if p == 90:
        return A(j) + 0 * i
    if p == 180:
        return L(i) + 0 * j
    dx = int(DERIV[p >> 1]) if p < 90 else (int(DERIV[(180 - p) >> 1]) if p < 180 else 0)
    dy = int(DERIV[(p - 90) >> 1]) if 90 < p < 180 else (int(DERIV[(270 - p) >> 1]) if p > 180 else 0)
    interp = lambda e, b, s: (e(b) * (32 - s) + e(b + 1) * s + 16) >> 5  # noqa: E731
    if p < 90:
        idx = (i + 1) * dx
        base = (idx >> (6 - upa)) + (j << upa)
        shift = ((idx << upa) >> 1) & 0x1F
        mx = (w + h - 1) << upa
        safe = np.minimum(base, mx - 1)
        return np.where(base < mx, interp(A, safe, shift), A(mx))
    if p > 180:
        idx = (j + 1) * dy
        base = (idx >> (6 - upl)) + (i << upl)
        shift = ((idx << upl) >> 1) & 0x1F
        my = (w + h - 1) << upl
        safe = np.minimum(base, my - 1)
        return np.where(base < my, interp(L, safe, shift), L(my))
    idx_a = (j << 6) - (i + 1) * dx
    base_a = idx_a >> (6 - upa)
    use_a = base_a >= -(1 << upa)
    sa = ((idx_a << upa) >> 1) & 0x1F
    idx_l = (i << 6) - (j + 1) * dy
    base_l = idx_l >> (6 - upl)
    sl = ((idx_l << upl) >> 1) & 0x1F
    va = interp(A, np.where(use_a, base_a, 0), sa)
    vl = interp(L, np.where(use_a, 0, base_l), sl)
    return np.where(use_a, va, vl)


def _plane(rng, fh, fw, bd, smooth=False):
    if smooth:
        base = rng.integers(0, 1 << bd)
        walk = np.cumsum(np.cumsum(rng.integers(-3, 4, (fh, fw)), 0), 1)
        return np.clip(base + walk, 0, (1 << bd) - 1).astype(np.uint16)
    return rng.integers(0, 1 << bd, (fh, fw)).astype(np.uint16)


# ---- tests -------------------------------------------------------------------------------

def test_edge_filter_strength_and_upsample_selection_match_spec_tables():
    """intra_edge_filter_strength / use_intra_edge_upsample (src/intrapred.rs:917, :1021)
    against the spec's selection tables, over every (w, h, type, delta) the encoder uses."""
    L = O.plib()
    for w in (4, 8, 16, 32, 64):
        for h in (4, 8, 16, 32, 64):
            for ft in (0, 1):
                for d in range(-120, 121):
                    assert L.orc_edge_filter_strength(w, h, ft, d) == spec_strength(w, h, ft, d)
                    assert L.orc_use_edge_upsample(w, h, ft, d) == spec_upsample_sel(w, h, ft, d)


def test_edge_filter_and_upsample_match_spec():
    """filter_intra_edge / upsample_intra_edge (src/intrapred.rs:977, :1026) against 7.11.2.12
    and 7.11.2.11 on random edges of every length the predictor uses."""
    L = O.plib()
    rng = np.random.default_rng(5)
    for size in range(2, 130):
        for strength in (1, 2, 3):
            e = rng.integers(0, 4096, 132).astype(np.int32)
            want = np.zeros(132 + 2 * OFF, np.int64)
            want[OFF - 2:OFF - 2 + 132] = e
            spec_filter_edge(want, size, strength)
            got = e.copy()
            L.orc_filter_edge(O.ptr(got, ctypes.c_int32), size, strength)
            assert (got == want[OFF - 2:OFF - 2 + 132]).all(), (size, strength)
    for num in range(1, 33):
        for bd in (8, 10, 12):
            e = rng.integers(0, 1 << bd, 132).astype(np.int32)
            want = np.zeros(132 + 2 * OFF, np.int64)
            want[OFF - 2:OFF - 2 + 132] = e
            spec_upsample(want, num, bd)
            got = e.copy()
            L.orc_upsample_edge(O.ptr(got, ctypes.c_int32), num, bd)
            assert (got == want[OFF - 2:OFF - 2 + 132]).all(), (num, bd)


def test_dr_intra_derivative_table():
    """DR_INTRA_DERIVATIVE (src/intrapred.rs:1081): every angle the predictor can index has a
    non-zero, strictly decreasing step, 64 at 45 degrees, within 8 % of 64*cot(angle) from 14
    degrees up (the table's values are those steps in 1/64 px)."""
    assert DERIV[45 >> 1] == 64
    used = sorted({a for base in (45, 67, 90, 113, 135, 157, 180, 203) for d in range(-3, 4)
                   for a in (base + 3 * d, 180 - base - 3 * d, base + 3 * d - 90,
                             270 - base - 3 * d) if 0 < a < 90})
    assert used == [3, 6, 9, 14, 17, 20, 23, 26, 29, 32, 36, 39, 42, 45, 48, 51, 54, 58, 61,
                    64, 67, 70, 73, 76, 81, 84, 87]
    vals = [int(DERIV[p >> 1]) for p in used]
    assert all(v > 0 for v in vals) and all(a > b for a, b in zip(vals, vals[1:]))
    for p, v in zip(used, vals):
        if p >= 14:
            assert abs(v - 64 / np.tan(np.radians(p))) <= max(0.08 * v, 0.6), p


def test_smooth_weights_shape():
    """sm_weights (src/intrapred.rs:1091): each size starts at 255, is non-increasing, and its
    last weight is the AV1 end point (64, 32, 16, 8, 4)."""
    for n, last in ((4, 64), (8, 32), (16, 16), (32, 8), (64, 4)):
        w = SMW[n:2 * n]
        assert w[0] == 255 and w[-1] == last and all(np.diff(w) <= 0)


def test_directional_core_matches_spec_per_pixel():
    """dr_predict_scalar (src/intrapred.rs:629) zones 1/2/3 with and without upsampled edges
    against the per-pixel 7.11.2.4 formulas, on random prepared edges."""
    rng = np.random.default_rng(11)
    for (bw, bh) in SIZES:
        n = bw + bh
        for p in sorted({b + 3 * d for b in NOMINAL.values() for d in range(-3, 4)} - {90, 180}):
            ups = (0, 1) if n <= 16 else (0,)
            for up in ups:
                e_a = rng.integers(0, 4096, 132).astype(np.int32)
                e_l = rng.integers(0, 4096, 132).astype(np.int32)
                zone = 1 if p < 90 else (2 if p < 180 else 3)
                dx = DERIV[p >> 1] if p < 90 else (DERIV[(180 - p) >> 1] if p < 180 else 0)
                dy = DERIV[(p - 90) >> 1] if 90 < p < 180 else (DERIV[(270 - p) >> 1] if p > 180 else 0)
                upa = up if zone in (1, 2) else 0
                upl = up if zone in (2, 3) else 0
                got = O.dr_predict(zone, bw, bh, n, int(dx), int(dy), upa, upl, e_a, e_l)
                A = lambda k: e_a[np.asarray(k) + 2].astype(np.int64)  # noqa: E731
                L = lambda k: e_l[np.asarray(k) + 2].astype(np.int64)  # noqa: E731
                i = np.arange(bh)[:, None]
                j = np.arange(bw)[None, :]
                interp = lambda e, b, s: (e(b) * (32 - s) + e(b + 1) * s + 16) >> 5  # noqa: E731
                if zone == 1:
                    idx = (i + 1) * int(dx)
                    base = (idx >> (6 - upa)) + (j << upa)
                    sh = ((idx << upa) >> 1) & 0x1F
                    mx = (n - 1) << upa
                    want = np.where(base < mx, interp(A, np.minimum(base, mx - 1), sh), A(mx))
                elif zone == 3:
                    idx = (j + 1) * int(dy)
                    base = (idx >> (6 - upl)) + (i << upl)
                    sh = ((idx << upl) >> 1) & 0x1F
                    my = (n - 1) << upl
                    want = np.where(base < my, interp(L, np.minimum(base, my - 1), sh), L(my))
                else:
                    ia = (j << 6) - (i + 1) * int(dx)
                    ba = ia >> (6 - upa)
                    ua = ba >= -(1 << upa)
                    il = (i << 6) - (j + 1) * int(dy)
                    bl = il >> (6 - upl)
                    want = np.where(ua, interp(A, np.where(ua, ba, 0), ((ia << upa) >> 1) & 31),
                                    interp(L, np.where(ua, 0, bl), ((il << upl) >> 1) & 31))
                assert (got == want).all(), (bw, bh, p, up)


AVAIL = [(0, 0, False, False), (8, 0, True, False), (0, 8, False, True), (8, 8, True, True),
         (8, 8, False, False), (8, 8, True, False), (8, 8, False, True)]


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_every_directional_mode_matches_spec(bd):
    """predict_nd_ad (src/intrapred.rs:359-1595) for V..D67 x angle delta -3..3 x every
    square and rectangular size x edge availability x smooth-neighbour filter type, from a
    recon plane, against the whole 7.11.2 process (edge preparation, corner filter, edge
    filter, upsample, projection).  Half the planes are smooth random walks so the filters
    see correlated edges."""
    rng = np.random.default_rng(100 + bd)
    planes = [_plane(rng, 160, 160, bd), _plane(rng, 160, 160, bd, smooth=True)]
    count = 0
    for (bw, bh) in SIZES:
        for mode in range(1, 9):
            for delta in range(-3, 4):
                ox, oy, tr, bl = AVAIL[count % len(AVAIL)]
                ft = bool((count // 7) & 1)
                recon = planes[count & 1]
                count += 1
                got = O.intra_pred(mode, delta, recon, ox, oy, bw, bh, tr, bl, 160, 160, ft, bd)
                want = spec_predict(mode, delta, recon, ox, oy, bw, bh, tr, bl, 160, 160, ft, bd)
                assert (got == want).all(), (bw, bh, mode, delta, ox, oy, tr, bl, ft)


def test_nondirectional_modes_match_spec():
    """DC (dc_pred_from_sum's 0x5556 / 0x3334 reciprocal multiply, :1772, against the
    spec's integer division), SMOOTH / SMOOTH_V / SMOOTH_H, PAETH, on every size, every
    availability, 8/10/12-bit, with extreme planes (all max, all zero, alternating)."""
    rng = np.random.default_rng(7)
    for bd in (8, 10, 12):
        mx = (1 << bd) - 1
        planes = [_plane(rng, 140, 140, bd), _plane(rng, 140, 140, bd, smooth=True),
                  np.full((140, 140), mx, np.uint16),
                  ((np.indices((140, 140)).sum(0) & 1) * mx).astype(np.uint16)]
        for (bw, bh) in SIZES:
            for mode in (0, 9, 10, 11, 12):
                for k, (ox, oy, tr, bl) in enumerate(AVAIL[:4]):
                    recon = planes[(k + mode) % len(planes)]
                    got = O.intra_pred(mode, 0, recon, ox, oy, bw, bh, tr, bl, 140, 140, 0, bd)
                    want = spec_predict(mode, 0, recon, ox, oy, bw, bh, tr, bl, 140, 140, 0, bd)
                    assert (got == want).all(), (bd, bw, bh, mode, ox, oy)


def test_dc_reciprocal_multiply_vs_exact_division():
    """dc_pred_from_sum (src/intrapred.rs:1772) uses dav1d's 8-bit multipliers (0x5556 for
    2:1, 0x3334 for 4:1, >> 16) at every bit depth.  They equal the spec's rounded division
    for every edge sum of an 8- or 10-bit block and of a 12-bit 2:1 block; for 12-bit 4:1
    blocks the result is one above the division once the sum reaches 16*4095 (4x16),
    32*4095 (8x32), 64*4095 (16x64) -- the reference's behaviour, kept."""
    L = O.plib()
    first_off = {(4, 16): 65526, (16, 4): 65526, (8, 32): 131052, (32, 8): 131052,
                 (16, 64): 262104, (64, 16): 262104}
    for bd in (8, 10, 12):
        mx = (1 << bd) - 1
        for (w, h) in [(4, 8), (8, 4), (8, 16), (16, 8), (16, 32), (32, 16), (32, 64), (64, 32),
                       (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]:
            n = w + h
            sums = list(range(0, mx * n + 1, max(1, mx * n // 4000))) + [mx * n]
            if bd == 12 and (w, h) in first_off:
                f = first_off[(w, h)]
                sums += list(range(f - 40, f + 40))
            for s in sums:
                got = L.orc_dc_pred_from_sum(s, w, h, 1, 1, bd)
                want = (s + (n >> 1)) // n
                if bd == 12 and (w, h) in first_off and s >= first_off[(w, h)]:
                    assert got - want in (0, 1), (w, h, s)
                else:
                    assert got == want, (bd, w, h, s)
            if bd == 12 and (w, h) in first_off:
                s = first_off[(w, h)]
                assert L.orc_dc_pred_from_sum(s, w, h, 1, 1, 12) == (s + (n >> 1)) // n + 1


def test_filter_intra_matches_spec():
    """filter_intra_predict_impl (src/intrapred.rs:473) for the 5 modes on every allowed
    size (max(W, H) <= 32) and availability, against 7.11.2.3 (Round2Signed, Clip1)."""
    rng = np.random.default_rng(3)
    for bd in (8, 10):
        recon = _plane(rng, 80, 80, bd)
        for (bw, bh) in SIZES:
            if max(bw, bh) > 32:
                continue
            for fm in range(5):
                for (ox, oy, _, _) in AVAIL[:4]:
                    got = O.intra_pred(16 + fm, 0, recon, ox, oy, bw, bh, 0, 0, 80, 80, 0, bd)
                    want = spec_predict(16 + fm, 0, recon, ox, oy, bw, bh, 0, 0, 80, 80, 0, bd)
                    assert (got == want).all(), (bd, bw, bh, fm, ox, oy)


def test_frame_edge_limits_top_right_and_bottom_left():
    """The px_have rules (src/intrapred.rs:1425-1459): a block near the right / bottom frame
    edge gets only the real samples left in the frame, then replication (spec aboveLimit /
    leftLimit)."""
    rng = np.random.default_rng(9)
    recon = _plane(rng, 128, 128, 10)
    for (bw, bh) in [(8, 8), (16, 8), (8, 16), (16, 16), (32, 16)]:
        for spare in (0, 1, 2, 3, 5, 40):
            fw, fh = 8 + bw + spare, 8 + bh + spare
            for mode in (3, 7, 8, 1, 2):
                for delta in (-3, 0, 2):
                    got = O.intra_pred(mode, delta, recon, 8, 8, bw, bh, 1, 1, fw, fh, 0, 10)
                    want = spec_predict(mode, delta, recon, 8, 8, bw, bh, 1, 1, fw, fh, 0, 10)
                    assert (got == want).all(), (bw, bh, spare, mode, delta)


def test_known_answers_diagonals_and_fills():
    """Closed forms: D45 on a 4x4 with filters off copies above[x+y+1] (including the top-right
    samples); D135 copies the diagonal (corner on it); with no neighbours, V gives 2^(bd-1)-1,
    H 2^(bd-1)+1 and DC 2^(bd-1); a constant plane predicts itself in every mode."""
    rng = np.random.default_rng(1)
    recon = _plane(rng, 32, 32, 10)
    got = O.intra_pred(3, 0, recon, 8, 8, 4, 4, 1, 1, 32, 32, 0, 10)
    top = recon[7, 8:16].astype(np.int64)
    for y in range(4):
        for x in range(4):
            assert got[y, x] == top[min(x + y + 1, 7)]
    got = O.intra_pred(4, 0, recon, 8, 8, 4, 4, 1, 1, 32, 32, 0, 10)
    for y in range(4):
        for x in range(4):
            want = (recon[7, 8 + x - y - 1] if x > y else
                    (recon[7, 7] if x == y else recon[8 + y - x - 1, 7]))
            assert got[y, x] == want, (y, x)
    for bd in (8, 10, 12):
        b = 1 << (bd - 1)
        assert (O.intra_pred(1, 0, recon, 0, 0, 8, 8, 0, 0, 32, 32, 0, bd) == b - 1).all()
        assert (O.intra_pred(2, 0, recon, 0, 0, 8, 8, 0, 0, 32, 32, 0, bd) == b + 1).all()
        assert (O.intra_pred(0, 0, recon, 0, 0, 8, 8, 0, 0, 32, 32, 0, bd) == b).all()
    const = np.full((160, 160), 777, np.uint16)
    for (bw, bh) in SIZES:
        for mode in list(range(13)) + ([16, 17, 18, 19, 20] if max(bw, bh) <= 32 else []):
            for delta in ((-3, -1, 0, 2, 3) if 1 <= mode <= 8 else (0,)):
                for ox, oy, tr, bl in AVAIL[1:]:
                    for ft in (0, 1):
                        p = O.intra_pred(mode, delta, const, ox, oy, bw, bh, tr, bl, 160, 160, ft, 12)
                        assert (p == 777).all(), (mode, delta, bw, bh)


# ---- CfL -----------------------------------------------------------------------------------

def test_cfl_ac_and_pred_closed_forms():
    """cfl_ac_444 / cfl_ac_sub (src/intrapred.rs:1112, :1133): the subsampled average scaled to
    1/8-pel (<<3 of the mean of the covered luma), minus the rounded block mean; cfl_pred_pixel
    (:1194) against the spec's Round2Signed(alpha * L, 6) + dc, clipped (7.11.5)."""
    L = O.plib()
    rng = np.random.default_rng(21)
    for (cw, ch) in [(4, 4), (8, 8), (16, 16), (32, 32), (4, 8), (16, 8), (8, 32)]:
        for sh, sv in ((0, 0), (1, 0), (0, 1), (1, 1)):
            lw, lh = cw << sh, ch << sv
            luma = rng.integers(0, 4096, (lh, lw + 3)).astype(np.uint16)
            ac = np.zeros(cw * ch, np.int32)
            L.orc_cfl_ac_sub(O.ptr(luma, ctypes.c_uint16), lw + 3, cw, ch, sh, sv,
                             O.ptr(ac, ctypes.c_int32))
            blk = luma[:lh, :lw].astype(np.int64).reshape(ch, 1 << sv, cw, 1 << sh).sum((1, 3))
            scaled = blk << (3 - sh - sv)
            n = cw * ch
            mean = (scaled.sum() + (n >> 1)) >> (n.bit_length() - 1)
            assert (ac.reshape(ch, cw) == scaled - mean).all()
            if sh == sv == 0:
                ac2 = np.zeros(cw * ch, np.int32)
                packed = np.ascontiguousarray(luma[:ch, :cw])
                L.orc_cfl_ac_444(O.ptr(packed, ctypes.c_uint16), cw, ch, O.ptr(ac2, ctypes.c_int32))
                assert (ac2 == ac).all()
    for bd in (8, 10, 12):
        mx = (1 << bd) - 1
        for a in range(-16, 17):
            for acv in list(range(-300, 301, 7)) + [-(mx << 3), mx << 3]:
                for dc in (0, mx >> 1, mx):
                    x = a * acv
                    r2 = (x + 32) >> 6 if x >= 0 else -((-x + 32) >> 6)
                    assert L.orc_cfl_pred_pixel(dc, acv, a, bd) == min(max(dc + r2, 0), mx)


def test_cfl_best_alpha_recovers_a_planted_alpha():
    """cfl_best_alpha_strided (src/intrapred.rs:1218): when the source IS an unclipped CfL
    prediction with alpha a* (|a*| <= 16), the search finds a zero-energy alpha; an all-zero
    AC gives 0."""
    L = O.plib()
    rng = np.random.default_rng(4)
    for trial in range(200):
        bd = (8, 10, 12)[trial % 3]
        w, h = [(4, 4), (8, 8), (16, 8), (32, 32)][trial % 4]
        # luma spread below 2^(bd-2) and dc at mid range: no prediction clips, so the
        # least-squares alpha lands within +/-1 of a*.
        luma = rng.integers(0, 1 << (bd - 2), (h, w)).astype(np.uint16)
        ac = np.zeros(w * h, np.int32)
        L.orc_cfl_ac_444(O.ptr(luma, ctypes.c_uint16), w, h, O.ptr(ac, ctypes.c_int32))
        a_star = int(rng.integers(-16, 17))
        dc = 1 << (bd - 1)
        src = np.array([L.orc_cfl_pred_pixel(dc, int(v), a_star, bd) for v in ac], np.uint16)
        got = L.orc_cfl_best_alpha(O.ptr(ac, ctypes.c_int32), w, O.ptr(src, ctypes.c_uint16), w,
                                   w, h, dc, bd)
        energy = sum((int(s) - L.orc_cfl_pred_pixel(dc, int(v), got, bd)) ** 2
                      for s, v in zip(src, ac))
        assert energy == 0, (trial, a_star, got)
    z = np.zeros(16, np.int32)
    s = np.full(16, 5, np.uint16)
    assert L.orc_cfl_best_alpha(O.ptr(z, ctypes.c_int32), 4, O.ptr(s, ctypes.c_uint16), 4, 4, 4, 0, 8) == 0


# ---- residual and distortion ----------------------------------------------------------------

def test_residual_sse_vis_brute_force():
    """residual_pred_scalar (src/rd_sse.rs:1179) and sse_recon over the visible corner
    (:454, :1262, :1290) against numpy."""
    L = O.plib()
    rng = np.random.default_rng(8)
    for (w, h) in SIZES[:13]:
        for bd in (8, 10, 12):
            stride = w + 5
            src = rng.integers(0, 1 << bd, (h, stride)).astype(np.uint16)
            pred = rng.integers(-50, (1 << bd) + 50, w * h).astype(np.int32)
            res = np.zeros(w * h, np.int32)
            L.orc_residual_pred(O.ptr(src, ctypes.c_uint16), stride, O.ptr(pred, ctypes.c_int32),
                                w, h, O.ptr(res, ctypes.c_int32))
            assert (res.reshape(h, w) == src[:, :w].astype(np.int64) - pred.reshape(h, w)).all()
            inv = rng.integers(-300, 300, w * h).astype(np.int32)
            for vw, vh in ((w, h), (w - 1, h), (w, 1), (1, 1), (0, h)):
                got = L.orc_sse_recon_vis(O.ptr(pred, ctypes.c_int32), O.ptr(inv, ctypes.c_int32),
                                          O.ptr(src, ctypes.c_uint16), stride, w, vw, vh, bd)
                rec = np.clip(pred + inv, 0, (1 << bd) - 1).reshape(h, w)
                d = src[:vh, :vw].astype(np.int64) - rec[:vh, :vw]
                assert got == int((d * d).sum())


H4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]])


def _satd_np(e):
    h, w = e.shape
    t = e.reshape(h // 4, 4, w // 4, 4).transpose(0, 2, 1, 3)
    c = np.einsum("ij,abjk,lk->abil", H4, t, H4)
    return int(np.abs(c).sum())


def test_satd_matches_hadamard_matrix_form():
    """luma_satd_scalar (src/partition_rd.rs:86) and satd_sad_proxy_scalar (src/rd_sse.rs:1143)
    against H E H^T with the 4x4 Hadamard matrix (rows in any order: the L1 norm is invariant);
    edge blocks mask invisible samples to zero error; exact reconstruction gives 0
    (src/partition_rd.rs:149)."""
    L = O.plib()
    rng = np.random.default_rng(12)
    for (w, h) in SIZES:
        for bd in (8, 10):
            src = rng.integers(0, 1 << bd, (h, w)).astype(np.uint16)
            pred = rng.integers(0, 1 << bd, w * h).astype(np.int32)
            inv = rng.integers(-40, 40, w * h).astype(np.int32)
            for vw, vh in ((w, h), (w - 3, h), (w, h - 2), (1, 1)):
                got = L.orc_luma_satd(O.ptr(src, ctypes.c_uint16), w, w, h, vw, vh, bd,
                                      O.ptr(pred, ctypes.c_int32), O.ptr(inv, ctypes.c_int32))
                rec = np.clip(pred + inv, 0, (1 << bd) - 1).reshape(h, w)
                e = src.astype(np.int64) - rec
                e[vh:, :] = 0
                e[:, vw:] = 0
                assert got == _satd_np(e)
                got2 = L.orc_satd_sad_proxy(O.ptr(src, ctypes.c_uint16), w, w, h, vw, vh,
                                            O.ptr(pred, ctypes.c_int32))
                e2 = src.astype(np.int64) - pred.reshape(h, w)
                e2[vh:, :] = 0
                e2[:, vw:] = 0
                assert got2 == int(np.abs(e2).sum()) + (_satd_np(e2) >> 2)
    src = (np.arange(64) * 3).astype(np.uint16)
    pred = src.astype(np.int32)
    assert L.orc_luma_satd(O.ptr(src, ctypes.c_uint16), 8, 8, 8, 8, 8, 8,
                           O.ptr(pred, ctypes.c_int32), None) == 0


def _f32(x):
    return np.float32(x)


def _fma32(a, b, c):
    """Correctly rounded f32 fused multiply-add, exactly, via rationals."""
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    r = np.float32(float(exact))
    cands = [r, np.nextafter(r, np.float32(np.inf)), np.nextafter(r, np.float32(-np.inf))]
    best = min(cands, key=lambda v: (abs(Fraction(float(v)) - exact),
                                     int(np.frombuffer(np.float32(v).tobytes(), np.uint32)[0]) & 1))
    return np.float32(best)


def test_weighted_sse_known_answers_and_f64():
    """weighted_sse_scalar / weighted_sse_avx2 (src/rd_sse.rs:1433, src/avx/rd.rs:767): the
    flat-DC known answer 1.8/0.4625*16e^2 (src/rd_sse.rs:1982), the normalised mean weight,
    and both within 1e-5 of the f64 weighted DCT energy (the reference's own tolerance
    between its kernels)."""
    d, wts = O.wsse_tables()
    assert abs(wts.mean() - 1.0) < 1e-5
    flat = np.full(16, 3, np.int32)
    want = 1.8 / 0.4625 * 144.0
    for avx in (0, 1):
        assert abs(O.weighted_sse(flat, 4, 4, avx) - want) < 1e-3 * want
    err = ((np.arange(64 * 64) * 7919) % 513 - 256).astype(np.int32)
    D = d.astype(np.float64)
    for (w, h) in [(4, 4), (8, 4), (4, 16), (16, 16), (32, 8), (64, 64), (12, 8), (20, 4), (28, 12)]:
        e = err[:w * h].reshape(h, w).astype(np.float64)
        tiles = e.reshape(h // 4, 4, w // 4, 4).transpose(0, 2, 1, 3)
        c = np.einsum("ki,abij,lj->abkl", D, tiles, D)
        ref = float((wts.astype(np.float64) * c * c).sum())
        for avx in (0, 1):
            got = O.weighted_sse(err[:w * h], w, h, avx)
            assert abs(got - ref) <= 1e-5 * ref, (w, h, avx, got, ref)


def test_weighted_sse_avx2_op_order_exact():
    """weighted_sse_avx2 (src/avx/rd.rs:767-849) bit for bit against a step-by-step replay of
    its vector program (lanes, fused multiply-adds rounded once, the reduction tree, the
    4-wide tail) with an exact rational fma, on 8x4, 12x4 and 4x8 errors."""
    d, wts = O.wsse_tables()
    rng = np.random.default_rng(2)
    for (w, h) in [(8, 4), (12, 4), (4, 8), (16, 8)]:
        err = rng.integers(-700, 700, w * h).astype(np.int32)
        e = err.reshape(h, w).astype(np.float32)
        acc = np.zeros((4, 2, 4), np.float32)  # [k][half][lane]
        pairs = w // 8 * 8

        def coef(r0, c0, k):
            t = []
            for j in range(4):
                v = _f32(e[r0, c0 + j] * d[k, 0])
                for i in (1, 2, 3):
                    v = _fma32(e[r0 + i, c0 + j], d[k, i], v)
                t.append(v)
            cs = []
            for lane in range(4):
                v = _f32(d[lane, 0] * t[0])
                for jj in (1, 2, 3):
                    v = _fma32(d[lane, jj], t[jj], v)
                cs.append(v)
            return cs
        for r0 in range(0, h, 4):
            for c0 in range(0, pairs, 8):
                for k in range(4):
                    for half in (0, 1):
                        cs = coef(r0, c0 + 4 * half, k)
                        for lane in range(4):
                            acc[k, half, lane] = _fma32(_f32(wts[k, lane] * cs[lane]), cs[lane],
                                                        acc[k, half, lane])
        s = [_f32(_f32(_f32(acc[0, 0, x] + acc[0, 1, x]) + _f32(acc[1, 0, x] + acc[1, 1, x])) +
                  _f32(_f32(acc[2, 0, x] + acc[2, 1, x]) + _f32(acc[3, 0, x] + acc[3, 1, x])))
             for x in range(4)]
        if pairs < w:
            for r0 in range(0, h, 4):
                for k in range(4):
                    cs = coef(r0, pairs, k)
                    for lane in range(4):
                        s[lane] = _fma32(_f32(wts[k, lane] * cs[lane]), cs[lane], s[lane])
        want = _f32(_f32(s[0] + s[2]) + _f32(s[1] + s[3]))
        got = np.float32(O.weighted_sse(err, w, h, 1))
        assert got.tobytes() == want.tobytes(), (w, h, got, want)


def test_weighted_dist_pads_and_rounds():
    """RdDispatch::weighted_dist (src/rd_sse.rs:244): the visible error zero-padded to whole
    4x4 tiles, scaled, rounded half away from zero."""
    L = O.plib()
    rng = np.random.default_rng(6)
    for vw, vh in ((5, 3), (4, 4), (13, 7), (64, 64), (1, 1)):
        e = rng.integers(-200, 200, vw * vh).astype(np.int32)
        pw, ph = (vw + 3) & ~3, (vh + 3) & ~3
        pad = np.zeros((ph, pw), np.int32)
        pad[:vh, :vw] = e.reshape(vh, vw)
        for scale in (1.0, 0.75, 2.5):
            raw = np.float32(O.weighted_sse(pad.ravel(), pw, ph, 1))
            want = int(np.sign(raw * np.float32(scale)) * np.floor(abs(raw * np.float32(scale)) + 0.5))
            got = L.orc_weighted_dist(O.ptr(e, ctypes.c_int32), vw, vh, scale, 1)
            assert got == want, (vw, vh, scale)
