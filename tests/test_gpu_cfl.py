"""Bit-exact parity of the GPU chroma-from-luma (mt_cfl) with the oracle's cfl_ac_sub /
cfl_ac_444, cfl_best_alpha_strided and cfl_pred_pixel (src/intrapred.rs:1112-1260), for every
chroma size up to 32x32, every subsampling, 8/10/12-bit, clipped source blocks and given
alphas."""
import ctypes
import sys

import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O

pytestmark = pytest.mark.gpu

DIMS = (4, 8, 16, 32)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_cfl_matches_oracle(dev, bd):
    rng = np.random.default_rng(70 + bd)
    mx = (1 << bd) - 1
    LW, LH, CW, CH = 256, 192, 160, 128
    walk = np.cumsum(np.cumsum(rng.integers(-3, 4, (LH, LW)), 0), 1)
    luma = np.clip(mx // 2 + walk + rng.integers(-20, 21, (LH, LW)), 0, mx).astype(np.uint16)
    src = rng.integers(0, mx + 1, (CH, CW)).astype(np.uint16)
    vis_w, vis_h = 150, 121
    recs = []
    for cw in DIMS:
        for ch in DIMS:
            for ss in range(4):
                for rep in range(3):
                    sh, sv = ss & 1, ss >> 1
                    lx = 2 * int(rng.integers(0, (LW - (cw << sh)) // 2 + 1))
                    ly = 2 * int(rng.integers(0, (LH - (ch << sv)) // 2 + 1))
                    plane = 1 if rep == 2 else 0
                    if plane:
                        x = min(max(vis_w - cw + int(rng.integers(1, cw)), 0), CW - cw)
                        y = min(max(vis_h - ch + int(rng.integers(1, ch)), 0), CH - ch)
                    else:
                        x = int(rng.integers(0, CW - cw + 1))
                        y = int(rng.integers(0, CH - ch + 1))
                    given = rep == 1 and ss == 3
                    alpha = int(rng.integers(-16, 17)) if given else 0
                    dc = int(rng.integers(0, mx + 1))
                    recs.append((lx, ly, x, y, cw, ch, sh, sv, plane, given, alpha, dc))
    n = len(recs)
    print(f"cfl bd={bd} blocks={n}", flush=True)
    desc = np.zeros(n, mt.CFLBLK_DTYPE)
    off = 0
    for i, (lx, ly, x, y, cw, ch, sh, sv, plane, given, alpha, dc) in enumerate(recs):
        flags = sh * mt.CFL_SS_HOR + sv * mt.CFL_SS_VER + (mt.CFL_ALPHA_GIVEN if given else 0)
        desc[i] = (off, lx, ly, x, y, cw, ch, flags, bd, 0, 1 + plane, alpha, 0, dc)
        off += cw * ch
    tl = torch.from_numpy(luma.view(np.int16)).to(dev)
    ts = torch.from_numpy(src.view(np.int16)).to(dev)
    planes = [(tl,), (ts,), (ts, CW, CH, vis_w, vis_h)]
    pred = torch.full((off,), -9, dtype=torch.int32, device=dev)
    ac = torch.full((off,), -9, dtype=torch.int32, device=dev)
    alphas = mt.cfl(torch.from_numpy(desc.view(np.uint8)).to(dev), planes, pred, ac).cpu().numpy()
    pred = pred.cpu().numpy()
    ac = ac.cpu().numpy()
    L = O.plib()
    bad = 0
    for i, (lx, ly, x, y, cw, ch, sh, sv, plane, given, alpha, dc) in enumerate(recs):
        lb = np.ascontiguousarray(luma[ly:ly + (ch << sv), lx:lx + (cw << sh)])
        want_ac = np.zeros(cw * ch, np.int32)
        L.orc_cfl_ac_sub(O.ptr(lb, ctypes.c_uint16), lb.shape[1], cw, ch, sh, sv,
                         O.ptr(want_ac, ctypes.c_int32))
        sb = np.ascontiguousarray(src[y:y + ch, x:x + cw])
        vw = min(max((vis_w if plane else CW) - x, 0), cw)
        vh = min(max((vis_h if plane else CH) - y, 0), ch)
        if given:
            want_a = alpha
        else:
            want_a = L.orc_cfl_best_alpha(O.ptr(want_ac, ctypes.c_int32), cw,
                                          O.ptr(sb, ctypes.c_uint16), cw, vw, vh, dc, bd)
        want_p = np.array([L.orc_cfl_pred_pixel(dc, int(a), want_a, bd) for a in want_ac])
        o = int(desc[i]["out_offset"])
        ok = (alphas[i] == want_a and np.array_equal(ac[o:o + cw * ch], want_ac) and
              np.array_equal(pred[o:o + cw * ch], want_p))
        if not ok:
            bad += 1
            if bad <= 5:
                print("mismatch", recs[i], alphas[i], want_a, file=sys.stderr)
    assert bad == 0, f"{bad} of {n} blocks differ"
