"""Bit-exact parity of the GPU distortion measures (mt_distortion) with the oracle:
sse_recon over the displayed part (src/rd_sse.rs:454), the weighted decision distortion in
the op order of weighted_sse_avx2 (src/rd_sse.rs:244, src/avx/rd.rs:767), luma_satd
(src/partition_rd.rs:86) and satd_sad_proxy (src/rd_sse.rs:1143), on every block size,
bit depth, clipped edge blocks and prediction-only blocks."""
import ctypes
import sys

import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [(4, 4), (4, 8), (8, 4), (4, 16), (16, 4), (8, 8), (8, 16), (16, 8), (16, 16),
         (16, 32), (32, 16), (32, 32), (64, 64), (8, 32), (32, 8), (16, 64), (64, 16),
         (32, 64), (64, 32)]
PW, PH = 256, 200


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_distortion_matches_oracle(dev, bd):
    rng = np.random.default_rng(500 + bd)
    mx = (1 << bd) - 1
    src = rng.integers(0, mx + 1, (PH, PW)).astype(np.uint16)
    # plane 0: fully displayed; plane 1: same samples, displayed area cut at (vis_w, vis_h)
    vis_w, vis_h = 181, 147
    recs = []
    for rep in range(7):
        for (bw, bh) in SIZES:
            plane = rep & 1
            if plane and rep in (5,):  # blocks straddling the displayed edge
                x = int(np.clip(vis_w - bw + int(rng.integers(1, bw)), 0, PW - bw))
                y = int(np.clip(vis_h - bh + int(rng.integers(1, bh)), 0, PH - bh))
            else:
                x = int(rng.integers(0, PW - bw + 1))
                y = int(rng.integers(0, PH - bh + 1))
            flags = mt.DIST_NO_RESID if rep == 3 else 0
            # scale 2^16 exposes every bit of the f32 weighted sum above 128 in the rounded value
            scale = [1.0, 0.0, 2.5, 1.0, 0.75, 1.0, 65536.0][rep]
            amp = [3, 40, 400, mx][rep % 4]
            recs.append((x, y, bw, bh, plane, flags, scale, amp))
    n = len(recs)
    print(f"distortion bd={bd} blocks={n}", flush=True)
    desc = np.zeros(n, mt.DISTBLK_DTYPE)
    preds, resids = [], []
    off = 0
    for i, (x, y, bw, bh, plane, flags, scale, amp) in enumerate(recs):
        s = src[y:y + bh, x:x + bw].astype(np.int64)
        p = np.clip(s + rng.integers(-amp, amp + 1, s.shape), -20, mx + 20).astype(np.int32)
        r = rng.integers(-amp, amp + 1, s.shape).astype(np.int32)
        preds.append(p.ravel())
        resids.append(r.ravel())
        desc[i]["pred_offset"] = off
        desc[i]["resid_offset"] = off
        desc[i]["x"], desc[i]["y"], desc[i]["bw"], desc[i]["bh"] = x, y, bw, bh
        desc[i]["bd"], desc[i]["plane"], desc[i]["flags"] = bd, plane, flags
        desc[i]["wsse_scale"] = scale
        off += bw * bh
    pred = np.concatenate(preds)
    resid = np.concatenate(resids)
    t_src = torch.from_numpy(src.view(np.int16)).to(dev)
    planes = [(t_src, PW, PH, PW, PH), (t_src, PW, PH, vis_w, vis_h)]
    out = mt.distortion(torch.from_numpy(desc.view(np.uint8)).to(dev), planes,
                        torch.from_numpy(pred).to(dev), torch.from_numpy(resid).to(dev))
    torch.cuda.synchronize()
    res = out.cpu().numpy().view(mt.DISTRES_DTYPE)
    L = O.plib()
    bad = 0
    for i, (x, y, bw, bh, plane, flags, scale, amp) in enumerate(recs):
        o = int(desc[i]["pred_offset"])
        p = np.ascontiguousarray(pred[o:o + bw * bh])
        r = np.ascontiguousarray(resid[o:o + bw * bh]) if not flags else np.zeros(bw * bh, np.int32)
        vw = min(max((vis_w if plane else PW) - x, 0), bw)
        vh = min(max((vis_h if plane else PH) - y, 0), bh)
        s = np.ascontiguousarray(src[y:y + bh, x:x + bw])
        sp = O.ptr(s, ctypes.c_uint16)
        pp, rp = O.ptr(p, ctypes.c_int32), O.ptr(r, ctypes.c_int32)
        sse = L.orc_sse_recon_vis(pp, rp, sp, bw, bw, vw, vh, bd)
        rec = np.clip(p + r, 0, mx).reshape(bh, bw)
        err = np.ascontiguousarray((s.astype(np.int64) - rec)[:vh, :vw].astype(np.int32))
        dist = L.orc_weighted_dist(O.ptr(err, ctypes.c_int32), vw, vh, scale, 1) if scale > 0 else sse
        satd = L.orc_luma_satd(sp, bw, bw, bh, vw, vh, bd, pp, rp)
        proxy = L.orc_satd_sad_proxy(sp, bw, bw, bh, vw, vh, pp)
        g = res[i]
        if (int(g["sse"]), int(g["dist"]), int(g["satd"]), int(g["proxy"])) != (sse, dist, satd, proxy):
            bad += 1
            if bad <= 5:
                print("mismatch", recs[i], (g["sse"], g["dist"], g["satd"], g["proxy"]),
                      (sse, dist, satd, proxy), file=sys.stderr)
    assert bad == 0, f"{bad} of {n} blocks differ"
