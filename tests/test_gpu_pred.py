"""Bit-exact parity of the GPU intra prediction (mt_intra_pred) with the oracle's
predict_nd_ad / dc_pred / filter_predict (src/intrapred.rs:232-415) on every mode, angle
delta, block size, edge availability, smooth-neighbour filter type and bit depth, from
random, smooth, extreme and frame-edge-limited planes."""
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
PW, PH = 192, 176          # plane allocation (stride, rows)
LIM_W, LIM_H = 150, 140    # frame size of the edge-limited plane


def make_planes(rng, bd):
    mx = (1 << bd) - 1
    rand = rng.integers(0, mx + 1, (PH, PW))
    walk = np.cumsum(np.cumsum(rng.integers(-3, 4, (PH, PW)), 0), 1)
    smooth = np.clip(mx // 2 + walk, 0, mx)
    alt = (np.indices((PH, PW)).sum(0) & 1) * mx
    lim = rng.integers(0, mx + 1, (PH, PW))
    return [a.astype(np.uint16) for a in (rand, smooth, alt, lim)]


def make_blocks(rng, bd):
    recs = []
    for (bw, bh) in SIZES:
        modes = list(range(13)) + ([16, 17, 18, 19, 20] if max(bw, bh) <= 32 else [])
        for mode in modes:
            deltas = range(-3, 4) if 1 <= mode <= 8 else (0,)
            for delta in deltas:
                for v in range(5):
                    plane = int(rng.integers(0, 4))
                    fw, fh = (LIM_W, LIM_H) if plane == 3 else (PW, PH)
                    if v == 0:
                        x, y = 0, 0
                    elif v == 1:
                        x, y = 8 * int(rng.integers(1, 4)), 0
                    elif v == 2:
                        x, y = 0, 8 * int(rng.integers(1, 4))
                    elif v == 3:  # near the right / bottom frame edge: clipped extensions
                        x = fw - bw - int(rng.integers(0, 6)) * 2
                        y = fh - bh - int(rng.integers(0, 6)) * 2
                    else:
                        x = 4 * int(rng.integers(1, (fw - bw) // 4))
                        y = 4 * int(rng.integers(1, (fh - bh) // 4))
                    flags = int(rng.integers(0, 8))
                    recs.append((x, y, bw, bh, mode, delta, flags, plane))
    return recs


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_intra_pred_matches_oracle(dev, bd):
    rng = np.random.default_rng(1000 + bd)
    planes = make_planes(rng, bd)
    recs = make_blocks(rng, bd)
    n = len(recs)
    print(f"intra pred bd={bd} blocks={n}", flush=True)
    desc = np.zeros(n, mt.PREDBLK_DTYPE)
    off = 0
    for i, (x, y, bw, bh, mode, delta, flags, plane) in enumerate(recs):
        desc[i] = (off, x, y, bw, bh, mode, delta, flags, bd, plane, 0)
        off += bw * bh
    tplanes = [torch.from_numpy(p.view(np.int16)).to(dev) for p in planes]
    views = [(tplanes[k], LIM_W, LIM_H) if k == 3 else (tplanes[k],) for k in range(4)]
    blocks = torch.from_numpy(desc.view(np.uint8)).to(dev)
    pred = torch.full((off,), -7, dtype=torch.int32, device=dev)
    status = torch.full((n,), 99, dtype=torch.int32, device=dev)
    mt.intra_pred(blocks, views, pred, status)
    torch.cuda.synchronize()
    got = pred.cpu().numpy()
    st = status.cpu().numpy()
    assert (st == 0).all()
    bad = 0
    for i, (x, y, bw, bh, mode, delta, flags, plane) in enumerate(recs):
        fw, fh = (LIM_W, LIM_H) if plane == 3 else (PW, PH)
        want = O.intra_pred(mode, delta, planes[plane], x, y, bw, bh, flags & 1, flags & 2,
                            fw, fh, flags & 4, bd).ravel()
        o = int(desc[i]["out_offset"])
        if not np.array_equal(got[o:o + bw * bh], want):
            bad += 1
            if bad <= 5:
                print("mismatch", recs[i], file=sys.stderr)
    assert bad == 0, f"{bad} of {n} blocks differ"


def test_intra_pred_rejects_invalid_descriptors(dev):
    """Descriptors outside their plane, of a bad size, mode, delta or bit depth get
    MT_ERR_ARG in status and leave their output untouched; valid ones still run."""
    plane = torch.zeros((64, 64), dtype=torch.int16, device=dev)
    bad = [(60, 0, 8, 8, 0, 0, 0, 10, 0),    # reaches past the plane width
           (0, 0, 6, 8, 0, 0, 0, 10, 0),     # not a power of two
           (0, 0, 8, 8, 13, 0, 0, 10, 0),    # mode 13 (CfL) is not a predictor here
           (0, 0, 64, 64, 16, 0, 0, 10, 0),  # filter intra above 32
           (0, 0, 8, 8, 1, 4, 0, 10, 0),     # delta out of range
           (0, 0, 8, 8, 1, 0, 0, 9, 0),      # bit depth
           (0, 0, 8, 8, 1, 0, 0, 10, 1)]     # plane index beyond the array
    good = (8, 8, 8, 8, 9, 0, 0, 10, 0)
    recs = bad + [good]
    desc = np.zeros(len(recs), mt.PREDBLK_DTYPE)
    for i, (x, y, bw, bh, mode, delta, flags, bd, pl) in enumerate(recs):
        desc[i] = (i * 4096, x, y, bw, bh, mode, delta, flags, bd, pl, 0)
    pred = torch.full((len(recs) * 4096,), -7, dtype=torch.int32, device=dev)
    status = torch.zeros(len(recs), dtype=torch.int32, device=dev)
    mt.intra_pred(torch.from_numpy(desc.view(np.uint8)).to(dev), [(plane,)], pred, status)
    torch.cuda.synchronize()
    st = status.cpu().numpy()
    assert (st[:-1] == -1).all() and st[-1] == 0
    p = pred.cpu().numpy().reshape(len(recs), 4096)
    assert (p[:-1] == -7).all()
    assert (p[-1, :64] == 0).all()
