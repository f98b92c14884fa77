"""The deblocking filter on the GPU (mt_deblock_frame, mt_deblock_plane, mt_rdo_deblock_maps:
frame_deblock src/coder/frame_deblock.rs:206, filter_plane_impl src/loopfilter.rs:483) against
the oracle's sequential deblocking (oracle/oracle_lf.cpp), bit for bit:
  - random quadtree partitions at 4:2:0 and 4:4:4, 8 / 10 / 12-bit, visible sizes below the
    coded size, partial tiles, every level class and sharpness, chroma levels 0 and luma 0;
  - the 3840 x 2160 4:2:0 frame of the bench workload;
  - the alignment rule of the chroma planes through mt_deblock_plane;
  - the maps of an RDO's decisions (mt_rdo_deblock_maps) equal record_blk of its leaves, and
    the frame filtered through them equals the oracle;
  - maps that break the parallel order's condition raise MT_DEBLOCK_ORDER."""
import numpy as np
import pytest
import torch

import lfgen
import oracle_lib as O
from maroontree_amd import deblock as D
from maroontree_amd import rdo as R

pytestmark = pytest.mark.gpu


def _t(a, dev):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        a = a.view(np.int16)
    return torch.from_numpy(a).to(dev)


def _u16(t):
    return t.cpu().numpy().view(np.uint16)


def _frame_case(dev, seed, w8, h8, sub, bd, lv, sharp, crop=True):
    print(f"deblock frame seed {seed} {w8}x{h8} sub {sub} bd {bd} levels {lv} sharp {sharp}", flush=True)
    rng = np.random.default_rng(seed)
    leaves = lfgen.random_leaves(rng, w8, h8)
    m = lfgen.maps_from_leaves(leaves, w8, h8)
    y = lfgen.blocky_plane(rng, leaves, w8, h8, bd)
    u = lfgen.blocky_plane(rng, leaves, w8 >> sub, h8 >> sub, bd, sub)
    v = lfgen.blocky_plane(rng, leaves, w8 >> sub, h8 >> sub, bd, sub)
    dw, dh = (w8 - int(rng.integers(0, 8)), h8 - int(rng.integers(0, 8))) if crop else (w8, h8)
    maps = _t(np.stack([m[k] for k in D.MAP_NAMES]), dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    got = D.frame(_t(y, dev), _t(u, dev), _t(v, dev), maps, dw, dh, sub, sharp, bd, *lv, status=status)
    torch.cuda.synchronize()
    want = [y.copy(), u.copy(), v.copy()]
    O.frame_deblock(want[0], want[1], want[2], dw, dh, m, sub, sub, sharp, bd, *lv)
    for pl in range(3):
        assert (_u16(got[pl]) == want[pl]).all(), pl
    assert int(status.item()) == 0


@pytest.mark.parametrize("seed,w8,h8,sub,bd,lv,sharp", [
    (31, 200, 136, 1, 8, (12, 4, 6), 7), (32, 136, 200, 0, 10, (40, 20, 0), 3), (33, 264, 72, 1, 12, (63, 32, 32), 0),
    (34, 72, 264, 0, 8, (0, 5, 5), 7), (35, 64, 64, 1, 8, (20, 0, 0), 5), (36, 328, 256, 1, 10, (25, 13, 13), 7)])
def test_frame_matches_oracle(dev, seed, w8, h8, sub, bd, lv, sharp):
    _frame_case(dev, seed, w8, h8, sub, bd, lv, sharp)


def test_4k_frame_matches_oracle(dev):
    _frame_case(dev, 40, 3840, 2160, 1, 8, (4, 2, 2), 7, crop=False)


def test_plane_alignment_rule(dev):
    rng = np.random.default_rng(41)
    w, h = 200, 136
    leaves = lfgen.random_leaves(rng, w, h)
    m = lfgen.maps_from_leaves(leaves, w, h)
    y = lfgen.blocky_plane(rng, leaves, w, h, 10)
    for luma in (0, 1):
        print(f"deblock plane alignment rule luma {luma}", flush=True)
        got = D.plane(_t(y, dev), _t(m["blk4"], dev), _t(m["blk4h"], dev), None, None, w - 5, h - 2, 30, 2, luma,
                      16, 10)
        torch.cuda.synchronize()
        want = y.copy()
        O.filter_plane(want, w - 5, h - 2, m["blk4"], m["blk4h"], None, None, 30, 2, luma, 16, 10)
        assert (_u16(got) == want).all(), luma


def test_overlapping_edges_are_reported(dev):
    w, h = 128, 64
    bw = np.full((h // 4, w // 4), 8, np.uint8)
    flags = np.ones_like(bw)
    y = np.full((h, w), 512, np.uint16)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    D.plane(_t(y, dev), _t(bw, dev), _t(bw, dev), _t(flags, dev), _t(flags, dev), w, h, 20, 0, 1, 16, 10,
            status=status)
    torch.cuda.synchronize()
    assert int(status.item()) & D.DEBLOCK_ORDER


def test_levels_match_oracle(dev):
    for q in range(256):
        for sub in (0, 1, 2):
            assert D.levels(q, sub) == O.lf_levels(q, sub), (q, sub)


@pytest.mark.parametrize("w,h,chroma,quality", [(200, 136, R.YUV420, 60), (136, 72, R.YUV444, 30)])
def test_rdo_maps_and_encode_deblock(dev, w, h, chroma, quality):
    print(f"encode + deblock {w}x{h} chroma {chroma} quality {quality}", flush=True)
    from maroontree_amd import workload as W
    rgb = W.synth_rgb(300 + w, w, h)
    r, g, b = rgb[..., 0], rgb[..., 1], rgb[..., 2]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    cfg = R.encode_config(chroma=chroma, quality=quality, speed=R.SPEED_FAST)
    out = R.encode_rgb8(t(g), t(b), t(r), cfg)
    torch.cuda.synchronize()
    nodes = R.nodes_to_np(out["nodes"])
    leaves = lfgen.leaves_from_nodes(nodes, w, h)
    m = lfgen.maps_from_leaves(leaves, w, h)
    maps = D.rdo_maps(out["nodes"], w, h)
    torch.cuda.synchronize()
    mg = maps.cpu().numpy()
    for i, k in enumerate(D.MAP_NAMES):
        assert (mg[i] == m[k]).all(), k
    c420 = chroma == R.YUV420
    sub = 2 if c420 else 0
    ly, luv, sharp = O.lf_levels(out["qidx"], sub)
    assert out["lf_levels"] == (ly, luv, luv, sharp) and out["dbk_status"] == 0
    assert "deblock" not in out["not_applied"]
    want = [_u16(out[k]).copy() for k in ("recon_y", "recon_u", "recon_v")]
    s = 1 if c420 else 0
    O.frame_deblock(want[0], want[1], want[2], w, h, m, s, s, sharp, 8, ly, luv, luv)
    for k, wv in zip(("dbk_y", "dbk_u", "dbk_v"), want):
        assert (_u16(out[k]) == wv).all(), k
    assert (want[0] != _u16(out["recon_y"])).any()
