"""Bit-exact parity of the whole-frame luma intra RDO (mt_intra_rdo) with the oracle's
restatement (oracle/oracle_rdo.cpp): the reconstruction plane, every quadtree node's
decision (cost, NONE cost, split, mode), and the final accounting (bits, SSE, candidate and
leaf counts).  Frames are seeded synthetic images (gradients, value noise, rectangles) with
sizes that are not multiples of 64, so superblocks cross the frame edge."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu

ALL = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4


def run_both(dev, img, bd, qidx, qm, flags, prune=None, wsse=1.0):
    lam = R.rdo_lambda(qidx, bd, 2)
    print(f"rdo {img.shape[1]}x{img.shape[0]} bd={bd} q={qidx} qm={qm} flags={flags} "
          f"prune={prune is not None}", flush=True)
    src = torch.from_numpy(img.view(np.int16)).to(dev)
    pr = None if prune is None else torch.from_numpy(prune.view(np.int16)).to(dev)
    rec, nodes, st = R.intra_rdo(src, bd, qidx, qm, flags, lam, wsse, prune=pr)
    torch.cuda.synchronize()
    g_rec = rec.cpu().numpy().view(np.uint16)
    g_nodes = R.nodes_to_np(nodes)
    o_rec, o_nodes, _, o_st = O.intra_rdo(img, bd, qidx, qm, flags, lam, wsse, prune=prune)
    return g_rec, g_nodes, st, o_rec, o_nodes, o_st


def assert_same(g_rec, g_nodes, st, o_rec, o_nodes, o_st):
    diff = np.argwhere(g_rec != o_rec)
    assert len(diff) == 0, f"{len(diff)} recon samples differ, first {diff[:4].tolist()}"
    for f in ("split", "mode"):
        bad = np.argwhere(g_nodes[f] != o_nodes[f])
        assert len(bad) == 0, (f, bad[:4].tolist())
    for f in ("cost", "none_cost"):
        a, b = g_nodes[f].view(np.uint32), o_nodes[f].view(np.uint32)
        bad = np.argwhere(a != b)
        assert len(bad) == 0, (f, bad[:4].tolist(), g_nodes[f][tuple(bad[0])], o_nodes[f][tuple(bad[0])])
    assert st["sse"] == o_st["sse"]
    assert st["candidates"] == o_st["candidates"]
    assert st["leaves"] == o_st["leaves"]
    assert st["bits"] == o_st["bits"], (st["bits"], o_st["bits"])


@pytest.mark.parametrize("bd,qidx,w,h,flags", [
    (10, 91, 200, 136, ALL),
    (8, 40, 136, 72, ALL),
    (8, 200, 128, 128, 0),
    (12, 150, 96, 80, R.MIN4),
    (10, 17, 160, 64, R.ANGLE_DELTAS | R.FILTER_INTRA),
])
def test_rdo_matches_oracle(dev, bd, qidx, w, h, flags):
    img = W.synth_image(100 + qidx, w, h, bd)
    qm = mt.qm_level_law(qidx, 0)
    out = run_both(dev, img, bd, qidx, qm, flags)
    assert_same(*out)
    g_rec, _, st, *_ = out
    sse = int(((g_rec.astype(np.int64) - img) ** 2).sum())
    assert st["sse"] == sse


def test_rdo_with_prune_masks_matches_oracle(dev):
    rng = np.random.default_rng(5)
    img = W.synth_image(77, 192, 128, 10)
    nsb = 3 * 2
    prune = rng.integers(0, 1 << 14, (nsb, R.NODES)).astype(np.uint16)
    prune |= (rng.random((nsb, R.NODES)) < 0.15).astype(np.uint16) * R.FORCE_SPLIT
    prune |= (rng.random((nsb, R.NODES)) < 0.15).astype(np.uint16) * R.NO_SPLIT
    out = run_both(dev, img, 10, 91, 8, ALL, prune=prune)
    assert_same(*out)
    full = run_both(dev, img, 10, 91, 8, ALL)
    assert out[2]["candidates"] < full[2]["candidates"]


def test_rdo_plain_sse_matches_oracle(dev):
    img = W.synth_image(9, 128, 64, 8)
    out = run_both(dev, img, 8, 120, 15, ALL, wsse=0.0)
    assert_same(*out)


def test_rdo_batch_of_uneven_frames_matches_oracle(dev):
    """Frames of different sizes, bit depths, quantizers and prune masks in one call: each
    frame's result equals the oracle's for that frame alone."""
    rng = np.random.default_rng(17)
    specs = [(200, 136, 10, 91), (64, 64, 8, 200), (136, 264, 12, 30), (72, 40, 8, 140), (128, 72, 10, 17)]
    frames, want = [], []
    for i, (w, h, bd, q) in enumerate(specs):
        img = W.synth_image(500 + i, w, h, bd)
        qm = mt.qm_level_law(q, 0)
        lam = R.rdo_lambda(q, bd, 2)
        prune = None
        if i == 3:
            nsb = ((w + 63) // 64) * ((h + 63) // 64)
            prune = rng.integers(0, 1 << 16, (nsb, R.NODES)).astype(np.uint16)
        fr = {"src": torch.from_numpy(img.view(np.int16)).to(dev), "bd": bd, "qidx": q, "qm_level": qm,
              "lam": lam}
        if prune is not None:
            fr["prune"] = torch.from_numpy(prune.view(np.int16)).to(dev)
        frames.append(fr)
        want.append((img, O.intra_rdo(img, bd, q, qm, ALL, lam, prune=prune)))
    print("rdo batch", len(frames), flush=True)
    got = R.intra_rdo_batch(frames, ALL)
    torch.cuda.synchronize()
    for (rec, nodes, st), (img, (o_rec, o_nodes, _, o_st)) in zip(got, want):
        assert_same(rec.cpu().numpy().view(np.uint16), R.nodes_to_np(nodes), st, o_rec, o_nodes, o_st)


def test_rdo_repeatable(dev):
    img = W.synth_image(31, 128, 128, 10)
    src = torch.from_numpy(img.view(np.int16)).to(dev)
    a = R.intra_rdo(src, 10, 91, 8, ALL)
    b = R.intra_rdo(src, 10, 91, 8, ALL)
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) and a[2] == b[2]
