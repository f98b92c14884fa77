"""MT_RDO_GRAPH: the RDO's wavefront steps captured into one HIP graph per call.  Through the
C ABI, every result (Y, U, V reconstructions, every node record, bits, SSE, candidate and
leaf counts) equals the eagerly launched RDO's, on single frames (luma only, 4:2:0, 4:4:4,
with the trellis) and on an uneven batch, over several calls in a row with eager calls in
between (each call captures, instantiates and launches its own graph)."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu

ALL = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL


def _t(dev, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(dev)


def _frames(dev, specs):
    out = []
    for (seed, w, h, bd, q, fmt) in specs:
        sub = 2 if fmt == R.CHROMA_420 else 0
        if fmt == R.LUMA_ONLY:
            fr = {"src": _t(dev, W.synth_image(seed, w, h, bd))}
        else:
            y, u, v = W.synth_yuv(seed, w, h, bd, sub)
            fr = {"src": _t(dev, y), "src_u": _t(dev, u), "src_v": _t(dev, v), "chroma": fmt}
        fr.update(bd=bd, qidx=q, qm_level=mt.qm_level_law(q, sub), lam=R.rdo_lambda(q, bd, sub if fmt else 2))
        out.append(fr)
    return out


def _snap(res):
    out = []
    for rec, nodes, st in res:
        item = [rec.cpu().numpy().copy(), nodes.cpu().numpy().copy(),
                (st["bits"], st["sse"], st["candidates"], st["leaves"])]
        if "recon_uv" in st:
            item += [p.cpu().numpy().copy() for p in st["recon_uv"]] + [(st["bits_uv"], st["sse_u"], st["sse_v"])]
        out.append(item)
    return out


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for p, q in zip(x, y):
            if isinstance(p, np.ndarray):
                assert (p == q).all()
            else:
                assert p == q


@pytest.mark.parametrize("specs,flags", [
    ([(1, 200, 136, 10, 91, R.LUMA_ONLY)], R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4),
    ([(2, 136, 72, 8, 53, R.CHROMA_420)], ALL),
    ([(3, 128, 128, 8, 140, R.CHROMA_444)], ALL | R.TRELLIS),
    ([(4, 264, 136, 8, 91, R.CHROMA_420), (5, 64, 200, 10, 30, R.CHROMA_444), (6, 72, 64, 12, 180, R.LUMA_ONLY)], ALL)])
def test_graph_rdo_equals_eager(dev, specs, flags):
    print(f"graph rdo: {specs} flags={flags}", flush=True)
    frames = _frames(dev, specs)
    eager = _snap(R.intra_rdo_batch(frames, flags))
    for rep in range(3):
        got = _snap(R.intra_rdo_batch(frames, flags | R.GRAPH))
        _same(got, eager)
        _same(_snap(R.intra_rdo_batch(frames, flags)), eager)  # eager work between the graphs
