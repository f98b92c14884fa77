"""The GPU RDO's level planes and the AV1 / AVIF bytes coded from them, through the C ABI:
the levels of every transform block (luma and chroma, with and without the trellis) equal
the oracle's RDO's; the files coded from the GPU's decisions and levels are byte-identical to
the files coded from the oracle's; and the independent AV1 decoder installed with Pillow
decodes them to the GPU's reconstruction (monochrome and 4:4:4 identity: every sample of
every plane), at the seeded 512x512 4:4:4 plain-encode workload among others."""
import io

import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import av1
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu

ALL = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL


def _t(dev, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(dev)


def _decode(data):
    from PIL import Image
    im = Image.open(io.BytesIO(data))
    im.load()
    return np.asarray(im).astype(np.int64)


@pytest.mark.parametrize("w,h,bd,q,fmt,flags", [
    (200, 136, 8, 91, R.LUMA_ONLY, R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4),
    (136, 72, 10, 30, R.CHROMA_420, ALL),
    (264, 136, 8, 60, R.CHROMA_444, ALL),
    (128, 200, 8, 140, R.CHROMA_444, ALL | R.TRELLIS),
    (200, 136, 12, 91, R.CHROMA_420, ALL | R.TRELLIS),
    (512, 512, 8, 70, R.CHROMA_444, ALL)])
def test_gpu_levels_and_bytes_match_the_oracle(dev, w, h, bd, q, fmt, flags):
    print(f"av1: {w}x{h} bd={bd} q={q} fmt={fmt} flags={flags}", flush=True)
    sub = 2 if fmt == R.CHROMA_420 else 0
    if fmt == R.LUMA_ONLY:
        y = W.synth_image(21, w, h, bd)
        u = v = None
    else:
        y, u, v = W.synth_yuv(21, w, h, bd, sub)
    qm = mt.qm_level_law(q, sub)
    lam = R.rdo_lambda(q, bd, sub if fmt else 2)
    fr = {"src": _t(dev, y), "bd": bd, "qidx": q, "qm_level": qm, "lam": lam, "chroma": fmt, "levels": True}
    dc = ac = 0
    uqm = 15
    if fmt:
        dc, ac, uqm = R.chroma_laws(q, sub)
        fr.update(src_u=_t(dev, u), src_v=_t(dev, v))
    rec, nodes, st = R.intra_rdo_batch([fr], flags)[0]
    torch.cuda.synchronize()
    g_lv = [None if a is None else a.cpu().numpy() for a in st["levels"]]
    g_rec = rec.cpu().numpy().view(np.uint16)
    o_rec, o_u, o_v, o_nodes, o_st, o_lv = O.intra_rdo_levels(y, bd, q, qm, flags, lam, chroma=fmt, src_u=u,
                                                              src_v=v, uv_dc=dc, uv_ac=ac, uv_qm=uqm, nthreads=16)
    assert (g_rec == o_rec).all()
    for name, a, b in zip("yuv", g_lv, o_lv):
        if b is None:
            continue
        bad = np.argwhere(a != b)
        assert len(bad) == 0, f"levels {name}: {len(bad)} differ, first {bad[:4].tolist()}"
    cicp = av1.SRGB_IDENTITY if (fmt == R.CHROMA_444 and bd == 8) else av1.BT601
    kw = dict(chroma=fmt, qm_level=qm, uv_qm_level=uqm, uv_dc=dc, uv_ac=ac,
              filter_intra=bool(flags & R.FILTER_INTRA), cicp=cicp)
    fg, kg = av1.frame(R.nodes_to_np(nodes), g_lv, w, h, bd, q, **kw)
    fo, ko = av1.frame(o_nodes, o_lv, w, h, bd, q, **kw)
    bg = av1.encode_avif(fg)
    assert bg == av1.encode_avif(fo), "the GPU's and the oracle's decisions code to different bytes"
    print(f"av1: {len(bg)} bytes (rate model {(st['bits']) / 8:.0f})", flush=True)
    if bd == 8 and fmt in (R.LUMA_ONLY, R.CHROMA_444):
        dec = _decode(bg)
        assert (dec[..., 1] == g_rec).all()
        if fmt == R.CHROMA_444:
            ru, rv = (p.cpu().numpy().view(np.uint16) for p in st["recon_uv"])
            assert (dec[..., 2] == ru).all() and (dec[..., 0] == rv).all()
