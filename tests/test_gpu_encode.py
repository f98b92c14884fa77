"""mt_encode_rgb8 (the encoder's entry: encode_rgb8, src/avif.rs:915, up to the bitstream)
against the oracle: the colour conversion bit for bit (orc_csc_rgb8), then the whole RDO
on those planes with the configuration's quantizer, lambda, chroma laws and QM (orc
intra_rdo_yuv): reconstructions of all planes, every node record, bits and SSE."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu


def rgb_image(seed, w, h):
    rgb = W.synth_rgb(seed, w, h)
    return rgb[..., 0], rgb[..., 1], rgb[..., 2]


@pytest.mark.parametrize("w,h,chroma,speed,quality,qml", [
    (136, 72, R.YUV420, R.SPEED_SLOW, 80, -1),
    (128, 64, R.YUV444, R.SPEED_FAST, 50, -1),
    (72, 136, R.YUV420, R.SPEED_MEDIUM, 30, 6),
])
def test_encode_rgb8_matches_oracle(dev, w, h, chroma, speed, quality, qml):
    r, g, b = rgb_image(100 + w, w, h)
    print(f"encode {w}x{h} chroma {chroma} speed {speed} quality {quality}", flush=True)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    cfg = R.encode_config(chroma=chroma, speed=speed, quality=quality, qmatrix_level=qml)
    out = R.encode_rgb8(t(g), t(b), t(r), cfg)
    torch.cuda.synchronize()
    c420 = chroma == R.YUV420
    y, cb, cr = O.csc_rgb8(r, g, b, c420)
    for name, want in (("y", y), ("u", cb), ("v", cr)):
        assert (out[name].cpu().numpy().view(np.uint16) == want).all(), name
    sub = 2 if c420 else 0
    q = O.quality_to_q(quality)
    assert out["qidx"] == q
    dc, ac, uvqm = R.chroma_laws(q, sub)
    qm = mt.qm_level_law(q, sub) if qml < 0 else qml
    uvqm = uvqm if qml < 0 else qml
    assert (out["qm_level"], out["uv_qm_level"]) == (qm, uvqm)
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL | (R.TRELLIS if speed <= R.SPEED_SLOW else 0)
    assert out["rdo_flags"] == flags
    lam = O.rdo_lambda(q, 8, sub)
    rec, ru, rv, nodes, st = O.intra_rdo_yuv(y, cb, cr, 8, q, qm, flags, lam,
                                             R.CHROMA_420 if c420 else R.CHROMA_444, dc, ac, uvqm)
    assert (out["recon_y"].cpu().numpy().view(np.uint16) == rec).all()
    assert (out["recon_u"].cpu().numpy().view(np.uint16) == ru).all()
    assert (out["recon_v"].cpu().numpy().view(np.uint16) == rv).all()
    g = R.nodes_to_np(out["nodes"])
    assert g.shape == nodes.shape
    for f in ("split", "mode", "uv_mode", "cfl_alpha_u", "cfl_alpha_v"):
        assert (g[f] == nodes[f]).all(), f
    for f in ("cost", "none_cost"):
        assert (g[f].view(np.uint32) == nodes[f].view(np.uint32)).all(), f
    for k in ("bits", "sse", "candidates", "leaves", "bits_uv", "sse_u", "sse_v"):
        assert out["stats"][k] == st[k], k
    assert "adaptive_quant" in out["not_applied"] and "intrabc" in out["not_applied"]


def test_encode_rgb8_refuses_what_it_cannot_run(dev):
    z = torch.zeros((64, 64), dtype=torch.uint8, device=dev)
    for chroma in (R.YUV422, R.MONOCHROME):
        with pytest.raises(mt.MtError):
            R.encode_rgb8(z, z, z, R.encode_config(chroma=chroma))
    z = torch.zeros((60, 64), dtype=torch.uint8, device=dev)
    with pytest.raises(mt.MtError):
        R.encode_rgb8(z, z, z)


@pytest.mark.parametrize("w,h,chroma,speed,quality", [(200, 136, R.YUV420, R.SPEED_SLOW, 40),
                                                      (136, 128, R.YUV444, R.SPEED_SLOW, 60),
                                                      (136, 72, R.YUV420, R.SPEED_FAST, 30)])
def test_encode_rgb8_cdef_matches_oracle(dev, w, h, chroma, speed, quality):
    """With the deblock and CDEF planes given, the encoder runs CDEF at Slow (and with cdef
    set): the skip flags of its coded levels, frame_cdef on the deblocked planes (Slow's
    candidate pairs at Slow, the full list otherwise); equal to the oracle's chain of the same
    steps: RDO levels, skip map, frame_deblock, frame_cdef."""
    import lfgen
    from maroontree_amd import av1
    r, g, b = rgb_image(300 + w, w, h)
    print(f"encode + cdef {w}x{h} chroma {chroma} speed {speed} quality {quality}", flush=True)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    cfg = R.encode_config(chroma=chroma, speed=speed, quality=quality, cdef=1 if speed > R.SPEED_SLOW else 0)
    out = R.encode_rgb8(t(g), t(b), t(r), cfg)
    torch.cuda.synchronize()
    assert "cdef" not in out["not_applied"]
    c420 = chroma == R.YUV420
    sub = 2 if c420 else 0
    fmt = R.CHROMA_420 if c420 else R.CHROMA_444
    y, cb, cr = O.csc_rgb8(r, g, b, c420)
    q = out["qidx"]
    dc, ac, uvqm = R.chroma_laws(q, sub)
    qm = mt.qm_level_law(q, sub)
    lam = O.rdo_lambda(q, 8, sub)
    rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, 8, q, qm, out["rdo_flags"], lam, chroma=fmt, src_u=cb,
                                                    src_v=cr, uv_dc=dc, uv_ac=ac, uv_qm=uvqm, wsse_scale=1.0)
    ly, luv, sharp = O.lf_levels(q, sub)
    m = lfgen.maps_from_leaves(lfgen.leaves_from_nodes(nodes, w, h), w, h)
    s = 1 if c420 else 0
    O.frame_deblock(rec, ru, rv, w, h, m, s, s, sharp, 8, ly, luv, luv)
    f, keep = av1.frame(nodes, lv, w, h, 8, q, chroma=fmt, qm_level=qm, uv_qm_level=uvqm, uv_dc=dc, uv_ac=ac)
    skip = av1.skip8(f)
    planes, prm, grid, _, _, _ = O.cdef_frame((rec, ru, rv), (y, cb, cr), skip, fmt, 8, q, lam, wsse_scale=1.0,
                                              slow=speed <= R.SPEED_SLOW)
    assert out["cdef"]["on"] == prm["on"], (out["cdef"], prm)
    for k in ("bits", "damping", "y_pri", "y_sec", "uv_pri", "uv_sec"):
        assert out["cdef"][k] == prm[k], k
    if prm["on"]:
        assert (out["cdef_grid"][:len(grid)] == grid).all()
    for k, want in zip(("cdef_y", "cdef_u", "cdef_v"), planes):
        assert (out[k].cpu().numpy().view(np.uint16) == want).all(), k
