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
