"""The oracle's CDEF (oracle/oracle_cdef.cpp): the direction search recovers the direction of
blocks built constant along each direction's lines (where the normalised projection cost
reaches its Cauchy-Schwarz maximum), the filter's identities, and end to end against an
independent AV1 decoder: the RDO's decisions with the oracle's CDEF decision signalled in the
frame header and per superblock decode to the oracle's filtered planes, sample for sample
(monochrome and 4:4:4 identity, with and without deblocking before it)."""
import io

import numpy as np
import pytest

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import av1
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

# the lines of each direction's partial sums (src/cdef.rs:117-126): pixel (i, j) -> line
LINES = [lambda i, j: i + j, lambda i, j: i + j // 2, lambda i, j: i, lambda i, j: 3 + i - j // 2,
         lambda i, j: 7 + i - j, lambda i, j: 3 - i // 2 + j, lambda i, j: j, lambda i, j: i // 2 + j]


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_direction_of_line_patterns(bd):
    rng = np.random.default_rng(bd)
    for d in range(8):
        for _ in range(20):
            vals = rng.integers(0, 1 << bd, 15)
            blk = np.array([[vals[LINES[d](i, j)] for j in range(8)] for i in range(8)], np.uint16)
            plane = np.zeros((16, 16), np.uint16)
            plane[4:12, 4:12] = blk
            got, var = O.cdef_dir(plane, 4, 4, bd)
            assert got == d, (d, got)
            assert var >= 0


def test_filter_identities():
    rng = np.random.default_rng(5)
    for bd in (8, 10):
        p = rng.integers(0, 1 << bd, (24, 24)).astype(np.uint16)
        for d in range(8):
            assert (O.cdef_filter(p, 8, 8, 8, 8, 0, 0, d, 5, bd) == p).all()  # no strength: untouched
            flat = np.full((24, 24), 77 << (bd - 8), np.uint16)
            assert (O.cdef_filter(flat, 8, 8, 8, 8, 15 << (bd - 8), 4 << (bd - 8), d, 6, bd) == flat).all()
        # the twelve pairs: pair (0, 0) is the block itself, each pair equals the filter at
        # its effective strengths (variance-adjusted primary; direction 0 without primary)
        var = 3000
        c = O.cdef_candidates(p, 8, 8, 8, 8, 3, var, 5, bd)
        assert (c[0] == p[8:16, 8:16]).all()
        for si, sec in enumerate((0, 1, 2)):
            for pi, pri in enumerate((0, 1, 2, 4)):
                pe = O.lib().orc_cdef_adjust_pri(pri << (bd - 8), var)
                f = O.cdef_filter(p, 8, 8, 8, 8, pe, sec << (bd - 8), 3 if pri else 0, 5, bd)
                assert (c[si * 4 + pi] == f[8:16, 8:16]).all(), (pri, sec)
    # adjust_pri (AV1 7.15.1 adjust_strength): zero variance disables the primary taps
    assert O.lib().orc_cdef_adjust_pri(4, 0) == 0
    assert O.lib().orc_cdef_adjust_pri(16, 64 << 12) == (16 * 16 + 8) >> 4


def _decode(data):
    from PIL import Image
    im = Image.open(io.BytesIO(data))
    im.load()
    return np.asarray(im).astype(np.int64)


def rdo_with_cdef(seed, w, h, q, fmt, deblock, wsse, slow=True, bd=8):
    sub = 2 if fmt == R.CHROMA_420 else 0
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | (R.CFL if fmt else 0)
    if fmt == R.LUMA_ONLY:
        y = W.synth_image(seed, w, h, bd)
        u = v = None
    else:
        y, u, v = W.synth_yuv(seed, w, h, bd, sub)
    qm = mt.qm_level_law(q, sub)
    lam = R.rdo_lambda(q, bd, sub if fmt else 2)
    dc = ac = 0
    uqm = 15
    if fmt:
        dc, ac, uqm = R.chroma_laws(q, sub)
    rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, bd, q, qm, flags, lam, chroma=fmt, src_u=u, src_v=v,
                                                    uv_dc=dc, uv_ac=ac, uv_qm=uqm, nthreads=8)
    cicp = av1.SRGB_IDENTITY if fmt == R.CHROMA_444 else av1.BT601
    kw = dict(chroma=fmt, qm_level=qm, uv_qm_level=uqm, uv_dc=dc, uv_ac=ac, filter_intra=True, cicp=cicp)
    planes = (rec, ru, rv)
    if deblock:  # the deblocked frame is CDEF's input (specification 7.14 then 7.15)
        import lfgen
        ly, luv, sharp = O.lf_levels(q, sub if fmt else 2)
        m = lfgen.maps_from_leaves(lfgen.leaves_from_nodes(nodes, w, h), w, h)
        planes = (rec.copy(), None if ru is None else ru.copy(), None if rv is None else rv.copy())
        O.frame_deblock(planes[0], planes[1], planes[2], w, h, m, sub // 2, sub // 2, sharp, bd, ly, luv, luv)
        kw["lf_levels"] = (ly, ly, luv, luv) if fmt else (ly, ly, 0, 0)
        kw["lf_sharpness"] = sharp
    f, keep = av1.frame(nodes, lv, w, h, bd, q, **kw)
    skip = av1.skip8(f)
    out, prm, grid, dirs, vars_, tabs = O.cdef_frame(planes, (y, u, v), skip, fmt, bd, q, lam,
                                                     wsse_scale=wsse, slow=slow)
    return (y, u, v), planes, out, prm, grid, skip, kw, (nodes, lv)


@pytest.mark.parametrize("fmt,w,h,q,deblock,wsse,slow", [
    (R.LUMA_ONLY, 200, 136, 180, False, 0.0, True),
    (R.LUMA_ONLY, 264, 136, 200, True, 1.0, True),
    (R.LUMA_ONLY, 136, 128, 120, True, 0.0, False),
    (R.CHROMA_444, 136, 128, 200, False, 1.0, True),
    (R.CHROMA_444, 200, 136, 160, True, 0.0, False),
    (R.CHROMA_444, 200, 136, 230, True, 1.0, True),
])
def test_cdef_decodes_to_the_oracle(fmt, w, h, q, deblock, wsse, slow):
    src, pre, out, prm, grid, skip, kw, (nodes, lv) = rdo_with_cdef(41 + q, w, h, q, fmt, deblock, wsse, slow)
    assert prm["on"] == 1, prm
    assert (out[0] != pre[0]).any(), "CDEF changed nothing"
    assert not skip.all()
    f, keep = av1.frame(nodes, lv, w, h, 8, q, cdef=prm, cdef_grid=grid, **kw)
    dec = _decode(av1.encode_avif(f))
    assert (dec[..., 1] == out[0]).all()
    if fmt == R.CHROMA_444:
        assert (dec[..., 2] == out[1]).all() and (dec[..., 0] == out[2]).all()
