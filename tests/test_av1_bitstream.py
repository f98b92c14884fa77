"""The AV1 bitstream of the RDO's decisions (libmtgpu's host coder: mt_av1_tokens,
mt_av1_range_code, mt_av1_encode, mt_avif_encode), checked without a GPU:

- the range coder against the specification's symbol decoder (tests/spec_msac.py): random
  bool, literal and CDF tokens, and every token of real frames, decode to the same symbols,
  and the tile ends with the exit process's trailing one bit;
- the whole file against an independent AV1 decoder: the oracle's RDO (decisions and the
  reconstruction's levels) coded by the library, decoded by the AVIF decoder installed with
  Pillow (libavif), equals the oracle's reconstruction sample for sample -- luma-only frames
  (monochrome), 4:4:4 with the identity matrix (decoded R, G, B are V, Y, U exactly), 4:2:0
  with neutral chroma (decoded gray is Y exactly), across sizes that are not multiples of 64,
  every RDO flag, quantizers, QM levels and the loop filter at level 0; 4:2:0 with real chroma
  and 10-bit frames within the decoder's 8-bit RGB conversion;
- the refusals: several tiles, the identity matrix without 4:4:4, lossless.
"""
import io

import numpy as np
import pytest

import oracle_lib as O
from maroontree_amd import av1, workload as W
from spec_msac import SymbolDecoder

PIL = pytest.importorskip("PIL")
from PIL import Image, features  # noqa: E402

if not features.check("avif"):
    pytest.skip("the installed Pillow has no AVIF decoder", allow_module_level=True)

ANG, FI, MIN4, CFL = 1, 2, 4, 8


def _decode(data):
    im = Image.open(io.BytesIO(data))
    im.load()
    return np.asarray(im).astype(np.int64)


def _check_tokens(toks, tile):
    d = SymbolDecoder(tile)
    for i, t in enumerate(toks):
        t = int(t)
        if t >> 31:
            got, want = d.boolean((t >> 1) & 0x7FFF), t & 1
        else:
            cdf, n = av1.token_cdf(t)
            got, want = d.symbol(cdf, n), t & 15
        assert got == want, f"token {i}"
    assert d.trailing_ok()


def test_range_coder_round_trips_random_tokens():
    rng = np.random.default_rng(11)
    # CDF tokens need stored CDFs: take them from a real frame's token stream
    img = W.synth_image(2, 64, 64, 8)
    rec, _, _, nodes, st, lv = O.intra_rdo_levels(img, 8, 91, 15, ANG | FI | MIN4, O.rdo_lambda(91, 8, 2))
    f, keep = av1.frame(nodes, lv, 64, 64, 8, 91)
    toks = av1.tokens(f)
    cdf_toks = np.unique(toks[(toks >> 31) == 0] & np.uint32(0xFFFFFFF0))
    for trial in range(40):
        n = int(rng.integers(1, 600))
        toks = []
        for _ in range(n):
            k = rng.random()
            if k < 0.3:
                toks.append((1 << 31) | (int(rng.integers(1, 32768)) << 1) | int(rng.integers(0, 2)))
            elif k < 0.45:
                toks.append((1 << 31) | (16384 << 1) | int(rng.integers(0, 2)))
            else:
                base = int(rng.choice(cdf_toks))
                nsym = ((base >> 4) & 15) + 1
                toks.append(base | int(rng.integers(0, nsym)))
        toks = np.array(toks, np.uint32)
        _check_tokens(toks, av1.range_code(toks))
    # extreme probabilities
    for p in (1, 2, 32766, 32767):
        toks = np.array([(1 << 31) | (p << 1) | (i % 2) for i in range(300)], np.uint32)
        _check_tokens(toks, av1.range_code(toks))


def _mono(w, h, q, flags, bd=8, seed=5, qm=15, lf=(0, 0, 0, 0)):
    img = W.synth_image(seed, w, h, bd)
    rec, _, _, nodes, st, lv = O.intra_rdo_levels(img, bd, q, qm, flags, O.rdo_lambda(q, bd, 2))
    f, keep = av1.frame(nodes, lv, w, h, bd, q, qm_level=qm, filter_intra=bool(flags & FI), lf_levels=lf,
                        cicp=av1.BT601)
    return rec, st, f, keep


@pytest.mark.parametrize("w,h,q,flags,qm", [
    (64, 64, 120, ANG | FI | MIN4, 15), (200, 136, 91, ANG | FI | MIN4, 15), (136, 72, 30, ANG | FI | MIN4, 15),
    (128, 200, 180, 0, 15), (264, 136, 60, ANG | FI | MIN4, 7), (72, 264, 10, MIN4, 15), (96, 64, 250, ANG, 3),
    (8, 8, 120, ANG | FI | MIN4, 15), (16, 136, 140, FI | MIN4, 15)])
def test_monochrome_frames_decode_to_the_reconstruction(w, h, q, flags, qm):
    rec, st, f, keep = _mono(w, h, q, flags, qm=qm)
    toks = av1.tokens(f)
    obus = av1.encode_obus(f)
    _check_tokens(toks, av1.range_code(toks))
    assert obus.endswith(av1.range_code(toks))
    dec = _decode(av1.encode_avif(f))
    assert dec.shape[:2] == (h, w)
    assert (dec[..., 1] == rec).all(), f"{int((dec[..., 1] != rec).sum())} samples differ"
    # the tile is within a few percent of the RDO's rate model (static CDFs on both sides)
    tile_bits = 8 * len(av1.range_code(toks))
    assert abs(tile_bits - st["bits"]) < 0.05 * st["bits"] + 64


@pytest.mark.parametrize("w,h,q,flags", [(200, 136, 91, ANG | FI | MIN4 | CFL), (136, 72, 30, ANG | FI | MIN4 | CFL),
                                         (128, 200, 180, ANG | FI | MIN4 | CFL), (64, 64, 20, ANG | MIN4 | CFL),
                                         (264, 136, 60, ANG | FI | MIN4)])
def test_444_identity_frames_decode_to_all_three_reconstructions(w, h, q, flags):
    y, u, v = W.synth_yuv(7, w, h, 8, 0)
    dc, ac, qm = O.chroma_laws(q, 0)
    rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, 8, q, 15, flags, O.rdo_lambda(q, 8, 0), chroma=2, src_u=u,
                                                    src_v=v, uv_dc=dc, uv_ac=ac, uv_qm=qm)
    f, keep = av1.frame(nodes, lv, w, h, 8, q, chroma=2, uv_qm_level=qm, uv_dc=dc, uv_ac=ac,
                        filter_intra=bool(flags & FI), cicp=av1.SRGB_IDENTITY)
    dec = _decode(av1.encode_avif(f))
    for name, got, want in (("Y", dec[..., 1], rec), ("U", dec[..., 2], ru), ("V", dec[..., 0], rv)):
        assert (got == want).all(), f"{name}: {int((got != want).sum())} samples differ"


@pytest.mark.parametrize("w,h,q,flags", [(200, 136, 91, ANG | FI | MIN4 | CFL), (136, 72, 53, ANG | FI | MIN4 | CFL),
                                         (72, 136, 140, ANG | MIN4 | CFL)])
def test_420_frames(w, h, q, flags):
    dc, ac, qm = O.chroma_laws(q, 2)
    y, u, v = W.synth_yuv(9, w, h, 8, 2)
    # neutral chroma: where the chroma around a sample reconstructs to 128 the decoder's gray is
    # Y exactly (directional modes at the frame edges predict from 127 / 129, so not everywhere)
    g = np.full_like(u, 128)
    rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, 8, q, 15, flags, O.rdo_lambda(q, 8, 2), chroma=1, src_u=g,
                                                    src_v=g, uv_dc=dc, uv_ac=ac, uv_qm=qm)
    f, keep = av1.frame(nodes, lv, w, h, 8, q, chroma=1, uv_qm_level=qm, uv_dc=dc, uv_ac=ac,
                        filter_intra=bool(flags & FI), cicp=av1.BT601)
    dec = _decode(av1.encode_avif(f))
    neutral = (_upsample(ru) == 128) & (_upsample(rv) == 128)
    assert neutral.mean() > 0.05
    for c in range(3):
        assert (dec[..., c][neutral] == rec[neutral]).all()
    # real chroma: the decoder's BT.601 RGB against the same conversion of the reconstruction
    rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, 8, q, 15, flags, O.rdo_lambda(q, 8, 2), chroma=1, src_u=u,
                                                    src_v=v, uv_dc=dc, uv_ac=ac, uv_qm=qm)
    f, keep = av1.frame(nodes, lv, w, h, 8, q, chroma=1, uv_qm_level=qm, uv_dc=dc, uv_ac=ac,
                        filter_intra=bool(flags & FI), cicp=av1.BT601)
    dec = _decode(av1.encode_avif(f))
    uu, vv = _upsample(ru) - 128, _upsample(rv) - 128
    yy = rec.astype(np.float64)
    rgb = np.stack([yy + 1.402 * vv, yy - 0.344136 * uu - 0.714136 * vv, yy + 1.772 * uu], -1)
    err = np.abs(dec - np.clip(rgb, 0, 255))
    # the decoder's float conversion rounds once; one chroma sample off by one moves the
    # result by up to 1.8 x 9 / 16
    assert err.max() < 2.0, err.max()


def _upsample(c):
    """4:2:0 chroma to the luma grid with the 9-3-3-1 bilinear weights of centred samples."""
    c = c.astype(np.float64)
    p = np.pad(c, 1, mode="edge")
    H, Wd = c.shape
    out = np.zeros((2 * H, 2 * Wd))
    for dy in (0, 1):
        for dx in (0, 1):
            ny, nx = 2 * dy, 2 * dx
            out[dy::2, dx::2] = (9 * p[1:-1, 1:-1] + 3 * p[ny:ny + H, 1:-1] + 3 * p[1:-1, nx:nx + Wd]
                                 + p[ny:ny + H, nx:nx + Wd]) / 16
    return out


def test_10bit_monochrome_within_the_decoders_8bit_output():
    rec, st, f, keep = _mono(200, 136, 91, ANG | FI | MIN4, bd=10)
    dec = _decode(av1.encode_avif(f))
    assert np.abs(dec[..., 1] - np.round(rec * 255.0 / 1023.0)).max() <= 1


def test_loop_filter_levels_are_signalled():
    rec, st, f, keep = _mono(136, 72, 91, ANG | FI | MIN4, lf=(0, 0, 0, 0))
    plain = av1.encode_obus(f)
    rec, st, f2, keep2 = _mono(136, 72, 91, ANG | FI | MIN4, lf=(9, 9, 0, 0))
    filtered = av1.encode_obus(f2)
    assert plain != filtered and len(plain) == len(filtered)
    dec = _decode(av1.encode_avif(f2))
    assert dec.shape[:2] == (72, 136) and (dec[..., 1] != rec).any()


def test_refusals():
    rec, st, f, keep = _mono(64, 64, 91, 0)
    f.chroma = 1
    f.cicp[0], f.cicp[1], f.cicp[2] = av1.SRGB_IDENTITY
    f.coef_u = f.coef_v = f.coef_y
    f.coef_uv_stride = 64
    with pytest.raises(av1._lib.MtError):
        av1.encode_obus(f)
    rec, st, f, keep = _mono(64, 64, 91, 0)
    f.qidx = 0
    with pytest.raises(av1._lib.MtError):
        av1.encode_obus(f)
    rec, st, f, keep = _mono(64, 64, 91, 0)
    f.width = 4160  # 65 superblock columns need two tiles
    with pytest.raises(av1._lib.MtError):
        av1.encode_obus(f)


@pytest.mark.parametrize("fmt,w,h,q", [(0, 200, 136, 91), (0, 136, 72, 160), (2, 200, 136, 120), (2, 264, 136, 60)])
def test_deblocked_reconstruction_equals_the_decoders_loop_filter(fmt, w, h, q):
    """The frame header signals the reference's loop-filter levels (loop_filter_levels at the
    shipped tuning); the decoder's in-loop deblocking of the decoded frame equals the oracle's
    frame_deblock (src/coder/frame_deblock.rs:206) of the oracle's reconstruction over the
    maps of its decisions: the AV1 loop filter of the specification's 7.14, run by an
    independent decoder, pins the deblocking restatement end to end."""
    import lfgen
    flags = ANG | FI | MIN4 | (CFL if fmt else 0)
    sub = 0 if fmt == 2 else 2
    ly, luv, sharp = O.lf_levels(q, sub)
    assert ly > 0
    if fmt == 0:
        y = W.synth_image(13, w, h, 8)
        rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, 8, q, 15, flags, O.rdo_lambda(q, 8, 2))
        f, keep = av1.frame(nodes, lv, w, h, 8, q, filter_intra=True, lf_levels=(ly, ly, 0, 0),
                            lf_sharpness=sharp, cicp=av1.BT601)
    else:
        y, u, v = W.synth_yuv(13, w, h, 8, 0)
        dc, ac, qm = O.chroma_laws(q, 0)
        rec, ru, rv, nodes, st, lv = O.intra_rdo_levels(y, 8, q, 15, flags, O.rdo_lambda(q, 8, 0), chroma=2,
                                                        src_u=u, src_v=v, uv_dc=dc, uv_ac=ac, uv_qm=qm)
        f, keep = av1.frame(nodes, lv, w, h, 8, q, chroma=2, uv_qm_level=qm, uv_dc=dc, uv_ac=ac, filter_intra=True,
                            lf_levels=(ly, ly, luv, luv), lf_sharpness=sharp, cicp=av1.SRGB_IDENTITY)
    m = lfgen.maps_from_leaves(lfgen.leaves_from_nodes(nodes, w, h), w, h)
    want = [rec.copy(), None if ru is None else ru.copy(), None if rv is None else rv.copy()]
    O.frame_deblock(want[0], want[1], want[2], w, h, m, 0, 0, sharp, 8, ly, luv, luv)
    assert (want[0] != rec).any()  # the filter changed something
    dec = _decode(av1.encode_avif(f))
    assert (dec[..., 1] == want[0]).all(), f"Y: {int((dec[..., 1] != want[0]).sum())} samples differ"
    if fmt:
        assert (dec[..., 2] == want[1]).all() and (dec[..., 0] == want[2]).all()
