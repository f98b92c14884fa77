"""The oracle's colour conversion and quality map (oracle/oracle_csc.cpp: encode_rgb8's
RGB -> YCbCr of src/encoder.rs:114-127, :398-410, :560-602, and quality_to_q of
src/avif.rs:637), pinned by what the arithmetic fixes:
  - the Q13 coefficients are the rounded BT.601 full-range matrix (each row of Y sums to 8192,
    each chroma row to 0), so gray maps to (v, 128, 128) exactly and every sample is within
    one of the real-valued BT.601 conversion, rounded;
  - 4:2:0 chroma is the rounded mean of the four Q13 values: a constant 2 x 2 quad gives the
    4:4:4 value, and the mean of four 4:4:4-representable quads is bracketed by their
    rounded values; past an odd width / height the last column / row repeats;
  - quality_to_q passes through the reference's anchors (quality 50 -> 91, 80 -> 35) and
    never rises with quality."""
import numpy as np

import oracle_lib as O


def bt601(r, g, b):
    r, g, b = (p.astype(np.float64) for p in (r, g, b))
    y = 0.299 * r + 0.587 * g + 0.114 * b
    cb = 128.0 - 0.168736 * r - 0.331264 * g + 0.5 * b
    cr = 128.0 + 0.5 * r - 0.418688 * g - 0.081312 * b
    return y, cb, cr


def test_gray_is_exact():
    v = np.arange(256, dtype=np.uint8).reshape(16, 16)
    for c420 in (False, True):
        y, cb, cr = O.csc_rgb8(v, v, v, c420)
        assert (y == v).all()
        assert (cb == 128).all() and (cr == 128).all()


def test_444_within_one_of_bt601():
    rng = np.random.default_rng(5)
    r, g, b = (rng.integers(0, 256, (48, 40), dtype=np.uint8) for _ in range(3))
    y, cb, cr = O.csc_rgb8(r, g, b, False)
    fy, fcb, fcr = bt601(r, g, b)
    for got, want in ((y, fy), (cb, fcb), (cr, fcr)):
        assert np.abs(got.astype(np.float64) - np.clip(want, 0, 255)).max() <= 1.0
    # the extremes of the cube saturate where the real value leaves 0..255
    assert O.csc_rgb8(np.full((2, 2), 255, np.uint8), np.zeros((2, 2), np.uint8), np.zeros((2, 2), np.uint8),
                      False)[2][0, 0] == 255


def test_420_is_the_rounded_quad_mean():
    rng = np.random.default_rng(6)
    # constant quads: the 4:4:4 value
    q = rng.integers(0, 256, (3, 12, 10), dtype=np.uint8)
    r, g, b = (np.repeat(np.repeat(p, 2, 0), 2, 1) for p in q)
    y4, cb4, cr4 = O.csc_rgb8(r, g, b, False)
    y2, cb2, cr2 = O.csc_rgb8(r, g, b, True)
    assert (y2 == y4).all()
    assert (cb2 == cb4[::2, ::2]).all() and (cr2 == cr4[::2, ::2]).all()
    # general quads: within the real-valued mean, rounded, +-1
    r, g, b = (rng.integers(0, 256, (24, 20), dtype=np.uint8) for _ in range(3))
    _, cb2, cr2 = O.csc_rgb8(r, g, b, True)
    _, fcb, fcr = bt601(r, g, b)
    for got, f in ((cb2, fcb), (cr2, fcr)):
        m = 0.25 * (f[0::2, 0::2] + f[1::2, 0::2] + f[0::2, 1::2] + f[1::2, 1::2])
        assert np.abs(got - np.clip(m, 0, 255)).max() <= 1.0


def test_420_odd_sizes_repeat_the_edge():
    rng = np.random.default_rng(7)
    r, g, b = (rng.integers(0, 256, (9, 7), dtype=np.uint8) for _ in range(3))
    _, cb, cr = O.csc_rgb8(r, g, b, True)
    assert cb.shape == (5, 4)
    pad = lambda p: np.pad(p, ((0, 1), (0, 1)), mode="edge")  # noqa: E731
    _, cbp, crp = O.csc_rgb8(pad(r), pad(g), pad(b), True)
    assert (cb == cbp).all() and (cr == crp).all()


def test_quality_map():
    anchors = [(1, 215), (30, 140), (40, 116), (50, 91), (60, 70), (70, 53), (80, 35), (85, 28), (90, 17),
               (93, 12), (96, 7), (98, 5), (99, 3), (100, 1)]
    for ql, q in anchors:
        assert O.quality_to_q(ql) == q
    qs = [O.quality_to_q(ql) for ql in range(0, 102)]
    assert all(a >= b for a, b in zip(qs, qs[1:]))
    assert O.quality_to_q(0) == 215 and O.quality_to_q(101) == 1
    # between anchors: q0 + (q1 - q0) * t / (l1 - l0), truncated toward zero: 116 - 12.5 -> 104
    assert O.quality_to_q(45) == 104
