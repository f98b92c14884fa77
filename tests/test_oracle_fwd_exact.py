"""Exact properties of the oracle's forward transform + quantization (oracle/oracle_tx.cpp, the
reference's src/dct.rs), beyond the float tolerances of tests/test_oracle.py:
  - levels are the rounded targets: with quantizer matrices the level is computed from the
    f32 target (apply_qmatrix, src/dct.rs:111), so level == round-half-away(target), 0 below
    one half, bit for bit; with flat quantizers the integer quant_q16 path and the f32 target
    path agree except within 1e-4 of a half (the reference's own bound is 0.5 + 1e-3,
    src/dct.rs:4285-4290); the 1-D and identity types have a full-step dead zone;
  - the separable butterflies' exact zeros: a residual with constant rows has no coefficient
    outside fx == 0 under a horizontal DCT (the reference's orientation test,
    src/dct.rs:4300-4315, for every size and type), constant columns none outside fy == 0
    under a vertical DCT; row-symmetric residuals have no odd-fx coefficient under a
    horizontal DCT, and row-antisymmetric ones no even-fx coefficient where the vertical pass
    is the identity (a rounded vertical pass is not exactly odd, so elsewhere the
    antisymmetry survives only to within its rounding);
  - the identity passes do not mix positions: an impulse stays at its coefficient (IDTX),
    a single nonzero column stays in its fx (V_DCT), a single nonzero row in its fy (H_DCT);
  - zero in, zero out.
A transposed operand, a swapped pass order or a wrong orientation of any (size, type) fails
these, and so does a rounding change on the integer or the f32 side of any coefficient."""
import numpy as np
import pytest

import oracle_lib as O

SIZES = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
TYPES = [0, 1, 2, 3, 9, 10, 11]
# (vertical kind, horizontal kind): 'D' DCT, 'A' ADST, 'I' identity
KINDS = {0: "DD", 1: "AD", 2: "DA", 3: "AA", 9: "II", 10: "DI", 11: "ID"}


def pairs():
    return [(s, t) for s in SIZES for t in TYPES if O.lib().orc_tx_supported(s, t)]


def wh(s):
    return O.tx_w(s), O.tx_h(s)


def coded(s):
    w, h = wh(s)
    return min(w, 32), min(h, 32)


def fwd(s, t, r, q):
    lv, tg = O.fwd(s, t, r.ravel(), q)
    cw, ch = coded(s)
    return lv.reshape(cw, ch), tg.reshape(cw, ch)  # [fx][fy] (x-major)


def round_half_away(v):
    v = v.astype(np.float64)
    return (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)


@pytest.mark.parametrize("s,t", pairs())
def test_levels_are_the_rounded_targets(s, t):
    rng = np.random.default_rng(s * 16 + t)
    w, h = wh(s)
    for bd in (8, 10, 12):
        for qml in (15, 0, 6):
            q = O.quant(int(rng.integers(0, 200)), bd, qml)
            for _ in range(40):
                r = rng.integers(-(1 << bd) + 1, 1 << bd, (h, w)) // int(rng.choice([1, 4, 32]))
                lv, tg = fwd(s, t, r, q)
                # half-step dead zone (quant_q16, apply_qmatrix); the 1-D and identity types
                # quantize full-step (quant_fullstep src/dct.rs:373, identity_quant_t_direct
                # :1586): below one step the level is 0
                dead = 1.0 if t >= 9 else 0.5
                want = np.where(np.abs(tg) < dead, 0, round_half_away(tg))
                if qml < 15 and t < 9:
                    assert (lv == want).all(), (s, t, bd, qml)
                else:
                    frac = np.abs(np.abs(tg.astype(np.float64)) % 1.0 - 0.5)
                    sure = frac > 1e-4
                    assert (lv[sure] == want[sure]).all(), (s, t, bd, qml)
                    assert np.abs(lv - tg).max() <= dead + 1e-3


@pytest.mark.parametrize("s,t", pairs())
def test_butterfly_exact_zeros(s, t):
    rng = np.random.default_rng(1000 + s * 16 + t)
    w, h = wh(s)
    vk, hk = KINDS[t]
    q = O.quant(4, 10, 15)  # a fine quantizer: nothing rounds away by accident
    for _ in range(30):
        lv, tg = fwd(s, t, np.zeros((h, w), np.int64), q)
        assert not lv.any() and not tg.any()
        col = rng.integers(-500, 501, (h, 1))
        row = rng.integers(-500, 501, (1, w))
        if hk == "D":
            lv, tg = fwd(s, t, np.repeat(col, w, axis=1), q)  # constant rows
            assert not lv[1:].any() and not tg[1:].any(), "energy outside fx == 0"
            half = rng.integers(-500, 501, (h, w // 2))
            sym = np.concatenate([half, half[:, ::-1]], axis=1)
            lv, tg = fwd(s, t, sym, q)
            assert not lv[1::2].any() and not tg[1::2].any(), "odd fx of a row-symmetric residual"
            if vk == "I":  # only an exact (identity) vertical pass keeps the rows antisymmetric
                anti = np.concatenate([half, -half[:, ::-1]], axis=1)
                lv, tg = fwd(s, t, anti, q)
                assert not lv[0::2].any() and not tg[0::2].any(), "even fx of a row-antisymmetric residual"
        if vk == "D":
            lv, tg = fwd(s, t, np.repeat(row, h, axis=0), q)  # constant columns
            assert not lv[:, 1:].any() and not tg[:, 1:].any(), "energy outside fy == 0"
        if hk == "I":
            x0 = int(rng.integers(0, w))
            r = np.zeros((h, w), np.int64)
            r[:, x0] = rng.integers(-500, 501, h)
            lv, tg = fwd(s, t, r, q)
            others = np.ones(lv.shape, bool)
            others[x0] = False
            assert not lv[others].any() and not tg[others].any(), "horizontal identity mixed columns"
            assert tg[x0].any()
        if vk == "I":
            y0 = int(rng.integers(0, h))
            r = np.zeros((h, w), np.int64)
            r[y0, :] = rng.integers(-500, 501, w)
            lv, tg = fwd(s, t, r, q)
            others = np.ones(lv.shape, bool)
            others[:, y0] = False
            assert not lv[others].any() and not tg[others].any(), "vertical identity mixed rows"
            assert tg[:, y0].any()
