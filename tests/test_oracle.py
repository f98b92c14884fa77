"""Pins of the CPU oracle (runs without a GPU).

Each test fixes the oracle against something independent of its own formulas: the
reference's known answers and round-trip bounds, closed forms of the transform bases, the
reference's table data (checksums), or brute force.  Reference file:line in each docstring.
"""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from refgen import itx_levels, pat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SIZES = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
TYPES = [0, 1, 2, 3, 9, 10, 11]


def supported(s, t):
    return bool(O.lib().orc_tx_supported(s, t))


# ------------------------------------------------------------------ quantizer / QM
def test_qm_level0_4x4_known_answer():
    """normative_level_zero_4x4_weights_match_av1 (src/quant.rs:498)."""
    exp = [32, 43, 73, 97, 43, 67, 94, 110, 73, 94, 137, 150, 97, 110, 150, 200]
    got = [O.lib().orc_qm_weight(0, 0, x * 4 + y, 4, 4) for y in range(4) for x in range(4)]
    assert got == exp


def test_qm_forward_inverse_reciprocal():
    """forward_and_inverse_matrix_weights_are_reciprocals (src/quant.rs:511)."""
    for rc in range(256):
        inv = O.lib().orc_qm_weight(10, 0, rc, 16, 16)
        fwd = (1024 + inv // 2) // inv
        assert abs(fwd * inv - 1024) <= inv // 2


def test_flat_level_preserves_scalar_quantizers():
    """flat_level_preserves_scalar_quantizers (src/quant.rs:488)."""
    q = O.quant(128, 8, 15)
    for rc in range(64):
        assert O.lib().orc_dequant_step(ctypes.byref(q), rc, 8, 8) == (q.dc if rc == 0 else q.ac)


def test_qm_level_law_known_answers():
    """high_quality_qm_taper_is_format_specific (src/quant.rs:472)."""
    law = O.lib().orc_qm_level_law
    for sub in (1, 2):
        for qi in (20, 12, 7, 1):
            assert law(qi, sub) == 8
    assert law(7, 0) == 15
    # the law's end points and monotone descent (src/quant.rs:137-154)
    assert law(128, 0) == 0 and law(255, 0) == 0 and law(48, 0) == 8
    vals = [law(q, 0) for q in range(21, 129)]
    assert all(a >= b for a, b in zip(vals, vals[1:]))


def test_dequant_tables_spot_values():
    """AV1 spec Dc_Qlookup / Ac_Qlookup end points at 8/10/12 bit (src/coef_q.rs:44-2690)."""
    for bd, (dc0, dc255, ac255) in {8: (4, 1336, 1828), 10: (4, 5347, 7312), 12: (4, 21387, 29247)}.items():
        q = O.quant(255, bd)
        assert q.ac == ac255
        q8 = O.quant(8, bd)  # DC index = 8 - 8 = 0
        assert q8.dc == dc0
        assert O.quant(255, bd).dc == O.quant(247 + 8, bd).dc
    assert O.quant(255, 8).dc == 1336 - (1336 - O.quant(255, 8).dc)  # delta -8 applied
    assert O.quant(263 - 8, 8).dc == O.quant(255, 8).dc


def test_quant_multipliers_and_clips():
    """Quant::new_with_qm (src/quant.rs:354): q_mult = round(65536/step); itx_clips (:243)."""
    for bd in (8, 10, 12):
        for qi in range(256):
            q = O.quant(qi, bd)
            assert q.q_mult_dc == int(math.floor(65536.0 / q.dc + 0.5))
            assert q.q_mult_ac == int(math.floor(65536.0 / q.ac + 0.5))
        q = O.quant(0, bd)
        if bd == 8:
            assert (q.rmin, q.rmax, q.cmin, q.cmax, q.cf_max) == (-32768, 32767, -32768, 32767, 32767)
        else:
            assert q.rmax == (1 << (bd + 7)) - 1 and q.rmin == -(1 << (bd + 7))
            assert q.cmax == (1 << (bd + 5)) - 1 and q.cmin == -(1 << (bd + 5))


def _fnv(vals, width):
    h = 0xCBF29CE484222325
    for v in vals:
        v &= (1 << (8 * width)) - 1
        for b in range(width):
            h ^= (v >> (8 * b)) & 0xFF
            h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_tables_match_reference_checksums():
    """The committed table header holds the reference's values (src/coef_q.rs, src/qm_tables.rs,
    src/dct.rs ADST matrices): FNV-1a of every table against tests/golden/tables_checksum.json,
    written from the reference's text by oracle/gen_tables.py."""
    hdr = open(os.path.join(ROOT, "maroontree_amd", "csrc", "av1_tables.h")).read()
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "tables_checksum.json")))

    def grab(name, count):
        s = hdr.index(name)
        body = hdr[hdr.index("{", s) + 1:hdr.index("};", s)]
        vals = [int(x) for x in re.findall(r"-?\d+", body)]
        assert len(vals) == count
        return vals

    assert _fnv(grab("MT_DC_QLOOKUP", 768), 2) == int(golden["dc_qlookup"], 16)
    assert _fnv(grab("MT_AC_QLOOKUP", 768), 2) == int(golden["ac_qlookup"], 16)
    assert _fnv(grab("MT_IQM", 100320), 1) == int(golden["iqm"], 16)
    assert _fnv(grab("MT_FADST4_Q12", 16), 4) == int(golden["adst4"], 16)
    assert _fnv(grab("MT_FADST8_Q12", 64), 4) == int(golden["adst8"], 16)
    assert _fnv(grab("MT_FADST16_Q12", 256), 4) == int(golden["adst16"], 16)


# ------------------------------------------------------------------ forward constants / 1-D
def _wc(k, n):
    return round_half_away(65536.0 / (2.0 * math.cos((2 * k + 1) * math.pi / (2 * n))))


def round_half_away(v):
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def _parse_int_array(path, name):
    src = open(path).read()
    m = re.search(name + r"\[\d*\]\s*=\s*\{([^}]*)\}", src)
    return [int(x) for x in re.findall(r"-?\d+", m.group(1))]


@pytest.mark.parametrize("path", ["oracle/oracle_tx.cpp", "maroontree_amd/csrc/mt_tx1d.h"])
def test_butterfly_constants_closed_form(path):
    """verify_wc{4,8,16,32}_coefficients / verify_sqrt2 (src/dct.rs:3818-3857):
    WC = round(65536 / (2 cos((2k+1) pi / 2N))), SQRT2 = round(65536 sqrt 2)."""
    p = os.path.join(ROOT, path)
    names = ("WC8", "WC16", "WC32") if "oracle" in path else ("kWC8", "kWC16", "kWC32")
    for name, n in zip(names, (8, 16, 32)):
        vals = _parse_int_array(p, name)
        assert vals == [_wc(k, n) for k in range(n // 2)], name
    src = open(p).read()
    assert str(_wc(0, 4)) in src and str(_wc(1, 4)) in src
    assert str(round_half_away(65536 * math.sqrt(2))) in src


def _impulse_matrix_fwd_dct(n, amp=1 << 10):
    m = np.zeros((n, n))
    for k in range(n):
        v = np.zeros(n, np.int32)
        v[k] = amp
        O.lib().orc_dct1d(n, O.ptr(v, ctypes.c_int32))
        m[:, k] = v / amp
    return m


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_forward_dct_closed_form(n):
    """The integer forward DCT is DCT-II with row scale 1 (DC) and sqrt 2 (AC) up to Q16
    rounding (test_dct1d_*_matches_reference, src/dct.rs:3985-4128)."""
    m = _impulse_matrix_fwd_dct(n)
    i = np.arange(n)[None, :]
    k = np.arange(n)[:, None]
    c = np.cos(np.pi * (2 * i + 1) * k / (2 * n)) * np.where(k == 0, 1.0, math.sqrt(2))
    assert np.abs(m - c).max() < 0.01


@pytest.mark.parametrize("n", [8, 16, 32])
def test_forward_dct_reference_tolerances(n):
    """test_dct1d_{8,16,32}_matches_reference and the zero / DC-only cases (src/dct.rs:3994-4128):
    the float Loeffler network of the test gives X[0] = sum, X[k] = sqrt2 sum cos."""
    tol = {8: 5, 16: 13, 32: 32}[n]
    sig = {8: [math.sin(i * 0.7 + 0.3) for i in range(8)],
           16: [math.cos(i * 0.4 + 1.0) for i in range(16)],
           32: [math.sin(i * 0.3) for i in range(32)]}[n]
    fixed = np.array([round_half_away(v * 128.0) for v in sig], np.int32)
    got = fixed.copy()
    O.lib().orc_dct1d(n, O.ptr(got, ctypes.c_int32))
    i = np.arange(n)[None, :]
    k = np.arange(n)[:, None]
    c = np.cos(np.pi * (2 * i + 1) * k / (2 * n)) * np.where(k == 0, 1.0, math.sqrt(2))
    exp = c @ fixed.astype(np.float64)
    assert np.abs(got - np.round(exp)).max() <= tol
    z = np.zeros(n, np.int32)
    O.lib().orc_dct1d(n, O.ptr(z, ctypes.c_int32))
    assert not z.any()
    dc = np.full(n, {8: 256, 16: 128, 32: 32}[n], np.int32)
    O.lib().orc_dct1d(n, O.ptr(dc, ctypes.c_int32))
    assert dc[0] != 0 and np.abs(dc[1:]).max() <= {8: 2, 16: 4, 32: 8}[n]


@pytest.mark.parametrize("n", [4, 8, 16])
def test_forward_adst_closed_form(n):
    """Forward ADST matrices (src/dct.rs:3178, :304, :445): sqrt 2 x DST-IV for 8/16 and
    4/3 x the sinpi(k+1)(2j+1)/9 basis for 4 (the inverse of dav1d's ADST-4)."""
    amp = 1 << 12
    m = np.zeros((n, n))
    for k in range(n):
        v = np.zeros(n, np.int32)
        v[k] = amp
        out = np.zeros(n, np.int32)
        O.lib().orc_fadst1d(n, O.ptr(v, ctypes.c_int32), O.ptr(out, ctypes.c_int32))
        m[:, k] = out / amp
    j = np.arange(n)[None, :]
    k = np.arange(n)[:, None]
    if n == 4:
        s = (4.0 / 3.0) * np.sin(np.pi * (j + 1) * (2 * k + 1) / 9.0)
    else:
        s = math.sqrt(2) * np.sin(np.pi * (2 * j + 1) * (2 * k + 1) / (4 * n))
    assert np.abs(m - s).max() < 3e-3


def _impulse_matrix_inv(fn, n, amp=1 << 12):
    m = np.zeros((n, n))
    for k in range(n):
        v = np.zeros(n, np.int32)
        v[k] = amp
        fn(n, O.ptr(v, ctypes.c_int32), -(1 << 30), (1 << 30) - 1)
        m[:, k] = v / amp
    return m


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_inverse_dct_closed_form(n):
    """dav1d inverse DCT-N (src/idct.rs:1432-2675): out[i] = in0/sqrt2 + sum in_k cos(pi(2i+1)k/2N)
    (64-point: the 32 coded inputs)."""
    m = _impulse_matrix_inv(O.lib().orc_inv_dct1d, n)
    i = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    a = np.cos(np.pi * (2 * i + 1) * k / (2 * n))
    a[:, 0] /= math.sqrt(2)
    if n == 64:
        a[:, 32:] = 0
        m[:, 32:] = 0
    assert np.abs(m - a).max() < 2e-3


@pytest.mark.parametrize("n", [4, 8, 16])
def test_inverse_adst_closed_form(n):
    """dav1d inverse ADST (src/idct.rs:2837, :1473, :1968): DST-IV for 8/16 and
    (2 sqrt2 / 3) sin(pi(i+1)(2k+1)/9) for 4."""
    m = _impulse_matrix_inv(O.lib().orc_inv_adst1d, n)
    i = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    if n == 4:
        a = np.sin(np.pi * (i + 1) * (2 * k + 1) / 9.0) * 2 * math.sqrt(2) / 3
    else:
        a = np.sin(np.pi * (2 * i + 1) * (2 * k + 1) / (4 * n))
    assert np.abs(m - a).max() < 2e-3


def test_mul_q16_and_quant_q16_brute_force():
    """mul_q16 / quant_q16 (src/dct.rs:60, :70) against exact integer arithmetic."""
    rng = np.random.default_rng(3)
    for _ in range(20000):
        d = int(rng.integers(-(1 << 22), 1 << 22))
        c = int(rng.integers(0, 700000))
        assert O.lib().orc_mul_q16(d, c) == (d * c + 32768) >> 16
        prod = d * c
        mag = abs(prod)
        exp = 0 if mag < 32768 else ((mag + 32768) >> 16) * (1 if prod >= 0 else -1)
        assert O.lib().orc_quant_q16(d, c) == exp


# ------------------------------------------------------------------ apply_qmatrix
def test_apply_qmatrix_reference_inputs():
    """assert_qmatrix_matches_scalar (src/dct.rs:3450-3501): on the reference's inputs the x86
    AVX2 rounding (the oracle's default) and the scalar FastRound rounding agree, and the
    levels round half away from zero."""
    for n in (1, 3, 4, 7, 8, 15, 16, 31, 64, 128, 1024):
        iwt = np.array([((i * 37 + 1) % 255 + 1) for i in range(n)], np.uint8)
        tg = np.array([[0.0, 0.499, -0.499, 0.5, -0.5, 31.25, -63.75][i % 8] if i % 8 < 7
                       else ((i * 977 + 13) % 8191) / 17.0 - 240.0 for i in range(n)], np.float32)
        a_l = np.full(n, -(1 << 31), np.int32)
        b_l = a_l.copy()
        a_t, b_t = tg.copy(), tg.copy()
        O.lib().orc_apply_qmatrix(O.ptr(a_l, ctypes.c_int32), O.ptr(a_t, ctypes.c_float),
                                  O.ptr(iwt, ctypes.c_uint8), n, 0)
        O.lib().orc_apply_qmatrix(O.ptr(b_l, ctypes.c_int32), O.ptr(b_t, ctypes.c_float),
                                  O.ptr(iwt, ctypes.c_uint8), n, 1)
        assert (a_l == b_l).all() and (a_t.view(np.uint32) == b_t.view(np.uint32)).all()
        scale = np.array([((1024 + w // 2) // w) / 32.0 for w in iwt.astype(np.int64)], np.float32)
        t = (tg * scale).astype(np.float32)
        exp = np.where(np.abs(t) < 0.5, 0, np.sign(t) * np.floor(np.abs(t).astype(np.float64) + 0.5))
        assert (a_l == exp.astype(np.int32)).all()


def test_fast_round_matches_reference_formula():
    """FastRound (src/util.rs:35-59) restated in numpy f32: one-ulp nudge, magic round."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-4000, 4000, 5000), np.arange(-20, 20, 0.25)]).astype(np.float32)
    magic = np.float32(12582912.0)
    nudged = (xs.view(np.uint32) + np.uint32(1)).view(np.float32)
    exp = (nudged + magic) - magic
    for x, e in zip(xs, exp):
        assert O.lib().orc_fast_round(float(x)) == float(e)


# ------------------------------------------------------------------ round trips (reference)
def _rt(s, t, resid, q):
    lv, tg = O.fwd(s, t, resid, q)
    return lv, O.inv(s, t, lv, q)


def test_adst4_roundtrip_highq():
    """adst4_roundtrip_highq / adstdct_dctadst_4x4_roundtrip (src/idct.rs:3025, :3074)."""
    q = O.quant_chroma(8, 0, 0, 8, 15)
    residuals = [
        [40, -32, 20, -12, 28, -24, 16, -8, 36, -28, 24, -4, 32, -20, 12, -16],
        [80, 70, -40, 30, -60, 50, 20, -10, 44, -36, 24, -16, 32, -28, 12, -4],
        [-50, 48, -46, 44, 42, -40, 38, -36, 34, 32, -30, 28, 26, -24, 22, -20],
    ]
    for r in residuals:
        r = np.array(r, np.int32)
        for t in (3, 1, 2):
            _, rec = _rt(0, t, r, q)
            assert np.abs(rec - r).sum() * 4 <= np.abs(r).sum() + 16


def test_adst4_dc_only_is_finite():
    """adst4_dc_only_is_finite (src/idct.rs:3060)."""
    q = O.quant_chroma(40, 0, 0, 8, 15)
    lv = np.zeros(16, np.int32)
    lv[0] = 12
    assert np.abs(O.inv(0, 3, lv, q)).max() < 4096


def test_dct16x4_and_4x16_pair_with_inverse():
    """dct16x4_and_4x16_pair_with_inverse (src/dct.rs:3755)."""
    q = O.quant(8, 8)
    r = pat(64)
    for s in (14, 13):
        cf, rec = _rt(s, 0, r, q)
        assert cf.any()
        assert ((r - rec).astype(np.float64) ** 2).sum() < (r.astype(np.float64) ** 2).sum() * 0.05


def test_dct16x8_pairs_with_inverse():
    """dct16x8_pairs_with_inverse (src/dct.rs:3781)."""
    q = O.quant(32, 8)
    r = pat(128)
    _, rec = _rt(8, 0, r, q)
    assert abs(r.mean() - rec.mean()) < 2.0
    assert ((r - rec).astype(np.float64) ** 2).sum() < (r.astype(np.float64) ** 2).sum() * 0.5


def test_vh_dct_8x8_roundtrip():
    """vh_dct_8x8_roundtrip_recovers_residual (src/idct.rs:3292)."""
    q = O.quant(40, 8)
    ve = np.array([60 if x < 4 else -60 for y in range(8) for x in range(8)], np.int32)
    he = np.array([60 if y < 4 else -60 for y in range(8) for x in range(8)], np.int32)
    for r, t in ((ve, 11), (he, 10)):
        cf, rec = _rt(1, t, r, q)
        assert cf.any()
        assert ((rec - r).astype(np.float64) ** 2).mean() < q.ac ** 2


def test_rect16_seven_type_roundtrips():
    """rect16_seven_type_roundtrips (src/idct.rs:3336)."""
    q = O.quant(40, 8)
    for w, s in ((16, 8), (8, 7)):
        h = 128 // w
        r = np.array([(40 if x < w // 2 else -40) + (y - h // 2) * 6 for y in range(h) for x in range(w)],
                     np.int32)
        for t in (1, 2, 10, 11, 9):
            cf, rec = _rt(s, t, r, q)
            assert cf.any()
            assert ((rec - r).astype(np.float64) ** 2).mean() < q.ac ** 2


def test_tx64_round_trip_matches_scale():
    """tx64_round_trip_matches_scale (src/dct.rs:4205)."""
    for qidx, max_mse in ((0, 6.0), (40, 12.0), (120, 60.0)):
        q = O.quant(max(qidx, 1), 8)
        s = (0x9E37 + qidx) & 0xFFFFFFFF
        r = np.zeros(4096, np.int32)
        for y in range(64):
            for x in range(64):
                s ^= (s << 13) & 0xFFFFFFFF
                s ^= s >> 17
                s ^= (s << 5) & 0xFFFFFFFF
                noise = (s % 9) - 4
                fx, fy = x / 64.0, y / 64.0
                v = 60.0 * math.sin(fx * 3.1) * math.cos(fy * 2.3) + 25.0 * fx - 30.0 * fy
                r[y * 64 + x] = round_half_away(v) + noise
        cf, rec = _rt(4, 0, r, q)
        d = (r - rec).astype(np.float64)
        assert (d ** 2).mean() < max_mse
        assert abs(d.mean()) < 0.6


def test_dct64_coefficients_vs_float_basis():
    """dct64x64_coeffs (src/dct.rs:4155) is 8x the orthonormal 2-D DCT of the residual: the
    oracle's levels at q_mult 65536 (unit step) equal numpy's f64 transform to +-1."""
    rng = np.random.default_rng(7)
    r = rng.integers(-300, 300, 4096).astype(np.int32)
    q = O.quant(0, 8)
    q.q_mult_dc = q.q_mult_ac = 65536
    lv, tg = O.fwd(4, 0, r, q)
    i = np.arange(64)
    b = np.array([(math.sqrt(1 / 64) if k == 0 else math.sqrt(2 / 64)) * np.cos(np.pi * (2 * i + 1) * k / 128)
                  for k in range(32)])
    coef = 8.0 * (b @ r.reshape(64, 64).astype(np.float64) @ b.T)  # [v][u]
    exp = np.round(coef.T.reshape(-1))  # x-major: u*32 + v
    assert np.abs(lv - exp).max() <= 1


@pytest.mark.parametrize("s", REF_SIZES)
def test_forward_2d_is_scaled_orthonormal_dct(s):
    """Every DCT_DCT forward lands on 8 x the orthonormal 2-D DCT of the residual divided by the
    step (the scale the inverse expects, src/dct.rs:2840-2846), in x-major order: the target
    of a flat-QM unit-step quantizer against numpy."""
    w, h = O.tx_w(s), O.tx_h(s)
    rng = np.random.default_rng(s)
    r = rng.integers(-200, 200, w * h).astype(np.int32)
    q = O.quant(0, 8)
    q.q_mult_dc = q.q_mult_ac = 65536
    lv, tg = O.fwd(s, 0, r, q)

    def orth(n, keep):
        i = np.arange(n)
        return np.array([(math.sqrt(1 / n) if k == 0 else math.sqrt(2 / n)) * np.cos(np.pi * (2 * i + 1) * k / (2 * n))
                         for k in range(keep)])

    cw, ch = min(w, 32), min(h, 32)
    coef = 8.0 * (orth(h, ch) @ r.reshape(h, w).astype(np.float64) @ orth(w, cw).T)  # [fy][fx]
    exp = coef.T.reshape(-1)
    scale = np.abs(exp).max()
    assert np.abs(tg - exp).max() < 0.02 * scale + 2.0


def test_itx_dispatch_levels_generator_roundtrip_bounded():
    """The inverse on the reference's dispatch-test levels (src/idct.rs:734, :750) stays inside
    the pixel-domain bound |out| <= 2^(bd+1) for every reference (size, type) and quantizer."""
    quants = [O.quant(8, 8), O.quant(96, 8), O.quant_chroma(40, 0, 0, 10, 15),
              O.quant(40, 8, 0), O.quant(96, 10, 7), O.quant(180, 12, 14),
              O.quant_chroma(128, -7, 11, 12, 3)]
    for s in REF_SIZES:
        n = O.ncoef(s)
        for t in TYPES:
            if not supported(s, t):
                continue
            for qi, q in enumerate(quants):
                lv = itx_levels(n, 0x12345678 ^ qi)
                out = O.inv(s, t, lv, q)
                assert np.abs(out).max() <= (1 << (q.bd + 5))


def test_sse_recon_brute_force():
    """sse_recon_scalar (src/rd_sse.rs:1262) against a numpy restatement on random blocks."""
    rng = np.random.default_rng(11)
    for w, h in ((4, 4), (8, 16), (32, 32)):
        pred = rng.integers(0, 1024, w * h).astype(np.int32)
        resid = rng.integers(-600, 600, w * h).astype(np.int32)
        src = rng.integers(0, 1024, w * h).astype(np.uint16)
        got = O.lib().orc_sse_recon(O.ptr(pred, ctypes.c_int32), O.ptr(resid, ctypes.c_int32),
                                    O.ptr(src, ctypes.c_uint16), w, w, h, 1023)
        rec = np.clip(pred + resid, 0, 1023)
        assert got == int(((src.astype(np.int64) - rec) ** 2).sum())


def test_cdf_tables_match_reference_checksums():
    """Every table of av1_cdf_tables.h (default coefficient CDFs, scans, context offsets, intra
    mode CDFs; reference src/coef_q.rs, src/tables.rs) against the FNV-1a the recipe recorded
    from the reference's text."""
    hdr = open(os.path.join(ROOT, "maroontree_amd", "csrc", "av1_cdf_tables.h")).read()
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "tables_checksum.json")))
    names = re.findall(r"static const \w+ MT_(\w+?)(?:\[\d+\])+ = \{", hdr)
    assert len(names) >= 70
    for name in names:
        s = hdr.index(f"MT_{name}[")
        body = hdr[hdr.index("{", s) + 1:hdr.index("};", s)]
        vals = [int(x) for x in re.findall(r"-?\d+", body)]
        assert _fnv(vals, 4) == int(golden["cdf_" + name.lower()], 16), name
