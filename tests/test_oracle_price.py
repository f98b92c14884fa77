"""The oracle's decision pricing helpers (oracle/oracle_price.cpp) against what fixes them
independently: the exp2 table against its definition, dirty_exp2f against exp within the
reference's own bound (src/coder/lossy_state.rs:4307-4326), dirty_log1pf against log1p,
the activities against f64 moments, the SATD price and the perceptual scale against their
closed forms evaluated in f64, with the tuning's knees and anchors."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle_lib as O

LN2 = math.log(2.0)


def test_exp2_table_is_its_definition():
    getcontext().prec = 50
    ln2 = Decimal(2).ln()
    want = np.array([float((ln2 * Decimal(i - 32) / Decimal(64)).exp()) for i in range(64)]).astype(np.float32)
    assert (O.exp2_table().view(np.uint32) == want.view(np.uint32)).all()
    assert O.exp2_table()[32] == 1.0


def test_dirty_exp2f_within_the_reference_bound():
    worst = 0.0
    for i in range(65537):
        x = np.float32(-LN2 + 2 * LN2 * (i / 65536))
        got = O.dirty_exp2f(np.float32(x * np.float32(1.4426950408889634)))
        worst = max(worst, abs(got - math.exp(float(x))) / math.exp(float(x)))
    assert worst <= 3.0e-7, worst
    assert O.dirty_exp2f(0.0) == 1.0 and O.dirty_exp2f(1.0) == 2.0 and O.dirty_exp2f(-1.0) == 0.5


def test_dirty_log1pf_against_log1p():
    assert O.dirty_log1pf(0.0) == 0.0
    worst = 0.0
    for d in np.concatenate([np.geomspace(1e-6, 1e7, 4000), np.linspace(0, 4, 2001)[1:]]):
        d = float(np.float32(d))
        worst = max(worst, abs(O.dirty_log1pf(d) - math.log1p(d)) / math.log1p(d))
    assert worst < 3e-6, worst


def _act64(block, bd):
    v = block.astype(np.float64)
    var = max(v.var(), 0.0) / (1 << (2 * (bd - 8)))
    return math.log1p(var)


@pytest.mark.parametrize("bd,w,h", [(8, 200, 136), (10, 64, 64), (12, 72, 130)])
def test_sb_activity_and_tile_mean(bd, w, h):
    rng = np.random.default_rng(bd * 7 + w)
    plane = rng.integers(0, 1 << bd, (h, w)).astype(np.uint16)
    plane[: h // 2, : w // 3] = 37  # a flat region
    acts = []
    for sy in range(0, h, 64):
        for sx in range(0, w, 64):
            got = O.sb_activity(plane, sy, sx, bd)
            want = _act64(plane[sy:sy + 64, sx:sx + 64], bd)
            assert abs(got - want) <= 3e-6 * max(want, 1.0) + 1e-6, (sy, sx, got, want)
            acts.append(got)
    ref = O.tile_ref_activity(plane, bd)
    assert abs(ref - np.mean(acts)) <= 1e-5 * max(np.mean(acts), 1.0)
    assert O.sb_activity(np.full((64, 64), 900, np.uint16), 0, 0, 10) == 0.0
    assert O.sb_activity(plane, h, 0, bd) == 0.0  # outside the tile
    assert O.tile_ref_activity(np.zeros((1, 1), np.uint16), 8, w=0, h=0) == 5.0


def test_satd_scale_knees_and_k_satd_anchors():
    for q in (0, 10, 20):
        assert O.luma_satd_scale(q, 0) == np.float32(2.921556)
    for q in (48, 100, 255):
        assert O.luma_satd_scale(q, 0) == np.float32(1.25)
    for q in (0, 34, 200):
        assert O.luma_satd_scale(q, 1) == np.float32(1.25)
    assert abs(O.luma_satd_scale(34, 0) - (2.921556 + 1.25) / 2) < 1e-6
    anchors = {0: (28, 81, 140), 1: (28, 81, 140)}
    vals = {0: (35.0, 21.2, 18.1), 1: (24.5, 21.3, 18.2)}
    for sub in (0, 1):
        for qa, v in zip(anchors[sub], vals[sub]):
            assert O.k_satd(qa, sub) == np.float32(v)
        assert O.k_satd(0, sub) == np.float32(vals[sub][0]) and O.k_satd(255, sub) == np.float32(vals[sub][2])
        mid = (28 + 81) / 2
        assert abs(O.k_satd(int(mid), sub) - (vals[sub][0] + (vals[sub][1] - vals[sub][0]) * (int(mid) - 28) / 53)) < 1e-5


def test_luma_partition_distortion_closed_form():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        satd = int(rng.integers(0, 1 << 24))
        q = int(rng.integers(0, 256))
        sub = int(rng.integers(0, 2))
        qstep = float(np.float32(rng.uniform(0.0, 2000.0)))
        want = satd * max(qstep, 1.0) * 0.25 * O.luma_satd_scale(q, sub) * 2.775 / O.k_satd(q, sub) * 7.130643
        got = O.luma_partition_distortion(satd, qstep, q, sub)
        assert abs(got - want) <= 1e-6 * want + 1e-30
    assert O.luma_partition_distortion(0, 100.0, 50, 0) == 0.0
    assert O.luma_partition_distortion(1000, 0.25, 50, 1) == O.luma_partition_distortion(1000, 1.0, 50, 1)
    assert O.proxy_wsse(2.0) == np.float32(5.55)


def _k(fmt, q, mono_large=0):
    k = 0.5 if fmt == 2 else 0.4
    if fmt == 1 or mono_large:
        k *= 1.0 - min(max((20 - q) / 12, 0.0), 1.0)
    if fmt not in (1, 3):
        t = min(max((q - 50) / 50, 0.0), 1.0)
        floor = 0.4 if fmt == 0 else 0.0
        k *= (1 - floor) * t + floor
    return k


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_perceptual_rd_scale_closed_form(fmt):
    rng = np.random.default_rng(fmt + 11)
    bd = 10
    w, h = 200, 136
    plane = (rng.integers(0, 1 << bd, (h, w)) * rng.uniform(0, 1, (h, w)) ** 3).astype(np.uint16)
    plane[64:128, 0:64] = 512
    for q in (10, 30, 50, 75, 100, 200):
        for ref_act in (0.0, 3.0):
            for _ in range(40):
                dim = int(rng.choice([4, 8, 16, 32, 64]))
                px = int(rng.integers(0, (w + dim - 1) // dim)) * dim
                py = int(rng.integers(0, (h + dim - 1) // dim)) * dim
                got = O.perceptual_rd_scale(plane, bd, q, fmt, ref_act, px, py, dim)
                k = _k(fmt, q)
                sb = plane[py & ~63:(py & ~63) + 64, px & ~63:(px & ~63) + 64]
                refa = 0.5 * ref_act + 0.5 * _act64(sb, bd)
                blk = plane[py:py + dim, px:px + dim]
                if k == 0 or refa <= 0 or blk.size == 0:
                    assert got == 1.0, (q, px, py, dim)
                    continue
                e = min(max(k * (_act64(blk, bd) - refa), -LN2), LN2)
                want = min(max(math.exp(e), 0.5), 1.5)
                assert abs(got - want) <= 2e-5 * want, (fmt, q, px, py, dim, got, want)
    # a flat superblock without AQ, and blocks off the tile, are unscaled
    assert O.perceptual_rd_scale(plane, bd, 100, fmt, 0.0, 0, 64, 16) == 1.0
    assert O.perceptual_rd_scale(plane, bd, 100, fmt, 3.0, w, 0, 16) == 1.0
