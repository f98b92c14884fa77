"""The oracle's deblocking filter (oracle/oracle_lf.cpp: limits, loop_filter_scalar and
filter_plane_impl of src/loopfilter.rs, chroma_deblock_geom / frame_deblock of
src/coder/frame_deblock.rs, loop_filter_levels of src/obu.rs:306), pinned against an independent
restatement of AV1's normative loop filter (tests/spec_lf.py):
  - the limits of every level and sharpness;
  - every filter width on windows built to reach each branch (filter mask, flat, flat2, high
    edge variance), at 8, 10 and 12 bits, including the reference's own mixed-edge pattern
    (src/loopfilter.rs:674-733);
  - whole frames on random quadtree partitions, luma and chroma at 4:2:0 and 4:4:4, partial
    superblocks and visible sizes below the coded size: the reference's per-superblock-row
    order (vertical then horizontal edges of each row) equals the specification's order (all
    vertical edges of the frame, then all horizontal ones);
  - hand-worked answers: a 4-sample step through the narrow filter, and the level law at the
    values where rounding half away from zero differs from rounding to even."""
import numpy as np
import pytest

import lfgen
import oracle_lib as O
import spec_lf as S


def test_limits_match_spec():
    for lvl in range(64):
        for sharp in range(8):
            assert O.lf_limits(lvl, sharp) == S.limits(lvl, sharp), (lvl, sharp)


def _windows(rng, bd, n):
    """Crossing lines of 16 samples (offsets -8..7) that reach every branch."""
    one = 1 << (bd - 8)
    mid = 1 << (bd - 1)
    out = []
    for t in range(n):
        kind = t % 6
        if kind == 0:    # flat both sides, small step: the wide filters
            a, b = rng.integers(-1, 2, 2) * one
            w = [mid + a + int(rng.integers(-1, 2)) * (one if rng.random() < 0.3 else 0) for _ in range(8)] + \
                [mid + b + int(rng.integers(-1, 2)) * (one if rng.random() < 0.3 else 0) for _ in range(8)]
        elif kind == 1:  # a step of a few units: narrow filter, with and without hev
            s = int(rng.integers(1, 12)) * one
            w = [mid + int(rng.integers(-3, 4)) * one for _ in range(8)] + \
                [mid + s + int(rng.integers(-3, 4)) * one for _ in range(8)]
        elif kind == 2:  # a ramp
            g = int(rng.integers(0, 4))
            w = [mid + (k - 8) * g * one for k in range(16)]
        elif kind == 3:  # random texture
            w = list(mid + rng.integers(-20, 21, 16) * one)
        elif kind == 4:  # strong edge: rejected
            w = [int(rng.integers(0, 8)) * one] * 8 + [(1 << bd) - 1 - int(rng.integers(0, 8)) * one] * 8
        else:            # flat inner, rough outer (flat2 fails)
            w = [mid + int(rng.integers(-30, 31)) * one for _ in range(4)] + [mid] * 8 + \
                [mid + int(rng.integers(-30, 31)) * one for _ in range(4)]
        out.append([int(min(max(v, 0), (1 << bd) - 1)) for v in w])
    return out


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_loop_filter_matches_spec(bd):
    rng = np.random.default_rng(100 + bd)
    wins = _windows(rng, bd, 600)
    for t, win in enumerate(wins):
        for wd in (4, 6, 8, 16):
            luma = wd != 6
            for lvl, sharp in ((int(rng.integers(1, 64)), int(rng.integers(0, 8))), (63, 0), (4, 7)):
                # four lines: the window and three shifted variants
                lines = [win, win[1:] + win[-1:], win[:1] + win[:-1], win[::-1]]
                px = np.array([v for ln in lines for v in ln], np.uint16)
                e, i, h = O.lf_limits(lvl, sharp)
                O.loop_filter(px, 8, e, i, h, 16, 1, wd, bd)
                for li, ln in enumerate(lines):
                    F = {k: ln[k + 8] for k in range(-8, 8)}
                    want = list(ln)
                    for k, v in S.filter_samples(F, wd, lvl, sharp, bd, luma).items():
                        want[k + 8] = v
                    got = [int(x) for x in px[li * 16:(li + 1) * 16]]
                    assert got == want, (bd, t, wd, lvl, sharp, li)


def _mixed_edge_lines(bd):
    """The reference's mixed-edge test pattern (src/loopfilter.rs:674-733): a flat wide-filter
    line, a rejected line, a filtered non-flat line without and with high edge variance."""
    s, mid = 1 << (bd - 8), 1 << (bd - 1)
    flat = [mid, mid + s, mid - s, mid, mid + s, mid - s, mid, mid + s, mid, mid + 2 * s, mid + s, mid,
            mid + 2 * s, mid + s]
    rej = [0] * 7 + [(1 << bd) - 1] * 7
    ramp = [mid + (k - 7) * 2 * s for k in range(14)]
    hev = [mid] * 14
    hev[5], hev[7], hev[8] = mid - 8 * s, mid + 2 * s, mid + 10 * s
    return [flat, rej, ramp, hev]


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_mixed_edge_pattern_matches_spec(bd):
    for vertical in (False, True):
        for wd in (4, 6, 8, 16):
            for lvl in (1, 16, 32, 63):
                lines = _mixed_edge_lines(bd)
                W = 48
                img = np.full((W, W), 1 << (bd - 1), np.uint16)
                for li, ln in enumerate(lines):
                    for k, v in enumerate(ln):
                        y, x = (20 + li, 20 + k - 7) if vertical else (20 + k - 7, 20 + li)
                        img[y, x] = v
                flat = img.reshape(-1).copy()
                e, i, h = O.lf_limits(lvl, 0)
                if vertical:
                    O.loop_filter(flat, 20 * W + 20, e, i, h, W, 1, wd, bd)
                else:
                    O.loop_filter(flat, 20 * W + 20, e, i, h, 1, W, wd, bd)
                got = flat.reshape(W, W)
                want = img.astype(np.int64).copy()
                for li in range(4):
                    get = (lambda k: int(img[20 + li, 20 + k])) if vertical else (lambda k: int(img[20 + k, 20 + li]))
                    F = {k: get(k) for k in range(-8, 8)}
                    for k, v in S.filter_samples(F, wd, lvl, 0, bd, wd != 6).items():
                        if vertical:
                            want[20 + li, 20 + k] = v
                        else:
                            want[20 + k, 20 + li] = v
                assert (got == want).all(), (bd, vertical, wd, lvl)


def test_narrow_step_known_answer():
    # [100, 100 | 104, 104] at level 10, sharpness 0 (E 34, I 10, H 0), 8-bit, width 4:
    # filter = 3 * 4 = 12; f1 = 16 >> 3 = 2, f2 = 15 >> 3 = 1; p0 101, q0 102; no high edge
    # variance, so (f1 + 1) >> 1 = 1 moves p1 to 101 and q1 to 103.
    px = np.array([100, 100, 104, 104] * 4, np.uint16)
    e, i, h = O.lf_limits(10, 0)
    assert (e, i, h) == (34, 10, 0)
    O.loop_filter(px, 2, e, i, h, 4, 1, 4, 8)
    assert list(px[:4]) == [101, 101, 102, 103]


def test_level_law_known_answers():
    # (q / 8) rounded for luma, clamp 40; chroma (q / 10) * 0.5 at 4:2:0 (1.0 otherwise),
    # rounded half away from zero, at least 1 while luma filters, at most 32; sharpness 7
    assert O.lf_levels(35, 2) == (4, 2, 7)      # 3 * 0.5 = 1.5 -> 2
    assert O.lf_levels(50, 2) == (6, 3, 7)      # 5 * 0.5 = 2.5 -> 3 (not 2)
    assert O.lf_levels(7, 2) == (0, 0, 7)       # luma off: chroma off
    assert O.lf_levels(8, 2) == (1, 1, 7)       # chroma floors at 1
    assert O.lf_levels(255, 2) == (31, 13, 7)   # 25 * 0.5 = 12.5 -> 13
    assert O.lf_levels(255, 0) == (31, 25, 7)
    assert O.lf_levels(91, 0) == (11, 9, 7)


def _spec_frame(planes, leaves, w8, h8, disp_w, disp_h, sub, sharp, bd, lvls):
    out = []
    for pl, px in enumerate(planes):
        if px is None:
            out.append(None)
            continue
        ph, pw = px.shape
        rows = [list(map(int, r)) for r in px]
        grid = S.tx_grid(leaves, w8, h8, pl, sub)
        if pl == 0 or (lvls[0] > 0 and (lvls[1] > 0 or lvls[2] > 0)):
            S.deblock_plane(rows, pw, ph, grid, disp_w, disp_h, 0 if pl == 0 else sub, lvls[pl], sharp, bd, pl == 0)
        out.append(np.array(rows, np.uint16))
    return out


@pytest.mark.parametrize("seed,w8,h8,sub,bd", [(1, 96, 80, 1, 8), (2, 136, 72, 0, 10), (3, 64, 128, 1, 12),
                                               (4, 72, 96, 0, 8), (5, 128, 64, 1, 10)])
def test_frame_matches_spec_order(seed, w8, h8, sub, bd):
    rng = np.random.default_rng(seed)
    leaves = lfgen.random_leaves(rng, w8, h8)
    maps = lfgen.maps_from_leaves(leaves, w8, h8)
    y = lfgen.blocky_plane(rng, leaves, w8, h8, bd)
    u = lfgen.blocky_plane(rng, leaves, w8 >> sub, h8 >> sub, bd, sub)
    v = lfgen.blocky_plane(rng, leaves, w8 >> sub, h8 >> sub, bd, sub)
    disp_w = w8 - int(rng.integers(0, 8))
    disp_h = h8 - int(rng.integers(0, 8))
    for lvls, sharp in (((int(rng.integers(8, 40)), int(rng.integers(1, 20)), int(rng.integers(1, 20))), 7),
                        ((63, 32, 0), 0), ((5, 0, 0), 3)):
        got = [p.copy() for p in (y, u, v)]
        O.frame_deblock(got[0], got[1], got[2], disp_w, disp_h, maps, sub, sub, sharp, bd, *lvls)
        want = _spec_frame([y, u, v], leaves, w8, h8, disp_w, disp_h, sub, sharp, bd, lvls)
        for pl in range(3):
            assert (got[pl] == want[pl]).all(), (pl, lvls, sharp)
        assert (got[0] != y).any()  # the filter did something


def test_chroma_geometry_from_prediction_map():
    rng = np.random.default_rng(9)
    w8, h8 = 128, 64
    leaves = lfgen.random_leaves(rng, w8, h8)
    m = lfgen.maps_from_leaves(leaves, w8, h8)
    for sub in (0, 1):
        cbw, cbh = O.chroma_deblock_geom(m["pblk4"], m["pblk4h"], sub, sub)
        size, _, _ = S.tx_grid(leaves, w8, h8, 1, sub)
        assert (cbw == np.array(size) // 4).all() and (cbh == cbw).all(), sub


def test_level_zero_and_flat_are_fixed_points():
    rng = np.random.default_rng(3)
    w8, h8 = 64, 64
    leaves = lfgen.random_leaves(rng, w8, h8)
    m = lfgen.maps_from_leaves(leaves, w8, h8)
    y = lfgen.blocky_plane(rng, leaves, w8, h8, 8)
    got = y.copy()
    O.filter_plane(got, w8, h8, m["blk4"], m["blk4h"], m["blk4v"], m["blk4t"], 0, 7, 1, 16, 8)
    assert (got == y).all()
    flat = np.full((h8, w8), 77, np.uint16)
    O.filter_plane(flat, w8, h8, m["blk4"], m["blk4h"], m["blk4v"], m["blk4t"], 63, 0, 1, 16, 8)
    assert (flat == 77).all()
