"""Pins of the oracle's coefficient rate model (src/rate.rs, src/cost.rs) — CPU only."""
import math

import numpy as np

import oracle_lib as O


def test_cost_table_is_q22_minus_log2():
    """cost_q_table (src/cost.rs:159): round(-log2(p/32768) * 2^22), within one Q22 unit of the
    f64 value (the table is built from f32 log2)."""
    O.scan(1)  # binds the rate signatures
    for p in list(range(1, 300)) + list(range(32000, 32769)) + [1024, 4096, 16384]:
        exact = -math.log2(p / 32768.0) * (1 << 22)
        assert abs(O.lib().orc_cost_q(p) - exact) <= 2.0 + exact * 2e-7, p
    assert O.lib().orc_cost_q(32768) == 0 and O.lib().orc_cost_q(16384) == 1 << 22


def test_all_zero_is_the_skip_flag():
    """all_zero_is_the_skip_flag (src/rate.rs:320)."""
    c = O.rate_ctx()
    assert O.block_rate(np.zeros(64, np.int32), O.scan(1), c) == O.skip_cost(0, 1, 0, 1)


def test_far_eob_costs_more_than_dc():
    """far_eob_costs_more_than_dc (src/rate.rs:332)."""
    c = O.rate_ctx()
    s8 = O.scan(1)
    dc = np.zeros(64, np.int32)
    dc[0] = 3
    far = np.zeros(64, np.int32)
    far[s8[40]] = 3
    assert O.block_rate(far, s8, c) > O.block_rate(dc, s8, c) + 4.0


def test_clustered_beats_scattered():
    """clustered_beats_scattered (src/rate.rs:353)."""
    c = O.rate_ctx()
    s8 = O.scan(1)
    a = np.zeros(64, np.int32)
    b = np.zeros(64, np.int32)
    for i in (0, 1, 2, 3):
        a[s8[i]] = 2
    for i in (0, 12, 24, 36):
        b[s8[i]] = 2
    assert O.block_rate(a, s8, c) < O.block_rate(b, s8, c)


def test_rectangular_orientation_goldens():
    """rectangular_orientations_use_their_own_contexts (src/rate.rs:371-400): the reference's
    golden rates 38.532284 (4x8) and 38.2208 (8x4) for Cdfs::new(0), chroma, skip ctx 7."""
    s48, s84 = O.scan(5), O.scan(6)
    a = np.zeros(32, np.int32)
    b = np.zeros(32, np.int32)
    for i, lvl in ((0, 3), (1, -2), (4, 1), (17, 4)):
        a[s48[i]] = lvl
        b[s84[i]] = lvl
    ca = O.rate_ctx(cls=1, plane=1, w=4, h=8, eob_bin=4, skip_ctx=7)
    cb = O.rate_ctx(cls=1, plane=1, w=8, h=4, eob_bin=4, skip_ctx=7)
    ar, br = O.block_rate(a, s48, ca), O.block_rate(b, s84, cb)
    assert abs(ar - 38.532284) < 1e-5, ar
    assert abs(br - 38.2208) < 1e-5, br
    assert ar != br


def test_bound_is_exact_early_out():
    """real_block_bits_bounded (src/rate.rs:149): either the exact rate (<= bound) or +inf, and
    +inf only when the exact rate exceeds the bound."""
    rng = np.random.default_rng(1)
    s = O.scan(2)
    c = O.rate_ctx(cls=2, w=16, h=16, eob_bin=11)
    for _ in range(200):
        cf = rng.integers(-20, 21, 256).astype(np.int32)
        cf[rng.random(256) < 0.7] = 0
        full = O.block_rate(cf, s, c)
        bound = float(rng.uniform(0, full * 1.5))
        got = O.block_rate(cf, s, c, bound)
        if math.isinf(got):
            assert full > bound
        else:
            assert got == full


def test_scans_are_permutations_in_diagonal_order():
    """Every scan is a permutation of the block (src/tables.rs:348-455, src/coef_q.rs), and
    its first entries start at DC."""
    for s in (0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3):
        sc = O.scan(s)
        assert sorted(sc.tolist()) == list(range(O.ncoef(s)))
        assert sc[0] == 0


# ---- decision CDF snapshots ------------------------------------------------------------
def test_default_snapshot_reloaded_prices_like_its_slot():
    """Cdfs::new(2) loaded into slot 5 prices every block exactly as slot 2 does."""
    O.rate_load_cdfs(5, O.rate_default_cdfs(2))
    rng = np.random.default_rng(3)
    s8 = O.scan(1)
    for _ in range(200):
        cf = np.where(rng.random(64) < 0.3, rng.integers(-40, 41, 64), 0).astype(np.int32)
        a = O.block_rate(cf, s8, O.rate_ctx(qctx=2))
        b = O.block_rate(cf, s8, O.rate_ctx(qctx=5))
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)


def test_snapshot_skip_symbol_is_priced_from_the_loaded_cdf():
    """An all-zero 8x8 block costs its txb_skip symbol: -log2((32768 - c) / 32768) for the
    loaded skip CDF value c (the reference's cost table, Q22)."""
    snap = O.rate_default_cdfs(0)
    for c in (1024, 16384, 30000):
        snap["skip"][1, 0] = c
        O.rate_load_cdfs(6, snap)
        got = O.block_rate(np.zeros(64, np.int32), O.scan(1), O.rate_ctx(qctx=6))
        want = -math.log2((32768 - c) / 32768.0)
        assert abs(got - want) < 1e-5, (c, got, want)


def test_library_default_snapshots_match_oracle():
    import maroontree_amd as mt
    for q in range(4):
        a, b = mt.rate_default_cdfs(q), O.rate_default_cdfs(q)
        assert a.tobytes() == b.tobytes()
