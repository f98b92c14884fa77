"""The oracle's trellis quantization (oracle/oracle_rate.cpp: trellis_optimize and
trellis_optimize_ctx of src/trellis.rs on the shipped switches):
  - the reference's own validity test (src/trellis.rs:1076-1101, its xorshift generator and
    arguments): levels only fall toward zero, never change sign, and nothing survives past
    the new end of block; for both variants;
  - limits: a vanishing lambda keeps the rounded levels (every change costs distortion), a
    huge one zeroes the block;
  - trellis_optimize's end-of-block choice is the brute-force argmin, in f64, of its cost
    over every candidate end of block and the all-zero block, given the round-down it made
    (itself checked against the per-coefficient rule); near-ties are skipped;
  - the round-down's x86 form: fused multiply-adds on chunks of 8 / 4 scan positions whose
    levels are all below 64, the unfused scalar form otherwise — a level of 64 in a chunk
    switches the whole chunk."""
import math

import numpy as np
import pytest

import oracle_lib as O


def xorshift(seed):
    s = [seed]

    def nxt():
        x = s[0]
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        s[0] = x
        return x
    return nxt


def random_block(rng, n=256):
    """random_block of src/trellis.rs:1063: 1..40 positions of magnitude 0..6 plus a fraction,
    levels the rounded targets."""
    tf = np.zeros(n, np.float32)
    cf = np.zeros(n, np.int32)
    for _ in range(1 + rng() % 40):
        rc = rng() % n
        mag = np.float32(rng() % 6) + np.float32(np.float32(rng()) / np.float32(18446744073709551615.0))
        sign = np.float32(1.0 if rng() & 1 == 0 else -1.0)
        tf[rc] = sign * mag
        v = float(sign * mag)
        cf[rc] = int(math.copysign(math.floor(abs(v) + 0.5), v))
    return tf, cf


def test_trellis_output_is_valid_reference_case():
    rng = xorshift(0x2545F4914F6CDD1D)
    scan = O.scan(2)  # SCAN_16X16
    ctx = O.rate_ctx(qctx=0, cls=2, plane=0, w=16, h=16, eob_bin=11, dcs_ctx=0)
    changed = 0
    for _ in range(1000):
        tf, cf = random_block(rng)
        for opt in (O.trellis_ctx(cf, tf, 200.0, 220.0, scan, 0.05, ctx, 15),
                    O.trellis(cf, tf, 200.0, 220.0, scan, 0.05)):
            nz = np.nonzero(opt[scan])[0]
            if len(nz):
                assert not opt[scan[nz[-1] + 1:]].any()
            assert (np.abs(opt) <= np.abs(cf)).all()
            assert ((opt == 0) | (np.sign(opt) == np.sign(cf))).all()
            changed += int((opt != cf).any())
    assert changed > 500  # the trellis does something on these blocks


@pytest.mark.parametrize("size,cls,eob_bin", [(0, 0, 0), (1, 1, 5), (2, 2, 11), (3, 3, 15)])
def test_limits_of_lambda(size, cls, eob_bin):
    n = O.ncoef(size)
    w = h = int(math.isqrt(n))
    scan = O.scan(size)
    ctx = O.rate_ctx(qctx=1, cls=cls, plane=0, w=w, h=h, eob_bin=eob_bin)
    rng = np.random.default_rng(size)
    for _ in range(60):
        tf = (rng.standard_normal(n) * rng.choice([0.5, 3, 20])).astype(np.float32)
        frac = np.abs(np.abs(tf) % 1.0 - 0.5)
        tf[frac < 1e-3] += np.float32(0.01)  # no rounding ties
        cf = (np.sign(tf) * np.floor(np.abs(tf) + 0.5)).astype(np.int32)
        for f in (lambda c, t, l: O.trellis(c, t, 60.0, 70.0, scan, l),
                  lambda c, t, l: O.trellis_ctx(c, t, 60.0, 70.0, scan, l, ctx, 15)):
            assert (f(cf, tf, 1e-12) == cf).all()
            assert not f(cf, tf, 1e9).any()


def _lambda(dc_q, lambda0):
    return 0.045025873597302174 * 0.5525 * 1.15 * dc_q * dc_q * (3.3 + 0.0015 * dc_q) * (lambda0 / 0.05)


def _rate_bits(l):
    if l < 64:
        return [0.9, 2.7, 3.6][l] if l < 3 else 4.0 + (min(l - 3, 12) // 3 + 1) * 1.3 + (
            2 * (l - 14).bit_length() - 1 if l >= 15 else 0)
    return 4.0 + 5 * 1.3 + 2 * (l - 14).bit_length() - 1


@pytest.mark.parametrize("size", [0, 1, 2, 3])
def test_eob_choice_is_the_bruteforce_argmin(size):
    n = O.ncoef(size)
    scan = O.scan(size)
    rng = np.random.default_rng(40 + size)
    dc_q, ac_q, lam0 = 52.0, 61.0, 0.05
    lam = _lambda(dc_q, lam0)
    checked = 0
    for _ in range(300):
        tf = np.zeros(n, np.float32)
        k = int(rng.integers(1, min(n, 24)))
        pos = scan[:int(rng.integers(k, n + 1))]
        pick = rng.choice(pos, k, replace=False)
        tf[pick] = (rng.standard_normal(k) * rng.choice([0.8, 2, 6])).astype(np.float32)
        cf = (np.sign(tf) * np.floor(np.abs(tf.astype(np.float64)) + 0.5)).astype(np.int32)
        out = O.trellis(cf, tf, dc_q, ac_q, scan, lam0)
        # the round-down the oracle made: each level kept or lowered by one, per the local rule
        nzs = np.nonzero(cf[scan])[0]
        if len(nzs) == 0:
            continue
        m = nzs[-1] + 1
        q2 = np.where(scan[:m] == 0, dc_q * dc_q, ac_q * ac_q)
        at = np.abs(tf[scan[:m]]).astype(np.float64)
        l = np.abs(cf[scan[:m]]).astype(np.int64)
        rd = l.copy()
        tie = False
        for i in range(m):
            if l[i] == 0:
                continue
            cl = q2[i] * (at[i] - l[i]) ** 2 + lam * _rate_bits(int(l[i]))
            cd = q2[i] * (at[i] - l[i] + 1) ** 2 + lam * _rate_bits(int(l[i]) - 1)
            tie |= abs(cl - cd) < 1e-3 * max(cl, 1.0)
            if cd < cl:
                rd[i] -= 1
        if tie:
            continue
        dcur = q2 * (at - rd) ** 2
        dzero = q2 * at ** 2
        costs = {}
        for e in range(m):
            if rd[e] == 0:
                continue
            b = e if e < 2 else int(e).bit_length()
            sig = b * 0.9 + ((b - 2 if b > 1 else 0) + 2.0)
            pre = sum(dcur[i] + lam * _rate_bits(int(rd[i])) for i in range(e))
            costs[e] = pre + dcur[e] + lam * sig + dzero[e + 1:].sum()
        skip = dzero.sum() + lam
        best = min(costs.values()) if costs else math.inf
        vals = sorted(list(costs.values()) + [skip])
        if len(vals) > 1 and vals[1] - vals[0] < 1e-3 * max(vals[0], 1.0):
            continue
        if skip < best:
            want = np.zeros(m, np.int64)
        else:
            e = min(costs, key=costs.get)
            want = rd.copy()
            want[e + 1:] = 0
        assert (np.abs(out[scan[:m]]) == want).all()
        checked += 1
    assert checked > 150


def test_round_down_chunks_switch_to_the_scalar_form_at_64():
    """A near-tie target at scan index 1 where the fused (one rounding) and unfused costs of
    level 2 vs 1 decide differently (found by a search with an independent numpy emulation
    of both roundings): the oracle takes the fused decision while every level of the chunk
    of 8 is below 64, and the unfused one once scan index 7 holds a level of 70."""
    scan = O.scan(1)  # 8x8
    dc_q = ac_q = 40.0
    f = np.float32
    aom = f(f(f(f(f(0.045025873597302174) * f(0.5525)) * f(1.15)) * f(dc_q)) * f(dc_q))
    aom = f(aom * f(f(3.3) + f(f(0.0015) * f(dc_q))))
    q2 = f(f(ac_q) * f(ac_q))
    r1, r2 = f(2.7), f(3.6)

    def decide(t, lam, fused):
        el, edn = f(t - f(2.0)), f(t - f(1.0))
        if fused:
            cl = f(float(q2) * float(f(el * el)) + float(f(lam * r2)))
            cd = f(float(q2) * float(f(edn * edn)) + float(f(lam * r1)))
        else:
            cl = f(f(q2 * f(el * el)) + f(lam * r2))
            cd = f(f(q2 * f(edn * edn)) + f(lam * r1))
        return cd < cl
    found = None
    lam0 = f(0.05)
    for _ in range(20000):
        lam0 = np.nextafter(lam0, f(1.0))
        lam = f(aom * f(lam0 / f(0.05)))
        t = f(1.5 + float(lam) * 0.9 / (2 * float(q2)))
        for _k in range(3):
            t = np.nextafter(t, f(0.0))
        for _k in range(7):
            if decide(t, lam, True) != decide(t, lam, False):
                found = (t, lam0, lam)
                break
            t = np.nextafter(t, f(3.0))
        if found:
            break
    assert found is not None
    t, lam0, lam = found
    tf = np.zeros(64, np.float32)
    tf[scan[1]] = t
    tf[scan[8:16]] = 30.0  # keeps the end of block past the chunk
    cf = np.floor(np.abs(tf.astype(np.float64)) + 0.5).astype(np.int32)
    assert cf[scan[1]] == 2
    for lvl7, fused in ((3, True), (70, False)):
        c, tt = cf.copy(), tf.copy()
        c[scan[7]] = lvl7
        tt[scan[7]] = np.float32(lvl7 + 0.2)
        out = O.trellis(c, tt, dc_q, ac_q, scan, float(lam0))
        assert out[scan[1]] == (1 if decide(t, lam, fused) else 2), (lvl7, fused)
