"""The oracle's dequantization and inverse transforms (oracle/oracle_itx.cpp, a transliteration
of the reference's dav1d-style src/idct.rs) against an independent restatement of the AV1
specification's dequantization and inverse transform process (tests/spec_itx.py), bit for
bit:
  - every 1-D inverse (DCT-4..64, ADST-4/8/16) on random inputs;
  - every 2-D (size, type) of the reference at 8/10/12-bit, flat and QM quantizers, on the
    levels the forward transform produces from random residuals and on sparse random levels,
    wherever the specification defines the result (its intermediate-range conformance holds);
  - planted 1-LSB mistakes in copies of the oracle (rounding offsets, a butterfly constant, a
    shift) each make the comparison fail.
The specification restatement is itself checked against the closed-form DCT / DST
bases (impulse responses).

Quantizer matrices: the table the oracle uses is the reference's AV1_IQM, which stores
dav1d's runtime tables in the transposed coefficient order (the reference's x-major layout),
so the restatement reads the block's table region x-major.  The 4x4..32x32 matrices are
symmetric, so for square sizes the orientation is pinned; for the rectangular sizes it rests
on that layout statement (the reference's src/quant.rs:97-100)."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import spec_itx as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
TYPES = [0, 1, 2, 3, 9, 10, 11]


def _iqm():
    hdr = open(os.path.join(ROOT, "maroontree_amd", "csrc", "av1_tables.h")).read()
    s = hdr.index("MT_IQM")
    body = hdr[hdr.index("{", s) + 1:hdr.index("};", s)]
    return np.array([int(x) for x in re.findall(r"-?\d+", body)], np.int64).reshape(15, 2, 3344)


IQM = _iqm()


def test_spec_lattices_reproduce_the_closed_form_bases():
    A = 4096
    for n in range(2, 7):
        N = 1 << n
        T = np.eye(N, dtype=np.int64) * A
        S.idct(T, n)
        i, k = np.arange(N)[:, None], np.arange(N)[None, :]
        a = np.cos(np.pi * (2 * i + 1) * k / (2 * N))
        a[:, 0] /= math.sqrt(2)
        assert np.abs(T.T / A - a).max() < 1e-3, N
    for n, f in ((2, S.iadst4), (3, S.iadst8), (4, S.iadst16)):
        N = 1 << n
        T = np.eye(N, dtype=np.int64) * A
        f(T)
        i, k = np.arange(N)[:, None], np.arange(N)[None, :]
        if N == 4:
            a = np.sin(np.pi * (i + 1) * (2 * k + 1) / 9.0) * 2 * math.sqrt(2) / 3
        else:
            a = np.sin(np.pi * (2 * i + 1) * (2 * k + 1) / (4 * N))
        assert np.abs(T.T / A - a).max() < 1e-3, N


def test_square_quantizer_matrices_are_symmetric():
    for lev in range(15):
        for c in range(2):
            for off, n in ((0, 4), (16, 8), (80, 16), (336, 32)):
                M = IQM[lev, c, off:off + n * n].reshape(n, n)
                assert (M == M.T).all(), (lev, c, n)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_oracle_inverse_dct_1d_equals_spec(n):
    rng = np.random.default_rng(n)
    X = rng.integers(-4000, 4001, (3000, n)).astype(np.int64)
    if n == 64:
        X[:, 32:] = 0
    T = X.copy()
    S.idct(T, int(math.log2(n)))
    for r in range(len(X)):
        c = X[r].astype(np.int32)
        O.lib().orc_inv_dct1d(n, O.ptr(c, ctypes.c_int32), -(1 << 20), (1 << 20) - 1)
        assert (c == T[r]).all(), (n, r)


@pytest.mark.parametrize("n", [4, 8, 16])
def test_oracle_inverse_adst_1d_equals_spec(n):
    rng = np.random.default_rng(100 + n)
    X = rng.integers(-4000, 4001, (3000, n)).astype(np.int64)
    T = X.copy()
    {4: S.iadst4, 8: S.iadst8, 16: S.iadst16}[n](T)
    for r in range(len(X)):
        c = X[r].astype(np.int32)
        O.lib().orc_inv_adst1d(n, O.ptr(c, ctypes.c_int32), -(1 << 20), (1 << 20) - 1)
        assert (c == T[r]).all(), (n, r)


def _levels(rng, s, t, q, bd, n):
    """x-major levels: half the forward transform of random residuals (of several scales),
    half sparse random levels of a few magnitudes."""
    W, H = S.TX_WH[s]
    cw, ch = min(W, 32), min(H, 32)
    out = []
    for i in range(n):
        if i % 2 == 0:
            r = rng.integers(-(1 << bd) + 1, 1 << bd, W * H) // int(rng.choice([1, 3, 16, 64]))
            out.append(O.fwd(s, t, r, q)[0])
        else:  # magnitudes up to about 2^(bd+1) after dequantization
            top = max(1, (1 << (bd + 1)) // max(q.ac, 1))
            amp = int(rng.integers(1, top + 1))
            lv = rng.integers(-amp, amp + 1, cw * ch) * (rng.random(cw * ch) < 0.15)
            lv[0] = rng.integers(-top, top + 1)
            out.append(lv.astype(np.int32))
    return np.array(out, np.int32)


def spec_inverse(lv_xm, s, t, q, bd, qml):
    """Dequantization (7.12.3) and the 2-D inverse (7.13.3) of x-major levels."""
    W, H = S.TX_WH[s]
    cw, ch = min(W, 32), min(H, 32)
    n = lv_xm.shape[0]
    lv_rm = lv_xm.reshape(n, cw, ch).transpose(0, 2, 1).reshape(n, ch * cw).astype(np.int64)
    qq = np.full(cw * ch, q.ac, np.int64)
    qq[0] = q.dc
    if qml < 15 and t < 9:  # quantizer matrices apply to the 2-D types only
        region = IQM[qml, 0, S.QM_OFFSET[s]:S.QM_OFFSET[s] + cw * ch]
        qq = S.round2(qq * region.reshape(cw, ch).T.ravel(), 5)
    denom = (1 if W * H > 256 else 0) + (1 if W * H > 1024 else 0)
    mag = ((np.abs(lv_rm) * qq) & 0xFFFFFF) >> denom
    lim = 1 << (7 + bd)
    dq = np.clip(np.where(lv_rm < 0, -mag, mag), -lim, lim - 1)
    D = np.zeros((n, H, W), np.int64)
    D[:, :ch, :cw] = dq.reshape(n, ch, cw)
    return S.inverse_2d(D.reshape(n, H * W), s, t, bd)


def compare(inv, s, t, bd, qml, seed, n=60):
    """(conforming blocks, mismatching conforming blocks) of the oracle-style inverse `inv`
    against the specification."""
    rng = np.random.default_rng(seed)
    q = O.quant(int(rng.integers(0, 256)), bd, qml)
    lv = _levels(rng, s, t, q, bd, n)
    R, ok = spec_inverse(lv, s, t, q, bd, qml)
    W, H = S.TX_WH[s]
    bad = 0
    for r in range(n):
        if ok[r] and (inv(s, t, lv[r], q).reshape(H, W) != R[r]).any():
            bad += 1
    return int(ok.sum()), bad


@pytest.mark.parametrize("s", SIZES)
def test_oracle_2d_inverse_equals_spec(s):
    checked = 0
    for t in TYPES:
        if not O.lib().orc_tx_supported(s, t):
            continue
        for bd in (8, 10, 12):
            for qml in (15, 0, 7):
                ok, bad = compare(O.inv, s, t, bd, qml, seed=s * 1000 + t * 37 + bd + qml)
                assert bad == 0, (S.TX_WH[s], t, bd, qml, bad)
                assert ok >= 30, (S.TX_WH[s], t, bd, qml, ok)  # at least the forward transform's levels conform
                checked += ok
    assert checked > 0


# ---- planted mistakes -------------------------------------------------------------------
MUTANTS = {
    "dct4 even-half rounding": ("int32_t t0 = ((in0 + in2) * 181 + 128) >> 8;",
                                "int32_t t0 = ((in0 + in2) * 181 + 127) >> 8;", [(0, 0)]),
    "dct16 odd-half constant": ("t9a = ((t14 * 1567 - t9", "t9a = ((t14 * 1568 - t9", [(2, 0)]),
    "adst4 output 2 rounding": ("int32_t o2 = (209 * (in0 - in2 + in3) + 128) >> 8;",
                                "int32_t o2 = (209 * (in0 - in2 + in3) + 127) >> 8;", [(0, 3)]),
    "rect2 prescale rounding": ("rect2 ? (v * 181 + 128) >> 8 : v", "rect2 ? (v * 181 + 127) >> 8 : v", [(7, 0)]),
    "inter-pass shift without rounding": ("if (mid_shift) t = (t + (1 << (mid_shift - 1))) >> mid_shift;",
                                          "if (mid_shift) t = t >> mid_shift;", [(1, 0)]),
    "identity-16 rounding": ("return 2 * v + ((v * 1697 + 1024) >> 11);",
                             "return 2 * v + ((v * 1697 + 1023) >> 11);", [(8, 10)]),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("itx_mutants")
    src = open(os.path.join(ROOT, "oracle", "oracle_itx.cpp")).read()
    libs = {}
    for name, (old, new, _) in MUTANTS.items():
        assert src.count(old) >= 1, name
        path = d / (re.sub(r"\W+", "_", name) + ".cpp")
        path.write_text(src.replace(old, new, 1))
        so = str(path) + ".so"
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fwrapv", "-ffp-contract=off", "-shared",
                        "-I", os.path.join(ROOT, "oracle"), "-o", so, str(path),
                        os.path.join(ROOT, "oracle", "oracle_tx.cpp"),
                        os.path.join(ROOT, "oracle", "oracle_rate.cpp")], check=True)  # scans, trellis
        L = ctypes.CDLL(so)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.orc_inv_txfm.argtypes = [ctypes.c_int, ctypes.c_int, i32p, ctypes.POINTER(O.OrcQuant), i32p]
        libs[name] = L
    return libs


def _sweep_levels(s, n):
    """3n x-major level blocks: coefficients 0..3 (the first n blocks) or all coefficients
    (the next n) hold the level i + 1, sign alternating by block, a sweep of dequantized
    magnitudes over the rounding ties of every shift and prescale; then n blocks of levels
    chosen to land on those ties under an 8-step quantizer."""
    W, H = S.TX_WH[s]
    cw, ch = min(W, 32), min(H, 32)
    out = np.zeros((3 * n, cw * ch), np.int32)
    rng = np.random.default_rng(n)
    for i in range(n):
        out[i, :4] = (i + 1) * (1 if i % 2 == 0 else -1)
        out[n + i, :] = (i + 1) * (1 if i % 2 == 0 else -1)  # every coefficient
    # levels whose products with an 8-step land on v = 128 (mod 256) or 1024 (mod 2048): the
    # ties of the x * 181 / 256 and x * 1697 / 2048 roundings, at varied magnitudes
    odd = 2 * rng.integers(0, 3, (n, cw * ch)) + 1
    base = np.where(rng.random((n, cw * ch)) < 0.5, 16, 128)
    sign = np.where(rng.random((n, cw * ch)) < 0.5, -1, 1)
    out[2 * n:] = (base * odd * sign * (rng.random((n, cw * ch)) < 0.3)).astype(np.int32)
    return out


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(mutant_libs, name):
    L = mutant_libs[name]
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.orc_inv_dct1d.argtypes = L.orc_inv_adst1d.argtypes = [ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int]

    def inv(s, t, levels, q):
        levels = np.ascontiguousarray(levels, dtype=np.int32)
        out = np.zeros(O.tx_w(s) * O.tx_h(s), np.int32)
        assert L.orc_inv_txfm(s, t, O.ptr(levels, ctypes.c_int32), ctypes.byref(q), O.ptr(out, ctypes.c_int32)) == 0
        return out
    caught = 0
    rng = np.random.default_rng(11)
    for s, t in MUTANTS[name][2]:
        W, H = S.TX_WH[s]
        # the 1-D passes on their own (no later shift to hide a 1-LSB slip)
        for n, kind, fn in ((W, S.TX_KINDS[t][1], None), (H, S.TX_KINDS[t][0], None)):
            if kind == S.IDENTITY:
                continue
            X = rng.integers(-3000, 3001, (4000, n)).astype(np.int64)
            T = X.copy()
            S.inverse_1d(T, kind, int(math.log2(n)))
            f = L.orc_inv_dct1d if kind == S.DCT else L.orc_inv_adst1d
            for r in range(len(X)):
                c = X[r].astype(np.int32)
                f(n, O.ptr(c, ctypes.c_int32), -(1 << 20), (1 << 20) - 1)
                caught += int((c != T[r]).any())
        # 2-D: random forward levels and a sweep of level magnitudes over several quantizers
        for bd in (8, 10):
            caught += compare(inv, s, t, bd, 15, seed=7 + bd, n=40)[1]
            for qidx in (1, 40, 120):
                q = O.quant(qidx, bd, 15)
                lv = _sweep_levels(s, 600)
                R, ok = spec_inverse(lv, s, t, q, bd, 15)
                for r in np.nonzero(ok)[0]:
                    caught += int((inv(s, t, lv[r], q).reshape(H, W) != R[r]).any())
    assert caught > 0, name
