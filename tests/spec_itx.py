"""TEST INFRASTRUCTURE ONLY: an independent restatement of AV1's normative dequantization and
inverse transform process, used to pin the oracle's transliteration of the reference's
dav1d-style inverse (src/idct.rs) bit for bit.  Nothing here is product code.

Written from the AV1 Bitstream & Decoding Process Specification, not from the reference:
  - 7.12.3 dequantization: dq = (|level| * q) & 0xFFFFFF, >> dqDenom (1 for 512- and
    1024-sample blocks, 2 above), sign, clip to 8 + BitDepth bits; with a quantizer matrix
    q = Round2(q * Quantizer_Matrix[level][plane][Qm_Offset[txSz] + pos], 5) for the 2-D
    (non-identity) types;
  - 7.13.2.2 the butterfly B(a, b, angle, flip) and Hadamard H(a, b, flip);
  - 7.13.2.3 the inverse DCT array permutation (bit reversal) and the inverse DCT lattice;
  - 7.13.2.5-7.13.2.7 the inverse ADST4 (sinpi), ADST8/16 input / output permutations and
    lattices;
  - 7.13.2.15 the identity transforms (5793/4096, 2, 11586/4096, 4);
  - 7.13.3 the 2-D process: the rect2 prescale Round2(T * 2896, 12), the row transform,
    Round2 by Transform_Row_Shift, the column clamp to Max(BitDepth + 6, 16) bits, the column
    transform, Round2 by 4.  Rows and columns past 32 of a 64-point transform are zero.

The specification leaves the result undefined when a value stored during the row transforms
needs more than 8 + BitDepth bits, or one stored during the column transforms more than
Max(BitDepth + 6, 16) bits (a conformance requirement, not a clamp).  Every call returns,
per block, whether it stayed inside those ranges; only such blocks pin anything.
All arithmetic is int64 numpy over a batch of blocks (one block per row).
"""
import math

import numpy as np

COS128 = np.array([int(round(4096.0 * math.cos(i * math.pi / 128.0))) for i in range(65)], np.int64)
SINPI = (1321, 2482, 3344, 3803)  # round(4096 * 2 * sqrt(2) / 3 * sin(k * pi / 9)), k = 1..4


def cos128(angle):
    a = angle & 255
    if a <= 64:
        return int(COS128[a])
    if a <= 128:
        return -int(COS128[128 - a])
    if a <= 192:
        return -int(COS128[a - 128])
    return int(COS128[256 - a])


def sin128(angle):
    return cos128(angle - 64)


def round2(x, n):
    return x if n == 0 else (x + (1 << (n - 1))) >> n


def brev(nbits, x):
    r = 0
    for i in range(nbits):
        r |= ((x >> i) & 1) << (nbits - 1 - i)
    return r


class Lattice:
    """Runs B / H steps on T (N, n) int64 and records each block's largest stored |value|."""

    def __init__(self, T):
        self.T = T
        self.peak = np.zeros(T.shape[0], np.int64)

    def _note(self, *cols):
        for c in cols:
            self.peak = np.maximum(self.peak, np.abs(self.T[:, c]))

    def B(self, a, b, angle, flip):
        T = self.T
        c, s = cos128(angle), sin128(angle)
        x = T[:, a] * c - T[:, b] * s
        y = T[:, a] * s + T[:, b] * c
        T[:, a] = round2(x, 12)
        T[:, b] = round2(y, 12)
        if flip:
            T[:, [a, b]] = T[:, [b, a]]
        self._note(a, b)

    def H(self, a, b, flip):
        if flip:
            a, b = b, a
        T = self.T
        x, y = T[:, a].copy(), T[:, b].copy()
        T[:, a] = x + y
        T[:, b] = x - y
        self._note(a, b)


def idct(T, n):
    """7.13.2.3: inverse DCT of length 2^n on the rows of T (in place); returns the peak."""
    L = Lattice(T)
    N = 1 << n
    T[:, :] = T[:, [brev(n, i) for i in range(N)]]
    B, H = L.B, L.H
    if n == 6:
        for i in range(16):
            B(32 + i, 63 - i, 63 - 4 * brev(4, i), 0)
    if n >= 5:
        for i in range(8):
            B(16 + i, 31 - i, 6 + (brev(3, 7 - i) << 3), 0)
    if n == 6:
        for i in range(16):
            H(32 + i * 2, 33 + i * 2, i & 1)
    if n >= 4:
        for i in range(4):
            B(8 + i, 15 - i, 12 + (brev(2, 3 - i) << 4), 0)
    if n >= 5:
        for i in range(8):
            H(16 + 2 * i, 17 + 2 * i, i & 1)
    if n == 6:
        for i in range(4):
            for j in range(2):
                B(62 - i * 4 - j, 33 + i * 4 + j, 60 - 16 * brev(2, i) + 64 * j, 1)
    if n >= 3:
        for i in range(2):
            B(4 + i, 7 - i, 56 - 32 * i, 0)
    if n >= 4:
        for i in range(4):
            H(8 + 2 * i, 9 + 2 * i, i & 1)
    if n >= 5:
        for i in range(2):
            for j in range(2):
                B(30 - 4 * i - j, 17 + 4 * i + j, 24 + (j << 6) + ((1 - i) << 5), 1)
    if n == 6:
        for i in range(8):
            for j in range(2):
                H(32 + i * 4 + j, 35 + i * 4 - j, i & 1)
    for i in range(2):
        B(2 * i, 2 * i + 1, 32 + 16 * i, 1 - i)
    if n >= 3:
        for i in range(2):
            H(4 + 2 * i, 5 + 2 * i, i)
    if n >= 4:
        for i in range(2):
            B(14 - i, 9 + i, 48 + 64 * i, 1)
    if n >= 5:
        for i in range(4):
            for j in range(2):
                H(16 + 4 * i + j, 19 + 4 * i - j, i & 1)
    if n == 6:
        for i in range(2):
            for j in range(4):
                B(61 - i * 8 - j, 34 + i * 8 + j, 56 - i * 32 + (j >> 1) * 64, 1)
    for i in range(2):
        H(i, 3 - i, 0)
    if n >= 3:
        B(6, 5, 32, 1)
    if n >= 4:
        for i in range(2):
            for j in range(2):
                H(8 + 4 * i + j, 11 + 4 * i - j, i)
    if n >= 5:
        for i in range(4):
            B(29 - i, 18 + i, 48 + (i >> 1) * 64, 1)
    if n == 6:
        for i in range(4):
            for j in range(4):
                H(32 + 8 * i + j, 39 + 8 * i - j, i & 1)
    if n >= 3:
        for i in range(4):
            H(i, 7 - i, 0)
    if n >= 4:
        for i in range(2):
            B(13 - i, 10 + i, 32, 1)
    if n >= 5:
        for i in range(2):
            for j in range(4):
                H(16 + i * 8 + j, 23 + i * 8 - j, i)
    if n == 6:
        for i in range(8):
            B(59 - i, 36 + i, 48 if i < 4 else 112, 1)
    if n >= 4:
        for i in range(8):
            H(i, 15 - i, 0)
    if n >= 5:
        for i in range(4):
            B(27 - i, 20 + i, 32, 1)
    if n == 6:
        for i in range(8):
            H(32 + i, 47 - i, 0)
            H(48 + i, 63 - i, 1)
    if n >= 5:
        for i in range(16):
            H(i, 31 - i, 0)
    if n == 6:
        for i in range(8):
            B(55 - i, 40 + i, 32, 1)
    if n == 6:
        for i in range(32):
            H(i, 63 - i, 0)
    return L.peak


def _adst_in_perm(n):
    N = 1 << n
    return [(i - 1) if (i & 1) else (N - i - 1) for i in range(N)]


def _adst_out(T, n):
    N = 1 << n
    src = T.copy()
    for i in range(N):
        a = (i >> 3) & 1
        b = ((i >> 2) & 1) ^ ((i >> 3) & 1)
        c = ((i >> 1) & 1) ^ ((i >> 2) & 1)
        d = (i & 1) ^ ((i >> 1) & 1)
        idx = ((d << 3) | (c << 2) | (b << 1) | a) >> (4 - n)
        T[:, i] = -src[:, idx] if (i & 1) else src[:, idx]


def iadst4(T):
    """7.13.2.6: the sinpi form; returns the peak of the stored values."""
    x = [T[:, i].copy() for i in range(4)]
    s0 = SINPI[0] * x[0]
    s1 = SINPI[1] * x[0]
    s2 = SINPI[2] * x[1]
    s3 = SINPI[3] * x[2]
    s4 = SINPI[0] * x[2]
    s5 = SINPI[1] * x[3]
    s6 = SINPI[3] * x[3]
    a7 = x[0] - x[2]
    b7 = a7 + x[3]
    s0 = s0 + s3
    s1 = s1 - s4
    s3 = s2
    s2 = SINPI[2] * b7
    s0 = s0 + s5
    s1 = s1 - s6
    x0 = s0 + s3
    x1 = s1 + s3
    x2 = s2
    x3 = s0 + s1
    x3 = x3 - s3
    T[:, 0] = round2(x0, 12)
    T[:, 1] = round2(x1, 12)
    T[:, 2] = round2(x2, 12)
    T[:, 3] = round2(x3, 12)
    return np.abs(T).max(axis=1)


def iadst8(T):
    """7.13.2.7."""
    T[:, :] = T[:, _adst_in_perm(3)]
    L = Lattice(T)
    for i in range(4):
        L.B(2 * i, 2 * i + 1, 60 - 16 * i, 1)
    for i in range(4):
        L.H(i, 4 + i, 0)
    for i in range(2):
        L.B(4 + 3 * i, 5 + i, 48 - 32 * i, 1)
    for i in range(2):
        for j in range(2):
            L.H(4 * j + i, 2 + 4 * j + i, 0)
    for i in range(2):
        L.B(2 + 4 * i, 3 + 4 * i, 32, 1)
    _adst_out(T, 3)
    return L.peak


def iadst16(T):
    """7.13.2.8."""
    T[:, :] = T[:, _adst_in_perm(4)]
    L = Lattice(T)
    for i in range(8):
        L.B(2 * i, 2 * i + 1, 62 - 8 * i, 1)
    for i in range(8):
        L.H(i, 8 + i, 0)
    for i in range(2):
        L.B(8 + 2 * i, 9 + 2 * i, 56 - 32 * i, 1)
        L.B(13 + 2 * i, 12 + 2 * i, 8 + 32 * i, 1)
    for i in range(4):
        L.H(i, 4 + i, 0)
        L.H(8 + i, 12 + i, 0)
    for i in range(2):
        L.B(4 + 8 * i, 5 + 8 * i, 48, 1)
        L.B(7 + 8 * i, 6 + 8 * i, 16, 1)
    for i in range(2):
        for j in range(2):
            L.H(8 * j + i, 2 + 8 * j + i, 0)
            L.H(4 + 8 * j + i, 6 + 8 * j + i, 0)
    for i in range(4):
        L.B(2 + 4 * i, 3 + 4 * i, 32, 1)
    _adst_out(T, 4)
    return L.peak


def iidentity(T, n):
    """7.13.2.15."""
    if n == 2:
        T[:, :] = round2(T * 5793, 12)
    elif n == 3:
        T[:, :] = T * 2
    elif n == 4:
        T[:, :] = round2(T * 11586, 12)
    else:
        T[:, :] = T * 4
    return np.abs(T).max(axis=1)


DCT, ADST, IDENTITY = 0, 1, 2


def inverse_1d(T, kind, n):
    if kind == DCT:
        return idct(T, n)
    if kind == IDENTITY:
        return iidentity(T, n)
    return {2: iadst4, 3: iadst8, 4: iadst16}[n](T)


# AV1 TxSize numbering, and Transform_Row_Shift
TX_WH = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
         (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]
ROW_SHIFT = [0, 1, 2, 2, 2, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2]
QM_OFFSET = [0, 16, 80, 336, 336, 1360, 1392, 1424, 1552, 1680, 2192, 336, 336, 2704, 2768, 2832, 3088, 1680, 2192]
# 2-D types (AV1 TxType numbering): (vertical 1-D kind, horizontal 1-D kind)
TX_KINDS = {0: (DCT, DCT), 1: (ADST, DCT), 2: (DCT, ADST), 3: (ADST, ADST), 9: (IDENTITY, IDENTITY),
            10: (DCT, IDENTITY), 11: (IDENTITY, DCT)}


def dequant(levels_rm, W, H, dc_q, ac_q, bd, qm=None):
    """7.12.3 on row-major levels (N, H*W) -> (N, H*W) int64.  qm: optional (H*W,) weights
    of the block in row-major order (None: flat)."""
    n = W * H
    q = np.full(n, ac_q, np.int64)
    q[0] = dc_q
    if qm is not None:
        q = round2(q * qm.astype(np.int64), 5)
    denom = (1 if n > 256 else 0) + (1 if n > 1024 else 0)
    lv = levels_rm.astype(np.int64)
    mag = ((np.abs(lv) * q) & 0xFFFFFF) >> denom
    dq = np.where(lv < 0, -mag, mag)
    lim = 1 << (7 + bd)
    return np.clip(dq, -lim, lim - 1)


def inverse_2d(dq_rm, tx_size, tx_type, bd):
    """7.13.3 on dequantized row-major coefficients (N, H*W).  Returns (residual (N, H, W),
    conforming (N,) bool)."""
    W, H = TX_WH[tx_size]
    N = dq_rm.shape[0]
    colk, rowk = TX_KINDS[tx_type]
    lw, lh = int(math.log2(W)), int(math.log2(H))
    D = dq_rm.reshape(N, H, W).astype(np.int64)
    R = np.zeros((N, H, W), np.int64)
    row_lim = 1 << (bd + 7)
    col_bits = max(bd + 6, 16)
    col_lim = 1 << (col_bits - 1)
    ok = np.ones(N, bool)
    for i in range(H):
        T = np.zeros((N, W), np.int64)
        if i < 32:
            T[:, :min(W, 32)] = D[:, i, :min(W, 32)]
        if abs(lw - lh) == 1:
            T = round2(T * 2896, 12)
        ok &= np.abs(T).max(axis=1) < row_lim  # the row transform's input range
        T = np.clip(T, -row_lim, row_lim - 1)
        peak = inverse_1d(T, rowk, lw)
        ok &= peak < row_lim
        R[:, i, :] = round2(T, ROW_SHIFT[tx_size])
    R = np.clip(R, -col_lim, col_lim - 1)
    for j in range(W):
        T = R[:, :, j].copy()
        peak = inverse_1d(T, colk, lh)
        ok &= peak < col_lim
        R[:, :, j] = round2(T, 4)
    return R, ok
