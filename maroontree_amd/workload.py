"""Synthetic workloads of the candidate evaluation: a seeded test image and the full luma
candidate set of its superblocks.

The image follows the reference benchmark's recipe: four linear gradients of random
orientation, three octaves of value noise (amplitude 40 at 8 bits), and 20 hard-edged
rectangles, axis-aligned or oblique, of uniform random colour; scaled to the bit depth.

The candidate set of one 64x64 superblock is every square block of the partition quadtree
(4x4 ... 64x64) under each of the eleven non-directional-plus-directional luma modes the
reference's intra search evaluates (ND_LUMA_MODES, src/intrapred.rs:1712) with DCT_DCT: the
work of one exhaustive luma mode decision per block.  Predictions read a synthetic
reconstruction plane (source plus small noise), so every candidate is independent.
"""
import numpy as np

from . import _lib

ND_LUMA_MODES = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 12]
SQUARE_TX = {4: 0, 8: 1, 16: 2, 32: 3, 64: 4}
EOB_BIN_LUMA = {16: 0, 64: 5, 256: 11, 1024: 15}


def _noise_octave(g, octave, height, width, ys, xs):
    """Bilinear value noise of the lattice g at cell size `octave` (g[iy, ix] gathered by
    repeating lattice values, which equals the direct gather)."""
    fy, fx = ys / octave, xs / octave
    iy, ix = fy.astype(int), fx.astype(int)
    ty, tx = fy - iy, fx - ix
    if g.ndim == 3:
        ty, tx = ty[..., None], tx[..., None]
    up = g.repeat(octave, axis=0).repeat(octave, axis=1)  # up[y, x] = g[y // octave, x // octave]
    g00, g01 = up[:height, :width], up[:height, octave:octave + width]
    g10, g11 = up[octave:octave + height, :width], up[octave:octave + height, octave:octave + width]
    return g00 * (1 - tx) * (1 - ty) + g01 * tx * (1 - ty) + g10 * (1 - tx) * ty + g11 * tx * ty


def _rect_mask(rng, width, height, xs, ys):
    """One rectangle of the recipe: centre, half sizes and (half the time) an oblique angle;
    returns (y0, x0, mask over its bounding box) or None when the box is empty."""
    cx, cy = rng.uniform(0, width), rng.uniform(0, height)
    hw, hh = rng.uniform(8, max(width / 6, 9)), rng.uniform(8, max(height / 6, 9))
    th = rng.uniform(0, np.pi) if rng.random() < 0.5 else 0.0
    r = np.hypot(hw, hh) + 2
    x0, x1 = max(int(cx - r), 0), min(int(cx + r) + 1, width)
    y0, y1 = max(int(cy - r), 0), min(int(cy + r) + 1, height)
    bx, by = xs[:, x0:x1], ys[y0:y1]
    u = (bx - cx) * np.cos(th) + (by - cy) * np.sin(th)
    v = -(bx - cx) * np.sin(th) + (by - cy) * np.cos(th)
    return y0, x0, (np.abs(u) <= hw) & (np.abs(v) <= hh)


def synth_image(seed, width, height, bd=10):
    """The seeded test image, uint16 (height, width) at bit depth bd."""
    rng = np.random.default_rng(seed)
    ys = np.arange(height, dtype=np.float64)[:, None]
    xs = np.arange(width, dtype=np.float64)[None, :]
    img = np.zeros((height, width))
    for _ in range(4):
        th = rng.uniform(0, 2 * np.pi)
        proj = xs * np.cos(th) + ys * np.sin(th)
        proj = (proj - proj.min()) / max(proj.max() - proj.min(), 1e-9)
        img += rng.uniform(20, 70) * proj
    for octave, amp in ((64, 40.0), (32, 20.0), (16, 10.0)):
        g = rng.uniform(-1, 1, (height // octave + 2, width // octave + 2))
        img += amp * _noise_octave(g, octave, height, width, ys, xs)
    for _ in range(20):
        y0, x0, m = _rect_mask(rng, width, height, xs, ys)
        img[y0:y0 + m.shape[0], x0:x0 + m.shape[1]][m] = rng.uniform(0, 255)
    img = np.clip(np.rint(img), 0, 255) * (1 << (bd - 8))
    return img.astype(np.uint16)


def synth_images(specs, threads=16):
    """synth_image over a list of (seed, width, height, bd) on a thread pool (numpy's array
    work releases the GIL); returns the images in order."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(lambda s: synth_image(*s), specs))


def synth_rgb(seed, width, height):
    """The same recipe on three 8-bit channels (rectangles of one shape get an RGB colour);
    returns uint8 (height, width, 3)."""
    rng = np.random.default_rng(seed)
    ys = np.arange(height, dtype=np.float64)[:, None]
    xs = np.arange(width, dtype=np.float64)[None, :]
    img = np.zeros((height, width, 3))
    for _ in range(4):
        th = rng.uniform(0, 2 * np.pi)
        proj = xs * np.cos(th) + ys * np.sin(th)
        proj = (proj - proj.min()) / max(proj.max() - proj.min(), 1e-9)
        img += rng.uniform(20, 70, 3) * proj[..., None]
    for octave, amp in ((64, 40.0), (32, 20.0), (16, 10.0)):
        g = rng.uniform(-1, 1, (height // octave + 2, width // octave + 2, 3))
        img += amp * _noise_octave(g, octave, height, width, ys, xs)
    for _ in range(20):
        y0, x0, m = _rect_mask(rng, width, height, xs, ys)
        img[y0:y0 + m.shape[0], x0:x0 + m.shape[1]][m] = rng.uniform(0, 255, 3)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def rgb_to_luma(rgb):
    """Full-range BT.601 luma (the matrix of the sRGB CICP's YCbCr), rounded, uint16."""
    r, g, b = (rgb[..., i].astype(np.float64) for i in range(3))
    return np.clip(np.rint(0.299 * r + 0.587 * g + 0.114 * b), 0, 255).astype(np.uint16)


def synth_yuv(seed, width, height, bd=8, sub=2):
    """The recipe's RGB image as full-range BT.601 Y, Cb, Cr at bit depth bd (8-bit values
    scaled by 2^(bd - 8)); sub 2: chroma 4:2:0 (2x2 means, rounded), 0: 4:4:4.  Returns
    three uint16 planes."""
    rgb = synth_rgb(seed, width, height).astype(np.float64)
    r, g, b = rgb[..., 0], rgb[..., 1], rgb[..., 2]
    y = 0.299 * r + 0.587 * g + 0.114 * b
    cb = 128.0 - 0.168736 * r - 0.331264 * g + 0.5 * b
    cr = 128.0 + 0.5 * r - 0.418688 * g - 0.081312 * b
    sc = float(1 << (bd - 8))
    mx = (1 << bd) - 1

    def q(p):
        return np.clip(np.rint(p * sc), 0, mx).astype(np.uint16)
    if sub == 2:
        cb = 0.25 * (cb[0::2, 0::2] + cb[1::2, 0::2] + cb[0::2, 1::2] + cb[1::2, 1::2])
        cr = 0.25 * (cr[0::2, 0::2] + cr[1::2, 0::2] + cr[0::2, 1::2] + cr[1::2, 1::2])
    return q(y), q(cb), q(cr)


def synth_yuvs(specs, threads=16):
    """synth_yuv over a list of (seed, width, height, bd, sub) on a thread pool."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(lambda s: synth_yuv(*s), specs))


def synth_recon(src, seed, bd=10, amp=6):
    """A stand-in for the decision-time reconstruction: the source plus small noise."""
    rng = np.random.default_rng(seed ^ 0x5EED)
    noise = rng.integers(-amp, amp + 1, src.shape) << max(bd - 8, 0)
    return np.clip(src.astype(np.int64) + noise, 0, (1 << bd) - 1).astype(np.uint16)


def sb_candidates(width, height, bd=10, qidx=91, qm_level=None, lam=64.0, wsse_scale=1.0):
    """Every (square block, luma mode) candidate of every full 64x64 superblock of a
    width x height luma plane, as CAND_DTYPE records; returns (records, pool elements).
    The order is the encoder's: superblocks in raster order, within one the blocks by size
    and then in raster order, and each block's modes one after another (the candidates of
    one block share its edges and source samples); the pool follows the same order."""
    bdi = {8: 0, 10: 1, 12: 2}[bd]
    if qm_level is None:
        qm_level = _lib.qm_level_law(qidx, 0)
    sbw, sbh = width // 64, height // 64
    ox, oy, bsz, mode, pf = [], [], [], [], []
    for bs in (4, 8, 16, 32, 64):
        n = 64 // bs
        for by in range(n):
            for bx in range(n):
                x, y = bx * bs, by * bs
                # top-right / bottom-left availability in quadtree decoding order: the row
                # above the superblock is decoded, and so is the right sibling's top edge for a
                # left child, but not the next superblock's; bottom-left only along the decoded
                # superblock to the left (the sample counts clip at the frame edge by themselves)
                tr = (y == 0 or bx % 2 == 0) and not (x + bs == 64 and y > 0)
                bl = x == 0 and y + bs < 64
                for m in ND_LUMA_MODES:
                    ox.append(x)
                    oy.append(y)
                    bsz.append(bs)
                    mode.append(m)
                    pf.append((1 if tr else 0) | (2 if bl else 0))
    ox, oy, bsz = np.array(ox), np.array(oy), np.array(bsz)
    sy, sx = np.mgrid[0:sbh, 0:sbw]
    sx, sy = sx.ravel() * 64, sy.ravel() * 64
    total = len(sx) * len(ox)
    c = np.zeros(total, _lib.CAND_DTYPE)
    c["x"] = (sx[:, None] + ox[None, :]).ravel()
    c["y"] = (sy[:, None] + oy[None, :]).ravel()
    c["tx_size"] = np.tile(np.array([SQUARE_TX[int(b)] for b in bsz]), len(sx))
    c["mode"] = np.tile(np.array(mode), len(sx))
    c["pflags"] = np.tile(np.array(pf), len(sx))
    px = np.tile(bsz.astype(np.uint64) ** 2, len(sx))
    off = np.concatenate([[0], np.cumsum(px)[:-1]]).astype(np.uint64)
    c["offset"] = off
    off = int(px.sum()) if total else 0
    c["planes"] = 1 << 2
    c["qidx"] = qidx
    c["bd_qm"] = (qm_level & 15) | (bdi << 4)
    ncoef = np.minimum(c["tx_size"].astype(np.int64), 3)
    c["eob_bin"] = np.array([EOB_BIN_LUMA[16], EOB_BIN_LUMA[64], EOB_BIN_LUMA[256],
                             EOB_BIN_LUMA[1024]])[ncoef]
    c["txtp_set"] = 0
    # coefficient CDF context of the frame's qindex (get_q_ctx: 20 / 60 / 120)
    c["qctx"] = 0 if qidx <= 20 else 1 if qidx <= 60 else 2 if qidx <= 120 else 3
    c["wsse_scale"] = wsse_scale
    c["lambda"] = lam
    assert off < (1 << 32)
    return c, off


# rate-table slot of the end-of-block class of a coded coefficient count, (luma, chroma)
EOB_BIN = {16: (0, 1), 32: (3, 4), 64: (5, 6), 128: (8, 9), 256: (11, 12), 512: (13, 14), 1024: (15, 16)}


def chain_trellis_descriptors(blocks, sub=2):
    """Slow-mode trellis descriptors (TRELLISBLK_DTYPE) for a chain batch of TXBLK_DTYPE
    records: trellis_optimize_ctx on every block (src/trellis.rs:55), the frame qindex's
    coefficient context (get_q_ctx), the luma/chroma end-of-block class of the coded size,
    the DC-sign context of a block without coded neighbours, and the lambda the reference's
    encoder derives per frame (trellis_lambda x hbd_trellis_scale at chroma subsampling
    `sub`: 0 4:4:4, 1 4:2:2, 2 4:2:0)."""
    n = len(blocks)
    tr = np.zeros(n, _lib.TRELLISBLK_DTYPE)
    for k in ("offset", "tx_size", "qidx", "bd_qm"):
        tr[k] = blocks[k]
    tr["flags"] = _lib.TRELLIS_CTX
    q = blocks["qidx"].astype(np.int64)
    tr["qctx"] = np.where(q <= 20, 0, np.where(q <= 60, 1, np.where(q <= 120, 2, 3)))
    chroma = (blocks["bd_qm"] & _lib.TXBLK_CHROMA) != 0
    tr["plane"] = chroma
    w = np.array([_lib.tx_width(s) for s in range(19)])
    h = np.array([_lib.tx_height(s) for s in range(19)])
    nc = np.minimum(w, 32) * np.minimum(h, 32)
    slot = np.array([EOB_BIN[int(c)] for c in nc], np.uint8)
    tr["eob_bin"] = slot[blocks["tx_size"], chroma.astype(np.int64)]
    bdi = (blocks["bd_qm"] >> 4) & 3
    lam = np.array([[_lib.trellis_lambda0(qi, (8, 10, 12)[b], sub) for qi in range(256)] for b in range(3)],
                   np.float32)
    tr["lambda0"] = lam[np.minimum(bdi, 2), q]
    return tr
