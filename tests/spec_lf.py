"""TEST INFRASTRUCTURE ONLY: an independent restatement of AV1's normative loop filter, used to
pin the oracle's transliteration of the reference's dav1d-style deblocking (src/loopfilter.rs,
src/coder/frame_deblock.rs).  Nothing here is product code.

Written from the AV1 Bitstream & Decoding Process Specification, not from the reference:
  - 7.14.1 / 7.14.2 the edge loop: per plane, pass 0 filters every vertical transform edge of
    the whole frame, then pass 1 every horizontal one; the frame's left column and top row
    are not edges; a 4 x 4 unit is on screen when its luma position lies inside the visible
    frame;
  - 7.14.3 the filter size: the smaller of the two transform dimensions across the edge,
    capped at 16 for luma and 6 for chroma;
  - 7.14.4 the limits of a level: shift = 2 for sharpness > 4, 1 for sharpness > 0, else 0;
    limit = Clip3(1, 9 - sharpness, lvl >> shift) with sharpness, else Max(1, lvl >> shift);
    blimit = 2 * (lvl + 2) + limit; thresh = lvl >> 4;
  - 7.14.6.2 the masks (filter, high edge variance, flat, flat2) at the bit depth;
  - 7.14.6.3 the narrow filter in the signed (sample - 0x80 << (BitDepth - 8)) domain with
    filter4_clamp and Round2;
  - 7.14.6.4 the wide filter: taps over a Clip3-extended window, weight 2 within n2 of the
    centre, Round2 by log2Size, n = 6 (log2Size 4), 3 (luma) or 2 (chroma), n2 = 0 for the
    luma 8-tap, else 1.
Pure Python integers over lists: small frames only.
"""


def limits(lvl, sharp):
    shift = 2 if sharp > 4 else (1 if sharp > 0 else 0)
    if sharp > 0:
        limit = min(max(lvl >> shift, 1), 9 - sharp)
    else:
        limit = max(1, lvl >> shift)
    return 2 * (lvl + 2) + limit, limit, lvl >> 4


def _narrow(F, hev, bd):
    half = 1 << (bd - 1)
    off = 0x80 << (bd - 8)

    def c4(x):
        return min(max(x, -half), half - 1)

    ps1, ps0, qs0, qs1 = F[-2] - off, F[-1] - off, F[0] - off, F[1] - off
    f = c4(ps1 - qs1) if hev else 0
    f = c4(f + 3 * (qs0 - ps0))
    f1 = c4(f + 4) >> 3
    f2 = c4(f + 3) >> 3
    out = {0: c4(qs0 - f1) + off, -1: c4(ps0 + f2) + off}
    if not hev:
        f = (f1 + 1) >> 1
        out[1] = c4(qs1 - f) + off
        out[-2] = c4(ps1 + f) + off
    return out


def _wide(F, log2size, luma):
    n = 6 if log2size == 4 else (3 if luma else 2)
    n2 = 0 if (log2size == 3 and luma) else 1
    out = {}
    for i in range(-n, n):
        t = 0
        for j in range(-n, n + 1):
            p = min(max(i + j, -(n + 1)), n)
            t += F[p] * (2 if abs(j) <= n2 else 1)
        out[i] = (t + (1 << (log2size - 1))) >> log2size
    return out


def filter_samples(F, size, lvl, sharp, bd, luma):
    """7.14.6.2: F maps offsets -8..7 to samples (p0 at -1, q0 at 0).  Returns the offsets
    written and their values."""
    blimit, limit, thresh = limits(lvl, sharp)
    sh = bd - 8
    blimit, limit, thresh, one = blimit << sh, limit << sh, thresh << sh, 1 << sh
    q = [F[k] for k in range(8)]
    p = [F[-1 - k] for k in range(8)]
    hev = abs(p[1] - p[0]) > thresh or abs(q[1] - q[0]) > thresh
    mask = abs(p[1] - p[0]) <= limit and abs(q[1] - q[0]) <= limit and \
        abs(p[0] - q[0]) * 2 + abs(p[1] - q[1]) // 2 <= blimit
    if size >= 6:
        mask = mask and abs(p[2] - p[1]) <= limit and abs(q[2] - q[1]) <= limit
    if size >= 8:
        mask = mask and abs(p[3] - p[2]) <= limit and abs(q[3] - q[2]) <= limit
    if not mask:
        return {}
    flat = False
    if size >= 6:
        flat = all(abs(p[k] - p[0]) <= one and abs(q[k] - q[0]) <= one for k in (1, 2))
        if size >= 8:
            flat = flat and abs(p[3] - p[0]) <= one and abs(q[3] - q[0]) <= one
    flat2 = size == 16 and all(abs(p[k] - p[0]) <= one and abs(q[k] - q[0]) <= one for k in (4, 5, 6))
    if size == 4 or not flat:
        return _narrow(F, hev, bd)
    if size in (6, 8) or not flat2:
        return _wide(F, 3, luma)
    return _wide(F, 4, luma)


def tx_grid(leaves, w8, h8, plane, sub):
    """Transform geometry of a plane from the luma leaves (x, y, size) of a square quadtree:
    per 4 x 4 unit of the plane, the (width, height) of its transform block and whether the
    unit starts that block horizontally / vertically.  Luma: one transform per leaf.  Chroma
    at 4:2:0 (sub 1): the leaf's half-size block, one 4 x 4 for each 8 x 8 area of 4 x 4 leaves;
    4:4:4 (sub 0): the leaf's own size.  Chroma transforms are at most 32 x 32."""
    pw, ph = (w8, h8) if plane == 0 else (w8 >> sub, h8 >> sub)
    nc, nr = pw // 4, ph // 4
    size = [[0] * nc for _ in range(nr)]
    sx = [[False] * nc for _ in range(nr)]
    sy = [[False] * nc for _ in range(nr)]
    blocks = set()
    for (x, y, s) in leaves:
        if plane == 0:
            blocks.add((x, y, s, s))
        elif sub == 0:
            blocks.add((x, y, s, min(s, 32)))
        elif s >= 8:
            blocks.add((x >> 1, y >> 1, s >> 1, min(s >> 1, 32)))
        else:
            blocks.add(((x >> 3) << 2, (y >> 3) << 2, 4, 4))
    for (bx, by, bs, ts) in blocks:
        for ty in range(by, by + bs, ts):
            for tx in range(bx, bx + bs, ts):
                for r in range(ty // 4, min((ty + ts) // 4, nr)):
                    for c in range(tx // 4, min((tx + ts) // 4, nc)):
                        size[r][c] = ts
                        sx[r][c] = c * 4 == tx
                        sy[r][c] = r * 4 == ty
    return size, sx, sy


def deblock_plane(px, pw, ph, grid, vis_w, vis_h, sub, lvl, sharp, bd, luma):
    """7.14.2 on one plane (px: list of rows, modified in place), edges in raster order within
    each pass.  vis_w / vis_h: the visible luma size; sub: the plane's subsampling shift."""
    if lvl == 0:
        return
    size, sx, sy = grid
    cap = 16 if luma else 6
    nc, nr = pw // 4, ph // 4
    for pass_ in (0, 1):
        for r in range(nr):
            for c in range(nc):
                if ((c * 4) << sub) >= vis_w or ((r * 4) << sub) >= vis_h:
                    continue
                if pass_ == 0:
                    if c == 0 or not sx[r][c]:
                        continue
                    fs = min(cap, size[r][c], size[r][c - 1])
                else:
                    if r == 0 or not sy[r][c]:
                        continue
                    fs = min(cap, size[r][c], size[r - 1][c])
                for line in range(4):
                    if pass_ == 0:
                        y, x0 = r * 4 + line, c * 4
                        F = {k: px[y][min(max(x0 + k, 0), pw - 1)] for k in range(-8, 8)}
                        for k, v in filter_samples(F, fs, lvl, sharp, bd, luma).items():
                            px[y][x0 + k] = v
                    else:
                        x, y0 = c * 4 + line, r * 4
                        F = {k: px[min(max(y0 + k, 0), ph - 1)][x] for k in range(-8, 8)}
                        for k, v in filter_samples(F, fs, lvl, sharp, bd, luma).items():
                            px[y0 + k][x] = v
