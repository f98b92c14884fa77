"""Test inputs for the deblocking filter: random square quadtree partitions of a frame, the
reference's deblock maps for them (record_blk, src/coder/superblock.rs:548-562: every 4 x 4
unit of a leaf holds the leaf's size in 4-units in the transform and prediction maps, and
flags the leaf's first column / row), and blocky reconstructions to filter."""
import numpy as np


def random_leaves(rng, w8, h8, p_split=(0.6, 0.55, 0.5, 0.45), min_size=4):
    """Leaves (x, y, size) of a random quadtree over the 64 x 64 superblocks of a w8 x h8
    frame (multiples of 8): nodes crossing the frame edge split, nodes outside are dropped.
    p_split: the split probability at 64, 32, 16 and 8."""
    leaves = []

    def walk(x, y, s, lvl):
        if x >= w8 or y >= h8:
            return
        crosses = x + s > w8 or y + s > h8
        if s > min_size and (crosses or rng.random() < p_split[lvl]):
            h = s // 2
            for dy in (0, h):
                for dx in (0, h):
                    walk(x + dx, y + dy, h, lvl + 1)
        else:
            leaves.append((x, y, s))

    for sy in range(0, h8, 64):
        for sx in range(0, w8, 64):
            walk(sx, sy, 64, 0)
    return leaves


def maps_from_leaves(leaves, w8, h8):
    nc4, nr4 = w8 // 4, h8 // 4
    m = {k: np.zeros((nr4, nc4), np.uint8) for k in ("blk4", "blk4h", "blk4v", "blk4t", "pblk4", "pblk4h")}
    for (x, y, s) in leaves:
        d, bx, by = s // 4, x // 4, y // 4
        for k in ("blk4", "blk4h", "pblk4", "pblk4h"):
            m[k][by:by + d, bx:bx + d] = d
        m["blk4v"][by:by + d, bx:bx + d] = 0
        m["blk4v"][by:by + d, bx] = 1
        m["blk4t"][by:by + d, bx:bx + d] = 0
        m["blk4t"][by, bx:bx + d] = 1
    return m


def blocky_plane(rng, leaves, pw, ph, bd, sub=0, step=6, noise=1):
    """A smooth ramp plus one random offset per leaf (the leaf's area at the plane's
    subsampling) plus small noise: block edges the filter smooths, flat runs for its wide
    taps, and some edges too strong to filter."""
    one = 1 << (bd - 8)
    yy, xx = np.mgrid[0:ph, 0:pw]
    base = (1 << (bd - 1)) + ((xx * 3 + yy * 2) % 97 - 48) * one // 2
    img = base.astype(np.int64)
    for (x, y, s) in leaves:
        x0, y0, sz = x >> sub, y >> sub, max(s >> sub, 1)
        img[y0:y0 + sz, x0:x0 + sz] += int(rng.integers(-step, step + 1)) * one
    img += rng.integers(-noise, noise + 1, img.shape) * one
    # a few sharp features (strong edges the masks reject)
    for _ in range(max(1, pw * ph // 2048)):
        x, y = int(rng.integers(0, pw)), int(rng.integers(0, ph))
        img[y:y + 3, x:x + 3] += int(rng.integers(-60, 61)) * one
    return np.clip(img, 0, (1 << bd) - 1).astype(np.uint16)


def leaves_from_nodes(nodes, w8, h8):
    """The leaves (x, y, size) of an RDO's node records (structured array (nsb, 341) with a
    split field: 0 leaf, 1 split; node n of level L at base[L] + iy * 2^L + ix)."""
    base = (0, 1, 5, 21, 85)
    sbc = (w8 + 63) // 64
    leaves = []
    for sb in range(nodes.shape[0]):
        sx, sy = (sb % sbc) * 64, (sb // sbc) * 64

        def walk(L, ix, iy):
            s = 64 >> L
            sp = int(nodes[sb, base[L] + iy * (1 << L) + ix]["split"])
            if sp == 0:
                leaves.append((sx + ix * s, sy + iy * s, s))
            elif sp == 1:
                for dy in (0, 1):
                    for dx in (0, 1):
                        walk(L + 1, 2 * ix + dx, 2 * iy + dy)

        walk(0, 0, 0)
    return leaves
