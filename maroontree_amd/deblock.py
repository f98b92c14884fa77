"""The in-loop deblocking filter (include/mtgpu.h: mt_deblock_frame, mt_deblock_plane,
mt_deblock_levels, mt_rdo_deblock_maps), mirroring the reference's frame_deblock
(src/coder/frame_deblock.rs:206), filter_plane (src/loopfilter.rs:442) and
loop_filter_levels (src/obu.rs:306).  Planes are uint16 (stored as int16) cuda tensors; the
filter is out of place."""
import ctypes

from ._lib import MtError, _check, _dev_ptr, _stream_ptr, init, lib

MAP_NAMES = ("blk4", "blk4h", "blk4v", "blk4t", "pblk4", "pblk4h")
DEBLOCK_ORDER, DEBLOCK_BAD_MAP = 1, 2


class MtLfPlane(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("stride", ctypes.c_int32),
                ("w", ctypes.c_int32), ("h", ctypes.c_int32), ("vis_w", ctypes.c_int32),
                ("vis_h", ctypes.c_int32), ("nc4", ctypes.c_int32), ("bw4", ctypes.c_void_p),
                ("bh4", ctypes.c_void_p), ("vedge4", ctypes.c_void_p), ("hedge4", ctypes.c_void_p)]


class MtDeblockGeom(ctypes.Structure):
    _fields_ = [("w8", ctypes.c_int32), ("h8", ctypes.c_int32), ("disp_w", ctypes.c_int32),
                ("disp_h", ctypes.c_int32)] + [(n, ctypes.c_void_p) for n in MAP_NAMES] + \
               [("sub_x", ctypes.c_uint8), ("sub_y", ctypes.c_uint8), ("mono", ctypes.c_uint8),
                ("sharp", ctypes.c_uint8), ("bd", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 3)]


_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        P = ctypes.POINTER
        L.mt_deblock_levels.argtypes = [i32, i32, P(i32), P(i32), P(i32)]
        L.mt_deblock_plane.argtypes = [P(MtLfPlane), i32, i32, i32, i32, i32, vp, vp]
        L.mt_deblock_frame.argtypes = [P(MtDeblockGeom), vp, vp, vp, ctypes.c_int32, ctypes.c_int32, vp, vp, vp,
                                       i32, i32, i32, vp, vp]
        L.mt_rdo_deblock_maps.argtypes = [vp, i32, i32, vp, vp]
        _bound = True
    return L


def levels(qidx, sub):
    """mt_deblock_levels -> (lvl_y, lvl_uv, sharpness) of a base_q_idx; sub 0 4:4:4, 1 4:2:2,
    2 4:2:0."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(_L().mt_deblock_levels(int(qidx), int(sub), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value, c.value


def _opt(t):
    return ctypes.c_void_p(None) if t is None else _dev_ptr(t)


def rdo_maps(nodes, width, height, out=None, stream=None):
    """mt_rdo_deblock_maps: the reference's deblock maps of an RDO's node records.  Returns a
    uint8 cuda tensor (6, height / 4, width / 4) in MAP_NAMES order."""
    import torch
    init(nodes.device.index or 0)
    if out is None:
        out = torch.empty((6, height // 4, width // 4), dtype=torch.uint8, device=nodes.device)
    _check(_L().mt_rdo_deblock_maps(_dev_ptr(nodes), int(width), int(height), _dev_ptr(out), _stream_ptr(stream)))
    return out


def frame(y, u, v, maps, disp_w, disp_h, sub, sharp, bd, lvl_y, lvl_u, lvl_v, out=None, status=None,
          stream=None):
    """mt_deblock_frame: y (h8, w8), u / v (h8 >> sub, w8 >> sub) or None (monochrome); maps a
    uint8 tensor (6, h8 / 4, w8 / 4) or a dict of MAP_NAMES.  Returns (dst_y, dst_u, dst_v)."""
    import torch
    init(y.device.index or 0)
    h8, w8 = y.shape
    mono = u is None
    g = MtDeblockGeom()
    g.w8, g.h8, g.disp_w, g.disp_h = w8, h8, int(disp_w), int(disp_h)
    for i, n in enumerate(MAP_NAMES):
        m = maps[n] if isinstance(maps, dict) else maps[i]
        setattr(g, n, _dev_ptr(m).value)
    g.sub_x = g.sub_y = int(sub)
    g.mono, g.sharp, g.bd = 1 if mono else 0, int(sharp), int(bd)
    if out is None:
        out = (torch.empty_like(y), None if mono else torch.empty_like(u), None if mono else torch.empty_like(v))
    uv_stride = 0 if mono else u.shape[1]
    _check(_L().mt_deblock_frame(ctypes.byref(g), _dev_ptr(y), _opt(u), _opt(v), w8, uv_stride, _dev_ptr(out[0]),
                                 _opt(out[1]), _opt(out[2]), int(lvl_y), int(lvl_u), int(lvl_v), _opt(status),
                                 _stream_ptr(stream)))
    return out


def plane(src, bw4, bh4, vedge4, hedge4, vis_w, vis_h, level, sharp, is_luma, sb_rows4, bd, out=None,
          status=None, stream=None):
    """mt_deblock_plane: src (h, w); maps uint8 (h / 4, w / 4), vedge4 / hedge4 None for the
    alignment rule.  Returns the filtered plane."""
    import torch
    init(src.device.index or 0)
    h, w = src.shape
    if out is None:
        out = torch.empty_like(src)
    p = MtLfPlane()
    p.src, p.dst = _dev_ptr(src).value, _dev_ptr(out).value
    p.stride, p.w, p.h, p.vis_w, p.vis_h, p.nc4 = w, w, h, int(vis_w), int(vis_h), w // 4
    p.bw4, p.bh4 = _dev_ptr(bw4).value, _dev_ptr(bh4).value
    p.vedge4 = None if vedge4 is None else _dev_ptr(vedge4).value
    p.hedge4 = None if hedge4 is None else _dev_ptr(hedge4).value
    _check(_L().mt_deblock_plane(ctypes.byref(p), int(level), int(sharp), int(is_luma), int(sb_rows4), int(bd),
                                 _opt(status), _stream_ptr(stream)))
    return out


__all__ = ["levels", "rdo_maps", "frame", "plane", "MAP_NAMES", "DEBLOCK_ORDER", "DEBLOCK_BAD_MAP", "MtError"]
