"""CDEF (include/mtgpu.h: mt_cdef_frame, mt_av1_skip8), mirroring the reference's frame_cdef
(src/coder/frame_cdef.rs:49) and its filter (src/cdef.rs).  Planes are uint16 (stored as
int16) cuda tensors; the filter is out of place."""
import ctypes

import numpy as np

from ._lib import _check, _dev_ptr, _stream_ptr, init, lib

LUMA_ONLY, CHROMA_420, CHROMA_444 = 0, 1, 2  # MT_RDO_LUMA_ONLY, MT_RDO_CHROMA_420, _444 (as rdo's)

N_UNITS_MAX_ENTRIES = 8


class MtCdefParams(ctypes.Structure):
    _fields_ = [("on", ctypes.c_int32), ("bits", ctypes.c_int32), ("damping", ctypes.c_int32),
                ("n", ctypes.c_int32), ("y_pri", ctypes.c_uint8 * 8), ("y_sec", ctypes.c_uint8 * 8),
                ("uv_pri", ctypes.c_uint8 * 8), ("uv_sec", ctypes.c_uint8 * 8)]

    def as_dict(self):
        n = self.n
        return {"on": self.on, "bits": self.bits, "damping": self.damping, "n": n,
                "y_pri": list(self.y_pri[:n]), "y_sec": list(self.y_sec[:n]),
                "uv_pri": list(self.uv_pri[:n]), "uv_sec": list(self.uv_sec[:n])}


class MtCdefFrameArgs(ctypes.Structure):
    vp = ctypes.c_void_p
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("bd", ctypes.c_uint8),
                ("chroma", ctypes.c_uint8), ("slow", ctypes.c_uint8), ("pad0", ctypes.c_uint8),
                ("qidx", ctypes.c_int32), ("wsse_scale", ctypes.c_float), ("lambda_", ctypes.c_float),
                ("rec_y", vp), ("rec_u", vp), ("rec_v", vp), ("y_stride", ctypes.c_int32),
                ("uv_stride", ctypes.c_int32), ("src_y", vp), ("src_u", vp), ("src_v", vp),
                ("src_y_stride", ctypes.c_int32), ("src_uv_stride", ctypes.c_int32), ("skip8", vp),
                ("dst_y", vp), ("dst_u", vp), ("dst_v", vp)]


_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        vp = ctypes.c_void_p
        L.mt_cdef_frame.argtypes = [ctypes.POINTER(MtCdefFrameArgs), ctypes.POINTER(MtCdefParams), vp, vp, vp]
        _bound = True
    return L


def _opt(t):
    return ctypes.c_void_p(None) if t is None else _dev_ptr(t)


def units(width, height):
    return ((width + 63) // 64) * ((height + 63) // 64)


def frame(rec, src, skip8, chroma, bd, qidx, lam, wsse_scale=0.0, slow=True, out=None, stream=None,
          with_tables=False):
    """mt_cdef_frame.  rec, src: (y, u, v) cuda tensors (u, v None for luma only); skip8: uint8
    cuda tensor (h / 8, w / 8).  Returns (dst planes, params dict, grid (units,) uint8[,
    tables (2, units, 12) int64])."""
    import torch
    y = rec[0]
    init(y.device.index or 0)
    h, w = y.shape
    mono = chroma == LUMA_ONLY
    a = MtCdefFrameArgs()
    a.width, a.height, a.bd, a.chroma, a.slow = w, h, int(bd), int(chroma), 1 if slow else 0
    a.qidx, a.wsse_scale, a.lambda_ = int(qidx), float(wsse_scale), float(lam)
    if out is None:
        out = tuple(None if p is None else torch.empty_like(p) for p in rec)
    a.rec_y, a.rec_u, a.rec_v = (_opt(p).value for p in rec)
    a.src_y, a.src_u, a.src_v = (_opt(p).value for p in src)
    a.dst_y, a.dst_u, a.dst_v = (_opt(p).value for p in out)
    a.y_stride, a.src_y_stride = y.stride(0), src[0].stride(0)
    a.uv_stride = 0 if mono else rec[1].stride(0)
    a.src_uv_stride = 0 if mono else src[1].stride(0)
    a.skip8 = _dev_ptr(skip8).value
    prm = MtCdefParams()
    grid = np.zeros(units(w, h), np.uint8)
    tabs = np.zeros((2, units(w, h), 12), np.int64) if with_tables else None
    _check(_L().mt_cdef_frame(ctypes.byref(a), ctypes.byref(prm), grid.ctypes.data,
                              None if tabs is None else tabs.ctypes.data, _stream_ptr(stream)))
    r = (out, prm.as_dict(), grid)
    return r + (tabs,) if with_tables else r


def sub_of(chroma):
    return 1 if chroma == CHROMA_420 else 0
