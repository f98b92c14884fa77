"""The AV1 bitstream and AVIF file of the RDO's decisions (host side of libmtgpu:
mt_av1_encode, mt_avif_encode, mt_av1_tokens, mt_av1_range_code, mt_av1_cdf).

The coder reads the node records and the level planes the RDO leaves (intra_rdo_batch with
levels=True) after they are copied to the host."""
import ctypes

import numpy as np

from . import _lib
from ._lib import lib, _check
from .rdo import NODE_DTYPE, LUMA_ONLY, CHROMA_420, CHROMA_444  # noqa: F401


class MtAv1Frame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("bd", ctypes.c_uint8),
                ("chroma", ctypes.c_uint8), ("qidx", ctypes.c_uint8), ("qm_level", ctypes.c_uint8),
                ("uv_qm_level", ctypes.c_uint8), ("uv_dc_delta", ctypes.c_int8), ("uv_ac_delta", ctypes.c_int8),
                ("filter_intra", ctypes.c_uint8), ("lf_level", ctypes.c_uint8 * 4), ("lf_sharpness", ctypes.c_uint8),
                ("cicp", ctypes.c_uint8 * 3), ("full_range", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 3),
                ("nodes", ctypes.c_void_p), ("coef_y", ctypes.c_void_p), ("coef_u", ctypes.c_void_p),
                ("coef_v", ctypes.c_void_p), ("coef_stride", ctypes.c_int32), ("coef_uv_stride", ctypes.c_int32),
                ("cdef_on", ctypes.c_uint8), ("cdef_damping", ctypes.c_uint8), ("cdef_bits", ctypes.c_uint8),
                ("cdef_pad", ctypes.c_uint8), ("cdef_y_pri", ctypes.c_uint8 * 8), ("cdef_y_sec", ctypes.c_uint8 * 8),
                ("cdef_uv_pri", ctypes.c_uint8 * 8), ("cdef_uv_sec", ctypes.c_uint8 * 8),
                ("cdef_grid", ctypes.c_void_p)]


SRGB_IDENTITY = (1, 13, 0)   # BT.709 primaries, sRGB transfer, identity matrix (4:4:4 only)
BT601 = (1, 13, 6)           # BT.709 primaries, sRGB transfer, BT.601 matrix
UNSPECIFIED = (2, 2, 2)


def _sigs():
    L = lib()
    if not getattr(L, "_av1_sigs", False):
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        for f in ("mt_av1_encode", "mt_avif_encode", "mt_av1_tokens"):
            getattr(L, f).argtypes = [ctypes.POINTER(MtAv1Frame), vp, sz, ctypes.POINTER(sz)]
        L.mt_av1_skip8.argtypes = [ctypes.POINTER(MtAv1Frame), vp]
        L.mt_av1_range_code.argtypes = [vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.mt_av1_cdf.argtypes = [ctypes.c_uint32, ctypes.c_int, vp]
        L._av1_sigs = True
    return L


def frame(nodes, levels, width, height, bd, qidx, chroma=LUMA_ONLY, qm_level=15, uv_qm_level=15, uv_dc=0,
          uv_ac=0, filter_intra=True, lf_levels=(0, 0, 0, 0), lf_sharpness=0, cicp=UNSPECIFIED, full_range=True,
          cdef=None, cdef_grid=None):
    """An mt_av1_frame over host arrays: nodes (NODE_DTYPE records, (nsb, 341) or flat), levels
    = (luma, u, v) int32 planes (u, v None for luma only); cdef: a CDEF parameter dict (the
    params of cdef.frame or the oracle's) with its per-superblock grid.  Returns (struct,
    keep-alive list)."""
    nodes = np.asarray(nodes)
    if nodes.dtype.names and nodes.dtype != NODE_DTYPE:  # another record layout with the same fields
        conv = np.zeros(nodes.size, NODE_DTYPE)
        for k in ("cost", "none_cost", "split", "mode", "uv_mode", "cfl_alpha_u", "cfl_alpha_v"):
            conv[k] = nodes.reshape(-1)[k]
        nodes = conv
    keep = [np.ascontiguousarray(nodes.view(NODE_DTYPE).reshape(-1)),
            np.ascontiguousarray(levels[0], np.int32)]
    f = MtAv1Frame()
    f.width, f.height, f.bd, f.chroma, f.qidx = width, height, bd, chroma, qidx
    f.qm_level, f.uv_qm_level, f.uv_dc_delta, f.uv_ac_delta = qm_level, uv_qm_level, uv_dc, uv_ac
    f.filter_intra = 1 if filter_intra else 0
    for i in range(4):
        f.lf_level[i] = lf_levels[i]
    f.lf_sharpness = lf_sharpness
    for i in range(3):
        f.cicp[i] = cicp[i]
    f.full_range = 1 if full_range else 0
    f.nodes = keep[0].ctypes.data
    f.coef_y = keep[1].ctypes.data
    f.coef_stride = keep[1].shape[1]
    if chroma != LUMA_ONLY:
        keep.append(np.ascontiguousarray(levels[1], np.int32))
        keep.append(np.ascontiguousarray(levels[2], np.int32))
        f.coef_u = keep[2].ctypes.data
        f.coef_v = keep[3].ctypes.data
        f.coef_uv_stride = keep[2].shape[1]
    if cdef is not None and cdef.get("on", 1):
        f.cdef_on, f.cdef_damping, f.cdef_bits = 1, cdef["damping"], cdef["bits"]
        for i in range(1 << cdef["bits"]):
            f.cdef_y_pri[i], f.cdef_y_sec[i] = cdef["y_pri"][i], cdef["y_sec"][i]
            f.cdef_uv_pri[i], f.cdef_uv_sec[i] = cdef["uv_pri"][i], cdef["uv_sec"][i]
        if cdef["bits"]:
            g = np.ascontiguousarray(cdef_grid, np.uint8)
            keep.append(g)
            f.cdef_grid = g.ctypes.data
    return f, keep


def _call_bytes(fn, f):
    n = ctypes.c_size_t()
    _check(fn(ctypes.byref(f), None, 0, ctypes.byref(n)))
    out = np.zeros(max(n.value, 1), np.uint8)
    _check(fn(ctypes.byref(f), out.ctypes.data, out.size, ctypes.byref(n)))
    return out[:n.value].tobytes()


def encode_obus(f):
    """Sequence header + frame OBU of an mt_av1_frame (bytes)."""
    return _call_bytes(_sigs().mt_av1_encode, f)


def encode_avif(f):
    """The AVIF file of an mt_av1_frame (bytes)."""
    return _call_bytes(_sigs().mt_avif_encode, f)


def tokens(f):
    """The frame's entropy tokens (uint32 numpy array)."""
    L = _sigs()
    n = ctypes.c_size_t()
    _check(L.mt_av1_tokens(ctypes.byref(f), None, 0, ctypes.byref(n)))
    out = np.zeros(max(n.value, 1), np.uint32)
    _check(L.mt_av1_tokens(ctypes.byref(f), out.ctypes.data, out.size, ctypes.byref(n)))
    return out[:n.value]


def range_code(toks):
    """Tile data of a token stream (bytes)."""
    L = _sigs()
    toks = np.ascontiguousarray(toks, np.uint32)
    n = ctypes.c_size_t()
    _check(L.mt_av1_range_code(toks.ctypes.data, toks.size, None, 0, ctypes.byref(n)))
    out = np.zeros(max(n.value, 1), np.uint8)
    _check(L.mt_av1_range_code(toks.ctypes.data, toks.size, out.ctypes.data, out.size, ctypes.byref(n)))
    return out[:n.value].tobytes()


def token_cdf(tok):
    """(cumulative values, nsym) of a CDF-symbol token."""
    L = _sigs()
    nsym = ((int(tok) >> 4) & 15) + 1
    out = np.zeros(nsym - 1, np.uint16)
    _check(L.mt_av1_cdf(int(tok) >> 8, nsym, out.ctypes.data))
    return out, nsym


def skip8(f):
    """mt_av1_skip8: the frame's luma 8x8 skip map (uint8, (height / 8, width / 8))."""
    out = np.zeros((f.height // 8, f.width // 8), np.uint8)
    _check(_sigs().mt_av1_skip8(ctypes.byref(f), out.ctypes.data))
    return out
