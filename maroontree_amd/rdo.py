"""Whole-frame luma intra RDO over the superblock wavefront (mt_intra_rdo).

The procedure (include/mtgpu.h, mt_intra_rdo) replaces the reference's candidate generators
and argmin (rd_cost_square src/coder/partition_search.rs:338, pick_luma16 :4755, pick_block8
src/coder/block8.rs:689, pick_luma32 src/coder/block32.rs:728, rd_pick_luma64
src/coder/block64.rs:1812, decode_sb src/coder/superblock.rs:813) with one exhaustive,
data-parallel search whose every candidate runs through the library's bit-exact kernels.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import MtPlane, _check, _dev_ptr, _stream_ptr, init, lib

ANGLE_DELTAS, FILTER_INTRA, MIN4, CFL, TRELLIS, GRAPH = 1, 2, 4, 8, 32, 64


def cfl_steps(s):
    """Flag bits of cfl_alpha_rd_refine's walk with s (0..15) steps each way (with CFL)."""
    if not 0 <= s <= 15:
        raise ValueError("CfL walk steps are 0..15")
    return s << 8
NONE, SPLIT, UNCODED, HORZ, VERT = 0, 1, 2, 3, 4  # node decisions (the split field; HORZ / VERT: oracle only)
LUMA_ONLY, CHROMA_420, CHROMA_444 = 0, 1, 2
NODES = 341
ALL_MODES, FORCE_SPLIT, NO_SPLIT = 0x3FFF, 0x4000, 0x8000
LEVEL_BASE = [0, 1, 5, 21, 85]
NODE_DTYPE = np.dtype([("cost", "<f4"), ("none_cost", "<f4"), ("split", "u1"), ("mode", "u1"),
                       ("uv_mode", "u1"), ("pad0", "u1"), ("cfl_alpha_u", "i1"), ("cfl_alpha_v", "i1"),
                       ("pad1", "u1", (2,))])


class MtRdoFrame(ctypes.Structure):
    _fields_ = [("src", MtPlane), ("recon", ctypes.c_void_p), ("recon_stride", ctypes.c_int32),
                ("bd", ctypes.c_uint8), ("qidx", ctypes.c_uint8), ("qm_level", ctypes.c_uint8),
                ("chroma", ctypes.c_uint8), ("lambda_", ctypes.c_float), ("wsse_scale", ctypes.c_float),
                ("prune", ctypes.c_void_p), ("nodes", ctypes.c_void_p),
                ("src_u", MtPlane), ("src_v", MtPlane), ("recon_u", ctypes.c_void_p),
                ("recon_v", ctypes.c_void_p), ("recon_uv_stride", ctypes.c_int32),
                ("uv_dc_delta", ctypes.c_int8), ("uv_ac_delta", ctypes.c_int8),
                ("uv_qm_level", ctypes.c_uint8), ("pad1", ctypes.c_uint8),
                ("coef", ctypes.c_void_p), ("coef_u", ctypes.c_void_p), ("coef_v", ctypes.c_void_p),
                ("coef_stride", ctypes.c_int32), ("coef_uv_stride", ctypes.c_int32)]


class MtRdoStats(ctypes.Structure):
    _fields_ = [("bits", ctypes.c_double), ("sse", ctypes.c_int64), ("candidates", ctypes.c_int64),
                ("leaves", ctypes.c_int32), ("sbs", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("bits_uv", ctypes.c_double), ("sse_u", ctypes.c_int64),
                ("sse_v", ctypes.c_int64)]


def _sigs():
    L = lib()
    if not getattr(L, "_rdo_chroma_sigs", False):
        L.mt_chroma_q_deltas.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
        L.mt_qm_chroma_level_law.argtypes = [ctypes.c_int, ctypes.c_int]
        L._rdo_chroma_sigs = True
    return L


def rdo_lambda(qidx, bd, sub=2):
    """The reference's decision lambda lambda_d (mt_rdo_lambda)."""
    return float(lib().mt_rdo_lambda(qidx, bd, sub))


def mode_lambda(qidx, bd, sub=2):
    return float(lib().mt_mode_lambda(qidx, bd, sub))


def chroma_laws(qidx, sub):
    """The reference's frame chroma quantizer (dc delta, ac delta, QM level) for sub 2 (4:2:0)
    or 0 (4:4:4): chroma_dc_delta, chroma_ac_delta, qm_chroma_level_law."""
    L = _sigs()
    dc, ac = ctypes.c_int(), ctypes.c_int()
    _check(L.mt_chroma_q_deltas(qidx, sub, ctypes.byref(dc), ctypes.byref(ac)))
    return dc.value, ac.value, int(L.mt_qm_chroma_level_law(qidx, sub))


def node_geometry(node):
    """(level, ix, iy, size) of quadtree node `node` of a superblock."""
    L = 4 if node >= 85 else 3 if node >= 21 else 2 if node >= 5 else 1 if node >= 1 else 0
    k = node - LEVEL_BASE[L]
    n = 1 << L
    return L, k % n, k // n, 64 >> L


def _plane(t, w, h):
    return MtPlane(_dev_ptr(t).value, t.stride(0), w, h, w, h, 0)


def intra_rdo(src, bd, qidx, qm_level=15, flags=ANGLE_DELTAS | FILTER_INTRA | MIN4, lam=None,
              wsse_scale=1.0, prune=None, recon=None, nodes=None, stream=None, **chroma):
    """RDO of one frame.  src: uint16 cuda tensor (H, W), W and H multiples of 8; prune:
    optional uint16 cuda tensor (nsb, 341) of node masks; chroma keywords as
    intra_rdo_batch's frame dicts.  Returns (recon uint16 (H, W) cuda tensor, nodes uint8
    cuda tensor (nsb, 341 * 16), stats dict)."""
    fr = dict(src=src, bd=bd, qidx=qidx, qm_level=qm_level, lam=lam, prune=prune, recon=recon, nodes=nodes)
    fr.update(chroma)
    return intra_rdo_batch([fr], flags, wsse_scale, stream)[0]


def intra_rdo_batch(frames, flags=ANGLE_DELTAS | FILTER_INTRA | MIN4, wsse_scale=1.0, stream=None):
    """RDO of several frames in one call (mt_intra_rdo_batch).  frames: list of dicts with
    src (16-bit cuda tensor (H, W)), bd, qidx and optionally qm_level (15), lam (the
    reference's decision lambda for 4:2:0), prune, recon, nodes; with chroma: chroma
    (CHROMA_420 or CHROMA_444), src_u and src_v (16-bit cuda tensors of the chroma size),
    optionally uv_dc, uv_ac, uv_qm (default: the reference's laws), recon_u, recon_v;
    levels=True adds the reconstruction's level planes (int32 cuda tensors of the plane sizes,
    zero where no level is stored).
    Returns a list of (recon, nodes, stats) in the same order; with chroma, stats holds
    recon_uv = (recon_u, recon_v), bits_uv, sse_u and sse_v; with levels, stats["levels"] =
    (luma, u, v) (u and v None without chroma)."""
    import torch
    n = len(frames)
    arr = (MtRdoFrame * max(n, 1))()
    outs = []
    for i, fr in enumerate(frames):
        src = fr["src"]
        if src.dtype != torch.int16 and src.dtype != torch.uint16:
            raise _lib.MtError("src must be a 16-bit tensor")
        H, W = src.shape
        dev = src.device
        init(dev.index or 0)
        nsb = ((W + 63) // 64) * ((H + 63) // 64)
        rec = fr.get("recon")
        if rec is None:
            rec = torch.empty((H, W), dtype=src.dtype, device=dev)
        nodes = fr.get("nodes")
        if nodes is None:
            nodes = torch.empty(nsb * NODES * NODE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        ch = int(fr.get("chroma", LUMA_ONLY))
        sub = 2 if ch == CHROMA_420 else 0
        lam = fr.get("lam")
        if lam is None:
            lam = rdo_lambda(fr["qidx"], fr["bd"], sub if ch else 2)
        f = arr[i]
        f.src = _plane(src, W, H)
        f.recon = _dev_ptr(rec).value
        f.recon_stride = rec.stride(0)
        f.bd, f.qidx, f.qm_level = fr["bd"], fr["qidx"], fr.get("qm_level", 15)
        f.lambda_ = lam
        f.wsse_scale = wsse_scale
        pr = fr.get("prune")
        if pr is not None:
            if tuple(pr.shape) != (nsb, NODES):
                raise _lib.MtError("prune must have shape (nsb, 341)")
            f.prune = _dev_ptr(pr).value
        f.nodes = _dev_ptr(nodes).value
        uv = None
        if ch:
            su, sv = fr["src_u"], fr["src_v"]
            Hc, Wc = su.shape
            ru = fr.get("recon_u")
            rv = fr.get("recon_v")
            if ru is None:
                ru = torch.empty((Hc, Wc), dtype=src.dtype, device=dev)
            if rv is None:
                rv = torch.empty((Hc, Wc), dtype=src.dtype, device=dev)
            if ru.stride(0) != rv.stride(0):
                raise _lib.MtError("recon_u and recon_v must share a stride")
            dc, ac, qm = chroma_laws(fr["qidx"], sub)
            f.chroma = ch
            f.src_u = _plane(su, Wc, Hc)
            f.src_v = _plane(sv, Wc, Hc)
            f.recon_u = _dev_ptr(ru).value
            f.recon_v = _dev_ptr(rv).value
            f.recon_uv_stride = ru.stride(0)
            f.uv_dc_delta = fr.get("uv_dc", dc)
            f.uv_ac_delta = fr.get("uv_ac", ac)
            f.uv_qm_level = fr.get("uv_qm", qm)
            uv = (ru, rv)
        lv = None
        if fr.get("levels"):
            ly = torch.zeros((H, W), dtype=torch.int32, device=dev)
            f.coef = _dev_ptr(ly).value
            f.coef_stride = W
            lu = lvv = None
            if ch:
                Hc, Wc = fr["src_u"].shape
                lu = torch.zeros((Hc, Wc), dtype=torch.int32, device=dev)
                lvv = torch.zeros((Hc, Wc), dtype=torch.int32, device=dev)
                f.coef_u = _dev_ptr(lu).value
                f.coef_v = _dev_ptr(lvv).value
                f.coef_uv_stride = Wc
            lv = (ly, lu, lvv)
        outs.append((rec, nodes, uv, lv))
    st = (MtRdoStats * max(n, 1))()
    _check(lib().mt_intra_rdo_batch(arr, n, flags, st, _stream_ptr(stream)))
    res = []
    for i in range(n):
        d = {"bits": st[i].bits, "sse": st[i].sse, "candidates": st[i].candidates, "leaves": st[i].leaves,
             "sbs": st[i].sbs, "steps": st[i].steps}
        if outs[i][2] is not None:
            d.update(recon_uv=outs[i][2], bits_uv=st[i].bits_uv, sse_u=st[i].sse_u, sse_v=st[i].sse_v)
        if outs[i][3] is not None:
            d["levels"] = outs[i][3]
        res.append((outs[i][0], outs[i][1], d))
    return res


def nodes_to_np(nodes):
    """Device node records -> structured numpy array (nsb, 341) of NODE_DTYPE."""
    a = nodes.cpu().numpy().view(NODE_DTYPE)
    return a.reshape(-1, NODES)


def psnr(sse, npx, bd):
    return 10.0 * np.log10(((1 << bd) - 1) ** 2 * npx / max(int(sse), 1))


# ---- the encoder's entry (mt_encode_rgb8) ----------------------------------------------
YUV420, YUV422, YUV444, MONOCHROME = 0, 1, 2, 3  # ChromaFormat (src/avif.rs:77)
SPEED_EXTRA_SLOW, SPEED_SLOW, SPEED_MEDIUM, SPEED_FAST = 0, 1, 2, 3  # Speed (src/avif.rs:91)
NOT_APPLIED = {1: "adaptive_quant", 2: "variance_boost", 4: "dark_aq", 8: "cdef", 16: "wiener",
               32: "screen_content", 64: "intrabc", 128: "updating_cdf", 256: "extra_slow", 512: "deblock"}


class MtEncodeConfig(ctypes.Structure):
    _fields_ = [("quality", ctypes.c_uint8), ("chroma", ctypes.c_uint8), ("speed", ctypes.c_uint8),
                ("quantization_matrices", ctypes.c_uint8), ("qmatrix_level", ctypes.c_int16),
                ("adaptive_quant", ctypes.c_uint8), ("variance_boost", ctypes.c_uint8),
                ("dark_aq", ctypes.c_uint8), ("cdef", ctypes.c_uint8), ("wiener", ctypes.c_uint8),
                ("screen_content", ctypes.c_uint8), ("intrabc", ctypes.c_uint8),
                ("updating_cdf", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 2)]


def _cdef_params_type():
    from .cdef import MtCdefParams
    return MtCdefParams


class MtEncodeOut(ctypes.Structure):
    _fields_ = [("y", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("recon_y", ctypes.c_void_p), ("recon_u", ctypes.c_void_p), ("recon_v", ctypes.c_void_p),
                ("nodes", ctypes.c_void_p), ("stats", MtRdoStats), ("qidx", ctypes.c_uint8),
                ("qm_level", ctypes.c_uint8), ("uv_qm_level", ctypes.c_uint8), ("pad0", ctypes.c_uint8),
                ("rdo_flags", ctypes.c_int32), ("not_applied", ctypes.c_uint32), ("pad1", ctypes.c_int32),
                ("dbk_y", ctypes.c_void_p), ("dbk_u", ctypes.c_void_p), ("dbk_v", ctypes.c_void_p),
                ("lf_level_y", ctypes.c_uint8), ("lf_level_u", ctypes.c_uint8), ("lf_level_v", ctypes.c_uint8),
                ("lf_sharpness", ctypes.c_uint8), ("dbk_status", ctypes.c_uint32),
                ("cdef_y", ctypes.c_void_p), ("cdef_u", ctypes.c_void_p), ("cdef_v", ctypes.c_void_p),
                ("cdef", _cdef_params_type()), ("cdef_grid", ctypes.c_void_p)]


def encode_config(**overrides):
    """EncodeConfig's defaults (mt_encode_config_default) with keyword overrides."""
    cfg = MtEncodeConfig()
    lib().mt_encode_config_default(ctypes.byref(cfg))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def quality_to_qindex(quality):
    return int(lib().mt_quality_to_qindex(int(quality)))


def encode_rgb8(g, b, r, cfg=None, stream=None, deblock=True, cdef=True):
    """mt_encode_rgb8: g, b, r uint8 cuda tensors (H, W) (PlanarImage's plane order), W and H
    multiples of 8.  Returns a dict: y, u, v (the converted planes), recon_y, recon_u,
    recon_v, nodes (uint8 (nsb, 341 * 16)), stats, qidx, qm_level, uv_qm_level, rdo_flags
    and not_applied (the names of the config switches this library does not run); with
    deblock, also dbk_y, dbk_u, dbk_v (the reconstruction after the deblocking filter),
    lf_levels (y, u, v, sharpness) and dbk_status; with deblock and cdef, also cdef_y,
    cdef_u, cdef_v (after CDEF, where the config runs it), cdef (the header values) and
    cdef_grid."""
    import torch
    cfg = cfg if cfg is not None else encode_config()
    H, W = g.shape
    dev = g.device
    init(dev.index or 0)
    c420 = cfg.chroma == YUV420
    Hc, Wc = (H // 2, W // 2) if c420 else (H, W)
    nsb = ((W + 63) // 64) * ((H + 63) // 64)
    t = {k: torch.empty(s, dtype=torch.int16, device=dev)
         for k, s in (("y", (H, W)), ("u", (Hc, Wc)), ("v", (Hc, Wc)), ("recon_y", (H, W)),
                      ("recon_u", (Hc, Wc)), ("recon_v", (Hc, Wc)))
                      + ((("dbk_y", (H, W)), ("dbk_u", (Hc, Wc)), ("dbk_v", (Hc, Wc))) if deblock else ())
                      + ((("cdef_y", (H, W)), ("cdef_u", (Hc, Wc)), ("cdef_v", (Hc, Wc))) if deblock and cdef
                         else ())}
    nodes = torch.empty(nsb * NODES * NODE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    out = MtEncodeOut()
    for k, v in t.items():
        setattr(out, k, _dev_ptr(v).value)
    out.nodes = _dev_ptr(nodes).value
    grid = np.zeros(nsb, np.uint8)
    out.cdef_grid = grid.ctypes.data
    _check(lib().mt_encode_rgb8(_dev_ptr(g), _dev_ptr(b), _dev_ptr(r), W, H, ctypes.byref(cfg), ctypes.byref(out),
                                _stream_ptr(stream)))
    st = out.stats
    res = dict(t)
    res.update(nodes=nodes, qidx=out.qidx, qm_level=out.qm_level, uv_qm_level=out.uv_qm_level,
               rdo_flags=out.rdo_flags, dbk_status=out.dbk_status, cdef=out.cdef.as_dict(), cdef_grid=grid,
               lf_levels=(out.lf_level_y, out.lf_level_u, out.lf_level_v, out.lf_sharpness),
               not_applied=[n for bit, n in NOT_APPLIED.items() if out.not_applied & bit],
               stats={"bits": st.bits, "sse": st.sse, "candidates": st.candidates, "leaves": st.leaves,
                      "sbs": st.sbs, "steps": st.steps, "bits_uv": st.bits_uv, "sse_u": st.sse_u,
                      "sse_v": st.sse_v})
    return res
