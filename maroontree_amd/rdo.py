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

ANGLE_DELTAS, FILTER_INTRA, MIN4 = 1, 2, 4
NODES = 341
ALL_MODES, FORCE_SPLIT, NO_SPLIT = 0x3FFF, 0x4000, 0x8000
LEVEL_BASE = [0, 1, 5, 21, 85]
NODE_DTYPE = np.dtype([("cost", "<f4"), ("none_cost", "<f4"), ("split", "u1"), ("mode", "u1"),
                       ("pad", "u1", (2,))])


class MtRdoFrame(ctypes.Structure):
    _fields_ = [("src", MtPlane), ("recon", ctypes.c_void_p), ("recon_stride", ctypes.c_int32),
                ("bd", ctypes.c_uint8), ("qidx", ctypes.c_uint8), ("qm_level", ctypes.c_uint8),
                ("pad0", ctypes.c_uint8), ("lambda_", ctypes.c_float), ("wsse_scale", ctypes.c_float),
                ("prune", ctypes.c_void_p), ("nodes", ctypes.c_void_p)]


class MtRdoStats(ctypes.Structure):
    _fields_ = [("bits", ctypes.c_double), ("sse", ctypes.c_int64), ("candidates", ctypes.c_int64),
                ("leaves", ctypes.c_int32), ("sbs", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


def rdo_lambda(qidx, bd, sub=2):
    """The reference's decision lambda lambda_d (mt_rdo_lambda)."""
    return float(lib().mt_rdo_lambda(qidx, bd, sub))


def mode_lambda(qidx, bd, sub=2):
    return float(lib().mt_mode_lambda(qidx, bd, sub))


def node_geometry(node):
    """(level, ix, iy, size) of quadtree node `node` of a superblock."""
    L = 4 if node >= 85 else 3 if node >= 21 else 2 if node >= 5 else 1 if node >= 1 else 0
    k = node - LEVEL_BASE[L]
    n = 1 << L
    return L, k % n, k // n, 64 >> L


def intra_rdo(src, bd, qidx, qm_level=15, flags=ANGLE_DELTAS | FILTER_INTRA | MIN4, lam=None,
              wsse_scale=1.0, prune=None, recon=None, nodes=None, stream=None):
    """RDO of one frame.  src: uint16 cuda tensor (H, W), W and H multiples of 8; prune:
    optional uint16 cuda tensor (nsb, 341) of node masks.  Returns (recon uint16 (H, W) cuda
    tensor, nodes uint8 cuda tensor (nsb, 341 * 12), stats dict)."""
    import torch
    if src.dtype != torch.int16 and src.dtype != torch.uint16:
        raise _lib.MtError("src must be a 16-bit tensor")
    H, W = src.shape
    dev = src.device
    init(dev.index or 0)
    nsb = ((W + 63) // 64) * ((H + 63) // 64)
    if recon is None:
        recon = torch.empty((H, W), dtype=src.dtype, device=dev)
    if nodes is None:
        nodes = torch.empty(nsb * NODES * NODE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    if lam is None:
        lam = rdo_lambda(qidx, bd, 2)
    f = MtRdoFrame()
    f.src = MtPlane(_dev_ptr(src).value, src.stride(0), W, H, W, H, 0)
    f.recon = _dev_ptr(recon).value
    f.recon_stride = recon.stride(0)
    f.bd, f.qidx, f.qm_level = bd, qidx, qm_level
    f.lambda_ = lam
    f.wsse_scale = wsse_scale
    if prune is not None:
        if tuple(prune.shape) != (nsb, NODES) or prune.dtype not in (torch.int16, torch.uint16):
            raise _lib.MtError("prune must be a 16-bit tensor of shape (nsb, 341)")
        f.prune = _dev_ptr(prune).value
    f.nodes = _dev_ptr(nodes).value
    st = MtRdoStats()
    _check(lib().mt_intra_rdo(ctypes.byref(f), flags, ctypes.byref(st), _stream_ptr(stream)))
    return recon, nodes, {"bits": st.bits, "sse": st.sse, "candidates": st.candidates,
                          "leaves": st.leaves, "sbs": st.sbs, "steps": st.steps}


def intra_rdo_batch(frames, flags=ANGLE_DELTAS | FILTER_INTRA | MIN4, wsse_scale=1.0, stream=None):
    """RDO of several frames in one call (mt_intra_rdo_batch).  frames: list of dicts with
    src (16-bit cuda tensor (H, W)), bd, qidx and optionally qm_level (15), lam (the
    reference's decision lambda for 4:2:0), prune, recon, nodes.  Returns a list of
    (recon, nodes, stats) in the same order."""
    import torch
    n = len(frames)
    arr = (MtRdoFrame * max(n, 1))()
    outs = []
    for i, fr in enumerate(frames):
        src = fr["src"]
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
        lam = fr.get("lam")
        if lam is None:
            lam = rdo_lambda(fr["qidx"], fr["bd"], 2)
        f = arr[i]
        f.src = MtPlane(_dev_ptr(src).value, src.stride(0), W, H, W, H, 0)
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
        outs.append((rec, nodes))
    st = (MtRdoStats * max(n, 1))()
    _check(lib().mt_intra_rdo_batch(arr, n, flags, st, _stream_ptr(stream)))
    return [(outs[i][0], outs[i][1], {"bits": st[i].bits, "sse": st[i].sse, "candidates": st[i].candidates,
                                       "leaves": st[i].leaves, "sbs": st[i].sbs, "steps": st[i].steps})
            for i in range(n)]


def nodes_to_np(nodes):
    """Device node records -> structured numpy array (nsb, 341) of NODE_DTYPE."""
    a = nodes.cpu().numpy().view(NODE_DTYPE)
    return a.reshape(-1, NODES)


def psnr(sse, npx, bd):
    return 10.0 * np.log10(((1 << bd) - 1) ** 2 * npx / max(int(sse), 1))
