"""ctypes binding of libmtgpu.so (include/mtgpu.h)."""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(PKG, "libmtgpu.so")

# AV1 TxSize numbering; the 13 sizes the reference has a forward transform for.
TX_NAMES = ["4x4", "8x8", "16x16", "32x32", "64x64", "4x8", "8x4", "8x16", "16x8", "16x32",
            "32x16", "32x64", "64x32", "4x16", "16x4", "8x32", "32x8", "16x64", "64x16"]
TX_SIZES_REF = [0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4]
TX_SIZES_ALL = list(range(19))
TXTP_NAMES = {0: "DCT_DCT", 1: "ADST_DCT", 2: "DCT_ADST", 3: "ADST_ADST", 9: "IDTX",
              10: "V_DCT", 11: "H_DCT"}
TX_TYPES_REF = [0, 1, 2, 3, 9, 10, 11]


class MtError(RuntimeError):
    pass


class MtQuant(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "dc", "ac", "q_mult_dc", "q_mult_ac", "rmin", "rmax", "cmin", "cmax", "cf_max",
        "qm_level", "qm_chroma", "qidx", "bd")]


class MtTxBlk(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint32), ("tx_size", ctypes.c_uint8),
                ("tx_type", ctypes.c_uint8), ("qidx", ctypes.c_uint8), ("bd_qm", ctypes.c_uint8)]


class MtTxRes(ctypes.Structure):
    _fields_ = [("sse", ctypes.c_int64), ("nnz", ctypes.c_uint32), ("hash", ctypes.c_uint32)]


RATEBLK_DTYPE = np.dtype([("offset", "<u4"), ("tx_size", "u1"), ("plane", "u1"), ("eob_bin", "u1"),
                          ("skip_ctx", "u1"), ("dcs_ctx", "u1"), ("txtp_set", "u1"), ("txtp_mode", "u1"),
                          ("txtp", "u1"), ("qctx", "u1"), ("flags", "u1"), ("pad0", "u1"), ("pad1", "u1"),
                          ("skip_one", "<f4"), ("skip_zero", "<f4"), ("bound", "<f4"), ("pad2", "<u4")])
BLK_DTYPE = np.dtype([("offset", "<u4"), ("tx_size", "u1"), ("tx_type", "u1"), ("qidx", "u1"),
                      ("bd_qm", "u1")])
RES_DTYPE = np.dtype([("sse", "<i8"), ("nnz", "<u4"), ("hash", "<u4")])
QUANT_DTYPE = np.dtype([(n, "<i4") for n, _ in MtQuant._fields_])
PREDBLK_DTYPE = np.dtype([("out_offset", "<u4"), ("x", "<u2"), ("y", "<u2"), ("bw", "u1"),
                          ("bh", "u1"), ("mode", "u1"), ("angle_delta", "i1"), ("flags", "u1"),
                          ("bd", "u1"), ("plane", "u1"), ("pad", "u1")])
DISTBLK_DTYPE = np.dtype([("pred_offset", "<u4"), ("resid_offset", "<u4"), ("x", "<u2"), ("y", "<u2"),
                          ("bw", "u1"), ("bh", "u1"), ("bd", "u1"), ("plane", "u1"), ("flags", "u1"),
                          ("pad", "u1", 3), ("wsse_scale", "<f4")])
DISTRES_DTYPE = np.dtype([("sse", "<i8"), ("dist", "<i8"), ("satd", "<u8"), ("proxy", "<u8")])
DIST_NO_RESID = 1
CFLBLK_DTYPE = np.dtype([("out_offset", "<u4"), ("lx", "<u2"), ("ly", "<u2"), ("x", "<u2"), ("y", "<u2"),
                         ("cw", "u1"), ("ch", "u1"), ("flags", "u1"), ("bd", "u1"), ("luma_plane", "u1"),
                         ("src_plane", "u1"), ("alpha", "i1"), ("pad", "u1"), ("dc", "<i4")])
CFL_SS_HOR, CFL_SS_VER, CFL_ALPHA_GIVEN = 1, 2, 4
CAND_DTYPE = np.dtype([("offset", "<u4"), ("x", "<u2"), ("y", "<u2"), ("tx_size", "u1"), ("tx_type", "u1"),
                       ("mode", "u1"), ("angle_delta", "i1"), ("pflags", "u1"), ("planes", "u1"),
                       ("qidx", "u1"), ("bd_qm", "u1"), ("rplane", "u1"), ("eob_bin", "u1"),
                       ("skip_ctx", "u1"), ("dcs_ctx", "u1"), ("txtp_set", "u1"), ("txtp_mode", "u1"),
                       ("txtp", "u1"), ("qctx", "u1"), ("rflags", "u1"), ("planes_hi", "u1"),
                       ("dc_delta", "i1"), ("ac_delta", "i1"),
                       ("skip_one", "<f4"), ("skip_zero", "<f4"), ("wsse_scale", "<f4"),
                       ("lambda", "<f4"), ("extra_bits", "<f4")])
CANDRES_DTYPE = np.dtype([("sse", "<i8"), ("dist", "<i8"), ("satd", "<u8"), ("bits", "<f4"),
                          ("cost", "<f4"), ("nnz", "<u4"), ("status", "<i4"), ("alpha", "<i4"),
                          ("pad", "<i4")])
COEF_CDFS_DTYPE = np.dtype([("skip", "<u2", (5, 13)), ("base_tok", "<u2", (5, 2, 42, 3)),
                            ("br_tok", "<u2", (4, 2, 21, 3)), ("eob_base", "<u2", (5, 2, 4, 2)),
                            ("eob_hi", "<u2", (5, 2, 11)), ("dc_sign", "<u2", (2, 3)),
                            ("eob_bin", "<u2", (17, 10)), ("txtp4", "<u2", (13, 6)),
                            ("txtp8", "<u2", (13, 6)), ("txtp16", "<u2", (13, 4)), ("pad", "<u2")])
EOB_BIN_K = [4, 4, 4, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 9, 9, 10, 10]
RATE_SLOTS_FIXED, RATE_SLOTS = 4, 8
PRED_HAVE_TR, PRED_HAVE_BL, PRED_SMOOTH_NB = 1, 2, 4
TXBLK_CHROMA = 0x40
PRED_FILTER0 = 16
PRED_CFL = 13
CAND_CFL_SS_HOR, CAND_CFL_SS_VER, CAND_CFL_ALPHA_GIVEN = 8, 16, 32
MAX_PLANES = 64


class MtPlane(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", ctypes.c_int32), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("vis_w", ctypes.c_int32), ("vis_h", ctypes.c_int32),
                ("pad", ctypes.c_int32)]

_lib = None


def lib():
    """The loaded library; raises MtError when libmtgpu.so has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise MtError(f"libmtgpu.so not built ({lib_path}); run __graft_entry__.build()")
        L = ctypes.CDLL(lib_path)
        P, vp, i32, u64 = ctypes.POINTER, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
        L.mt_init.argtypes = [i32]
        L.mt_last_error.restype = ctypes.c_char_p
        L.mt_version.restype = ctypes.c_char_p
        L.mt_quant_new.argtypes = [i32, i32, i32, P(MtQuant)]
        L.mt_quant_new_chroma.argtypes = [i32, i32, i32, i32, i32, P(MtQuant)]
        L.mt_qm_level_law.argtypes = [i32, i32]
        L.mt_tx_width.argtypes = L.mt_tx_height.argtypes = L.mt_tx_ncoef.argtypes = [i32]
        L.mt_tx_supported.argtypes = [i32, i32]
        L.mt_fwd_txfm.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp]
        L.mt_inv_txfm.argtypes = [i32, i32, i32, vp, vp, vp, vp]
        L.mt_tx_chain.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp]
        L.mt_tx_chain_noalloc.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp]
        L.mt_trellis.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.mt_tx_chain_trellis.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, vp]
        L.mt_trellis_lambda0.argtypes = [i32, i32, i32]
        L.mt_trellis_lambda0.restype = ctypes.c_float
        L.mt_profile_end.argtypes = [P(ctypes.c_double), P(ctypes.c_int), i32]
        f32 = ctypes.c_float
        L.mt_luma_satd_scale.argtypes = L.mt_k_satd.argtypes = [i32, i32]
        L.mt_luma_satd_scale.restype = L.mt_k_satd.restype = L.mt_proxy_wsse.restype = f32
        L.mt_proxy_wsse.argtypes = [f32]
        L.mt_luma_partition_distortion.argtypes = [vp, vp, f32, i32, i32, i32, vp, vp]
        L.mt_sb_activity.argtypes = [MtPlane, i32, vp, vp, vp]
        L.mt_perceptual_rd_scale.argtypes = [vp, i32, MtPlane, i32, i32, i32, i32, vp, vp, vp, vp]
        L.mt_tx_schedule_set.argtypes = [i32, i32, i32, i32]
        L.mt_tx_schedule_get.argtypes = [i32, i32, P(ctypes.c_int), P(ctypes.c_int)]
        L.mt_block_rate.argtypes = [vp, i32, vp, vp, vp]
        L.mt_intra_pred.argtypes = [vp, i32, P(MtPlane), i32, vp, vp, vp]
        L.mt_distortion.argtypes = [vp, i32, P(MtPlane), i32, vp, vp, vp, vp]
        L.mt_cfl.argtypes = [vp, i32, P(MtPlane), i32, vp, vp, vp, vp]
        L.mt_eval_candidates.argtypes = [vp, i32, P(MtPlane), i32, u64, vp, vp]
        L.mt_eval_candidates_trellis.argtypes = [vp, i32, P(MtPlane), i32, u64, vp, vp]
        L.mt_gen_pool_size.argtypes = [u64, i32, i32]
        L.mt_gen_pool_size.restype = u64
        L.mt_gen_tx_blocks.argtypes = [u64, i32, i32, vp, vp, u64]
        L.mt_intra_rdo.argtypes = [vp, i32, vp, vp]
        L.mt_rate_default_cdfs.argtypes = [i32, vp]
        L.mt_rate_load_cdfs.argtypes = [i32, vp]
        L.mt_intra_rdo_batch.argtypes = [vp, i32, i32, vp, vp]
        L.mt_rdo_lambda.argtypes = L.mt_mode_lambda.argtypes = [i32, i32, i32]
        L.mt_rdo_lambda.restype = L.mt_mode_lambda.restype = ctypes.c_float
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MtError(f"mtgpu error {rc}: {lib().mt_last_error().decode()}")


_inited = set()


def init(device=0):
    """mt_init on `device` (idempotent); requires a GPU."""
    if device not in _inited:
        _check(lib().mt_init(device))
        _inited.add(device)


def quant_new(qidx, bd=8, qm_level=15):
    q = MtQuant()
    _check(lib().mt_quant_new(qidx, bd, qm_level, ctypes.byref(q)))
    return q


def quant_new_chroma(qidx, dc_delta, ac_delta, bd, qm_level):
    q = MtQuant()
    _check(lib().mt_quant_new_chroma(qidx, dc_delta, ac_delta, bd, qm_level, ctypes.byref(q)))
    return q


def quant_to_np(qs):
    """List of MtQuant -> structured numpy array (QUANT_DTYPE)."""
    arr = np.zeros(len(qs), QUANT_DTYPE)
    for i, q in enumerate(qs):
        arr[i] = tuple(getattr(q, n) for n, _ in MtQuant._fields_)
    return arr


def qm_level_law(qidx, sub=0):
    return lib().mt_qm_level_law(qidx, sub)


def tx_width(s):
    return lib().mt_tx_width(s)


def tx_height(s):
    return lib().mt_tx_height(s)


def tx_ncoef(s):
    return lib().mt_tx_ncoef(s)


def tx_supported(s, t):
    return bool(lib().mt_tx_supported(s, t))


def _dev_ptr(t):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise MtError("expected a torch tensor on the GPU")
    if not t.is_contiguous():
        raise MtError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _quants_tensor(quants, device):
    import torch
    arr = quants if isinstance(quants, np.ndarray) else quant_to_np(quants)
    return torch.from_numpy(arr.view(np.int32).reshape(-1).copy()).to(device)


def fwd_txfm(tx_size, tx_type, resid, quants, stream=None):
    """Forward transform + quant of n blocks: resid int32 (n, W*H) cuda tensor, quants a list
    of MtQuant (or QUANT_DTYPE array).  Returns (levels int32 (n, ncoef), targets f32)."""
    import torch
    n = resid.shape[0]
    nc = tx_ncoef(tx_size)
    dev = resid.device
    init(dev.index or 0)
    qt = _quants_tensor(quants, dev)
    lv = torch.empty((n, nc), dtype=torch.int32, device=dev)
    tg = torch.empty((n, nc), dtype=torch.float32, device=dev)
    _check(lib().mt_fwd_txfm(tx_size, tx_type, n, _dev_ptr(resid.contiguous()), _dev_ptr(qt),
                             _dev_ptr(lv), _dev_ptr(tg), _stream_ptr(stream)))
    return lv, tg


def inv_txfm(tx_size, tx_type, levels, quants, stream=None):
    """Dequant + inverse of n blocks: levels int32 (n, ncoef) -> residual int32 (n, W*H)."""
    import torch
    n = levels.shape[0]
    dev = levels.device
    init(dev.index or 0)
    qt = _quants_tensor(quants, dev)
    out = torch.empty((n, tx_width(tx_size) * tx_height(tx_size)), dtype=torch.int32, device=dev)
    _check(lib().mt_inv_txfm(tx_size, tx_type, n, _dev_ptr(levels.contiguous()), _dev_ptr(qt),
                             _dev_ptr(out), _stream_ptr(stream)))
    return out


def tx_chain(blocks, pool, res, levels=None, targets=None, recon=None, stream=None):
    """Run the batched chain.  blocks: uint8 cuda tensor of n*8 bytes (BLK_DTYPE records),
    pool: int16 cuda tensor, res: uint8 cuda tensor of n*16 bytes (RES_DTYPE)."""
    n = blocks.numel() // BLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    opt = lambda t: _dev_ptr(t) if t is not None else None  # noqa: E731
    _check(lib().mt_tx_chain(_dev_ptr(blocks), n, _dev_ptr(pool), _dev_ptr(res), opt(levels),
                             opt(targets), opt(recon), _stream_ptr(stream)))


NOALLOC_MAX_BLOCKS = 1 << 20
TRELLIS_CTX = 1
CAND_TRELLIS = 2
TRELLISBLK_DTYPE = np.dtype([("offset", "<u4"), ("tx_size", "u1"), ("qidx", "u1"), ("bd_qm", "u1"),
                             ("flags", "u1"), ("dc_delta", "i1"), ("ac_delta", "i1"), ("qctx", "u1"),
                             ("eob_bin", "u1"), ("dcs_ctx", "u1"), ("plane", "u1"), ("pad0", "u1"),
                             ("pad1", "u1"), ("lambda0", "<f4"), ("band_tilt", "<f4")])


def trellis(blocks, levels, targets, work=None, status=None, stream=None):
    """mt_trellis: blocks uint8 cuda tensor of TRELLISBLK_DTYPE records; levels int32 (rewritten
    in place) and targets float32 cuda tensors laid out alike; work is not used."""
    n = blocks.numel() // TRELLISBLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    _check(lib().mt_trellis(_dev_ptr(blocks), n, _dev_ptr(levels), _dev_ptr(targets),
                            _dev_ptr(work) if work is not None else None,
                            _dev_ptr(status) if status is not None else None, _stream_ptr(stream)))


def tx_chain_trellis(blocks, trellis_blocks, pool, res, levels, targets, recon, stream=None):
    """mt_tx_chain_trellis: the chain with the trellis between quantization and
    dequantization; trellis_blocks: uint8 cuda tensor of TRELLISBLK_DTYPE records (one per
    block); levels / targets / recon required (laid out like the pool)."""
    n = blocks.numel() // BLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    _check(lib().mt_tx_chain_trellis(_dev_ptr(blocks), _dev_ptr(trellis_blocks), n, _dev_ptr(pool), _dev_ptr(res),
                                     _dev_ptr(levels), _dev_ptr(targets), _dev_ptr(recon), _stream_ptr(stream)))


def trellis_lambda0(qidx, bd, sub):
    return float(lib().mt_trellis_lambda0(qidx, bd, sub))


def tx_chain_noalloc(blocks, pool, res, levels=None, targets=None, recon=None, stream=None):
    """mt_tx_chain_noalloc: tx_chain on the library's fixed device workspace (no allocation;
    at most NOALLOC_MAX_BLOCKS blocks; calls overlapping on one device race)."""
    n = blocks.numel() // BLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    opt = lambda t: _dev_ptr(t) if t is not None else None  # noqa: E731
    _check(lib().mt_tx_chain_noalloc(_dev_ptr(blocks), n, _dev_ptr(pool), _dev_ptr(res), opt(levels),
                                     opt(targets), opt(recon), _stream_ptr(stream)))


PRDOBLK_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("dim", "<u2"), ("pad", "<u2")])


def luma_satd_scale(base_q_idx, subsampled):
    return lib().mt_luma_satd_scale(base_q_idx, int(subsampled))


def k_satd(base_q_idx, subsampled):
    return lib().mt_k_satd(base_q_idx, int(subsampled))


def proxy_wsse(d):
    return lib().mt_proxy_wsse(d)


def luma_partition_distortion(satd, qstep, base_q_idx, subsampled, stream=None):
    """SATD partition prices (float32 cuda tensor) of an int64/uint64 cuda tensor of SATD
    values; qstep is a float or a float32 cuda tensor of the same length."""
    import torch
    n = satd.numel()
    init(satd.device.index or 0)
    out = torch.empty(n, dtype=torch.float32, device=satd.device)
    qd = _dev_ptr(qstep) if isinstance(qstep, torch.Tensor) else None
    qs = 0.0 if isinstance(qstep, torch.Tensor) else float(qstep)
    _check(lib().mt_luma_partition_distortion(_dev_ptr(satd), qd, qs, n, base_q_idx, int(subsampled),
                                              _dev_ptr(out), _stream_ptr(stream)))
    return out


def _one_plane(t, width=None, height=None):
    h, stride = t.shape
    w = width or stride
    hh = height or h
    return MtPlane(_dev_ptr(t), stride, w, hh, w, hh, 0)


def sb_activity(plane, bd, width=None, height=None, stream=None):
    """(sb_activity per superblock, tile_ref_activity) of a 2-D uint16 cuda plane, both
    float32 cuda tensors."""
    import torch
    init(plane.device.index or 0)
    pl = _one_plane(plane, width, height)
    nsb = ((pl.width + 63) // 64) * ((pl.height + 63) // 64)
    act = torch.empty(nsb, dtype=torch.float32, device=plane.device)
    ref = torch.empty(1, dtype=torch.float32, device=plane.device)
    _check(lib().mt_sb_activity(pl, bd, _dev_ptr(act), _dev_ptr(ref), _stream_ptr(stream)))
    return act, ref


def perceptual_rd_scale(blocks, plane, bd, base_q_idx, fmt, sb_act, ref_act=None, mono_large=False,
                        width=None, height=None, stream=None):
    """perceptual_rd_scale of PRDOBLK_DTYPE blocks (uint8 cuda tensor) as float32 cuda tensor."""
    import torch
    n = blocks.numel() // PRDOBLK_DTYPE.itemsize
    init(plane.device.index or 0)
    out = torch.empty(n, dtype=torch.float32, device=plane.device)
    _check(lib().mt_perceptual_rd_scale(_dev_ptr(blocks), n, _one_plane(plane, width, height), bd, base_q_idx,
                                        fmt, int(mono_large), _dev_ptr(ref_act) if ref_act is not None else None,
                                        _dev_ptr(sb_act), _dev_ptr(out), _stream_ptr(stream)))
    return out


def tx_schedule_set(tx_size, tile_shift, wg_per_cu, device=0):
    """Installs the chain's launch schedule of one transform size (mt_tx_schedule_set)."""
    init(device)
    _check(lib().mt_tx_schedule_set(device, tx_size, tile_shift, wg_per_cu))


def tx_schedule_get(tx_size, device=0):
    """(tile_shift, wg_per_cu) of one transform size on `device`."""
    init(device)
    ts, wg = ctypes.c_int(), ctypes.c_int()
    _check(lib().mt_tx_schedule_get(device, tx_size, ctypes.byref(ts), ctypes.byref(wg)))
    return ts.value, wg.value


PROFILE_SLOTS = ["bucket", "4x4", "4x8", "8x4", "4x16", "16x4", "8x8", "8x16", "16x8", "16x16",
                 "16x32", "32x16", "32x32", "64x64", "rate", "intra_pred", "distortion", "eval",
                 "rdo_decide", "rdo_recon", "trellis"]


def profile_begin():
    _check(lib().mt_profile_begin())


def profile_end():
    """{slot name: (total ms, launches)} of the chain launches since profile_begin()."""
    n = len(PROFILE_SLOTS)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_int * n)()
    _check(lib().mt_profile_end(ms, cnt, n))
    return {PROFILE_SLOTS[i]: (ms[i], cnt[i]) for i in range(n)}


def block_rate(blocks, levels, stream=None):
    """Coefficient rate of n blocks: blocks is a uint8 cuda tensor of n*32 bytes
    (RATEBLK_DTYPE records), levels an int32 cuda tensor; returns float32 bits (n,)."""
    import torch
    n = blocks.numel() // RATEBLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    out = torch.empty(n, dtype=torch.float32, device=blocks.device)
    _check(lib().mt_block_rate(_dev_ptr(blocks), n, _dev_ptr(levels), _dev_ptr(out), _stream_ptr(stream)))
    return out


def rate_default_cdfs(qctx):
    """The coefficient CDFs of Cdfs::new(qctx) as a COEF_CDFS_DTYPE record."""
    out = np.zeros(1, COEF_CDFS_DTYPE)
    _check(lib().mt_rate_default_cdfs(qctx, out.ctypes.data_as(ctypes.c_void_p)))
    return out[0]


def rate_load_cdfs(slot, cdfs, device=0):
    """Loads a COEF_CDFS_DTYPE snapshot into rate slot 4..7 of `device`."""
    init(device)
    rec = np.ascontiguousarray(np.asarray(cdfs, COEF_CDFS_DTYPE).reshape(1))
    _check(lib().mt_rate_load_cdfs(slot, rec.ctypes.data_as(ctypes.c_void_p)))


def gen_tx_blocks(seed, n, all_sizes=False):
    """Synthetic microbench input on the host: (blocks structured array, int16 pool)."""
    L = lib()
    total = L.mt_gen_pool_size(seed, n, int(all_sizes))
    blocks = np.zeros(n, BLK_DTYPE)
    pool = np.zeros(max(total, 1), np.int16)
    _check(L.mt_gen_tx_blocks(seed, n, int(all_sizes), blocks.ctypes.data_as(ctypes.c_void_p),
                              pool.ctypes.data_as(ctypes.c_void_p), total))
    return blocks, pool


def plane_views(planes):
    """ctypes array of mt_plane from [(uint16 cuda tensor 2-D, width, height, vis_w, vis_h)].
    width/height default to the tensor's shape, vis_w/vis_h to width/height."""
    arr = (MtPlane * max(1, len(planes)))()
    for i, p in enumerate(planes):
        t = p[0]
        h, stride = t.shape
        w = p[1] if len(p) > 1 else stride
        hh = p[2] if len(p) > 2 else h
        vw = p[3] if len(p) > 3 else w
        vh = p[4] if len(p) > 4 else hh
        arr[i] = MtPlane(_dev_ptr(t), stride, w, hh, vw, vh, 0)
    return arr


def intra_pred(blocks, planes, pred, status=None, stream=None):
    """Intra prediction of n blocks.  blocks: uint8 cuda tensor of n*16 bytes (PREDBLK_DTYPE
    records); planes: list as plane_views(); pred: int32 cuda tensor receiving each block's
    prediction at its out_offset; status: optional int32 cuda tensor (n,)."""
    n = blocks.numel() // PREDBLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    pv = plane_views(planes)
    _check(lib().mt_intra_pred(_dev_ptr(blocks), n, pv, len(planes), _dev_ptr(pred),
                               _dev_ptr(status) if status is not None else None,
                               _stream_ptr(stream)))


def distortion(blocks, planes, pred, resid=None, stream=None):
    """Distortion measures of n blocks.  blocks: uint8 cuda tensor of n*24 bytes
    (DISTBLK_DTYPE records); returns a uint8 cuda tensor of n*32 bytes (DISTRES_DTYPE)."""
    import torch
    n = blocks.numel() // DISTBLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    out = torch.empty(n * DISTRES_DTYPE.itemsize, dtype=torch.uint8, device=blocks.device)
    pv = plane_views(planes)
    _check(lib().mt_distortion(_dev_ptr(blocks), n, pv, len(planes), _dev_ptr(pred),
                               _dev_ptr(resid) if resid is not None else None, _dev_ptr(out),
                               _stream_ptr(stream)))
    return out


def eval_candidates(cands, planes, pool_elems, stream=None, out=None, trellis=False):
    """Evaluate n intra candidates.  cands: uint8 cuda tensor of n*48 bytes (CAND_DTYPE
    records); planes: list as plane_views(); returns a uint8 cuda tensor of n*48 bytes
    (CANDRES_DTYPE).  trellis: mt_eval_candidates_trellis (candidates with MT_CAND_TRELLIS
    in rflags run the trellis)."""
    import torch
    n = cands.numel() // CAND_DTYPE.itemsize
    init(cands.device.index or 0)
    if out is None:
        out = torch.empty(n * CANDRES_DTYPE.itemsize, dtype=torch.uint8, device=cands.device)
    pv = plane_views(planes)
    fn = lib().mt_eval_candidates_trellis if trellis else lib().mt_eval_candidates
    _check(fn(_dev_ptr(cands), n, pv, len(planes), int(pool_elems), _dev_ptr(out), _stream_ptr(stream)))
    return out


def cfl(blocks, planes, pred, ac=None, stream=None):
    """Chroma from luma for n blocks (CFLBLK_DTYPE records in a uint8 cuda tensor); writes
    the predictions (and the AC) at each out_offset; returns the alphas (int32 cuda (n,))."""
    import torch
    n = blocks.numel() // CFLBLK_DTYPE.itemsize
    init(blocks.device.index or 0)
    alpha = torch.empty(n, dtype=torch.int32, device=blocks.device)
    pv = plane_views(planes)
    _check(lib().mt_cfl(_dev_ptr(blocks), n, pv, len(planes), _dev_ptr(pred),
                        _dev_ptr(ac) if ac is not None else None, _dev_ptr(alpha), _stream_ptr(stream)))
    return alpha
