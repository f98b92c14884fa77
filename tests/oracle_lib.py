"""ctypes binding of the CPU oracle (oracle/liboracle.so).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class OrcQuant(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "dc", "ac", "q_mult_dc", "q_mult_ac", "rmin", "rmax", "cmin", "cmax", "cf_max",
        "qm_level", "qm_chroma", "qidx", "bd")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        i32p, f32p = P(ctypes.c_int32), P(ctypes.c_float)
        L.orc_quant_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(OrcQuant)]
        L.orc_quant_new_chroma.argtypes = [ctypes.c_int] * 5 + [P(OrcQuant)]
        L.orc_qm_level_law.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_qm_weight.argtypes = [ctypes.c_int] * 5
        L.orc_dequant_step.argtypes = [P(OrcQuant), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_fast_round.argtypes = [ctypes.c_float]
        L.orc_fast_round.restype = ctypes.c_float
        L.orc_fwd_txfm.argtypes = [ctypes.c_int, ctypes.c_int, i32p, P(OrcQuant), i32p, f32p]
        L.orc_inv_txfm.argtypes = [ctypes.c_int, ctypes.c_int, i32p, P(OrcQuant), i32p]
        L.orc_chain_block.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_int16), P(OrcQuant),
                                      i32p, f32p, i32p]
        L.orc_chain_block.restype = ctypes.c_int64
        L.orc_dct1d.argtypes = [ctypes.c_int, i32p]
        L.orc_fadst1d.argtypes = [ctypes.c_int, i32p, i32p]
        L.orc_inv_dct1d.argtypes = [ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int]
        L.orc_inv_adst1d.argtypes = [ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int]
        L.orc_mul_q16.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.orc_mul_q16.restype = ctypes.c_int32
        L.orc_quant_q16.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.orc_quant_q16.restype = ctypes.c_int32
        L.orc_apply_qmatrix.argtypes = [i32p, f32p, P(ctypes.c_uint8), ctypes.c_int, ctypes.c_int]
        L.orc_set_fullstep_scalar_target.argtypes = [ctypes.c_int]
        L.orc_tx_w.argtypes = L.orc_tx_h.argtypes = L.orc_tx_ncoef.argtypes = [ctypes.c_int]
        L.orc_tx_supported.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_sse_recon.argtypes = [i32p, i32p, P(ctypes.c_uint16), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
        L.orc_sse_recon.restype = ctypes.c_int64
        L.orc_block_hash.argtypes = [ctypes.c_int, i32p, f32p, ctypes.c_int, i32p]
        L.orc_block_hash.restype = ctypes.c_uint32
        u8p = P(ctypes.c_uint8)
        L.orc_chain_batch.argtypes = [ctypes.c_int, u8p, u8p, u8p, u8p, u8p, P(ctypes.c_uint64),
                                      P(ctypes.c_int16), P(ctypes.c_int64), P(ctypes.c_uint32),
                                      ctypes.c_int]
        _lib = L
    return _lib


def ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def quant(qidx, bd=8, qm_level=15):
    q = OrcQuant()
    lib().orc_quant_new(qidx, bd, qm_level, ctypes.byref(q))
    return q


def quant_chroma(qidx, dc_delta, ac_delta, bd, qm_level):
    q = OrcQuant()
    lib().orc_quant_new_chroma(qidx, dc_delta, ac_delta, bd, qm_level, ctypes.byref(q))
    return q


def tx_w(s):
    return lib().orc_tx_w(s)


def tx_h(s):
    return lib().orc_tx_h(s)


def ncoef(s):
    return lib().orc_tx_ncoef(s)


def fwd(s, t, resid, q):
    resid = np.ascontiguousarray(resid, dtype=np.int32)
    n = ncoef(s)
    lv = np.zeros(n, np.int32)
    tg = np.zeros(n, np.float32)
    rc = lib().orc_fwd_txfm(s, t, ptr(resid, ctypes.c_int32), ctypes.byref(q),
                            ptr(lv, ctypes.c_int32), ptr(tg, ctypes.c_float))
    if rc != 0:
        raise ValueError(f"unsupported size {s} type {t}")
    return lv, tg


def inv(s, t, levels, q):
    levels = np.ascontiguousarray(levels, dtype=np.int32)
    out = np.zeros(tx_w(s) * tx_h(s), np.int32)
    rc = lib().orc_inv_txfm(s, t, ptr(levels, ctypes.c_int32), ctypes.byref(q),
                            ptr(out, ctypes.c_int32))
    if rc != 0:
        raise ValueError(f"unsupported size {s} type {t}")
    return out


def chain_block(s, t, resid16, q):
    resid16 = np.ascontiguousarray(resid16, dtype=np.int16)
    n = ncoef(s)
    lv = np.zeros(n, np.int32)
    tg = np.zeros(n, np.float32)
    rec = np.zeros(tx_w(s) * tx_h(s), np.int32)
    sse = lib().orc_chain_block(s, t, ptr(resid16, ctypes.c_int16), ctypes.byref(q),
                                ptr(lv, ctypes.c_int32), ptr(tg, ctypes.c_float),
                                ptr(rec, ctypes.c_int32))
    return sse, lv, tg, rec


def chain_batch(sizes, types, qidx, bd, qm, offsets, pool, nthreads=8, with_hash=True):
    n = len(sizes)
    sse = np.zeros(n, np.int64)
    h = np.zeros(n, np.uint32) if with_hash else None
    a = [np.ascontiguousarray(x, dtype=np.uint8) for x in (sizes, types, qidx, bd, qm)]
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    pool = np.ascontiguousarray(pool, dtype=np.int16)
    u8 = ctypes.c_uint8
    lib().orc_chain_batch(n, *(ptr(x, u8) for x in a), ptr(offsets, ctypes.c_uint64),
                          ptr(pool, ctypes.c_int16), ptr(sse, ctypes.c_int64),
                          ptr(h, ctypes.c_uint32) if with_hash else None, nthreads)
    return sse, h


class OrcRateCtx(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "qctx", "cls", "plane", "w", "h", "eob_bin", "skip_ctx", "dcs_ctx", "txtp_set",
        "txtp_mode", "txtp", "use_skip_bits")] + [("skip_one", ctypes.c_float),
                                                   ("skip_zero", ctypes.c_float)]


def _rate_sigs(L):
    P = ctypes.POINTER
    L.orc_block_rate.argtypes = [P(ctypes.c_int32), P(ctypes.c_uint32), P(OrcRateCtx), ctypes.c_float]
    L.orc_block_rate.restype = ctypes.c_float
    L.orc_cost_q.argtypes = [ctypes.c_int]
    L.orc_cost_q.restype = ctypes.c_uint32
    L.orc_cdf_cost_raw.argtypes = [P(ctypes.c_uint16), ctypes.c_int, ctypes.c_int]
    L.orc_cdf_cost_raw.restype = ctypes.c_float
    L.orc_skip_cost.argtypes = [ctypes.c_int] * 4
    L.orc_skip_cost.restype = ctypes.c_float
    L.orc_scan.argtypes = [ctypes.c_int]
    L.orc_scan.restype = P(ctypes.c_uint32)


def scan(tx_size):
    L = lib()
    if not getattr(L, "_rate_sigs", False):
        _rate_sigs(L)
        L._rate_sigs = True
    p = L.orc_scan(tx_size)
    n = ncoef(tx_size)
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def rate_ctx(qctx=0, cls=1, plane=0, w=8, h=8, eob_bin=5, skip_ctx=0, dcs_ctx=0, txtp_set=0,
             txtp_mode=0, txtp=0, skip_bits=None):
    c = OrcRateCtx(qctx, cls, plane, w, h, eob_bin, skip_ctx, dcs_ctx, txtp_set, txtp_mode, txtp,
                   1 if skip_bits else 0, *(skip_bits or (0.0, 0.0)))
    return c


def block_rate(cf, scan_tab, ctx, bound=float("inf")):
    L = lib()
    if not getattr(L, "_rate_sigs", False):
        _rate_sigs(L)
        L._rate_sigs = True
    cf = np.ascontiguousarray(cf, dtype=np.int32)
    sc = np.ascontiguousarray(scan_tab, dtype=np.uint32)
    return L.orc_block_rate(ptr(cf, ctypes.c_int32), ptr(sc, ctypes.c_uint32), ctypes.byref(ctx), bound)


def skip_cost(qctx, cls, ctx, s):
    L = lib()
    if not getattr(L, "_rate_sigs", False):
        _rate_sigs(L)
        L._rate_sigs = True
    return L.orc_skip_cost(qctx, cls, ctx, s)
