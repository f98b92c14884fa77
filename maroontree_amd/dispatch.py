"""The reference's kernel-dispatch tables, batched on the GPU.

`DctDispatch` mirrors the forward methods of the reference's DctDispatch (src/dct.rs:2416,
methods :2624-2950) and `IdctDispatch` the inverse fields of its IdctDispatch
(src/idct.rs:119, macro :91) plus idct_dequant_64x64 (src/idct.rs:2677).  Each method takes
a batch: residual / levels as an int32 cuda tensor of shape (n, W*H) / (n, ncoef), and one
quantizer (MtQuant) per block or a single one for all.  Layout and values are the
reference's: residual row-major, levels and targets x-major (rc = x*H + y).
"""
from . import _lib

_TX = {"4x4": 0, "8x8": 1, "16x16": 2, "32x32": 3, "64x64": 4, "4x8": 5, "8x4": 6, "8x16": 7,
       "16x8": 8, "16x32": 9, "32x16": 10, "4x16": 13, "16x4": 14}
_TP = {"dct": 0, "adstdct": 1, "dctadst": 2, "adst": 3, "idtx": 9, "fvdct": 10, "fhdct": 11}

# DctDispatch method -> (TxSize, TxType)
FWD_METHODS = {}
# IdctDispatch field -> (TxSize, TxType)
INV_METHODS = {}
for _name, _tp in _TP.items():
    for _sz, _s in _TX.items():
        if _lib.TX_NAMES[_s] != _sz:
            raise AssertionError("size table out of order")
        FWD_METHODS[f"{_name}{_sz}_t"] = (_s, _tp)
_INV_PREFIX = {0: "idct", 1: "iadstdct", 2: "idctadst", 3: "iadst", 10: "ivdct", 11: "ihdct"}
for (_s_name, _s) in _TX.items():
    for _tp, _pre in _INV_PREFIX.items():
        INV_METHODS[f"{_pre}_dequant_{_s_name}"] = (_s, _tp)
    # identity inverses are named iidentity_* (square) / iidtx_* (rectangular) in the reference
    _id = "iidentity" if _s_name in ("4x4", "8x8", "16x16") else "iidtx"
    INV_METHODS[f"{_id}_dequant_{_s_name}"] = (_s, 9)


def _quants(q, n):
    return [q] * n if isinstance(q, _lib.MtQuant) else list(q)


class DctDispatch:
    """Forward transform + quantization, one method per reference method name."""

    def __getattr__(self, name):
        if name not in FWD_METHODS:
            raise AttributeError(name)
        s, t = FWD_METHODS[name]
        if not _lib.tx_supported(s, t):
            raise AttributeError(f"{name}: the reference has no such forward transform")

        def method(residual, quant, stream=None):
            return _lib.fwd_txfm(s, t, residual, _quants(quant, residual.shape[0]), stream)

        method.__name__ = name
        return method

    @staticmethod
    def methods():
        return sorted(k for k, (s, t) in FWD_METHODS.items() if _lib.tx_supported(s, t))


class IdctDispatch:
    """Dequantization + inverse transform, one method per reference field name."""

    def __getattr__(self, name):
        if name not in INV_METHODS:
            raise AttributeError(name)
        s, t = INV_METHODS[name]
        if not _lib.tx_supported(s, t):
            raise AttributeError(f"{name}: the reference has no such inverse pairing")

        def method(levels, quant, stream=None):
            return _lib.inv_txfm(s, t, levels, _quants(quant, levels.shape[0]), stream)

        method.__name__ = name
        return method

    @staticmethod
    def methods():
        return sorted(k for k, (s, t) in INV_METHODS.items() if _lib.tx_supported(s, t))
