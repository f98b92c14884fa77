"""The C-ABI library loads and exports every entry point of include/mtgpu.h; its host-side
helpers agree with the oracle (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import numpy as np

import maroontree_amd as mt
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "mtgpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(mt_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = mt.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in mt.lib().mt_version()


def test_quant_new_matches_oracle():
    """mt_quant_new / mt_quant_new_chroma == Quant::new_with_qm / new_chroma_with_delta_qm
    (src/quant.rs:354, :394) restated in the oracle, for every qindex and bit depth."""
    fields = [n for n, _ in mt.MtQuant._fields_]
    for bd in (8, 10, 12):
        for qi in range(256):
            for qm in (0, 7, 15):
                a, b = mt.quant_new(qi, bd, qm), O.quant(qi, bd, qm)
                assert [getattr(a, f) for f in fields] == [getattr(b, f) for f in fields]
        for qi, dd, ad in ((40, 0, 0), (128, -7, 11), (250, 9, -20)):
            a, b = mt.quant_new_chroma(qi, dd, ad, bd, 3), O.quant_chroma(qi, dd, ad, bd, 3)
            assert [getattr(a, f) for f in fields] == [getattr(b, f) for f in fields]


def test_device_q_mult_formula_equals_f64_round():
    """The device derives round(65536/step) as (131072 + step) / (2 step); equal to the
    reference's f64 round for every step of every table."""
    for bd in (8, 10, 12):
        for qi in range(256):
            q = O.quant(qi, bd)
            for step, qm in ((q.dc, q.q_mult_dc), (q.ac, q.q_mult_ac)):
                assert (131072 + step) // (2 * step) == qm


def test_qm_level_law_matches_oracle():
    for sub in (0, 1, 2):
        for qi in range(256):
            assert mt.qm_level_law(qi, sub) == O.lib().orc_qm_level_law(qi, sub)


def test_tx_geometry_and_coverage():
    for s in range(19):
        assert mt.tx_width(s) == O.tx_w(s) and mt.tx_height(s) == O.tx_h(s)
        assert mt.tx_ncoef(s) == O.ncoef(s)
        for t in range(16):
            assert mt.tx_supported(s, t) == bool(O.lib().orc_tx_supported(s, t))
    n = sum(mt.tx_supported(s, t) for s in range(19) for t in range(16))
    assert n == 47  # the reference's forward methods (src/dct.rs:2416-2950)


def test_dispatch_mirror_names():
    """DctDispatch / IdctDispatch expose the reference's method names for every supported pair."""
    fwd = mt.DctDispatch.methods()
    inv = mt.IdctDispatch.methods()
    assert len(fwd) == 47 and len(inv) == 47
    for name in ("dct4x4_t", "adstdct8x16_t", "fvdct16x8_t", "idtx16x16_t", "dct64x64_t"):
        assert name in fwd
    for name in ("idct_dequant_32x16", "iadst_dequant_8x4", "ihdct_dequant_8x16",
                 "iidentity_dequant_16x16", "iidtx_dequant_16x8", "idct_dequant_64x64"):
        assert name in inv


def test_generator_is_deterministic_and_shaped():
    """The microbench generator: 13 reference sizes, the sizes' reference types, residual
    clipped to the bit depth, about half 10-bit and half QM."""
    b1, p1 = mt.gen_tx_blocks(42, 4000)
    b2, p2 = mt.gen_tx_blocks(42, 4000)
    assert (b1 == b2).all() and (p1 == p2).all()
    sizes = set(int(s) for s in b1["tx_size"])
    assert sizes == {0, 5, 6, 13, 14, 1, 7, 8, 2, 9, 10, 3, 4}
    assert all(mt.tx_supported(int(s), int(t)) for s, t in zip(b1["tx_size"], b1["tx_type"]))
    bdi = (b1["bd_qm"] >> 4) & 3
    assert 0.4 < (bdi == 1).mean() < 0.6 and set(bdi) <= {0, 1}
    qm = b1["bd_qm"] & 15
    assert 0.4 < (qm == 15).mean() < 0.8
    off = b1["offset"].astype(np.int64)
    npx = np.array([mt.tx_width(int(s)) * mt.tx_height(int(s)) for s in b1["tx_size"]])
    assert (np.diff(off) == npx[:-1]).all() and off[-1] + npx[-1] == len(p1)
    for i in range(0, 4000, 97):
        blk = p1[off[i]:off[i] + npx[i]]
        lim = 1023 if bdi[i] == 1 else 255
        assert np.abs(blk).max() <= lim
    # discrete Laplace(12): mean |x| = 2 e^(1/12) / (e^(1/6) - 1) / ... ~ 12
    assert 10.0 < np.abs(p1[:200000]).mean() < 14.0


def test_descriptor_layouts():
    """The numpy / ctypes records match the C structs of include/mtgpu.h byte for byte."""
    assert mt.PREDBLK_DTYPE.itemsize == 16
    assert mt.DISTBLK_DTYPE.itemsize == 24 and mt.DISTRES_DTYPE.itemsize == 32
    assert mt.CFLBLK_DTYPE.itemsize == 24
    assert mt.CAND_DTYPE.itemsize == 48 and mt.CANDRES_DTYPE.itemsize == 48
    assert mt.CAND_DTYPE.fields["rplane"][1] == 16 and mt.CAND_DTYPE.fields["skip_one"][1] == 28
    assert ctypes.sizeof(mt.MtPlane) == 32
    assert mt.RATEBLK_DTYPE.itemsize == 32
    assert mt.BLK_DTYPE.itemsize == 8 and mt.RES_DTYPE.itemsize == 16
    assert [mt.PREDBLK_DTYPE.fields[f][1] for f in ("out_offset", "x", "y", "bw", "bh", "mode",
                                                     "angle_delta", "flags", "bd", "plane")] == \
        [0, 4, 6, 8, 9, 10, 11, 12, 13, 14]


def test_encode_config_defaults_and_quality_map():
    """mt_encode_config_default mirrors EncodeConfig::default (src/avif.rs:320-344);
    mt_quality_to_qindex == the oracle's quality_to_q for every quality."""
    from maroontree_amd import rdo as R
    cfg = R.encode_config()
    assert (cfg.quality, cfg.chroma, cfg.speed, cfg.quantization_matrices, cfg.qmatrix_level) == \
        (80, R.YUV420, R.SPEED_SLOW, 1, -1)
    assert (cfg.adaptive_quant, cfg.variance_boost, cfg.dark_aq, cfg.cdef, cfg.wiener, cfg.screen_content,
            cfg.intrabc, cfg.updating_cdf) == (1, 1, 1, 0, 0, 1, 1, 1)
    for q in range(-2, 104):
        assert R.quality_to_qindex(q) == O.quality_to_q(q), q


def test_deblock_levels_and_layouts():
    """mt_deblock_levels (host code) == the oracle's loop_filter_levels for every base_q_idx and
    subsampling; the deblock records match include/mtgpu.h (mt_lf_plane 72 bytes,
    mt_deblock_geom 72, mt_encode_out 168 with dbk_y at 136)."""
    from maroontree_amd import deblock as D
    from maroontree_amd import rdo as R
    for q in range(256):
        for sub in (0, 1, 2):
            assert D.levels(q, sub) == O.lf_levels(q, sub), (q, sub)
    assert ctypes.sizeof(D.MtLfPlane) == 72 and ctypes.sizeof(D.MtDeblockGeom) == 72
    assert ctypes.sizeof(R.MtEncodeOut) == 248 and R.MtEncodeOut.dbk_y.offset == 136
    assert R.MtEncodeOut.cdef_y.offset == 168 and R.MtEncodeOut.cdef.offset == 192
    from maroontree_amd import av1, cdef
    assert ctypes.sizeof(av1.MtAv1Frame) == 120 and ctypes.sizeof(cdef.MtCdefParams) == 48


def test_modules_import_in_any_order():
    """Each Python module imports first in a fresh interpreter (no import cycle)."""
    import subprocess
    import sys
    for m in ("cdef", "rdo", "av1", "deblock", "prune"):
        r = subprocess.run([sys.executable, "-c", f"import maroontree_amd.{m}"], capture_output=True, text=True,
                           cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        assert r.returncode == 0, (m, r.stderr[-500:])
