"""maroontree_amd: MI355X (gfx950) intra rate-distortion candidate evaluation for maroontree.

The compute path is the HIP library ``libmtgpu.so`` (C ABI in ``include/mtgpu.h``); this
package loads it with ctypes and passes device pointers of torch tensors through.  There is
no CPU fallback: every entry point raises when the library or a GPU is missing.
"""
from ._lib import (  # noqa: F401
    MtQuant, MtTxBlk, MtTxRes, MtError, lib, lib_path, init, quant_new, quant_new_chroma,
    qm_level_law, tx_width, tx_height, tx_ncoef, tx_supported, fwd_txfm, inv_txfm, tx_chain,
    tx_chain_noalloc, NOALLOC_MAX_BLOCKS, trellis, trellis_lambda0, TRELLISBLK_DTYPE, TRELLIS_CTX,
    tx_chain_trellis, CAND_TRELLIS, tx_schedule_set, tx_schedule_get,
    PRDOBLK_DTYPE, luma_satd_scale, k_satd, proxy_wsse, luma_partition_distortion, sb_activity,
    perceptual_rd_scale,
    gen_tx_blocks, block_rate, RATEBLK_DTYPE, BLK_DTYPE, RES_DTYPE, PREDBLK_DTYPE, MtPlane,
    plane_views, intra_pred, distortion, DISTBLK_DTYPE, DISTRES_DTYPE, DIST_NO_RESID,
    PRED_HAVE_TR, PRED_HAVE_BL, PRED_SMOOTH_NB, PRED_FILTER0, PRED_CFL, CAND_CFL_SS_HOR, CAND_CFL_SS_VER,
    CAND_CFL_ALPHA_GIVEN, MAX_PLANES,
    eval_candidates, CAND_DTYPE, CANDRES_DTYPE, TXBLK_CHROMA, cfl, CFLBLK_DTYPE, CFL_SS_HOR, CFL_SS_VER,
    CFL_ALPHA_GIVEN, COEF_CDFS_DTYPE, EOB_BIN_K, RATE_SLOTS, RATE_SLOTS_FIXED, rate_default_cdfs,
    rate_load_cdfs, profile_begin, profile_end, PROFILE_SLOTS, TX_SIZES_REF, TX_SIZES_ALL, TX_TYPES_REF, TX_NAMES, TXTP_NAMES,
)
from .dispatch import DctDispatch, IdctDispatch  # noqa: F401
