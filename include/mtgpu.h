/*
 * mtgpu -- MI355X (gfx950) intra rate-distortion candidate evaluation for maroontree.
 *
 * C ABI of the HIP library libmtgpu.so.  Plain pointers and sizes only; every pointer
 * named *_d is device memory, `stream` is a hipStream_t (NULL = the default stream).
 *
 * Conventions of the reference encoder (src/quant.rs:95-99, src/rate.rs:93-95):
 *   residual / reconstruction blocks are row-major [y*W + x];
 *   levels and targets are x-major ("transposed") [x*H + y].
 * Transform sizes and types use the AV1 TxSize / TxType numbering (MT_TX_*, MT_TXTP_*).
 *
 * Return codes: 0 ok, negative MT_ERR_* on failure (no call aborts the process);
 * mt_last_error() describes the last failure of the calling thread.
 */
#ifndef MTGPU_H
#define MTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MT_OK = 0,
    MT_ERR_ARG = -1,          /* bad argument (size, pointer, count) */
    MT_ERR_UNSUPPORTED = -2,  /* (size, type) has no reference forward transform */
    MT_ERR_DEVICE = -3,       /* HIP runtime error */
    MT_ERR_NOT_INIT = -4,     /* mt_init not called on this device */
};

/* AV1 TxSize */
enum {
    MT_TX_4X4 = 0, MT_TX_8X8 = 1, MT_TX_16X16 = 2, MT_TX_32X32 = 3, MT_TX_64X64 = 4,
    MT_TX_4X8 = 5, MT_TX_8X4 = 6, MT_TX_8X16 = 7, MT_TX_16X8 = 8, MT_TX_16X32 = 9,
    MT_TX_32X16 = 10, MT_TX_32X64 = 11, MT_TX_64X32 = 12, MT_TX_4X16 = 13, MT_TX_16X4 = 14,
    MT_TX_8X32 = 15, MT_TX_32X8 = 16, MT_TX_16X64 = 17, MT_TX_64X16 = 18
};
/* AV1 TxType (the ones the reference encoder evaluates) */
enum {
    MT_TXTP_DCT_DCT = 0, MT_TXTP_ADST_DCT = 1, MT_TXTP_DCT_ADST = 2, MT_TXTP_ADST_ADST = 3,
    MT_TXTP_IDTX = 9, MT_TXTP_V_DCT = 10, MT_TXTP_H_DCT = 11
};

/* Quantizer state, field for field the reference's `Quant` (src/quant.rs:327-345),
 * plus the bit depth it was built for. */
typedef struct {
    int32_t dc, ac;                          /* dequant steps (dc_q / ac_q) */
    int32_t q_mult_dc, q_mult_ac;            /* round(65536 / step) */
    int32_t rmin, rmax, cmin, cmax, cf_max;  /* inverse-transform clips (itx_clips) */
    int32_t qm_level;                        /* 15 = flat */
    int32_t qm_chroma;                       /* 0 luma, 1 chroma */
    int32_t qidx;                            /* AC-side qindex */
    int32_t bd;                              /* 8, 10 or 12 */
} mt_quant;

/* One block of the batched transform chain: 8 bytes in HBM. */
typedef struct {
    uint32_t offset;   /* element offset of the block's int16 residual in the pool */
    uint8_t tx_size;   /* AV1 TxSize */
    uint8_t tx_type;   /* AV1 TxType */
    uint8_t qidx;      /* base_q_idx of Quant::new_with_qm */
    uint8_t bd_qm;     /* bits 0-3: QM level (15 = flat); bits 4-5: 0 8-bit, 1 10-bit, 2 12-bit;
                          bit 6 (MT_TXBLK_CHROMA): chroma quantizer, new_chroma_with_delta_qm
                          (src/quant.rs:394) with zero qindex deltas; else the luma new_with_qm */
} mt_txblk;
#define MT_TXBLK_CHROMA 0x40

/* Per-block result of the chain: 16 bytes in HBM. */
typedef struct {
    int64_t sse;       /* sse_recon(P, inverse, clamp(P + residual)), P = 1 << (bd - 1); -1 unsupported */
    uint32_t nnz;      /* nonzero levels */
    uint32_t hash;     /* order-independent checksum of levels, targets and reconstruction */
} mt_txres;

/* ---- setup ------------------------------------------------------------------------ */
/* Uploads the quantizer tables and the 64-point basis to `device` and makes it current.
 * Replaces the reference's per-encode dispatch construction (EncodingContext::new,
 * src/encoding_context.rs:99). */
int mt_init(int device);
const char* mt_last_error(void);
const char* mt_version(void);

/* ---- quantizer state (host) --------------------------------------------------------- */
/* Quant::new_with_qm (src/quant.rs:354) */
int mt_quant_new(int base_q_idx, int bd, int qm_level, mt_quant* out);
/* Quant::new_chroma_with_delta_qm (src/quant.rs:394) */
int mt_quant_new_chroma(int base_q_idx, int dc_delta, int ac_delta, int bd, int qm_level,
                        mt_quant* out);
/* qm_level_law (src/quant.rs:137) */
int mt_qm_level_law(int base_q_idx, int sub);
/* width/height of an AV1 TxSize; coded coefficient count (64-point sizes code 32) */
int mt_tx_width(int tx_size);
int mt_tx_height(int tx_size);
int mt_tx_ncoef(int tx_size);
/* 1 when the reference has a forward transform for (tx_size, tx_type) */
int mt_tx_supported(int tx_size, int tx_type);

/* ---- per-kernel entry points (the dispatch tables, batched) ------------------------ */
/* Forward transform + quantization (+QM) of n blocks of one (size, type):
 * DctDispatch::*_t (src/dct.rs:2624-2950).  resid_d: n * W*H int32 row-major;
 * quants_d: n mt_quant; levels_d / targets_d: n * ncoef x-major. */
int mt_fwd_txfm(int tx_size, int tx_type, int n, const int32_t* resid_d,
                const mt_quant* quants_d, int32_t* levels_d, float* targets_d, void* stream);
/* Dequantization + inverse transform of n blocks of one (size, type):
 * IdctDispatch::* (src/idct.rs:91-728) and idct_dequant_64x64 (src/idct.rs:2677).
 * levels_d: n * ncoef x-major; out_d: n * W*H int32 row-major. */
int mt_inv_txfm(int tx_size, int tx_type, int n, const int32_t* levels_d,
                const mt_quant* quants_d, int32_t* out_d, void* stream);

/* ---- batched candidate chain ------------------------------------------------------- */
/* fwd -> quant(+QM) -> dequant -> inverse -> SSE for n blocks of mixed size/type/q/bd.
 * pool_d holds the int16 residuals.  res_d: n results.  levels_d / targets_d / recon_d
 * are optional (NULL) full outputs laid out like pool_d: block i's levels, targets
 * (ncoef, x-major) and reconstructed residual (W*H, int32) start at blocks[i].offset;
 * targets_d needs levels_d.
 * Blocks of unsupported (size, type) get sse = -1.  Uses a library-owned workspace on
 * the device (allocated on first use, grown as needed). */
int mt_tx_chain(const mt_txblk* blocks_d, int n, const int16_t* pool_d, mt_txres* res_d,
                int32_t* levels_d, float* targets_d, int32_t* recon_d, void* stream);

/* mt_tx_chain without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as mt_tx_chain;
 * its workspace is a fixed device array of the library (4 MiB per device, held until the
 * process ends), so n is at most MT_NOALLOC_MAX_BLOCKS (larger n: MT_ERR_ARG, nothing
 * launched).  Every call on one device shares that array: calls that overlap on one
 * device (two streams, or two host threads) race and must be ordered by the caller. */
#define MT_NOALLOC_MAX_BLOCKS 1048576
int mt_tx_chain_noalloc(const mt_txblk* blocks_d, int n, const int16_t* pool_d, mt_txres* res_d,
                        int32_t* levels_d, float* targets_d, int32_t* recon_d, void* stream);

/* ---- coefficient rate ------------------------------------------------------------- */
/* One transform block to price: 32 bytes.  The rate is real_block_bits_bounded
 * (src/rate.rs:149) against the frame-initial coefficient CDFs Cdfs::new(qctx)
 * (src/coder/cdfs.rs:400): the exact symbols the coefficient coder would emit.
 * eob_bin picks the eob_bin CDF as the reference's caller does (numbering: 0 16 luma,
 * 1 16 chroma, 2 16 luma 1-D, 3 32 luma, 4 32 chroma, 5 64 luma, 6 64 chroma, 7 64 luma 1-D,
 * 8 128 luma, 9 128 chroma, 10 128 luma 1-D, 11 256 luma, 12 256 chroma, 13 512 luma,
 * 14 512 chroma, 15 1024 luma, 16 1024 chroma).  txtp_set: 0 none, 1 TX_4X4 intra set,
 * 2 TX_8X8 set, 3 TX_16X16 set (txtp_mode = luma intra mode row, txtp = index in the set).
 * flags bit 0: use skip_one / skip_zero (adaptive all-zero price) instead of txb_skip.
 * A descriptor whose eob_bin table is not the one of the block's coded coefficient count,
 * or whose indices are out of range, gets NaN. */
typedef struct {
    uint32_t offset;   /* element offset of the block's levels (ncoef, x-major) */
    uint8_t tx_size, plane, eob_bin, skip_ctx;
    uint8_t dcs_ctx, txtp_set, txtp_mode, txtp;
    uint8_t qctx;      /* rate slot: 0..3 Cdfs::new(qctx), 4..7 loaded snapshots */
    uint8_t flags, pad0, pad1;
    float skip_one, skip_zero;
    float bound;       /* early-out bound (bits); +inf to always price fully */
    uint32_t pad2;
} mt_rateblk;

/* Bits of n blocks (bits_d[i] = +inf when the walk exceeded blocks[i].bound). */
int mt_block_rate(const mt_rateblk* blocks_d, int n, const int32_t* levels_d, float* bits_d,
                  void* stream);

/* The coefficient part of a decision CDF snapshot (the reference's Cdfs, src/coder/cdfs.rs:
 * 108-172) in the raw AV1 form of the default tables: each row holds the cumulative
 * probabilities of the first k symbols (the last symbol's 32768 implied).  Rate slots
 * (the qctx field of mt_rateblk / mt_cand): 0..3 the frame-initial Cdfs::new(qctx),
 * MT_RATE_SLOTS_FIXED..MT_RATE_SLOTS-1 snapshots the caller loads, e.g. adapted or
 * pilot-warmed CDFs (switch_decision_cdfs, src/coder/lossy_state.rs:319). */
#define MT_RATE_SLOTS_FIXED 4
#define MT_RATE_SLOTS 8
typedef struct {
    uint16_t skip[5][13];             /* txb_skip [tx class][ctx] (k = 1) */
    uint16_t base_tok[5][2][42][3];   /* coeff_base [class][plane][ctx] (41 ctx at 4x4 / 8x8) */
    uint16_t br_tok[4][2][21][3];     /* coeff_br [class][plane][ctx] */
    uint16_t eob_base[5][2][4][2];    /* coeff_base_eob [class][plane][ctx] */
    uint16_t eob_hi[5][2][11];        /* eob_extra [class][plane][eob bin] (k = 1) */
    uint16_t dc_sign[2][3];           /* [plane][ctx] (k = 1) */
    uint16_t eob_bin[17][10];         /* eob_bin CDFs in the numbering of mt_rateblk.eob_bin;
                                         row e holds 4,4,4,5,5,6,6,6,7,7,7,8,8,9,9,10,10 values */
    uint16_t txtp4[13][6], txtp8[13][6], txtp16[13][4];  /* intra tx-type sets [intra mode] */
    uint16_t pad;
} mt_coef_cdfs;

/* The coefficient CDFs of Cdfs::new(qctx), qctx 0..3 (host). */
int mt_rate_default_cdfs(int qctx, mt_coef_cdfs* out);
/* Builds the rate tables of `cdfs` (CoefCostTables::build, src/rate.rs:58, with the
 * reference's cost_q_table) into slot MT_RATE_SLOTS_FIXED..MT_RATE_SLOTS-1 of the current
 * device; synchronous, and no rate evaluation may be in flight on the device meanwhile. */
int mt_rate_load_cdfs(int slot, const mt_coef_cdfs* cdfs);

/* ---- trellis quantization ---------------------------------------------------------- */
/* One transform block for mt_trellis: 24 bytes.  The quantizer is the chain's (qidx, bd_qm
 * as in mt_txblk, the dc/ac qindex deltas with MT_TXBLK_CHROMA); its dc / ac steps are the
 * trellis's dc_q / ac_q.  lambda0 is the reference's trellis lambda0 at the call site
 * (trellis_lambda() times hbd_trellis_scale and the perceptual factor; mt_trellis_lambda0
 * gives the first two); <= 0 leaves the block as it is.  The context variant prices against
 * the CDFs of snapshot qctx with the eob_bin CDF, dc-sign context and plane as in
 * mt_rateblk; band_tilt is Cdfs::band_tilt (0 off). */
enum { MT_TRELLIS_CTX = 1 };
typedef struct {
    uint32_t offset;        /* the block's first level / target / workspace element (x-major) */
    uint8_t tx_size, qidx, bd_qm, flags;
    int8_t dc_delta, ac_delta;
    uint8_t qctx, eob_bin, dcs_ctx, plane, pad0, pad1;
    float lambda0, band_tilt;
} mt_trellisblk;
/* Trellis quantization of n blocks in place: levels_d (int32) from the f32 targets targets_d
 * of the forward transform (both x-major at each block's offset, ncoef per block; 32x32 for
 * 64x64).  Without MT_TRELLIS_CTX: trellis_optimize (src/trellis.rs:944) with the x86
 * AVX2+FMA round-down (src/avx/trellis.rs:196); with it: trellis_optimize_ctx
 * (src/trellis.rs:55) on the shipped switches.  work_d: unused (may be null; the
 * trellis works in on-chip memory).  status_d (optional, n int32): 0, or MT_ERR_ARG for an
 * invalid descriptor (its levels untouched).  Uses a per-device sorting list that the
 * library grows on demand: calls that overlap on one device race. */
int mt_trellis(const mt_trellisblk* blocks_d, int n, int32_t* levels_d, const float* targets_d,
               float* work_d, int32_t* status_d, void* stream);
/* The chain with the trellis between quantization and dequantization, as the reference's
 * candidate loops run it at Speed::Slow (src/coder/partition_search.rs:509-513): forward +
 * quantization, mt_trellis with trellis_d[i] (the same offset, size and quantizer as
 * blocks_d[i]), then dequantization + inverse + SSE / nnz / checksum of the final levels as
 * mt_tx_chain reports them.  levels_d, targets_d and recon_d are required; blocks of an unsupported (size, type) get sse = -1
 * and unspecified levels. */
int mt_tx_chain_trellis(const mt_txblk* blocks_d, const mt_trellisblk* trellis_d, int n,
                        const int16_t* pool_d, mt_txres* res_d, int32_t* levels_d, float* targets_d,
                        int32_t* recon_d, void* stream);
/* The reference's trellis lambda0 of a frame: trellis_lambda() (src/cost.rs:110) times
 * hbd_trellis_scale (src/coder/lossy_state.rs:158, shipped tuning). sub: 0 4:4:4, 1 4:2:2,
 * 2 4:2:0. */
float mt_trellis_lambda0(int qidx, int bd, int sub);

/* ---- planes and intra prediction -------------------------------------------------- */
/* A plane of uint16 samples in device memory.  width/height: the frame size the intra
 * edge rules clip to (predict_nd_ad's fw/fh, src/intrapred.rs:359); every block must lie
 * inside it.  vis_w/vis_h: the displayed size distortion is measured over (SrcBlock::new's
 * plane_vis_w/h, src/rd_sse.rs:76).  The allocation holds at least stride*height samples. */
typedef struct {
    const uint16_t* data;
    int32_t stride;
    int32_t width, height;
    int32_t vis_w, vis_h;
    int32_t pad;
} mt_plane;
#define MT_MAX_PLANES 64

/* Intra modes (the reference's numbering, src/intrapred.rs:54-82), chroma-from-luma (a
 * candidate mode of mt_eval_candidates only) and filter intra. */
enum {
    MT_PRED_DC = 0, MT_PRED_V = 1, MT_PRED_H = 2, MT_PRED_D45 = 3, MT_PRED_D135 = 4,
    MT_PRED_D113 = 5, MT_PRED_D157 = 6, MT_PRED_D203 = 7, MT_PRED_D67 = 8,
    MT_PRED_SMOOTH = 9, MT_PRED_SMOOTH_V = 10, MT_PRED_SMOOTH_H = 11, MT_PRED_PAETH = 12,
    MT_PRED_CFL = 13,
    MT_PRED_FILTER0 = 16, MT_PRED_FILTER4 = 20  /* filter intra modes 0..4 */
};
enum { MT_PRED_HAVE_TR = 1, MT_PRED_HAVE_BL = 2, MT_PRED_SMOOTH_NB = 4 };

/* One prediction: 16 bytes in HBM. */
typedef struct {
    uint32_t out_offset;  /* int32 element offset of the bw*bh row-major prediction */
    uint16_t x, y;        /* block origin in the plane */
    uint8_t bw, bh;       /* 4, 8, 16, 32, 64 (filter intra: <= 32) */
    uint8_t mode;         /* MT_PRED_* */
    int8_t angle_delta;   /* -3..3, directional modes */
    uint8_t flags;        /* MT_PRED_HAVE_TR | MT_PRED_HAVE_BL | MT_PRED_SMOOTH_NB (filter_type) */
    uint8_t bd;           /* 8, 10, 12 */
    uint8_t plane;        /* index into the plane array */
    uint8_t pad;
} mt_predblk;

/* Intra prediction of n blocks from reconstructed planes:
 *   modes 1..12: IntraPredDispatch::predict_nd_ad (src/intrapred.rs:359, edge filter on);
 *   MT_PRED_DC: the block filled with IntraPredDispatch::dc_pred (src/intrapred.rs:232);
 *   filter intra k: IntraPredDispatch::filter_predict (src/intrapred.rs:399).
 * planes: host array of nplanes (1..MT_MAX_PLANES) views.  pred_d receives each block's
 * prediction at its out_offset.  status_d (optional, n int32): 0, or MT_ERR_ARG for a
 * descriptor that is invalid or reaches outside its plane (its output is left untouched). */
int mt_intra_pred(const mt_predblk* blocks_d, int n, const mt_plane* planes, int nplanes,
                  int32_t* pred_d, int32_t* status_d, void* stream);

/* ---- chroma from luma ------------------------------------------------------------- */
enum { MT_CFL_SS_HOR = 1, MT_CFL_SS_VER = 2, MT_CFL_ALPHA_GIVEN = 4 };

/* One CfL block: 24 bytes. */
typedef struct {
    uint32_t out_offset;    /* int32 element offset of the cw*ch prediction (and AC) */
    uint16_t lx, ly;        /* origin of the covered luma in the luma plane */
    uint16_t x, y;          /* chroma block origin in the source plane */
    uint8_t cw, ch;         /* 4, 8, 16, 32 */
    uint8_t flags;          /* MT_CFL_SS_HOR | MT_CFL_SS_VER | MT_CFL_ALPHA_GIVEN */
    uint8_t bd;
    uint8_t luma_plane, src_plane;  /* indices into the plane array */
    int8_t alpha;           /* used with MT_CFL_ALPHA_GIVEN */
    uint8_t pad;
    int32_t dc;             /* the chroma DC prediction the AC is added to */
} mt_cflblk;

/* Chroma-from-luma for n blocks: the luma AC of IntraPredDispatch::cfl_ac_sub (cfl_ac_444
 * when not subsampled; src/intrapred.rs:1112, :1133) from the reconstructed luma plane,
 * the alpha of cfl_best_alpha (analytic optimum, then the +/-3 energy search over the
 * displayed part of the source, :1218) unless given, and cfl_pred_pixel (:1194).
 * pred_d: the prediction at out_offset; ac_d (optional): the AC at out_offset;
 * alpha_d: n alphas (-128 for an invalid descriptor, whose outputs are left untouched). */
int mt_cfl(const mt_cflblk* blocks_d, int n, const mt_plane* planes, int nplanes,
           int32_t* pred_d, int32_t* ac_d, int32_t* alpha_d, void* stream);

/* ---- distortion ---------------------------------------------------------------------- */
enum { MT_DIST_NO_RESID = 1 };

/* One block to measure: 24 bytes in HBM. */
typedef struct {
    uint32_t pred_offset;   /* int32 element offset of the bw*bh row-major prediction */
    uint32_t resid_offset;  /* int32 element offset of the bw*bh decoded residual */
    uint16_t x, y;          /* block origin in the source plane */
    uint8_t bw, bh;         /* 4, 8, 16, 32, 64 */
    uint8_t bd, plane;      /* bit depth; index into the plane array */
    uint8_t flags;          /* MT_DIST_NO_RESID: measure the prediction alone */
    uint8_t pad[3];
    float wsse_scale;       /* RdDispatch::wsse_scale: > 0 weighted decision distortion */
} mt_distblk;

/* Measures of one block: 32 bytes. */
typedef struct {
    int64_t sse;     /* sum (src - clamp(pred + resid))^2 over the visible part:
                        sse_recon / chroma_sse with plain SSE (src/rd_sse.rs:454, :682); -1 invalid */
    int64_t dist;    /* what sse_recon returns under wsse_scale: round(scale * weighted SSE of the
                        visible error zero-padded to 4x4 tiles), in the op order of
                        weighted_sse_avx2 (src/rd_sse.rs:244, src/avx/rd.rs:767); = sse if scale <= 0 */
    uint64_t satd;   /* luma_satd (src/rd_sse.rs:635, src/partition_rd.rs:86): 4x4 Hadamard L1
                        of the same error, invisible samples at zero error */
    uint64_t proxy;  /* satd_sad_proxy (src/rd_sse.rs:552, :1143): SAD + SATD/4 of src - pred */
} mt_distres;

/* Distortion of n reconstructed candidates against their source plane (planes: host array
 * of views; the vis_w / vis_h of a plane clip every measure, as SrcBlock does). resid_d may
 * be NULL when every block has MT_DIST_NO_RESID. */
int mt_distortion(const mt_distblk* blocks_d, int n, const mt_plane* planes, int nplanes,
                  const int32_t* pred_d, const int32_t* resid_d, mt_distres* out_d, void* stream);

/* ---- batched candidate evaluation -------------------------------------------------- */
/* One intra candidate: 48 bytes in HBM.  The block is one transform block of tx_size.
 * A chroma-from-luma candidate (mode MT_PRED_CFL, square 4..32 chroma blocks) predicts
 * cfl_pred_pixel(dc_pred, alpha, luma AC) (src/intrapred.rs:1194) with the DC prediction
 * from the reconstruction plane and the AC of cfl_ac_sub / cfl_ac_444 (:1112, :1133) from
 * the luma reconstruction plane at (x << ss_x, y << ss_y); the alpha is cfl_best_alpha's
 * (:1218, against the source) unless MT_CAND_CFL_ALPHA_GIVEN; with MT_CAND_CFL_ALPHA_STEP it
 * is the alpha angle_delta steps from cfl_best_alpha's along its sign (< 0: toward zero;
 * clipped to -16..16 and never past zero), the alphas cfl_alpha_rd_refine's walk visits
 * (src/cost.rs:332); mt_candres.alpha reports the alpha used.
 * For it, pflags holds MT_CAND_CFL_* (not the MT_PRED_* flags), angle_delta the given
 * alpha, and txtp_mode the luma reconstruction plane's index (txtp_set must be 0: chroma
 * codes no transform type). */
enum { MT_CAND_CFL_SS_HOR = 8, MT_CAND_CFL_SS_VER = 16, MT_CAND_CFL_ALPHA_GIVEN = 32,
       MT_CAND_CFL_ALPHA_STEP = 64 };
typedef struct {
    uint32_t offset;      /* element offset of the candidate's W*H slots in the work pools */
    uint16_t x, y;        /* block origin (the source and reconstruction planes share it) */
    uint8_t tx_size;      /* AV1 TxSize, one of the 13 sizes the reference transforms */
    uint8_t tx_type;      /* AV1 TxType */
    uint8_t mode;         /* MT_PRED_* */
    int8_t angle_delta;   /* -3..3 (MT_PRED_CFL: the given alpha, -16..16) */
    uint8_t pflags;       /* MT_PRED_HAVE_TR | MT_PRED_HAVE_BL | MT_PRED_SMOOTH_NB (MT_PRED_CFL:
                             MT_CAND_CFL_*) */
    uint8_t planes;       /* bits 0-1: source plane index, bits 2-3: reconstruction plane index
                             (low bits; the high bits are in planes_hi) */
    uint8_t qidx;         /* base_q_idx of the quantizer */
    uint8_t bd_qm;        /* as mt_txblk: bits 0-3 QM level, bits 4-5 bit depth (0 8, 1 10, 2 12),
                             bit 6 chroma: new_chroma_with_delta_qm with dc_delta / ac_delta */
    uint8_t rplane, eob_bin, skip_ctx, dcs_ctx;  /* rate context, as mt_rateblk */
    uint8_t txtp_set, txtp_mode, txtp, qctx;
    uint8_t rflags;       /* bit 0: adaptive skip price (skip_one / skip_zero, as mt_rateblk);
                             MT_CAND_TRELLIS: trellis_optimize_ctx in the rate context above
                             between quantization and dequantization, lambda0 of
                             mt_trellis_lambda0(qidx, bd, (rflags >> 2) & 3) (honoured by
                             mt_eval_candidates_trellis) */
    uint8_t planes_hi;    /* bits 0-3: source plane index >> 2, bits 4-7: reconstruction >> 2
                             (0 with up to 4 planes) */
    int8_t dc_delta, ac_delta;  /* chroma qindex deltas (src/quant.rs:394) */
    float skip_one, skip_zero;
    float wsse_scale;     /* RdDispatch::wsse_scale: > 0 prices the weighted distortion */
    float lambda;         /* rd_cost multiplier */
    float extra_bits;     /* added to the coefficient bits (mode, angle delta, tx type ...) */
} mt_cand;

/* Result of one candidate: 48 bytes. */
typedef struct {
    int64_t sse;      /* plain SSE of the reconstruction over the displayed part */
    int64_t dist;     /* decision distortion (weighted when wsse_scale > 0), as sse_recon */
    uint64_t satd;    /* luma_satd of the reconstruction error */
    float bits;       /* real_block_bits_bounded + extra_bits */
    float cost;       /* rd_cost(dist, lambda, bits) = (float)dist + lambda * bits; +inf if invalid */
    uint32_t nnz;     /* nonzero levels */
    int32_t status;   /* 0; MT_ERR_ARG invalid candidate; MT_ERR_UNSUPPORTED (size, type) */
    int32_t alpha;    /* MT_PRED_CFL: the CfL alpha used; 0 otherwise */
    int32_t pad;
} mt_candres;

/* The reference's candidate loop body (src/coder/partition_search.rs:471-520) for n
 * candidates at once: prediction from the reconstruction plane, residual against the source
 * plane, forward transform + quant (+QM; luma or chroma quantizer), dequant + inverse,
 * reconstruction distortion
 * (mt_distortion), coefficient rate (mt_block_rate) and rd_cost (src/partition_rd.rs:140).
 * Trellis is not applied (the Speed::Fast path).  pool_elems bounds every candidate's
 * offset + W*H; the stage pools (predictions, residuals, levels, reconstructed residuals)
 * are library-owned device memory of that many elements, grown on demand.  Calls on one
 * device share them: issue them on one stream. */
int mt_eval_candidates(const mt_cand* cands_d, int n, const mt_plane* planes, int nplanes,
                       uint64_t pool_elems, mt_candres* res_d, void* stream);
/* mt_eval_candidates with the trellis: candidates with MT_CAND_TRELLIS run the reference's
 * Speed::Slow chain (forward + quant, trellis_optimize_ctx, dequant + inverse; src/coder/
 * partition_search.rs:4839-4860), the others the plain chain; the levels the rate prices and
 * nnz reports are the trellis's. */
enum { MT_CAND_TRELLIS = 2 };
int mt_eval_candidates_trellis(const mt_cand* cands_d, int n, const mt_plane* planes, int nplanes,
                               uint64_t pool_elems, mt_candres* res_d, void* stream);

/* ---- whole-frame luma intra RDO over the superblock wavefront ----------------------- */
/* The candidate generators and argmin of the reference's decision code (A19:
 * rd_cost_square src/coder/partition_search.rs:338, pick_luma16 :4755, pick_block8
 * src/coder/block8.rs:689, pick_luma32 src/coder/block32.rs:728, rd_pick_luma64
 * src/coder/block64.rs:1812, decode_sb src/coder/superblock.rs:813) replaced by one
 * exhaustive, data-parallel procedure of this library:
 *
 *  - 64x64 superblocks (SB) run along the anti-diagonal wavefront: SB (r, c) in step
 *    c + 2r, so its left, above and above-right SBs are reconstructed before it starts;
 *  - in a step, every quadtree node (64 .. 8, or .. 4 with MT_RDO_MIN4) of every SB of
 *    the step that lies inside the frame evaluates every allowed mode as one DCT_DCT
 *    candidate of mt_eval_candidates: the 13 intra modes, with MT_RDO_ANGLE_DELTAS the 6
 *    non-zero angle deltas of the 8 directional modes at >= 8x8, with MT_RDO_FILTER_INTRA
 *    the 5 filter-intra modes at <= 32x32; it predicts from the reconstruction, whose
 *    samples inside the SB still hold the source at decision time;
 *  - rate = coefficient bits (real_block_bits_bounded, txb skip / dc sign context 0, the
 *    intra tx-set symbol of DCT_DCT at <= 16x16) + luma_intra_syntax_bits
 *    (src/coder/superblock.rs:192: kf_y mode, angle delta, use_filter_intra / mode);
 *    cost = weighted distortion + lambda * rate (rd_cost);
 *  - bottom-up, a node keeps PARTITION_NONE (its cheapest mode; ties to the lower mode
 *    id) or SPLIT (its children's sum), each adding lambda * the partition symbol bits
 *    (partition_bits_bl, src/coder/lossy_state.rs:1718, frame-initial CDFs); nodes that
 *    cross the frame edge split without a symbol;
 *  - the chosen leaves are reconstructed in coding order, one wave per SB;
 *  - decision-time contexts (kf_y mode context, edge-filter type, partition context) read
 *    the final maps outside the SB and count neighbours inside it as DC_PRED / unsplit;
 *  - the final accounting re-prices every leaf and partition symbol with the final
 *    contexts against the final reconstruction.
 * Mode ids: 0..12 the intra modes (MT_PRED_*, delta 0); 13 + 6 (m - 1) + j directional
 * mode m in 1..8 with delta {-3, -2, -1, 1, 2, 3}[j]; 61..65 filter intra 0..4.
 *
 * Chroma (mt_rdo_frame.chroma 4:2:0 or 4:4:4), in the same wavefront step right after the
 * SB's luma reconstruction:
 *  - chroma blocks: 4:2:0: each luma leaf of >= 8x8 at (x, y), size s, owns the s/2 chroma
 *    block at (x/2, y/2); an 8x8 node split into 4x4 leaves owns one 4x4 chroma block, coded
 *    with (and taking the luma mode of) its last 4x4 leaf.  4:4:4: each leaf owns the chroma
 *    block of its own size and place; a 64x64 one is four 32x32 transform blocks in raster
 *    order (the largest chroma transform), each predicted in turn;
 *  - per chroma block, the uv modes 0..12 (angle delta 0) and, with MT_RDO_CFL where CfL is
 *    allowed (luma block <= 32x32), CfL (13): DC prediction + alpha * the AC of the final
 *    luma reconstruction, each plane's alpha from cfl_best_alpha; with S =
 *    MT_RDO_CFL_STEPS_OF(flags) > 0 also the 2S candidates per plane of the alphas S steps
 *    toward and away from zero, and cfl_alpha_rd_refine's walk (src/cost.rs:332; the
 *    reference's cfl_alpha_rd_steps, shipped 0, src/tuning.rs:830) picks each plane's alpha,
 *    U first against V's energy optimum, then V against U's result, scoring an alpha by
 *    dist + lambda * (coefficient bits + uv_mode_bits of the pair) (src/cost.rs:324); CfL
 *    with both alphas 0 is DC_PRED and is not a CfL candidate.  Each (mode, plane, transform block) is one
 *    candidate: the transform type the decoder derives from the uv mode
 *    (chroma_tx_for_mode, src/coder/chroma_tx.rs:46; DCT_DCT at 32x32), the chroma
 *    quantizer (qidx with uv_dc_delta / uv_ac_delta, uv_qm_level), plain SSE, the plane-1
 *    coefficient rate (the chroma eob_bin CDFs) with txb skip context 7 and dc sign
 *    context 0;
 *  - cost(mode) = the rd_costs summed over U then V, each over its transform blocks in
 *    raster order, + lambda * uv_mode_bits (src/coder/lossy_state.rs:1302: the uv_mode
 *    symbol under the luma mode, from the CfL-allowed or the no-CfL CDF as AV1 codes it; the
 *    delta-0 angle symbol of a directional mode at >= 8x8 blocks; the CfL sign and alpha
 *    symbols); the lowest cost wins, ties to the lower uv mode;
 *  - decision-time chroma samples inside the SB are the source, outside the final
 *    reconstruction; the chroma edge-filter type reads the final uv modes outside the SB
 *    and counts neighbours inside it as not smooth; the chosen chroma blocks are then
 *    reconstructed in coding order, and the final accounting re-prices them with the final
 *    contexts (mt_rdo_stats bits_uv, sse_u, sse_v). */
/* MT_RDO_TRELLIS: the Speed::Slow chain in every candidate and in the reconstruction:
 * trellis_optimize_ctx (src/trellis.rs:55) between quantization and dequantization in the
 * candidate's rate context, lambda0 = mt_trellis_lambda0(qidx, bd, sub) with sub 0 for
 * 4:4:4 frames and 2 for 4:2:0 and luma-only frames (their decision lambda's). */
enum { MT_RDO_ANGLE_DELTAS = 1, MT_RDO_FILTER_INTRA = 2, MT_RDO_MIN4 = 4, MT_RDO_CFL = 8,
       MT_RDO_TRELLIS = 32, MT_RDO_GRAPH = 64 };
/* CfL alpha walk steps (0..15) in flag bits 8-11, with MT_RDO_CFL. */
#define MT_RDO_CFL_STEPS(s) (((s) & 15) << 8)
#define MT_RDO_CFL_STEPS_OF(flags) (((flags) >> 8) & 15)
/* MT_RDO_GRAPH: the wavefront steps' launches (candidate generation, evaluation, argmin,
 * decisions, reconstruction of every step) are captured from the stream into one HIP graph,
 * instantiated and launched as a whole; the results are the same.  No counterpart in the
 * reference (its wavefront runs host threads). */
enum { MT_RDO_LUMA_ONLY = 0, MT_RDO_CHROMA_420 = 1, MT_RDO_CHROMA_444 = 2 };
#define MT_RDO_NODES 341          /* quadtree nodes of one SB: 1 + 4 + 16 + 64 + 256 */
#define MT_RDO_ALL_MODES 0x3FFF   /* prune mask bits 0-12: intra modes; bit 13: filter intra */
#define MT_RDO_FORCE_SPLIT 0x4000 /* node must split (ignored at the minimum size) */
#define MT_RDO_NO_SPLIT 0x8000    /* node is not split (ignored where it crosses the frame) */

/* One frame.  Node n of an SB is level L = 0 (64x64) .. 4 (4x4), n = base[L] + iy * 2^L + ix
 * with base = {0, 1, 5, 21, 85}; SBs in raster order. */
typedef struct {
    mt_plane src;          /* luma source (device), width and height multiples of 8 */
    uint16_t* recon;       /* out: reconstruction (device), recon_stride samples per row */
    int32_t recon_stride;
    uint8_t bd;            /* 8, 10, 12 */
    uint8_t qidx;          /* base_q_idx */
    uint8_t qm_level;      /* 0..15 (15 flat) */
    uint8_t chroma;        /* MT_RDO_LUMA_ONLY, MT_RDO_CHROMA_420, MT_RDO_CHROMA_444 */
    float lambda;          /* > 0 (mt_rdo_lambda gives the reference's decision lambda) */
    float wsse_scale;      /* RdDispatch::wsse_scale (1.0 shipped; 0 plain SSE) */
    const uint16_t* prune; /* optional (device): [nsb][MT_RDO_NODES] node masks; NULL = full */
    void* nodes;           /* out (device): [nsb][MT_RDO_NODES] mt_rdo_node */
    /* chroma (ignored when luma only): sources of the chroma size (width / 2 x height / 2
     * at 4:2:0), reconstructions out, the chroma quantizer (mt_chroma_q_deltas and
     * mt_qm_chroma_level_law give the reference's laws) */
    mt_plane src_u, src_v;
    uint16_t *recon_u, *recon_v;
    int32_t recon_uv_stride;
    int8_t uv_dc_delta, uv_ac_delta;
    uint8_t uv_qm_level, pad1;
    /* optional outputs (device; NULL = none): the quantized levels of the reconstruction's
     * transform blocks (what the bitstream codes; after the trellis with MT_RDO_TRELLIS):
     * the level of frequency (fx, fy) of the block at (x, y) is stored at sample
     * (x + fx, y + fy) of an int32 plane of the frame size (coef_stride per row; chroma:
     * the chroma size, coef_uv_stride); 64x64 blocks fill their low 32x32 frequencies and
     * leave the rest of their area untouched. */
    int32_t* coef;
    int32_t* coef_u;
    int32_t* coef_v;
    int32_t coef_stride, coef_uv_stride;
} mt_rdo_frame;

typedef struct {
    float cost;       /* the node's decided rd cost (+inf when the node was not searched) */
    float none_cost;  /* PARTITION_NONE cost incl. its partition symbol (+inf if none) */
    uint8_t split;    /* 0 leaf, 1 split, 2 not coded (outside the frame / not searched) */
    uint8_t mode;     /* mode id of the cheapest NONE candidate */
    uint8_t uv_mode;  /* the chroma block coded at this node: uv mode 0..12, 13 CfL; 255 none
                         (chroma blocks sit at leaves, and at 4:2:0 8x8 nodes split into 4x4) */
    uint8_t pad0;
    int8_t cfl_alpha_u, cfl_alpha_v;  /* CfL alphas (0 unless uv_mode is 13) */
    uint8_t pad1[2];
} mt_rdo_node;

typedef struct {
    double bits;         /* final rate: coefficients + mode syntax + partition symbols (luma
                            and chroma) */
    int64_t sse;         /* luma SSE of the reconstruction against the source */
    int64_t candidates;  /* decision candidates evaluated (luma and chroma) */
    int32_t leaves, sbs;
    int32_t steps, pad;
    double bits_uv;      /* the chroma part of bits (coefficients and uv mode syntax) */
    int64_t sse_u, sse_v;
} mt_rdo_stats;

/* Runs the RDO of one frame on `stream` and waits for it.  Returns MT_OK, MT_ERR_ARG for
 * invalid parameters, or a device error. */
int mt_intra_rdo(const mt_rdo_frame* frame, int flags, mt_rdo_stats* stats, void* stream);
/* The RDO of several frames of any sizes, bit depths, chroma formats and quantizers at once
 * (stats: one per frame): up to 16 frames (and MT_MAX_PLANES planes: 2 per luma-only frame,
 * 6 with chroma) share every wavefront step, so each step's batch evaluation holds the
 * candidates of all of them; larger counts run in such groups.  Each frame's result equals
 * its own mt_intra_rdo. */
int mt_intra_rdo_batch(const mt_rdo_frame* frames, int nframes, int flags, mt_rdo_stats* stats,
                       void* stream);

/* The reference's decision lambda lambda_d = mlam / K_PART_LAMBDA
 * (src/coder/lossy_state.rs:1769-1790): mode_lambda_q(dc_q, bd) (src/cost.rs:285) times
 * mode_lambda_weight(qidx, sub) (src/cost.rs:132, shipped tuning) over 2.775.
 * sub: 0 4:4:4, 1 4:2:2, 2 4:2:0.  mt_mode_lambda returns mlam. */
float mt_rdo_lambda(int qidx, int bd, int sub);
float mt_mode_lambda(int qidx, int bd, int sub);
/* The reference's frame chroma quantizer laws (shipped tuning): chroma_dc_delta /
 * chroma_ac_delta (src/quant.rs:204-240) and qm_chroma_level_law (:155).  sub as above. */
int mt_chroma_q_deltas(int qidx, int sub, int* dc_delta, int* ac_delta);
int mt_qm_chroma_level_law(int qidx, int sub);

/* ---- the encoder's entry: an 8-bit RGB image to the RDO's decisions ---------------- */
/* EncodeConfig (src/avif.rs:257-344) as far as this path goes; mt_encode_config_default
 * gives its defaults.  chroma: ChromaFormat (src/avif.rs:75-86); speed: Speed (:91). */
enum { MT_ENC_YUV420 = 0, MT_ENC_YUV422 = 1, MT_ENC_YUV444 = 2, MT_ENC_MONOCHROME = 3 };
enum { MT_SPEED_EXTRA_SLOW = 0, MT_SPEED_SLOW = 1, MT_SPEED_MEDIUM = 2, MT_SPEED_FAST = 3 };
typedef struct {
    uint8_t quality;                /* 1..100 (80) -> base_q_idx by quality_to_q */
    uint8_t chroma;                 /* MT_ENC_* (MT_ENC_YUV420) */
    uint8_t speed;                  /* MT_SPEED_* (MT_SPEED_SLOW) */
    uint8_t quantization_matrices;  /* 1: QM on (1) */
    int16_t qmatrix_level;          /* 0..15 for every plane, or -1: the level laws (-1) */
    uint8_t adaptive_quant, variance_boost, dark_aq, cdef, wiener, screen_content, intrabc,
        updating_cdf;               /* 1, 1, 1, 0, 0, 1, 1, 1 */
    uint8_t pad[2];
} mt_encode_config;
/* The CDEF decision of a frame (mt_cdef_frame, the encoder entry). */
typedef struct {
    int32_t on;        /* 1: the frame is filtered (enable_cdef with these entries) */
    int32_t bits;      /* cdef_bits (0..3) */
    int32_t damping;   /* the signalled damping, 3 + min(qidx / 64, 3) */
    int32_t n;         /* 1 << bits strength entries */
    uint8_t y_pri[8], y_sec[8], uv_pri[8], uv_sec[8];  /* signalled strengths per entry */
} mt_cdef_params;
/* Config switches this library does not run, reported in mt_encode_out.not_applied: the
 * per-superblock quantizer modulation (adaptive_quant / variance_boost / dark_aq), the
 * frame filters (cdef, also implied by Speed::Slow; wiener), palette (screen_content),
 * IntraBC, CDF adaptation (the rates price the frame-initial CDFs), ExtraSlow's extra
 * searches (it runs as Slow), and the deblocking filter when no dbk planes are given. */
enum { MT_ENC_NA_AQ = 1, MT_ENC_NA_VARIANCE_BOOST = 2, MT_ENC_NA_DARK_AQ = 4, MT_ENC_NA_CDEF = 8,
       MT_ENC_NA_WIENER = 16, MT_ENC_NA_PALETTE = 32, MT_ENC_NA_INTRABC = 64,
       MT_ENC_NA_UPDATING_CDF = 128, MT_ENC_NA_EXTRA_SLOW = 256, MT_ENC_NA_DEBLOCK = 512 };
typedef struct {
    /* out (device, caller-allocated): the planes after the colour conversion and their
     * reconstructions, width x height for luma, the chroma size for chroma (width / 2 x
     * height / 2 at 4:2:0, width x height at 4:4:4), rows packed */
    uint16_t *y, *u, *v;
    uint16_t *recon_y, *recon_u, *recon_v;
    void* nodes;                    /* out (device): [nsb][MT_RDO_NODES] mt_rdo_node */
    mt_rdo_stats stats;             /* out */
    uint8_t qidx, qm_level, uv_qm_level, pad0;
    int32_t rdo_flags;              /* the MT_RDO_* search run */
    uint32_t not_applied;           /* MT_ENC_NA_* */
    int32_t pad1;
    /* out (device, caller-allocated, optional): the reconstruction after the in-loop
     * deblocking filter (frame_deblock at loop_filter_levels, src/coder/frame_encode.rs:
     * 303-346), the planes' sizes as above; NULL skips it (MT_ENC_NA_DEBLOCK) */
    uint16_t *dbk_y, *dbk_u, *dbk_v;
    uint8_t lf_level_y, lf_level_u, lf_level_v, lf_sharpness;  /* out: the levels applied */
    uint32_t dbk_status;            /* out: the deblock's MT_DEBLOCK_* status (0) */
    /* out (device, caller-allocated, optional, with the dbk planes): the reconstruction after
     * CDEF (frame_cdef, src/coder/frame_cdef.rs:49), run when cdef is set or the speed is
     * Slow / ExtraSlow (src/avif.rs:449-451), with Slow's two-stage candidate pairs at those
     * speeds and the full list otherwise, over the skip flags of the frame's coded levels;
     * NULL skips it (MT_ENC_NA_CDEF) */
    uint16_t *cdef_y, *cdef_u, *cdef_v;
    mt_cdef_params cdef;            /* out: the CDEF header values (cdef.on 0: not filtered) */
    uint8_t* cdef_grid;             /* out (host, optional): each 64x64 unit's entry, raster */
} mt_encode_out;
void mt_encode_config_default(mt_encode_config* cfg);
/* quality_to_q (src/avif.rs:637): quality 1..100 -> base_q_idx 215..1. */
int mt_quality_to_qindex(int quality);
/* encode_rgb8 (src/avif.rs:915) up to the bitstream: the colour conversion of
 * encode_still_lossy_420_with_cdf / encode_still_lossy_with_cdf (src/encoder.rs:530-605,
 * :398-410: BT.601 full range in Q13 integers, 4:2:0 chroma as the rounded 2 x 2 mean of the
 * Q13 values) of the device planes g, b, r (PlanarImage's plane order, width x height bytes
 * each, rows packed), then mt_intra_rdo at quality_to_q(quality) with the reference's decision
 * lambda, chroma quantizer and QM laws (or the explicit level, or flat with
 * quantization_matrices = 0).  Speed: Slow and ExtraSlow search with MT_RDO_TRELLIS, Medium
 * and Fast without; every speed searches angle deltas, filter intra, 4x4 and CfL.
 * width and height: multiples of 8.  MT_ERR_UNSUPPORTED for 4:2:2 and monochrome. */
int mt_encode_rgb8(const uint8_t* g_d, const uint8_t* b_d, const uint8_t* r_d, int width, int height,
                   const mt_encode_config* cfg, mt_encode_out* out, void* stream);

/* ---- frame filters: the in-loop deblocking filter ------------------------------------- */
/* Deblocking is out of place: dst planes must not overlap src planes.  Planes are device
 * uint16, 8-byte aligned, with a stride (samples per row) that is a multiple of 4.  The
 * library filters every vertical edge of a plane at once, then every horizontal edge: on
 * the maps of a partition (edges at transform-block starts, each map holding the size of the
 * transform it lies in) the result equals the reference's order (per 64-row superblock row,
 * vertical then horizontal edges).  status_d (device, optional) receives the OR of
 * MT_DEBLOCK_ORDER where two edges of one direction come within each other's reach, an edge
 * reaches past the plane, or a horizontal edge reaches into the next superblock row (the
 * result may then differ from the reference's), and MT_DEBLOCK_BAD_MAP where a size is 0
 * under the alignment rule.  Asynchronous on `stream`. */
enum { MT_DEBLOCK_ORDER = 1, MT_DEBLOCK_BAD_MAP = 2 };
typedef struct {
    const uint16_t* src;   /* device: the plane before filtering */
    uint16_t* dst;         /* device: the filtered plane */
    int32_t stride;        /* samples per row of src and dst */
    int32_t w, h;          /* the plane's coded size, multiples of 4 */
    int32_t vis_w, vis_h;  /* its visible size: edges at or past it are not filtered */
    int32_t nc4;           /* row stride of the maps, >= w / 4 */
    const uint8_t* bw4;    /* device, h / 4 rows: transform width in 4-sample units */
    const uint8_t* bh4;    /* transform height in 4-sample units */
    const uint8_t* vedge4; /* device or NULL: nonzero where a vertical edge starts; NULL: where
                              the 4-unit column is a multiple of bw4 */
    const uint8_t* hedge4; /* the same for horizontal edges, rows and bh4 */
} mt_lf_plane;
/* filter_plane_impl (src/loopfilter.rs:483) of one plane: level 1..63 (<= 0 copies the plane),
 * sharpness 0..7, luma (widths 4 / 8 / 16) or chroma (4 / 6), sb_rows4 the superblock height
 * in 4-units (16 for luma, 16 >> sub_y for chroma), bd 8 / 10 / 12. */
int mt_deblock_plane(const mt_lf_plane* plane, int level, int sharp, int is_luma, int sb_rows4, int bd,
                     uint32_t* status_d, void* stream);
/* DeblockGeom (src/coder/frame_deblock.rs:38-57): the luma transform maps with their
 * edge-start flags and the luma prediction maps the chroma geometry derives from
 * (chroma_deblock_geom :91: the prediction size >> subsampling, 1..8 4-units, edges where the
 * position is a multiple of it).  Maps: device, (h8 / 4) rows of w8 / 4. */
typedef struct {
    int32_t w8, h8;          /* coded luma size, multiples of 8 */
    int32_t disp_w, disp_h;  /* visible size */
    const uint8_t *blk4, *blk4h, *blk4v, *blk4t;
    const uint8_t *pblk4, *pblk4h;
    uint8_t sub_x, sub_y, mono, sharp;
    uint8_t bd, pad[3];
} mt_deblock_geom;
/* frame_deblock (src/coder/frame_deblock.rs:206) in one launch: luma at lvl_y; U and V at
 * lvl_u / lvl_v only when luma is filtered and one of them is > 0 (otherwise copied).  Chroma
 * planes are (w8 >> sub_x) x (h8 >> sub_y); ignored when mono. */
int mt_deblock_frame(const mt_deblock_geom* geom, const uint16_t* src_y, const uint16_t* src_u,
                     const uint16_t* src_v, int32_t y_stride, int32_t uv_stride, uint16_t* dst_y,
                     uint16_t* dst_u, uint16_t* dst_v, int lvl_y, int lvl_u, int lvl_v, uint32_t* status_d,
                     void* stream);
/* ---- CDEF (src/cdef.rs, src/coder/frame_cdef.rs) ------------------------------------ */
typedef struct {
    int32_t width, height;  /* luma size, multiples of 8 (the coded area is the frame) */
    uint8_t bd, chroma;     /* chroma: MT_RDO_LUMA_ONLY, MT_RDO_CHROMA_420, MT_RDO_CHROMA_444 */
    uint8_t slow;           /* 1: Speed::Slow's two-stage candidate pairs, 0: the full list */
    uint8_t pad0;
    int32_t qidx;           /* base_q_idx: the damping */
    float wsse_scale;       /* > 0: the decision distortion is the scaled weighted SSE */
    float lambda;           /* rate multiplier of the header and cdef_idx bits (the frame's
                               lambda_d, mt_rdo_lambda) */
    const uint16_t *rec_y, *rec_u, *rec_v;  /* device: the deblocked reconstruction */
    int32_t y_stride, uv_stride;            /* of rec_* and dst_* */
    const uint16_t *src_y, *src_u, *src_v;  /* device: the source */
    int32_t src_y_stride, src_uv_stride;
    const uint8_t* skip8;   /* device: mt_av1_skip8's map */
    uint16_t *dst_y, *dst_u, *dst_v;        /* device: the filtered planes (not the inputs) */
} mt_cdef_frame_args;
/* frame_cdef (src/coder/frame_cdef.rs:49) on the GPU: per luma 8x8 the direction and
 * variance (src/cdef.rs:103-179), the distortion of all twelve signalled strength pairs
 * (primary 0, 1, 2, 4 x secondary 0, 1, 2) per 64x64 unit for luma and for U + V, the
 * frame decision on those tables (the directional-variance gate, Slow / Fast candidate
 * pairs, greedy strength set, mildest entry clearing a 40 per-mille margin, cdef_bits on
 * rate; shipped tuning, src/tuning.rs:772-776), and the filter of every non-skip 8x8 from
 * the pre-CDEF planes into dst (copied where nothing is filtered).  prm: the header values;
 * grid (host, optional): each 64x64 unit's entry, raster order; tables (host, optional):
 * [2][units][12] int64, luma then U + V.  Synchronizes the stream (the decision runs on
 * the host between the table and filter launches). */
int mt_cdef_frame(const mt_cdef_frame_args* args, mt_cdef_params* prm, uint8_t* grid, int64_t* tables,
                  void* stream);

/* loop_filter_levels / loop_filter_sharpness (src/obu.rs:306-334) at the shipped tuning:
 * lvl_y = round(qidx / 8) (<= 40); lvl_uv = max(round(qidx / 10 * s), 1) (<= 32) while lvl_y
 * > 0, with s = 0.5 at 4:2:0 (sub 2) and 1 otherwise; sharpness 7. */
int mt_deblock_levels(int qidx, int sub, int* lvl_y, int* lvl_uv, int* sharp);
/* The deblock maps of an RDO's decisions (nodes as mt_rdo_frame.nodes of a width x height
 * frame): record_blk (src/coder/superblock.rs:548) of every leaf, its transform being the
 * block (the RDO's square partitions).  maps_d (device): six (height / 4) x (width / 4) planes
 * in the order blk4, blk4h, blk4v, blk4t, pblk4, pblk4h. */
int mt_rdo_deblock_maps(const void* nodes_d, int width, int height, uint8_t* maps_d, void* stream);

/* ---- learned candidate pruning (bf16 CNN on MFMA) ---------------------------------- */
/* A small network replaces / feeds the reference's hand pruning (proxy_mode_beam
 * src/coder/partition_search.rs:187, split_breakout_k :135-160, rank_luma_modes
 * src/coder/superblock.rs:102): 64x64 luma superblock + qindex -> conv 3x3 s2 (1->16) ->
 * conv 3x3 s2 (16->32) -> conv 3x3 s2 (32->64), ReLU each -> 4x4 average pool ++ qindex/255
 * -> fc 257->32 ReLU -> fc 32->863: outputs 10 n .. 10 n + 9 the logits of the 10 AV1
 * partition types (NONE, HORZ, VERT, SPLIT, HORZ_A, HORZ_B, VERT_A, VERT_B, HORZ_4, VERT_4)
 * of quadtree node n = 0..84 (64x64 .. 8x8), outputs 850..862 the scores of intra modes 0..12
 * being used by the superblock's leaves (the masks keep the top-k).  Parameters: one flat
 * fp32 array (weights [K][N] row-major then biases, layer by layer: conv1 9x16 + 16, conv2
 * 144x32 + 32 (k = tap * 16 + channel), conv3 288x64 + 64, fc1 257x32 + 32, fc2 32x863 + 863);
 * every product runs in bf16 with fp32 accumulation.  Labels and loss terms are [batch][98]:
 * column n < 85 the partition type of node n (its softmax cross-entropy), 85 + j mode j
 * (binary cross-entropy). */
#define MT_PNET_PARAMS 60031
#define MT_PNET_OUTPUTS 863
#define MT_PNET_LABELS 98

/* C[m][n] = sum_k A[m*sam + k*sak] * B[k*sbk + n*sbn], A and B bf16 (device), C fp32 [M][N]:
 * the MFMA GEMM every layer uses. */
int mt_gemm_bf16(int M, int N, int K, const uint16_t* A_d, long sam, long sak, const uint16_t* B_d,
                 long sbk, long sbn, float* C_d, void* stream);
/* Device workspace (bytes) of a batch: bf16 weights, activations, columns, gradients. */
size_t mt_pnet_workspace_bytes(int batch);
/* The frame's 64x64 superblocks in raster order into pix_d [nsb][4096] (edge samples
 * replicated past the frame). */
int mt_pnet_extract(const mt_plane* plane, uint16_t* pix_d, void* stream);
/* Forward of `batch` superblocks: sample b = pix_d[idx_d[b]] (idx_d NULL: b) with qindex
 * qidx_d[idx_d[b]]; logits_d (optional): [batch][863].  Keeps the activations in work_d
 * for mt_pnet_backward. */
int mt_pnet_forward(const float* params_d, const uint16_t* pix_d, const int32_t* idx_d,
                    const uint8_t* qidx_d, int bd, int batch, void* work_d, float* logits_d,
                    void* stream);
/* Backward of the last forward against targets_d / mask_d [batch][98]: per node n < 85 the
 * masked softmax cross-entropy of its 10 partition logits with the partition type targets[n]
 * (0..9), weighted by norm_split; per mode j the masked binary cross-entropy of its score with
 * targets[85 + j], weighted by norm_mode.  grads_d: MT_PNET_PARAMS fp32 (overwritten);
 * loss_terms_d (optional): the [batch][98] weighted loss terms. */
int mt_pnet_backward(const float* params_d, const float* targets_d, const float* mask_d, int batch,
                     float norm_split, float norm_mode, void* work_d, float* grads_d,
                     float* loss_terms_d, void* stream);
/* One Adam step (bias-corrected; `step` counts from 1) on the fp32 parameters. */
int mt_pnet_adam(float* params_d, const float* grads_d, float* m_d, float* v_d, int step, float lr,
                 float beta1, float beta2, float eps, void* stream);
/* mt_intra_rdo prune masks [batch][MT_RDO_NODES] from the logits: on a node of level L
 * (0: 64x64 .. 4: 4x4) the superblock's topk5[L] best-scored intra modes (DC_PRED and
 * filter intra always kept), and for nodes 0..84 MT_RDO_NO_SPLIT when the probability that
 * the node is partitioned, 1 - softmax(partition logits)[NONE], is < lo4[L],
 * MT_RDO_FORCE_SPLIT when > hi4[L] (host arrays).  Between the thresholds a node
 * keeps both choices: the full search there.  Modes fall back the same way: when the
 * product over the dropped modes j of (1 - sigmoid(mode score j)) -- the estimated
 * probability that the kept modes cover every mode the superblock uses -- is below
 * mode_cover (0..1; 0 = never), the node keeps all 13 modes. */
int mt_pnet_prune_masks(const float* logits_d, int batch, const int* topk5, const float* lo4,
                        const float* hi4, float mode_cover, uint16_t* prune_d, void* stream);

/* ---- the AV1 bitstream of the RDO's decisions (host) -------------------------------- */
/* One frame as the RDO left it, in host memory.  The coder writes the syntax of AV1's intra
 * frame (specification 5.9-5.11) with one tile, static frame-initial CDFs
 * (disable_cdf_update = 1, the CDFs the RDO prices against), TX_MODE_LARGEST (every block's
 * transform is its size; 4:4:4 64x64 chroma in four 32x32), DCT_DCT luma and the decoder-
 * derived chroma types, CDEF as given, no restoration / segmentation, the reference's luma DC
 * offset DeltaQYDc = -8 (src/obu.rs:443-451).  The reference writes the same syntax with its
 * own decisions (src/obu.rs:396-528, src/coder/tile_encode.rs). */
typedef struct {
    int32_t width, height;      /* frame size, multiples of 8; one tile: at most 4096 wide and
                                   4096 * 2304 samples (else MT_ERR_UNSUPPORTED) */
    uint8_t bd;                 /* 8, 10, 12 */
    uint8_t chroma;             /* MT_RDO_LUMA_ONLY (monochrome), MT_RDO_CHROMA_420, _444 */
    uint8_t qidx;               /* base_q_idx (1..255) */
    uint8_t qm_level;           /* luma QM level (15 flat) */
    uint8_t uv_qm_level;        /* chroma QM level */
    int8_t uv_dc_delta, uv_ac_delta;
    uint8_t filter_intra;       /* 1: the RDO ran with MT_RDO_FILTER_INTRA (enable_filter_intra) */
    uint8_t lf_level[4];        /* loop_filter_level[0..3] (0..63; luma 0, 0 = no filtering) */
    uint8_t lf_sharpness;       /* 0..7 */
    uint8_t cicp[3];            /* colour primaries, transfer, matrix (2, 2, 2: unspecified);
                                   the identity matrix needs 4:4:4 */
    uint8_t full_range;
    uint8_t pad[3];
    const mt_rdo_node* nodes;   /* [nsb][MT_RDO_NODES], superblocks in raster order */
    const int32_t* coef_y;      /* mt_rdo_frame's level planes, copied to the host */
    const int32_t* coef_u;
    const int32_t* coef_v;
    int32_t coef_stride, coef_uv_stride;
    /* CDEF (5.9.19 cdef_params, 5.11.56 read_cdef), as mt_cdef_frame decided it: cdef_on 0
     * codes enable_cdef = 0; else the damping (3..6), cdef_bits (0..3), the 1 << cdef_bits
     * strength entries (primary 0..15, secondary 0, 1, 2 or 4; uv ignored for monochrome)
     * and, with cdef_bits > 0, each superblock's entry (raster order) coded at its first
     * non-skip block. */
    uint8_t cdef_on, cdef_damping, cdef_bits, cdef_pad;
    uint8_t cdef_y_pri[8], cdef_y_sec[8], cdef_uv_pri[8], cdef_uv_sec[8];
    const uint8_t* cdef_grid;   /* [sb_rows][sb_cols] entry indices; NULL with cdef_bits 0 */
} mt_av1_frame;

/* The luma 8x8 skip map of a frame's decisions (host, [height / 8][width / 8]): 1 where
 * every block covering the 8x8 codes no coefficient in any plane (the AV1 skip flag; CDEF
 * leaves such 8x8s alone, specification 7.15). */
int mt_av1_skip8(const mt_av1_frame* f, uint8_t* skip8);

/* The frame's sequence header OBU and frame OBU (an AV1 temporal unit without its temporal
 * delimiter).  out may be NULL to ask for the size (*len); a smaller cap is MT_ERR_ARG with
 * *len set. */
int mt_av1_encode(const mt_av1_frame* f, uint8_t* out, size_t cap, size_t* len);
/* The same OBUs as an AVIF still image (ftyp, meta: hdlr, pitm, iloc, iinf, iprp with ispe,
 * av1C, pixi, colr; mdat), the container of src/isobmff.rs. */
int mt_avif_encode(const mt_av1_frame* f, uint8_t* out, size_t cap, size_t* len);
/* The frame's entropy tokens in coding order (the wire format between the decisions and the
 * range coder, the role of the reference's EntropyTokens, src/odec.rs:96-141): one uint32 per
 * coded symbol.  Bit 31 clear: a symbol of a stored CDF, offset << 8 | (nsym - 1) << 4 |
 * symbol (mt_av1_cdf returns the CDF); bit 31 set: a bool, cdf0 << 1 | value with cdf0 =
 * 32768 * P(0) (literal bits: cdf0 = 16384).  tokens may be NULL to ask for the count. */
int mt_av1_tokens(const mt_av1_frame* f, uint32_t* tokens, size_t cap, size_t* n);
/* The range coder alone: the tile data of a token stream (specification 8.2.6 inverted,
 * with the exit process's trailing one bit). */
int mt_av1_range_code(const uint32_t* tokens, size_t n, uint8_t* out, size_t cap, size_t* len);
/* The nsym - 1 cumulative values (AV1 form, 32768 * P(symbol <= k)) of a token's CDF. */
int mt_av1_cdf(uint32_t offset, int nsym, uint16_t* cdf);

/* ---- decision pricing ---------------------------------------------------------------- */
/* luma_satd_scale (src/partition_rd.rs:30), k_satd (:54) and proxy_wsse
 * (src/coder/lossy_state.rs:1829) at the shipped tuning (host functions). */
float mt_luma_satd_scale(int base_q_idx, int subsampled);
float mt_k_satd(int base_q_idx, int subsampled);
float mt_proxy_wsse(float d);
/* luma_partition_distortion (src/coder/lossy_state.rs:2385): the SATD partition price of n
 * blocks from their luma_satd (mt_distortion's / mt_eval_candidates' satd), in the proxy
 * unit (the shipped proxy_luma_bridge): satd * max(qstep, 1) / 4 * luma_satd_scale *
 * K_PART_LAMBDA / k_satd * proxy_luma_scale, f32 in that order.  qstep_d (n floats, the
 * ac_q the caller passes) may be NULL: then every block uses `qstep`.  subsampled: 4:2:0,
 * 4:2:2 or monochrome. */
int mt_luma_partition_distortion(const uint64_t* satd_d, const float* qstep_d, float qstep, int n,
                                 int base_q_idx, int subsampled, float* price_d, void* stream);
/* sb_activity (src/coder/aq.rs:140) of every 64x64 superblock of a luma plane (act_d:
 * ceil(W/64) * ceil(H/64) floats, raster order) and, when ref_act_d is not NULL,
 * tile_ref_activity (:163): their f32 mean in raster order (one float on the device). */
int mt_sb_activity(mt_plane src, int bd, float* act_d, float* ref_act_d, void* stream);
/* One block of mt_perceptual_rd_scale: 8 bytes. */
typedef struct {
    uint16_t x, y;     /* luma position */
    uint16_t dim;      /* the block size the caller prices (4..64) */
    uint16_t pad;
} mt_prdoblk;
/* perceptual_rd_scale (src/coder/lossy_state.rs:4189): the per-block lambda scale of n blocks
 * of the tile whose luma source is `src` (its width x height is the tile), at the shipped
 * tuning.  format: MT_ENC_*; mono_large: LossyTile::with_mono_large; ref_act_d: the AQ
 * state's tile activity (one float on the device, e.g. mt_sb_activity's ref_act_d), or NULL
 * for AQ off (0); sb_act_d: mt_sb_activity of the same plane. */
int mt_perceptual_rd_scale(const mt_prdoblk* blocks_d, int n, mt_plane src, int bd, int base_q_idx,
                           int format, int mono_large, const float* ref_act_d, const float* sb_act_d,
                           float* scale_d, void* stream);

/* Launch schedule of the chain's kernel for one transform size on one device (no
 * counterpart in the reference, which runs one block per call on the CPU; the schedule
 * autotuner, maroontree_amd/autotune.py, measures and installs these).  tile_shift scales
 * the blocks a 256-thread workgroup holds in LDS: 1 doubles, 0 keeps, -1 halves the
 * default max(256 / min(W, H), 4096 / (W * H)) (64x64: one wave per block, 0 only);
 * wg_per_cu (1..64) caps the persistent grid at wg_per_cu workgroups per CU.  Applies to
 * mt_tx_chain and mt_tx_chain_noalloc; results never depend on it.  Defaults: 0 / 8, and
 * 0 / 16 at 64x64.  Sizes without a chain kernel give MT_ERR_UNSUPPORTED. */
int mt_tx_schedule_set(int device, int tx_size, int tile_shift, int wg_per_cu);
int mt_tx_schedule_get(int device, int tx_size, int* tile_shift, int* wg_per_cu);

/* Per-kernel timing (no counterpart in the reference; measurement hook).  Between begin
 * and end every launch is bracketed by HIP events on its stream; end synchronizes them and
 * returns, per slot, the summed milliseconds and launch count.  Slot 0 = chain bucketing
 * (3 small kernels), slots 1..13 = the chain's size kernels in the order 4x4, 4x8, 8x4,
 * 4x16, 16x4, 8x8, 8x16, 16x8, 16x16, 16x32, 32x16, 32x32, 64x64; 14 = rate, 15 = intra
 * prediction, 16 = distortion, 17 = candidate evaluation, 18 = the RDO's candidate
 * generation / argmin / decision kernels, 19 = the RDO's superblock reconstruction. */
int mt_profile_begin(void);
int mt_profile_end(double* kernel_ms, int* launches, int max_slots);

/* Synthetic input of the batched transform microbench (host memory):
 * n blocks, TX size uniform over the 13 reference sizes (or all 19 when all_sizes),
 * type uniform over the size's reference types, residual discrete Laplace(12) clipped to
 * +-255 (8-bit) / +-1023 (10-bit, half the blocks), qindex uniform 0-255, QM level flat or
 * qm_level_law(q, 0).  blocks: n entries; pool: pool_elems int16 (mt_gen_pool_size). */
uint64_t mt_gen_pool_size(uint64_t seed, int n, int all_sizes);
int mt_gen_tx_blocks(uint64_t seed, int n, int all_sizes, mt_txblk* blocks, int16_t* pool,
                     uint64_t pool_elems);

#ifdef __cplusplus
}
#endif
#endif /* MTGPU_H */
