/*
 * CPU oracle for the maroontree intra candidate-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar restatement of the reference encoder's
 * kernels, written from the Rust source text, used as the checker for the HIP library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (maroontree_amd / libmtgpu) never links or calls it.
 *
 * Conventions (reference src/quant.rs:95-99, src/rate.rs:93-95):
 *   residual / prediction / reconstruction blocks are row-major  [y*W + x]
 *   levels and targets are x-major ("transposed")                [x*H + y]
 * Sizes and transform types use the AV1 TxSize / TxType numbering.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quantizer state, field for field the reference's `Quant` (src/quant.rs:327-345). */
typedef struct {
    int32_t dc, ac;               /* dequant steps */
    int32_t q_mult_dc, q_mult_ac; /* round(65536 / step), f64 */
    int32_t rmin, rmax, cmin, cmax, cf_max; /* itx clips */
    int32_t qm_level;             /* 15 = flat */
    int32_t qm_chroma;            /* 0 luma, 1 chroma */
    int32_t qidx;
    int32_t bd;
} orc_quant;

/* Quant::new_with_qm (src/quant.rs:354) */
void orc_quant_new(int base_q_idx, int bd, int qm_level, orc_quant* out);
/* Quant::new_chroma_with_delta_qm (src/quant.rs:394) */
void orc_quant_new_chroma(int base_q_idx, int dc_delta, int ac_delta, int bd, int qm_level,
                          orc_quant* out);
/* qm_level_law (src/quant.rs:137) */
int orc_qm_level_law(int base_q_idx, int sub);
int orc_qm_chroma_level_law(int base_q_idx, int sub);
int orc_chroma_dc_delta(int base_q_idx, int sub);
int orc_chroma_ac_delta(int base_q_idx, int sub);
/* qm_weight (src/quant.rs:92); 32 when flat */
int orc_qm_weight(int level, int chroma, int rc, int w, int h);
/* Dct::dequant_step (src/quant.rs:281) */
int orc_dequant_step(const orc_quant* q, int rc, int w, int h);
/* FastRound::fast_round (src/util.rs:35) */
float orc_fast_round(float x);

/* TX size helpers: width/height of an AV1 TxSize, and whether the reference has a
 * forward transform for (size, type).  Returns 0/1. */
int orc_tx_w(int tx_size);
int orc_tx_h(int tx_size);
int orc_tx_supported(int tx_size, int tx_type);
/* number of coded coefficients (64x64 codes its low 32x32) */
int orc_tx_ncoef(int tx_size);

/* Forward transform + quantization + QM, i.e. the reference's DctDispatch::*_t methods
 * (src/dct.rs:2624-2950) on the x86 AVX2 selection.  resid: W*H row-major.
 * levels/targets: ncoef, x-major.  Returns 0, or -1 for an unsupported (size, type). */
int orc_fwd_txfm(int tx_size, int tx_type, const int32_t* resid, const orc_quant* q,
                 int32_t* levels, float* targets);
/* Dequant + inverse transform, the reference's IdctDispatch fields plus idct_dequant_64x64
 * (src/idct.rs:168-728, :1338, :2677).  levels: ncoef x-major.  out: W*H row-major. */
int orc_inv_txfm(int tx_size, int tx_type, const int32_t* levels, const orc_quant* q,
                 int32_t* out);

/* sse_recon_scalar (src/rd_sse.rs:1262): sum over a w x h block of
 * (src - clamp(pred + resid, 0, maxv))^2; pred/resid packed w-wide, src strided. */
int64_t orc_sse_recon(const int32_t* pred, const int32_t* resid, const uint16_t* src,
                      int stride, int w, int h, int maxv);

/* apply_qmatrix in the selected x86 form (src/avx/qmatrix.rs:34): exact round half away
 * from zero; scalar_fastround=1 uses the scalar FastRound form (src/dct.rs:111). */
void orc_apply_qmatrix(int32_t* levels, float* targets, const uint8_t* iwt, int n,
                       int scalar_fastround);

/* 1-D primitives, exported for closed-form and brute-force tests. */
void orc_dct1d(int n, int32_t* buf);          /* dct1d_{4,8,16,32}_i32 (src/dct.rs:171-633) */
void orc_fadst1d(int n, const int32_t* in, int32_t* out); /* fwd_adst{4,8,16}_1d */
void orc_inv_dct1d(int n, int32_t* c, int mn, int mx);    /* inv_dct{4..64}_1d, stride 1 */
void orc_inv_adst1d(int n, int32_t* c, int mn, int mx);   /* inv_adst{4,8,16}_1d */
int32_t orc_mul_q16(int32_t data, int32_t coeff);
int32_t orc_quant_q16(int32_t data, int32_t coeff);

/* The batched transform/quant microbench chain for one block:
 * resid (int16 row-major W*H) -> fwd+quant(+QM) -> dequant -> inverse -> SSE, with the
 * SSE convention sse_recon(pred = P, resid_inv, src = clamp(P + resid)), P = 1 << (bd-1).
 * Any of levels/targets/recon may be NULL.  Returns the SSE, or -1 if unsupported. */
int64_t orc_chain_block(int tx_size, int tx_type, const int16_t* resid, const orc_quant* q,
                        int32_t* levels, float* targets, int32_t* recon);

/* Whole batch, threads over blocks (nthreads <= 0: one thread).  desc/offsets as in the
 * library's microbench layout; sse_out[i] = SSE of block i, -1 if unsupported. */
int orc_chain_batch(int n, const uint8_t* tx_size, const uint8_t* tx_type, const uint8_t* qidx,
                    const uint8_t* bd, const uint8_t* qm_level, const uint64_t* offset,
                    const int16_t* pool, int64_t* sse_out, uint32_t* hash_out, int nthreads);

/* Order-independent per-block checksum over levels, targets and reconstructed residual,
 * the same mix the library computes on the device. */
uint32_t orc_block_hash(int ncoef, const int32_t* levels, const float* targets, int npx,
                        const int32_t* recon);

/* ---- coefficient rate (src/rate.rs:149, src/cost.rs:159-262) ---------------------- */
/* eob_bin: which eob_bin CDF the caller prices with (the reference picks it by size/plane):
 *  0 16 luma, 1 16 chroma, 2 16 luma 1-D, 3 32 luma, 4 32 chroma, 5 64 luma, 6 64 chroma,
 *  7 64 luma 1-D, 8 128 luma, 9 128 chroma, 10 128 luma 1-D, 11 256 luma, 12 256 chroma,
 *  13 512 luma, 14 512 chroma, 15 1024 luma, 16 1024 chroma.
 * txtp_set: 0 none (chroma), 1 the TX_4X4 intra set, 2 the TX_8X8 set, 3 the TX_16X16 set;
 * txtp_mode the luma intra mode row, txtp the set index of the type. */
typedef struct {
    int32_t qctx, cls, plane, w, h, eob_bin, skip_ctx, dcs_ctx;
    int32_t txtp_set, txtp_mode, txtp;
    int32_t use_skip_bits;
    float skip_one, skip_zero;
} orc_rate_ctx;

/* The coefficient part of a decision CDF snapshot in raw form: the library's mt_coef_cdfs
 * layout. */
typedef struct {
    uint16_t skip[5][13];
    uint16_t base_tok[5][2][42][3];
    uint16_t br_tok[4][2][21][3];
    uint16_t eob_base[5][2][4][2];
    uint16_t eob_hi[5][2][11];
    uint16_t dc_sign[2][3];
    uint16_t eob_bin[17][10];
    uint16_t txtp4[13][6], txtp8[13][6], txtp16[13][4];
    uint16_t pad;
} orc_coef_cdfs;
/* The raw snapshot of Cdfs::new(qctx), and loading one into slot 4..7 (the qctx of
 * orc_rate_ctx selects the slot; 0..3 are the frame-initial snapshots). */
void orc_rate_default_cdfs(int qctx, orc_coef_cdfs* out);
int orc_rate_load_cdfs(int slot, const orc_coef_cdfs* cdfs);

/* real_block_bits_bounded against Cdfs::new(qctx); cf x-major, scan of the block size. */
float orc_block_rate(const int32_t* cf, const uint32_t* scan, const orc_rate_ctx* c, float bound);
/* cost_q_table entry (Q22 bits) and the cost of one symbol of a raw AV1 CDF */
uint32_t orc_cost_q(int p);
float orc_cdf_cost_raw(const uint16_t* raw_cdf, int nsym, int s);
float orc_skip_cost(int qctx, int cls, int ctx, int s);
/* the scan table of an AV1 TxSize (64x64 uses the 32x32 scan); NULL if none */
const uint32_t* orc_scan(int tx_size);

/* ---- trellis quantization (src/trellis.rs) ---------------------------------------------
 * cf (levels, x-major) is rewritten in place from the f32 targets tf.  orc_trellis is
 * trellis_optimize (:944) with the x86 AVX2+FMA round-down; orc_trellis_ctx is
 * trellis_optimize_ctx (:55) on the shipped switches, pricing against the CDFs of snapshot
 * c->qctx with c->cls / plane / w / h / eob_bin / dcs_ctx (w, h: the coded block, 32 x 32
 * for 64 x 64). */
void orc_trellis(int32_t* cf, const float* tf, float dc_q, float ac_q, const uint32_t* scan, int n,
                 float lambda0);
void orc_trellis_ctx(int32_t* cf, const float* tf, float dc_q, float ac_q, const uint32_t* scan,
                     float lambda0, const orc_rate_ctx* c, int qm_level, float band_tilt);
/* trellis_lambda() (src/cost.rs:110) times hbd_trellis_scale (src/coder/lossy_state.rs:158,
 * shipped tuning); sub 0 4:4:4, 1 4:2:2, 2 4:2:0. */
float orc_trellis_lambda0(int qidx, int bd, int sub);
/* Trellis parameters of one block, the library's mt_trellisblk layout. */
typedef struct {
    uint32_t offset;
    uint8_t tx_size, qidx, bd_qm, flags;
    int8_t dc_delta, ac_delta;
    uint8_t qctx, eob_bin, dcs_ctx, plane, pad0, pad1;
    float lambda0, band_tilt;
} orc_trellisblk;
/* The chain with the trellis between quantization and dequantization: per block i the
 * chain of orc_chain_batch, with trellis tr[i] (flags bit 0: the context variant) applied
 * to the levels before the inverse; the checksum covers the final levels. */
int orc_chain_batch_trellis(int n, const uint8_t* tx_size, const uint8_t* tx_type, const uint8_t* qidx,
                            const uint8_t* bd, const uint8_t* qm_level, const uint64_t* offset,
                            const int16_t* pool, const orc_trellisblk* tr, int64_t* sse_out,
                            uint32_t* hash_out, int nthreads);


/* ---- intra prediction (src/intrapred.rs) ------------------------------------------- */
/* Edges are passed as arrays of 132 samples whose element 2 is logical index 0 (the
 * reference's IntraEdge, origin 2). */
int orc_edge_filter_strength(int w, int h, int filter_type, int delta); /* :917 */
int orc_use_edge_upsample(int w, int h, int filter_type, int delta);    /* :1021 */
void orc_filter_edge(int32_t* edge_origin2, int size, int strength);    /* :977 */
void orc_upsample_edge(int32_t* edge_origin2, int num_px, int bd);      /* :1026 */
/* dr_predict_scalar (:629); zone 1, 2 or 3 */
void orc_dr_predict(int zone, int bw, int bh, int edge_len, int dx, int dy, int up_above,
                    int up_left, const int32_t* above_origin2, const int32_t* left_origin2,
                    int32_t* out);
int32_t orc_dc_pred_from_sum(int32_t edge_sum, int w, int h, int have_top, int have_left, int bd);
int32_t orc_dc_pred(const uint16_t* recon, int stride, int ox, int oy, int w, int h, int bd);
/* filter intra mode 0..4 (:473); w multiple of 4, h even, both <= 32 */
int orc_filter_intra(int fmode, const uint16_t* recon, int stride, int ox, int oy, int w, int h,
                     int bd, int32_t* out);
/* predict_nd_ad (:359) for modes 1..12; mode 0 fills dc_pred; 16+k filter intra mode k.
 * recon: plane (stride samples per row); (ox, oy) block origin; fw/fh the frame size of the
 * px_have rules.  out: bw*bh row-major.  Returns 0, or -1 for an invalid call. */
int orc_intra_pred(int mode, int angle_delta, const uint16_t* recon, int stride, int ox, int oy,
                   int bw, int bh, int have_tr, int have_bl, int fw, int fh, int filter_type,
                   int bd, int32_t* out);

/* ---- CfL (src/intrapred.rs:1112-1260) ---------------------------------------------- */
void orc_cfl_ac_444(const uint16_t* luma, int w, int h, int32_t* ac);
void orc_cfl_ac_sub(const uint16_t* luma, int lstride, int cw, int ch, int ss_hor, int ss_ver,
                    int32_t* ac);
int32_t orc_cfl_pred_pixel(int32_t dc, int32_t ac, int32_t alpha, int bd);
int32_t orc_cfl_best_alpha(const int32_t* ac, int ac_stride, const uint16_t* src, int src_stride,
                           int w, int h, int32_t dc, int bd);
/* cfl_alpha_rd_refine's walk (src/cost.rs:332) over walk costs cost[a + 16], and the alpha
 * s steps from a0 along a0's sign (s < 0 toward zero; clipped to -16..16, never past 0). */
int32_t orc_cfl_alpha_step(int32_t a0, int32_t s);
int32_t orc_cfl_alpha_rd_refine(int32_t a0, float c0, int steps, const float* cost);

/* ---- CDEF (src/cdef.rs, src/coder/frame_cdef.rs) ------------------------------------ */
typedef struct {
    int32_t bits, damping, n;  /* cdef_bits, the signalled damping (3..6), 1 << bits entries */
    uint8_t y_pri[8], y_sec[8], uv_pri[8], uv_sec[8];
} orc_cdef_params;
int orc_cdef_dir(const uint16_t* p, int stride, int w, int h, int x, int y, int bd, int32_t* var);
int32_t orc_cdef_adjust_pri(int32_t pri, int32_t var);
void orc_cdef_filter(const uint16_t* p, int stride, int w, int h, int x, int y, int bw, int bh, int pri, int sec,
                     int dir, int damping, int bd, uint16_t* dst, int dstride);
void orc_cdef_candidates(const uint16_t* p, int stride, int w, int h, int x, int y, int bw, int bh, int dir,
                         int32_t var, int damping, int bd, int32_t* out);
/* frame_cdef: decision and apply in place (planes w x h, chroma (w >> sub_x) x (h >> sub_y),
 * w and h multiples of 8).  skip8: [h/8][w/8] luma 8x8 skip flags.  Returns 1 when CDEF is
 * applied, 0 when off, -1 on bad arguments.  Optional outputs: grid [units], dirs / vars
 * [h/8][w/8], tables [2][units][12] (luma, then U + V). */
int orc_cdef_frame(uint16_t* rec_y, uint16_t* rec_u, uint16_t* rec_v, const uint16_t* src_y, const uint16_t* src_u,
                   const uint16_t* src_v, int w, int h, int sub_x, int sub_y, int mono, const uint8_t* skip8,
                   int qidx, int bd, int slow, float wsse_scale, float lambda, orc_cdef_params* prm,
                   uint8_t* grid_out, int32_t* dirs_out, int32_t* vars_out, int64_t* tables_out);

/* ---- residual and distortion (src/rd_sse.rs, src/partition_rd.rs) ----------------- */
void orc_residual_pred(const uint16_t* src, int stride, const int32_t* pred, int w, int h,
                       int32_t* dst);
/* plain SSE of clamp(pred + resid) over the visible vw x vh corner; resid may be NULL */
int64_t orc_sse_recon_vis(const int32_t* pred, const int32_t* resid, const uint16_t* src,
                          int stride, int w, int vw, int vh, int bd);
void orc_wsse_tables(float* dct16, float* weights16);
float orc_weighted_sse_scalar(const int32_t* err, int w, int h);
float orc_weighted_sse_avx2(const int32_t* err, int w, int h);
/* weighted_dist: err_vis packed vw wide; avx2 selects the op order of weighted_sse_avx2 */
int64_t orc_weighted_dist(const int32_t* err_vis, int vw, int vh, float scale, int avx2);
uint64_t orc_luma_satd(const uint16_t* src, int stride, int w, int h, int vw, int vh, int bd,
                       const int32_t* pred, const int32_t* resid);
uint64_t orc_satd_sad_proxy(const uint16_t* src, int stride, int w, int h, int vw, int vh,
                            const int32_t* pred);


/* ---- one intra candidate end to end (src/coder/partition_search.rs:471-520) ---------- */
/* Same layout as the library's candidate record (48 bytes).  Mode 13 is chroma-from-luma:
 * DC prediction + alpha * the luma AC of plane txtp_mode at (x << ss_x, y << ss_y); pflags
 * 8 / 16 subsampling, 32 alpha given (in angle_delta). */
typedef struct {
    uint32_t offset;
    uint16_t x, y;
    uint8_t tx_size, tx_type, mode;
    int8_t angle_delta;
    uint8_t pflags, planes, qidx, bd_qm;
    uint8_t rplane, eob_bin, skip_ctx, dcs_ctx;
    uint8_t txtp_set, txtp_mode, txtp, qctx;
    uint8_t rflags, planes_hi; /* high bits of the plane indices (bits 0-3 source, 4-7 recon) */
    int8_t dc_delta, ac_delta; /* chroma (bd_qm bit 6) qindex deltas */
    float skip_one, skip_zero, wsse_scale, lambda, extra_bits;
} orc_cand;
typedef struct {
    int64_t sse, dist;
    uint64_t satd;
    float bits, cost;
    uint32_t nnz;
    int32_t status; /* 0, -1 invalid, -2 unsupported (size, type) */
    int32_t alpha;  /* CfL candidates (mode 13): the alpha used */
    int32_t pad;
} orc_candres;
typedef struct {
    const uint16_t* data;
    int32_t stride, width, height, vis_w, vis_h, pad;
} orc_plane;
void orc_eval_candidate(const orc_cand* c, const orc_plane* planes, int nplanes, orc_candres* r);
void orc_eval_candidate_recon_levels(const orc_cand* c, const orc_plane* planes, int nplanes, orc_candres* r,
                                     int32_t* recon, int32_t* levels);
void orc_eval_batch(const orc_cand* c, int n, const orc_plane* planes, int nplanes,
                    orc_candres* r, int nthreads);
/* orc_eval_candidate that also returns the reconstruction clamp(pred + inverse), w*h */
void orc_eval_candidate_recon(const orc_cand* c, const orc_plane* planes, int nplanes,
                              orc_candres* r, int32_t* recon);

/* ---- whole-frame luma intra RDO over the superblock wavefront ---------------------- */
/* The library's mt_intra_rdo procedure (see include/mtgpu.h) on the CPU, candidate by
 * candidate through orc_eval_candidate. */
typedef struct {
    int32_t width, height;  /* multiples of 8 */
    int32_t bd, qidx, qm_level, flags;  /* 1 angle deltas, 2 filter intra, 4 to 4x4, 8 CfL, 16 HORZ / VERT */
    float lambda, wsse_scale;
    int32_t chroma;                     /* 0 luma only, 1 4:2:0, 2 4:4:4 */
    int32_t uv_dc_delta, uv_ac_delta, uv_qm_level;
} orc_rdo_params;
typedef struct {
    float cost, none_cost;
    uint8_t split, mode, rmode0, rmode1; /* split: 0 NONE, 1 SPLIT, 2 not coded, 3 HORZ, 4 VERT */
    uint8_t uv_mode, uv_mode1;
    int8_t cfl_alpha_u, cfl_alpha_v, cfl_alpha_u1, cfl_alpha_v1;
    uint8_t pad[2];
} orc_rdo_node;
typedef struct {
    double bits;
    int64_t sse, candidates;
    int32_t leaves, sbs;
    double bits_uv;
    int64_t sse_u, sse_v;
} orc_rdo_stats;
/* src: width x height samples, src_stride per row; prune: NULL or [nsb][341] node masks
 * (nsb in SB raster order); recon: width x height out (stride width); nodes: [nsb][341]
 * out; modes (optional): (width/4) x (height/4) out, the final luma mode map (filter intra
 * as DC_PRED).  Returns 0, or -1 for invalid parameters. */
int orc_intra_rdo(const uint16_t* src, int src_stride, const orc_rdo_params* p,
                  const uint16_t* prune, uint16_t* recon, orc_rdo_node* nodes,
                  uint8_t* modes, orc_rdo_stats* stats, int nthreads);
/* with chroma (p->chroma 1 or 2): src_u / src_v of the chroma size (uv_stride per row),
 * recon_u / recon_v out (stride: the chroma width) */
int orc_intra_rdo_yuv(const uint16_t* src, int src_stride, const uint16_t* src_u, const uint16_t* src_v,
                      int uv_stride, const orc_rdo_params* p, const uint16_t* prune, uint16_t* recon,
                      uint16_t* recon_u, uint16_t* recon_v, orc_rdo_node* nodes, uint8_t* modes,
                      orc_rdo_stats* stats, int nthreads);
/* the decision lambda of the reference (src/coder/lossy_state.rs:1769-1790, src/cost.rs:132,
 * :285): mode_lambda_q(dc_q, bd) * mode_lambda_weight(qidx, sub) / K_PART_LAMBDA */
/* orc_intra_rdo_yuv with the reconstruction's levels (optional, NULL = none): the level of
 * frequency (fx, fy) of the transform block at (x, y) stored at sample (x + fx, y + fy) of a
 * W x H int32 plane (chroma: the chroma size); 64x64 blocks fill their low 32x32. */
int orc_intra_rdo_yuv_coef(const uint16_t* src, int src_stride, const uint16_t* src_u, const uint16_t* src_v,
                           int uv_stride, const orc_rdo_params* p, const uint16_t* prune, uint16_t* recon,
                           uint16_t* recon_u, uint16_t* recon_v, orc_rdo_node* nodes, uint8_t* modes,
                           orc_rdo_stats* stats, int nthreads, int32_t* coef, int32_t* coef_u, int32_t* coef_v);
float orc_rdo_lambda(int qidx, int bd, int sub);
float orc_mode_lambda(int qidx, int bd, int sub);
/* the RDO's top-right / bottom-left availability of an s x s block at (xin, yin) of its SB */
int orc_rdo_have_tr(int xin, int yin, int s);
int orc_rdo_have_bl(int xin, int yin, int s);
/* half k (0 / 1 top / bottom of HORZ, 2 / 3 left / right of VERT) of the s x s node at
 * (xin, yin): bit 0 top-right, bit 1 bottom-left coded before it */
int orc_rdo_rect_avail(int xin, int yin, int s, int k);

/* the reference's RGB -> YCbCr conversion (c420: 4:2:0 chroma of the rounded-up half size,
 * else 4:4:4) and quality_to_q (oracle_csc.cpp) */
void orc_csc_rgb8(const uint8_t* r, const uint8_t* g, const uint8_t* b, int w, int h, int c420, uint16_t* y,
                  uint16_t* cb, uint16_t* cr);
int orc_quality_to_q(int quality);

/* ---- in-loop deblocking (oracle_lf.cpp) ---- */
/* limits (src/loopfilter.rs:33) */
void orc_lf_limits(int level, int sharp, int* e, int* i, int* h);
/* loop_filter_scalar (src/loopfilter.rs:106): four lines from px[base], stride_a between
 * lines, stride_b across the edge */
void orc_loop_filter(uint16_t* px, int64_t base, int e, int i, int h, int64_t stride_a, int64_t stride_b,
                     int wd, int bd);
/* filter_plane_impl (src/loopfilter.rs:483): px packed w x h; vedge4 / hedge4 may be NULL
 * (edges where the position is a multiple of the size) */
void orc_filter_plane(uint16_t* px, int w, int h, int vis_w, int vis_h, const uint8_t* bw4, const uint8_t* bh4,
                      const uint8_t* vedge4, const uint8_t* hedge4, int nc4, int level, int sharp, int is_luma,
                      int sb_rows4, int bd);
/* chroma_deblock_geom (src/coder/frame_deblock.rs:91) */
void orc_chroma_deblock_geom(const uint8_t* pblk4, const uint8_t* pblk4h, int nc4, int cw8, int ch8, int sub_x,
                             int sub_y, uint8_t* cbw4, uint8_t* cbh4);
/* frame_deblock (src/coder/frame_deblock.rs:206): planes packed, chroma (w8 >> sub_x) wide */
void orc_frame_deblock(uint16_t* y, uint16_t* u, uint16_t* v, int w8, int h8, int disp_w, int disp_h,
                       const uint8_t* blk4, const uint8_t* blk4h, const uint8_t* blk4v, const uint8_t* blk4t,
                       const uint8_t* pblk4, const uint8_t* pblk4h, int sub_x, int sub_y, int mono, int sharp,
                       int bd, int lvl_y, int lvl_u, int lvl_v);
/* loop_filter_levels / loop_filter_sharpness (src/obu.rs:306-334), shipped tuning */
void orc_lf_levels(int qidx, int sub, int* lvl_y, int* lvl_uv, int* sharp);

/* ---- decision pricing helpers (oracle_price.cpp) ---- */
float orc_dirty_log1pf(float d);              /* src/aq_common.rs:90 */
float orc_dirty_exp2f(float d);               /* src/coder/lossy_state.rs:4284 */
float orc_exp2_table(int i);                  /* its table, 2^((i - 32) / 64) */
void orc_block_moments(const uint16_t* plane, int stride, int px, int py, int w, int h,
                       int64_t* sum, int64_t* sum_sq);
float orc_sb_activity(const uint16_t* yp, int pw, int sb_y, int sb_x, int width, int height, int bd);
float orc_tile_ref_activity(const uint16_t* yp, int pw, int w, int h, int bd);
float orc_luma_satd_scale(int base_q_idx, int subsampled);   /* src/partition_rd.rs:30 */
float orc_k_satd(int base_q_idx, int subsampled);            /* src/partition_rd.rs:54 */
float orc_luma_partition_distortion(uint64_t satd, float qstep, int base_q_idx, int subsampled);
float orc_proxy_wsse(float d);
float orc_perceptual_rd_scale(const uint16_t* src, int pw, int w, int h, int bd, int base_q_idx,
                              int fmt, int mono_large, float ref_act, int px, int py, int dim);

#ifdef __cplusplus
}
#endif
