// Trellis quantization of one transform block: trellis_optimize / trellis_optimize_ctx
// (reference src/trellis.rs:944, :55), shared by the batched trellis kernel (mt_rate.hip) and
// the reconstruction of the intra RDO (mt_sbrecon.h, in the mtgpu.hip unit).
//
// One thread walks the block's scan sequentially with every f32 operation in the
// reference's order.  The context variant's level map (stride h, the reference's shipped
// layout) is the block's own x-major levels: every neighbour a context reads lies later in
// the scan, so it holds its final Step-A level when read, and positions past the block read
// as zero.  The suffix sums of the zero-level distortion take a float workspace of one
// element per coefficient.  The tables live in the rate unit's device memory and are reached
// through their addresses (TrellisTabs), so every unit reads the same copy.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/mtgpu.h"
#include "av1_tables.h"

namespace mtlib {
// Costs (bits) of every symbol the rate walk prices, for one qctx.
struct RateTables {
    float skip[5][13][2];
    float bt[5][2][42][4];    // CoefCostTables::bt (src/rate.rs:50)
    float br[4][2][21][13];   // CoefCostTables::br, cumulative base-range ladders
    float eob_base[5][2][4][3];
    float eob_hi[5][2][11][2];
    float dc_sign[2][3][2];
    float eob_bin[17][11];
    float txtp[3][13][7];     // intra tx-type sets of TX_4X4 / TX_8X8 / TX_16X16
};
// Device addresses of the trellis's tables on one device (mt_rate.hip fills them at init).
struct TrellisTabs {
    const RateTables* rate;  // [MT_RATE_SLOTS]
    const uint16_t* scan;    // scans concatenated per size (scan_offset)
    const uint8_t* lo_off;   // [3][5][5]: square, w < h, w > h (src/tables.rs:354, :456, :466)
    const uint16_t* dcq;     // [3][256] dc_q by bit depth
    const uint16_t* acq;     // [3][256] ac_q
    const uint8_t* iqm;      // [MT_QM_LEVELS][2][MT_QM_USED_SIZE]
};
TrellisTabs trellis_tabs(int device);
}  // namespace mtlib

namespace {

// Offset of a TxSize's scan in the concatenated scans (64x64 codes its 32x32 with the 32x32
// scan).
__host__ __device__ inline int scan_offset(int s) {
    switch (s) {
        case MT_TX_4X4: return 0;
        case MT_TX_4X8: return 16;
        case MT_TX_8X4: return 48;
        case MT_TX_8X8: return 80;
        case MT_TX_4X16: return 144;
        case MT_TX_16X4: return 208;
        case MT_TX_8X16: return 272;
        case MT_TX_16X8: return 400;
        case MT_TX_16X16: return 528;
        case MT_TX_16X32: return 784;
        case MT_TX_32X16: return 1296;
        case MT_TX_32X32: case MT_TX_64X64: return 1808;
        default: return -1;
    }
}

__device__ inline uint32_t uabs32(int32_t v) { return v < 0 ? 0u - (uint32_t)v : (uint32_t)v; }

// level_byte (src/tables.rs:365)
__device__ inline uint32_t level_byte(uint32_t m) {
    return m == 0 ? 0u : m <= 2 ? m * 0x41u : (m < 15u ? m : 15u) + 0xC0u;
}

__device__ inline float golomb_bits(uint32_t v) {
    return (float)(2 * (32 - __clz((int)(v + 1))) - 1);
}

// COEF_RATE_LUT (src/cost.rs:73) and the closed form above it (coef_rate_bits_slow, :52;
// f_fmlaf unfused as on the default x86-64 build)
__constant__ float c_coef_rate[64] = {
    0.9f, 2.7f, 3.6f, 5.3f, 5.3f, 5.3f, 6.6f, 6.6f, 6.6f, 7.9f, 7.9f, 7.9f, 9.2f, 9.2f, 9.2f, 11.5f,
    13.5f, 13.5f, 15.5f, 15.5f, 15.5f, 15.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f,
    19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f,
    19.5f, 19.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f,
    21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f};

__device__ inline float coef_rate_bits(uint32_t l) {
    if (l < 64) return c_coef_rate[l];
    float b = 4.0f;
    b = __fadd_rn(b, __fmul_rn((float)(12 / 3 + 1), 1.3f));
    b = __fadd_rn(b, __fadd_rn(__fmul_rn(2.0f, (float)(32 - __clz((int)(l - 15 + 1)))), -1.0f));
    return b;
}

// scaled_trellis_lambda (src/trellis.rs:41-52): TRELLIS_AOM_CALIB (an f32 product) times
// trellis_calib 1.15 (src/tuning.rs:953) times dc_q^2 (3.3 + 0.0015 dc_q), left to right,
// times lambda0 / TRELLIS_LAMBDA0 (0.05)
__device__ inline float trellis_lambda(float dc_q, float lambda0) {
    const float calib = __fmul_rn(0.045025873597302174f, 0.5525f);
    float a = __fmul_rn(calib, 1.15f);
    a = __fmul_rn(a, dc_q);
    a = __fmul_rn(a, dc_q);
    a = __fmul_rn(a, __fadd_rn(3.3f, __fmul_rn(0.0015f, dc_q)));
    return __fmul_rn(a, __fdiv_rn(lambda0, 0.05f));
}

__device__ inline int bitlen32(uint32_t v) { return v ? 32 - __clz((int)v) : 0; }

struct TrellisGeom {
    int so, cls, cb, w, h, N, log2h, oi, coded_w;
};

__device__ inline bool trellis_geom(int s, TrellisGeom& g) {
    g.so = s < 19 ? scan_offset(s) : -1;
    if (g.so < 0) return false;
    const int wf = s == MT_TX_4X4 || s == MT_TX_4X8 || s == MT_TX_4X16 ? 4
                 : s == MT_TX_8X8 || s == MT_TX_8X4 || s == MT_TX_8X16 ? 8
                 : s == MT_TX_16X16 || s == MT_TX_16X8 || s == MT_TX_16X4 || s == MT_TX_16X32 ? 16
                 : s == MT_TX_64X64 ? 64 : 32;
    const int hf = s == MT_TX_4X4 || s == MT_TX_8X4 || s == MT_TX_16X4 ? 4
                 : s == MT_TX_8X8 || s == MT_TX_4X8 || s == MT_TX_16X8 ? 8
                 : s == MT_TX_16X16 || s == MT_TX_8X16 || s == MT_TX_4X16 || s == MT_TX_32X16 ? 16
                 : s == MT_TX_64X64 ? 64 : 32;
    g.cls = ((31 - __clz(wf)) - 2 + (31 - __clz(hf)) - 2 + 1) >> 1;
    g.cb = g.cls < 3 ? g.cls : 3;
    g.w = wf > 32 ? 32 : wf;
    g.h = hf > 32 ? 32 : hf;
    g.N = g.w * g.h;
    g.log2h = 31 - __clz(g.h);
    g.oi = g.w < g.h ? 1 : (g.w > g.h ? 2 : 0);
    return true;
}

// QM level-table offset of a coded w x h block (the 64 x 64 block's coded 32 x 32)
__device__ inline int trellis_qm_offset(int w, int h) {
    if (w == h) return w == 4 ? 0 : w == 8 ? 16 : w == 16 ? 80 : 336;
    if (w == 4 && h == 8) return 1360;
    if (w == 8 && h == 4) return 1392;
    if (w == 8 && h == 16) return 1424;
    if (w == 16 && h == 8) return 1552;
    if (w == 16 && h == 32) return 1680;
    if (w == 32 && h == 16) return 2192;
    if (w == 4 && h == 16) return 2704;
    return 2768;  // 16 x 4
}

// trellis_round_down_scan_avx2 (src/avx/trellis.rs:196): chunks of 8, then 4 scan positions
// take the fused costs fma(q2, e^2, lambda * LUT[l]) unless a level of the chunk is 64 or
// more, in which case all of its positions take the scalar (unfused) form; the tail is scalar.
__device__ inline void round_down_one(int32_t* cf, const float* tf, uint32_t rc, float q2, float lam) {
    const int32_t c = cf[rc];
    if (c == 0) return;
    const uint32_t l = uabs32(c);
    const float at = fabsf(tf[rc]);
    const float el = __fsub_rn(at, (float)l), edn = __fsub_rn(at, (float)(l - 1));
    const float cl = __fadd_rn(__fmul_rn(q2, __fmul_rn(el, el)), __fmul_rn(lam, coef_rate_bits(l)));
    const float cd = __fadd_rn(__fmul_rn(q2, __fmul_rn(edn, edn)), __fmul_rn(lam, coef_rate_bits(l - 1)));
    if (cd < cl) cf[rc] = c < 0 ? c + 1 : c - 1;
}

__device__ inline void round_down_chunk(int32_t* cf, const float* tf, const uint16_t* scan, int i0, int len,
                                        float q2dc, float q2ac, float lam) {
    bool big = false;
    for (int k = 0; k < len; k++) big |= uabs32(cf[scan[i0 + k]]) > 63;
    if (big) {
        for (int k = 0; k < len; k++) {
            const uint32_t rc = scan[i0 + k];
            round_down_one(cf, tf, rc, rc == 0 ? q2dc : q2ac, lam);
        }
        return;
    }
    uint32_t lower = 0;
    for (int k = 0; k < len; k++) {
        const uint32_t rc = scan[i0 + k];
        const uint32_t l = uabs32(cf[rc]), ldn = l ? l - 1 : 0;
        const float at = fabsf(tf[rc]);
        const float q2 = (i0 == 0 && k == 0) ? q2dc : q2ac;
        const float el = __fsub_rn(at, (float)l), edn = __fsub_rn(at, (float)ldn);
        const float cl = __builtin_fmaf(q2, __fmul_rn(el, el), __fmul_rn(lam, c_coef_rate[l]));
        const float cd = __builtin_fmaf(q2, __fmul_rn(edn, edn), __fmul_rn(lam, c_coef_rate[ldn]));
        lower |= (cd < cl ? 1u : 0u) << k;
    }
    for (int k = 0; k < len; k++) {
        const uint32_t rc = scan[i0 + k];
        const int32_t c = cf[rc];
        if (((lower >> k) & 1) && c != 0) cf[rc] = c < 0 ? c + 1 : c - 1;
    }
}

// trellis_lambda() * hbd_trellis_scale (mt_trellis_lambda0): f32, the same operations.
__device__ inline float trellis_lambda0(int qidx, int bd, int sub) {
    float scale = 1.0f;
    if (bd > 8 && (sub == 1 || sub == 2)) {
        const float top = sub == 2 ? 0.174f : 1.0f, hi = sub == 2 ? 36.0f : 24.0f, lo = sub == 2 ? 20.0f : 10.0f;
        float wgt = __fdiv_rn(__fsub_rn(hi, (float)qidx), fmaxf(__fsub_rn(hi, lo), 1.0f));
        wgt = fminf(fmaxf(wgt, 0.0f), 1.0f);
        scale = __fadd_rn(__fmul_rn(__fsub_rn(top, 1.0f), wgt), 1.0f);
    }
    return __fmul_rn(0.05f, scale);
}

// The trellis of one block d over its levels cf and targets tf (x-major; generic pointers:
// global or LDS), with suf a float workspace of the block's coefficient count.  Returns
// MT_ERR_ARG (levels untouched) for an invalid descriptor, else MT_OK.
__device__ int trellis_block(const mt_trellisblk& d, int32_t* cf, const float* tf, float* suf,
                             const mtlib::TrellisTabs& T) {
    TrellisGeom G;
    const bool ctxv = d.flags & MT_TRELLIS_CTX;
    const int bdi = (d.bd_qm >> 4) & 3;
    bool ok = trellis_geom(d.tx_size, G) && bdi < 3 && !(d.flags & ~MT_TRELLIS_CTX);
    if (ok && ctxv) {
        const int ebn = d.eob_bin < 3 ? 16 : d.eob_bin < 5 ? 32 : d.eob_bin < 8 ? 64 : d.eob_bin < 11 ? 128
                      : d.eob_bin < 13 ? 256 : d.eob_bin < 15 ? 512 : 1024;
        ok = d.eob_bin <= 16 && ebn == G.N && d.dcs_ctx <= 2 && d.plane <= 1 && d.qctx < MT_RATE_SLOTS;
    }
    if (!ok) return MT_ERR_ARG;
    if (!(d.lambda0 > 0.0f)) return MT_OK;
    // the quantizer steps (Quant::new_with_qm / new_chroma_with_delta_qm)
    const bool chroma_q = d.bd_qm & MT_TXBLK_CHROMA;
    const int qi = d.qidx;
    const int dci = chroma_q ? min(255, max(0, qi + d.dc_delta)) : max(0, qi - 8);
    const int aci = chroma_q ? min(255, max(0, qi + d.ac_delta)) : qi;
    const float dc_q = (float)T.dcq[bdi * 256 + dci], ac_q = (float)T.acq[bdi * 256 + aci];
    const int qm_level = d.bd_qm & 15;
    const uint16_t* scan = T.scan + G.so;
    const int N = G.N, h = G.h, log2h = G.log2h;
    const float lam = trellis_lambda(dc_q, d.lambda0);
    const float q2dc = __fmul_rn(dc_q, dc_q), q2ac = __fmul_rn(ac_q, ac_q);
    int eob = N - 1;
    while (eob >= 0 && cf[scan[eob]] == 0) eob--;
    if (eob < 0) return MT_OK;
    const int m = eob + 1;

    if (!ctxv) {
        // trellis_optimize (src/trellis.rs:944)
        int i = 0;
        for (; i + 8 <= m; i += 8) round_down_chunk(cf, tf, scan, i, 8, q2dc, q2ac, lam);
        for (; i + 4 <= m; i += 4) round_down_chunk(cf, tf, scan, i, 4, q2dc, q2ac, lam);
        for (; i < m; i++) round_down_one(cf, tf, scan[i], scan[i] == 0 ? q2dc : q2ac, lam);
        float acc = 0.0f;
        for (int k = m - 1; k >= 0; k--) {
            const uint32_t rc = scan[k];
            const float at = fabsf(tf[rc]);
            acc = __fadd_rn(acc, __fmul_rn(rc == 0 ? q2dc : q2ac, __fmul_rn(at, at)));
            suf[k] = acc;
        }
        int best_e = -1;
        float best = INFINITY, pre = 0.0f;
        for (int e = 0; e < m; e++) {
            const uint32_t rc = scan[e];
            const uint32_t l = uabs32(cf[rc]);
            const float q2 = rc == 0 ? q2dc : q2ac;
            const float er = __fsub_rn(fabsf(tf[rc]), (float)l);
            const float dcur = __fmul_rn(q2, __fmul_rn(er, er));
            if (l != 0) {
                const int bin = e < 2 ? e : bitlen32((uint32_t)e);
                const int extra = bin > 1 ? bin - 2 : 0;
                const float sig = __fadd_rn(__fmul_rn((float)bin, 0.9f), __fadd_rn((float)extra, 2.0f));
                const float c = __fadd_rn(__fadd_rn(__fadd_rn(pre, dcur), __fmul_rn(lam, sig)),
                                          e + 1 < m ? suf[e + 1] : 0.0f);
                if (c < best) {
                    best = c;
                    best_e = e;
                }
            }
            pre = __fadd_rn(__fadd_rn(pre, dcur), __fmul_rn(lam, coef_rate_bits(l)));
        }
        const float skip = __fadd_rn(suf[0], __fmul_rn(lam, 1.0f));
        const int keep = (best_e < 0 || skip < best) ? 0 : best_e + 1;
        for (int k = keep; k < m; k++) cf[scan[k]] = 0;
        return MT_OK;
    }

    // trellis_optimize_ctx (src/trellis.rs:55), shipped switches
    const mtlib::RateTables& RT = T.rate[d.qctx];
    const int cls = G.cls, cb = G.cb, pl = d.plane;
    const float(*bt)[4] = RT.bt[cls][pl];
    const float(*br)[13] = RT.br[cb][pl];
    const float(*eob_base)[3] = RT.eob_base[cls][pl];
    const float(*eob_hi)[2] = RT.eob_hi[cls][pl];
    const float* dc_sign = RT.dc_sign[pl][d.dcs_ctx];
    const float* eob_bin = RT.eob_bin[d.eob_bin];
    const bool qm_on = qm_level < 15, qm_full = qm_level >= 5;
    const uint8_t* qmt = T.iqm + ((qm_on ? qm_level : 0) * 2 + (pl ? 1 : 0)) * MT_QM_USED_SIZE + trellis_qm_offset(G.w, h);
    const float tilt = d.band_tilt;
    const float n_inv = __fdiv_rn(1.0f, (float)N);
    auto qm_at = [&](uint32_t rc) {
        const float wq = __fdiv_rn((float)qmt[rc], 32.0f);
        return qm_full ? __fmul_rn(wq, wq) : wq;
    };
    auto distw = [&](int idx, uint32_t rc, uint32_t lev) {
        const float e = __fsub_rn(fabsf(tf[rc]), (float)lev);
        float dd = __fmul_rn(rc == 0 ? q2dc : q2ac, __fmul_rn(e, e));
        if (qm_on) dd = __fmul_rn(dd, qm_at(rc));
        if (tilt != 0.0f && fabsf(tf[rc]) < 2.0f)
            dd = __fdiv_rn(dd, __fadd_rn(1.0f, __fmul_rn(tilt, __fmul_rn((float)idx, n_inv))));
        return dd;
    };
    auto hi_cost = [&](uint32_t mm, int bc) {
        float v = br[bc][(mm - 3) < 12 ? (mm - 3) : 12];
        if (mm >= 15) v = __fadd_rn(v, golomb_bits(mm - 15));
        return v;
    };
    auto lvl = [&](int k) -> uint32_t { return k < N ? level_byte(uabs32(cf[k])) : 0u; };
    auto interior_ctx = [&](uint32_t rc, int& ctx, int& bc) {
        const int x = (int)(rc >> log2h), y = (int)(rc & (h - 1));
        const int base = x * h + y;
        const uint32_t hm = lvl(base + 1) + lvl(base + h) + lvl(base + h + 1);
        const uint32_t mag = hm + lvl(base + 2) + lvl(base + 2 * h);
        ctx = T.lo_off[(G.oi * 5 + (y < 4 ? y : 4)) * 5 + (x < 4 ? x : 4)] + (mag > 512 ? 4 : (int)((mag + 64) >> 7));
        const uint32_t mg = hm & 63;
        bc = ((y | x) > 1 ? 14 : 7) + (mg > 12 ? 6 : (int)((mg + 1) >> 1));
    };
    auto dc_brc = [&]() {
        const uint32_t mg = (lvl(1) + lvl(h) + lvl(h + 1)) & 63;
        return mg > 12 ? 6 : (int)((mg + 1) >> 1);
    };
    auto interior_rate = [&](int ctx, int bc, uint32_t k) {
        if (k == 0) return bt[ctx][0];
        const uint32_t tok = k < 3 ? k : 3;
        float r = bt[ctx][tok];
        if (tok == 3) r = __fadd_rn(r, hi_cost(k, bc));
        return __fadd_rn(r, 1.0f);
    };
    auto set_mag = [&](uint32_t rc, uint32_t k) { cf[rc] = cf[rc] < 0 ? -(int32_t)k : (int32_t)k; };

    // Step A: interior positions eob-1 .. 1, then DC
    for (int i = eob - 1; i >= 1; i--) {
        const uint32_t rc = scan[i], l = uabs32(cf[rc]);
        if (l == 0) continue;
        int ctx, bc;
        interior_ctx(rc, ctx, bc);
        auto rate_k = [&](uint32_t k) {
            return k == 0 ? bt[ctx][0] : k == 1 ? __fadd_rn(bt[ctx][1], 1.0f) : k == 2 ? __fadd_rn(bt[ctx][2], 1.0f)
                 : __fadd_rn(__fadd_rn(bt[ctx][3], hi_cost(k, bc)), 1.0f);
        };
        uint32_t best_k = l;
        float best_c = __fadd_rn(distw(i, rc, l), __fmul_rn(lam, rate_k(l)));
        for (int k = (int)l - 1; k >= 0; k--) {
            const float dk = distw(i, rc, (uint32_t)k);
            if (dk >= best_c) break;
            const float c = __fadd_rn(dk, __fmul_rn(lam, rate_k((uint32_t)k)));
            if (c < best_c) {
                best_c = c;
                best_k = (uint32_t)k;
            }
        }
        if (best_k != l) set_mag(rc, best_k);
    }
    {
        const uint32_t rc = scan[0], l = uabs32(cf[rc]);
        if (l != 0) {
            const int bc = dc_brc();
            const int sgn = cf[rc] < 0 ? 1 : 0;
            auto dc_rate = [&](uint32_t k) {
                if (k == 0) return bt[0][0];
                const uint32_t tok = k < 3 ? k : 3;
                float r = bt[0][tok];
                if (tok == 3) r = __fadd_rn(r, hi_cost(k, bc));
                return __fadd_rn(r, dc_sign[sgn]);
            };
            uint32_t best_k = l;
            float best_c = __fadd_rn(distw(0, rc, l), __fmul_rn(lam, dc_rate(l)));
            for (int k = (int)l - 1; k >= 0; k--) {
                const float dk = distw(0, rc, (uint32_t)k);
                if (dk >= best_c) break;
                const float c = __fadd_rn(dk, __fmul_rn(lam, dc_rate((uint32_t)k)));
                if (c < best_c) {
                    best_c = c;
                    best_k = (uint32_t)k;
                }
            }
            if (best_k != l) set_mag(rc, best_k);
        }
    }

    // Step B: the distortions (current level, zero) with QM and tilt weights, their suffix
    auto dists = [&](int i, float& dcur, float& dzero) {
        const uint32_t rc = scan[i];
        const float q2 = rc == 0 ? q2dc : q2ac;
        const float at = fabsf(tf[rc]);
        const float e = __fsub_rn(at, (float)uabs32(cf[rc]));
        dcur = __fmul_rn(q2, __fmul_rn(e, e));
        dzero = __fmul_rn(q2, __fmul_rn(at, at));
        if (qm_on) {
            const float w2 = qm_at(rc);
            dcur = __fmul_rn(dcur, w2);
            dzero = __fmul_rn(dzero, w2);
        }
        if (tilt != 0.0f && at < 2.0f) {
            const float wgt = __fdiv_rn(1.0f, __fadd_rn(1.0f, __fmul_rn(tilt, __fmul_rn((float)i, n_inv))));
            dzero = __fmul_rn(dzero, wgt);
            dcur = __fmul_rn(dcur, wgt);
        }
    };
    {
        float sacc = 0.0f;
        for (int i = m - 1; i >= 1; i--) {
            float dc_, dz;
            dists(i, dc_, dz);
            sacc = __fadd_rn(sacc, dz);
            suf[i] = sacc;
        }
    }
    auto suf_at = [&](int i) { return i < m ? suf[i] : 0.0f; };
    const uint32_t dc_rc = scan[0], dc_m = uabs32(cf[dc_rc]);
    float dcur0, dzero0;
    dists(0, dcur0, dzero0);
    auto eob_pt = [&](int e) {
        const int bin = e < 2 ? e : bitlen32((uint32_t)e);
        float c = eob_bin[bin];
        if (bin > 1) {
            const int nbits = bin - 2;
            c = __fadd_rn(c, eob_hi[bin][(e >> nbits) & 1]);
            c = __fadd_rn(c, (float)nbits);
        }
        return c;
    };
    auto eob_cand = [&](int e, uint32_t mm) {
        const int ctx_e = 1 + (e > N / 8 ? 1 : 0) + (e > N / 4 ? 1 : 0);
        const uint32_t tok = mm < 3 ? mm : 3;
        float c = eob_base[ctx_e][tok - 1];
        if (tok == 3) {
            const uint32_t rc = scan[e];
            const int ex = (int)(rc >> log2h), ey = (int)(rc & (h - 1));
            c = __fadd_rn(c, hi_cost(mm, (ex | ey) > 1 ? 14 : 7));
        }
        return __fadd_rn(__fmul_rn(lam, __fadd_rn(c, 1.0f)), distw(e, scan[e], mm));
    };
    float dc_cost;
    if (dc_m == 0) {
        dc_cost = __fmul_rn(lam, bt[0][0]);
    } else {
        const uint32_t tok = dc_m < 3 ? dc_m : 3;
        float r = bt[0][tok];
        if (tok == 3) r = __fadd_rn(r, hi_cost(dc_m, dc_brc()));
        r = __fadd_rn(r, dc_sign[cf[dc_rc] < 0 ? 1 : 0]);
        dc_cost = __fmul_rn(lam, r);
    }
    dc_cost = __fadd_rn(dc_cost, dcur0);
    int best_e = -1;
    uint32_t best_m = 0;
    float best = INFINITY, pre = 0.0f;
    for (int e = 1; e < m; e++) {
        const uint32_t rc = scan[e], mm = uabs32(cf[rc]);
        float dcur, dz;
        dists(e, dcur, dz);
        if (mm != 0) {
            const float base = __fadd_rn(__fadd_rn(__fadd_rn(dc_cost, pre), __fmul_rn(lam, eob_pt(e))), suf_at(e + 1));
            uint32_t mb = mm;
            float c = __fadd_rn(base, eob_cand(e, mm));
            if (mm >= 2) {
                const float lo = __fadd_rn(base, eob_cand(e, mm - 1));
                if (lo < c) {
                    c = lo;
                    mb = mm - 1;
                }
            }
            if (c < best) {
                best = c;
                best_e = e;
                best_m = mb;
            }
        }
        int ctx, bc;
        interior_ctx(rc, ctx, bc);
        pre = __fadd_rn(pre, __fadd_rn(__fmul_rn(lam, interior_rate(ctx, bc, mm)), dcur));
    }
    if (dc_m != 0) {
        const uint32_t tok = dc_m < 3 ? dc_m : 3;
        float c0 = __fadd_rn(eob_bin[0], eob_base[0][tok - 1]);
        if (tok == 3) c0 = __fadd_rn(c0, hi_cost(dc_m, 0));
        c0 = __fadd_rn(c0, dc_sign[cf[dc_rc] < 0 ? 1 : 0]);
        const float total0 = __fadd_rn(__fadd_rn(__fmul_rn(lam, c0), dcur0), suf_at(1));
        if (total0 < best) {
            best = total0;
            best_e = 0;
            best_m = dc_m;
        }
    }
    const float skip = __fadd_rn(__fadd_rn(suf_at(1), dzero0), __fmul_rn(lam, 1.0f));
    if (best_e < 0 || skip < best) {
        for (int k = 0; k < m; k++) cf[scan[k]] = 0;
        return MT_OK;
    }
    const uint32_t rc = scan[best_e];
    if (best_m != uabs32(cf[rc])) {
        const bool neg = cf[rc] != 0 ? cf[rc] < 0 : tf[rc] < 0.0f;
        cf[rc] = neg ? -(int32_t)best_m : (int32_t)best_m;
    }
    for (int k = best_e + 1; k < m; k++) cf[scan[k]] = 0;
    return MT_OK;
}

}  // namespace
