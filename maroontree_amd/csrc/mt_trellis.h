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

// Whether descriptor d names a block the trellis can run on (its coded geometry in g).
__device__ inline bool trellis_valid(const mt_trellisblk& d, TrellisGeom& g) {
    const int bdi = (d.bd_qm >> 4) & 3;
    if (!trellis_geom(d.tx_size, g) || bdi >= 3 || (d.flags & ~MT_TRELLIS_CTX)) return false;
    if (!(d.flags & MT_TRELLIS_CTX)) return true;
    const int ebn = d.eob_bin < 3 ? 16 : d.eob_bin < 5 ? 32 : d.eob_bin < 8 ? 64 : d.eob_bin < 11 ? 128
                  : d.eob_bin < 13 ? 256 : d.eob_bin < 15 ? 512 : 1024;
    return d.eob_bin <= 16 && ebn == g.N && d.dcs_ctx <= 2 && d.plane <= 1 && d.qctx < MT_RATE_SLOTS;
}

// Maximum / bitwise or over each aligned group of G lanes of the wave (G a power of two, at
// most 64).
template <int G>
__device__ inline int group_max(int v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
template <int G>
__device__ inline unsigned long long group_or(unsigned long long v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    return v;
}

// Orders the LDS accesses of the lanes of one wave around this point.  A wave's LDS
// instructions execute in issue order, so lanes exchange data through LDS once the compiler
// keeps each access on its side of the point; unlike a workgroup barrier this does not wait
// for the wave's outstanding global loads.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First scan index of the anti-diagonal x + y == dd of a w x h block whose scan visits the
// anti-diagonals one after another (every scan the trellis reads does).
__device__ inline int diag_start(int dd, int w, int h) {
    const int a = min(w, h), b = max(w, h);
    if (dd <= a) return dd * (dd + 1) / 2;
    if (dd <= b) return a * (a + 1) / 2 + (dd - a) * a;
    const int k = dd - b;
    return a * (a + 1) / 2 + (b - a) * a + k * (a - 1) - k * (k - 1) / 2;
}

// The trellis of one block per aligned group of G lanes (G >= 2, a power of two), the whole
// workgroup being one wave that calls this together: d the group's descriptor (have false:
// no block), cf / tf its x-major levels and targets in LDS, sa / sb LDS workspaces of one
// float per coefficient (16-byte aligned), lb LDS bytes for the coefficients and 2 h + 1
// more (h the coded height), scan the block's scan (LDS or global).  Every f32
// operation is the one of the reference's sequential walk, in its order; only independent
// work is spread over the lanes:
//   - trellis_optimize_ctx's Step A visits the scan backwards and prices each level against
//     the final levels of the context neighbours (x+1, y), (x, y+1), (x+1, y+1), (x+2, y),
//     (x, y+2) of the stride-h map, or, where y + 1 or y + 2 runs past the column, the next
//     column's first entries.  The former lie on later anti-diagonals, the latter on earlier
//     ones, and the scans run anti-diagonal by anti-diagonal, so the positions of one
//     anti-diagonal read no level another position of it writes: the anti-diagonals run
//     one after another from the last, their positions in parallel, and anti-diagonals
//     without a nonzero level (where the walk changes nothing) are skipped.
//   - the eob search's running sums (the suffix of the zero-level distortions, the prefix of
//     the kept-levels cost) are sequential f32 sums: one lane of the group runs each, over
//     per-position terms the group computed in parallel and laid out in the order of the
//     sum; the candidate costs that read them are again per position, and the earliest
//     strict minimum is a group reduction.
// Returns -1 without a block, MT_ERR_ARG (levels untouched) for an invalid descriptor, else
// MT_OK.  cf / tf must be visible to the group on entry; cf holds the result on return.
template <int G, bool TILT>
__device__ int trellis_group_impl(const mt_trellisblk& d, bool have, int32_t* cf, const float* tf, float* sa,
                                  float* sb, uint8_t* lb, const uint16_t* scan, const mtlib::TrellisTabs& T) {
    static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group of 2..64 lanes");
    const int gl = threadIdx.x & (G - 1);
    TrellisGeom g{};
    const bool ok = have && trellis_valid(d, g);
    const int status = !have ? -1 : ok ? MT_OK : MT_ERR_ARG;
    bool act = ok && d.lambda0 > 0.0f;
    const bool ctxv = d.flags & MT_TRELLIS_CTX;
    const int N = act ? g.N : 0, h = act ? g.h : 1, log2h = act ? g.log2h : 0;
    // the quantizer steps (Quant::new_with_qm / new_chroma_with_delta_qm) and lambda
    float dc_q = 1.0f, ac_q = 1.0f;
    if (act) {
        const int bdi = (d.bd_qm >> 4) & 3, qi = d.qidx;
        const bool chroma_q = d.bd_qm & MT_TXBLK_CHROMA;
        const int dci = chroma_q ? min(255, max(0, qi + d.dc_delta)) : max(0, qi - 8);
        const int aci = chroma_q ? min(255, max(0, qi + d.ac_delta)) : qi;
        dc_q = (float)T.dcq[bdi * 256 + dci];
        ac_q = (float)T.acq[bdi * 256 + aci];
    }
    const float lam = trellis_lambda(dc_q, act ? d.lambda0 : 0.05f);
    const float q2dc = __fmul_rn(dc_q, dc_q), q2ac = __fmul_rn(ac_q, ac_q);
    wave_lds_sync();
    // the end of block, the level bytes of the context sums, the anti-diagonals holding a
    // nonzero level
    int e_loc = -1;
    unsigned long long dm = 0;
    for (int i0 = gl; i0 < N; i0 += 4 * G) {  // four positions per lane and pass
        uint32_t rc[4];
        int32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * G;
            rc[u] = i < N ? scan[i] : 0u;
            c[u] = i < N ? cf[rc[u]] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * G;
            if (i >= N) continue;
            lb[rc[u]] = (uint8_t)level_byte(uabs32(c[u]));
            if (c[u] != 0) {
                e_loc = i;
                dm |= 1ull << ((rc[u] >> log2h) + (rc[u] & (h - 1)));
            }
        }
    }
    for (int k = N + gl; k <= N + 2 * h; k += G) lb[k] = 0;  // the context reads past the block
    const int eob = group_max<G>(e_loc);
    act = act && eob >= 0;
    const int m = eob + 1;
    dm = group_or<G>(dm);
    if (!(act && ctxv && eob >= 2)) dm = 0ull;

    // context-variant tables and weights
    const int qm_level = d.bd_qm & 15, pl = d.plane ? 1 : 0;
    const bool qm_on = qm_level < 15, qm_full = qm_level >= 5;
    const mtlib::RateTables& RT = T.rate[act && ctxv ? d.qctx : 0];
    const int cls = act ? g.cls : 0, cb = act ? g.cb : 0;
    const float(*bt)[4] = RT.bt[cls][pl];
    const float(*br)[13] = RT.br[cb][pl];
    const float(*eob_base)[3] = RT.eob_base[cls][pl];
    const float(*eob_hi)[2] = RT.eob_hi[cls][pl];
    const float* dc_sign = RT.dc_sign[pl][act && ctxv ? d.dcs_ctx : 0];
    const float* eob_bin = RT.eob_bin[act && ctxv ? d.eob_bin : 0];
    const uint8_t* qmt = T.iqm + ((qm_on ? qm_level : 0) * 2 + pl) * MT_QM_USED_SIZE +
                         (act ? trellis_qm_offset(g.w, h) : 0);
    const uint8_t* lo_off = T.lo_off + g.oi * 25;
    const float tilt = d.band_tilt;
    const float n_inv = __fdiv_rn(1.0f, (float)(N > 0 ? N : 1));
    auto qm_at = [&](uint32_t rc) {
        const float wq = __fdiv_rn((float)qmt[rc], 32.0f);
        return qm_full ? __fmul_rn(wq, wq) : wq;
    };
    auto distw = [&](int idx, uint32_t rc, uint32_t lev) {
        const float e = __fsub_rn(fabsf(tf[rc]), (float)lev);
        float dd = __fmul_rn(rc == 0 ? q2dc : q2ac, __fmul_rn(e, e));
        if (qm_on) dd = __fmul_rn(dd, qm_at(rc));
        if (TILT && tilt != 0.0f && fabsf(tf[rc]) < 2.0f)
            dd = __fdiv_rn(dd, __fadd_rn(1.0f, __fmul_rn(tilt, __fmul_rn((float)idx, n_inv))));
        return dd;
    };
    // the distortions (current level, zero) with QM and tilt weights
    auto dists = [&](int i, uint32_t rc, float& dcur, float& dzero) {
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
        if (TILT && tilt != 0.0f && at < 2.0f) {
            const float wgt = __fdiv_rn(1.0f, __fadd_rn(1.0f, __fmul_rn(tilt, __fmul_rn((float)i, n_inv))));
            dzero = __fmul_rn(dzero, wgt);
            dcur = __fmul_rn(dcur, wgt);
        }
    };
    auto hi_cost = [&](uint32_t mm, int bc) {
        float v = br[bc][(mm - 3) < 12 ? (mm - 3) : 12];
        if (mm >= 15) v = __fadd_rn(v, golomb_bits(mm - 15));
        return v;
    };
    auto lvl = [&](int k) -> uint32_t { return lb[k]; };  // zero past the block
    auto interior_ctx = [&](uint32_t rc, int& ctx, int& bc) {
        const int x = (int)(rc >> log2h), y = (int)(rc & (h - 1));
        const int base = x * h + y;
        const uint32_t hm = lvl(base + 1) + lvl(base + h) + lvl(base + h + 1);
        const uint32_t mag = hm + lvl(base + 2) + lvl(base + 2 * h);
        ctx = lo_off[(y < 4 ? y : 4) * 5 + (x < 4 ? x : 4)] + (mag > 512 ? 4 : (int)((mag + 64) >> 7));
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
    auto set_mag = [&](uint32_t rc, uint32_t k) {
        cf[rc] = cf[rc] < 0 ? -(int32_t)k : (int32_t)k;
        lb[rc] = (uint8_t)level_byte(k);
    };

    // trellis_optimize's round-down (src/trellis.rs:944): chunks of 8 scan positions, then one
    // of 4, then the scalar tail; each chunk reads and writes only its own levels
    if (act && !ctxv) {
        const int m8 = m & ~7;
        for (int c = gl; c * 8 < m8; c += G) round_down_chunk(cf, tf, scan, c * 8, 8, q2dc, q2ac, lam);
        if (gl == 0) {
            int i = m8;
            if (i + 4 <= m) {
                round_down_chunk(cf, tf, scan, i, 4, q2dc, q2ac, lam);
                i += 4;
            }
            for (; i < m; i++) round_down_one(cf, tf, scan[i], scan[i] == 0 ? q2dc : q2ac, lam);
        }
    }
    // trellis_optimize_ctx (src/trellis.rs:55) Step A: interior positions eob-1 .. 1 by
    // anti-diagonal, then DC.  What pricing a position reads besides its neighbours' levels
    // (its scan position, level, target, weights, context offset) Step A does not change: a
    // lane loads it for its position of the next anti-diagonal before the barrier.
    struct StepPos {
        uint32_t rc, l, base;
        int i, lo_ctx, bcb;
        float at, q2, w2, tdiv;
        bool tilted;
    };
    auto fill = [&](int i) {
        StepPos p;
        const uint32_t rc = scan[i];
        const int x = (int)(rc >> log2h), y = (int)(rc & (h - 1));
        p.rc = rc;
        p.l = uabs32(cf[rc]);
        p.i = i;
        p.base = (uint32_t)(x * h + y);
        p.lo_ctx = lo_off[(y < 4 ? y : 4) * 5 + (x < 4 ? x : 4)];
        p.bcb = (y | x) > 1 ? 14 : 7;
        p.at = fabsf(tf[rc]);
        p.q2 = rc == 0 ? q2dc : q2ac;
        p.w2 = qm_on ? qm_at(rc) : 1.0f;
        p.tilted = TILT && tilt != 0.0f && p.at < 2.0f;
        p.tdiv = p.tilted ? __fadd_rn(1.0f, __fmul_rn(tilt, __fmul_rn((float)i, n_inv))) : 1.0f;
        return p;
    };
    auto stage = [&](int dd) {
        StepPos p{};
        if (dd < 0 || !((dm >> dd) & 1ull)) return p;
        const int lo = max(1, diag_start(dd, g.w, h)), hi = min(eob - 1, diag_start(dd + 1, g.w, h) - 1);
        if (lo + gl <= hi) p = fill(lo + gl);
        return p;
    };
    // the interior level decision of one position (distw and the rates of the walk)
    auto price = [&](const StepPos& p) {
        if (p.l == 0) return;
        const int base = (int)p.base;
        const uint32_t hm = lvl(base + 1) + lvl(base + h) + lvl(base + h + 1);
        const uint32_t mag = hm + lvl(base + 2) + lvl(base + 2 * h);
        const int ctx = p.lo_ctx + (mag > 512 ? 4 : (int)((mag + 64) >> 7));
        const uint32_t mg = hm & 63;
        const int bc = p.bcb + (mg > 12 ? 6 : (int)((mg + 1) >> 1));
        const float b0 = bt[ctx][0], b1 = bt[ctx][1], b2 = bt[ctx][2], b3 = bt[ctx][3];
        auto rate_k = [&](uint32_t k) {
            return k == 0 ? b0 : k == 1 ? __fadd_rn(b1, 1.0f) : k == 2 ? __fadd_rn(b2, 1.0f)
                 : __fadd_rn(__fadd_rn(b3, hi_cost(k, bc)), 1.0f);
        };
        auto dist = [&](uint32_t lev) {
            const float e = __fsub_rn(p.at, (float)lev);
            float dd = __fmul_rn(p.q2, __fmul_rn(e, e));
            if (qm_on) dd = __fmul_rn(dd, p.w2);
            if (TILT && p.tilted) dd = __fdiv_rn(dd, p.tdiv);
            return dd;
        };
        uint32_t best_k = p.l;
        float best_c = __fadd_rn(dist(p.l), __fmul_rn(lam, rate_k(p.l)));
        for (int k = (int)p.l - 1; k >= 0; k--) {
            const float dk = dist((uint32_t)k);
            if (dk >= best_c) break;
            const float c = __fadd_rn(dk, __fmul_rn(lam, rate_k((uint32_t)k)));
            if (c < best_c) {
                best_c = c;
                best_k = (uint32_t)k;
            }
        }
        if (best_k != p.l) set_mag(p.rc, best_k);
    };
    {
        unsigned long long wm = group_or<64>(dm) & ~1ull;  // the same on every lane
        wm = ((unsigned long long)__builtin_amdgcn_readfirstlane((int)(wm >> 32)) << 32) |
             (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)wm);
        int dd = wm ? 63 - __clzll((long long)wm) : -1;
        StepPos cur = stage(dd);
        while (dd >= 0) {
            wm &= ~(1ull << dd);
            const int dn = wm ? 63 - __clzll((long long)wm) : -1;
            const StepPos nxt = stage(dn);
            wave_lds_sync();
            price(cur);
            if ((dm >> dd) & 1ull) {  // positions past the group's width (a group narrower than
                                      // the anti-diagonal)
                const int lo = max(1, diag_start(dd, g.w, h)), hi = min(eob - 1, diag_start(dd + 1, g.w, h) - 1);
                for (int i = lo + gl + G; i <= hi; i += G) {
                    price(fill(i));
                }
            }
            cur = nxt;
            dd = dn;
        }
    }
    wave_lds_sync();
    if (act && ctxv && gl == 0) {
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
    wave_lds_sync();

    // the terms of the two running sums, in the order of the sums.  Context variant (i in
    // 1..m-1): sa[m-1-i] the zero-level distortion of i, sb[i-1] lambda * rate of its kept
    // level plus its distortion.  Plain variant (i in 0..m-1): sa[m-1-i] q2 * |t|^2,
    // sb[i] the kept level's distortion.
    if (act) {
        auto term = [&](int i, float& ta, float& tb) {
            const uint32_t rc = scan[i];
            if (ctxv) {
                float dcur;
                dists(i, rc, dcur, ta);
                int ctx, bc;
                interior_ctx(rc, ctx, bc);
                tb = __fadd_rn(__fmul_rn(lam, interior_rate(ctx, bc, uabs32(cf[rc]))), dcur);
            } else {
                const float q2 = rc == 0 ? q2dc : q2ac;
                const float at = fabsf(tf[rc]);
                const float er = __fsub_rn(at, (float)uabs32(cf[rc]));
                ta = __fmul_rn(q2, __fmul_rn(at, at));
                tb = __fmul_rn(q2, __fmul_rn(er, er));
            }
        };
        const int first = ctxv ? 1 : 0;
        // two positions per lane and pass, both priced before either is stored
        for (int i0 = first + gl; i0 < m; i0 += 2 * G) {
            const int i1 = i0 + G;
            float a0, b0, a1 = 0.0f, b1 = 0.0f;
            term(i0, a0, b0);
            if (i1 < m) term(i1, a1, b1);
            sa[m - 1 - i0] = a0;
            sb[i0 - first] = b0;
            if (i1 < m) {
                sa[m - 1 - i1] = a1;
                sb[i1 - first] = b1;
            }
        }
    }
    wave_lds_sync();
    // the running sums, inclusive, in place: lane 0 of the group over sa (the suffix sums of
    // the walk, sa[m-1-i] = the sum over i..m-1), lane 1 over sb (the prefix sums: the plain
    // variant adds lambda * rate of each kept level after its distortion)
    {
        const int cnt = act ? (ctxv ? m - 1 : m) : 0;
        const bool run = act && gl < 2;
        float* arr = gl == 0 ? sa : sb;
        if (__ballot(act && !ctxv) == 0ull) {
            // four terms per 16-byte LDS access, the next quad loaded ahead of the sums
            const int nq = run ? (cnt + 3) >> 2 : 0;
            const int kq = group_max<64>(nq);
            float4* a4 = reinterpret_cast<float4*>(arr);
            float4 nxt = nq > 0 ? a4[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float acc = 0.0f;
            for (int q = 0; q < kq; q++) {
                float4 v = nxt;
                if (q + 1 < nq) nxt = a4[q + 1];
                acc = __fadd_rn(acc, v.x);
                v.x = acc;
                acc = __fadd_rn(acc, v.y);
                v.y = acc;
                acc = __fadd_rn(acc, v.z);
                v.z = acc;
                acc = __fadd_rn(acc, v.w);
                v.w = acc;
                if (q < nq) a4[q] = v;
            }
        } else {
            const int kmax = group_max<64>(run ? cnt : 0);
            const bool lr = gl == 1 && !ctxv;
            float acc = 0.0f;
            for (int k = 0; k < kmax; k++) {
                if (run && k < cnt) {
                    const float v = arr[k];
                    const float x = lr ? __fmul_rn(lam, coef_rate_bits(uabs32(cf[scan[k]]))) : 0.0f;
                    acc = lr ? __fadd_rn(__fadd_rn(acc, v), x) : __fadd_rn(acc, v);
                    arr[k] = acc;
                }
            }
        }
    }
    wave_lds_sync();

    // the candidate ends of block: each lane keeps its earliest strict minimum, the group
    // reduces them to the earliest strict minimum of the walk
    float best = INFINITY;
    int best_e = 0x7fffffff;
    uint32_t best_m = 0;
    float dcur0 = 0.0f, dzero0 = 0.0f, dc_cost = 0.0f;
    uint32_t dc_m = 0;
    auto suf_at = [&](int i) { return i < m ? sa[m - 1 - i] : 0.0f; };  // sum of i..m-1
    if (act && ctxv) {
        const uint32_t dc_rc = scan[0];
        dc_m = uabs32(cf[dc_rc]);
        dists(0, dc_rc, dcur0, dzero0);
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
        auto eob_cand = [&](int e, uint32_t rc, uint32_t mm) {
            const int ctx_e = 1 + (e > N / 8 ? 1 : 0) + (e > N / 4 ? 1 : 0);
            const uint32_t tok = mm < 3 ? mm : 3;
            float c = eob_base[ctx_e][tok - 1];
            if (tok == 3) {
                const int ex = (int)(rc >> log2h), ey = (int)(rc & (h - 1));
                c = __fadd_rn(c, hi_cost(mm, (ex | ey) > 1 ? 14 : 7));
            }
            return __fadd_rn(__fmul_rn(lam, __fadd_rn(c, 1.0f)), distw(e, rc, mm));
        };
        for (int e = 1 + gl; e < m; e += G) {
            const uint32_t rc = scan[e];
            const uint32_t mm = uabs32(cf[rc]);
            if (mm == 0) continue;
            const float pre = e >= 2 ? sb[e - 2] : 0.0f;  // the kept-levels cost of 1..e-1
            const float base = __fadd_rn(__fadd_rn(__fadd_rn(dc_cost, pre), __fmul_rn(lam, eob_pt(e))), suf_at(e + 1));
            uint32_t mb = mm;
            float c = __fadd_rn(base, eob_cand(e, rc, mm));
            if (mm >= 2) {
                const float lo = __fadd_rn(base, eob_cand(e, rc, mm - 1));
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
    } else if (act) {
        for (int e = gl; e < m; e += G) {
            const uint32_t rc = scan[e];
            const uint32_t l = uabs32(cf[rc]);
            if (l == 0) continue;
            const float q2 = rc == 0 ? q2dc : q2ac;
            const float er = __fsub_rn(fabsf(tf[rc]), (float)l);
            const float dcur = __fmul_rn(q2, __fmul_rn(er, er));
            const int bin = e < 2 ? e : bitlen32((uint32_t)e);
            const int extra = bin > 1 ? bin - 2 : 0;
            const float sig = __fadd_rn(__fmul_rn((float)bin, 0.9f), __fadd_rn((float)extra, 2.0f));
            const float pre = e >= 1 ? sb[e - 1] : 0.0f;
            const float c = __fadd_rn(__fadd_rn(__fadd_rn(pre, dcur), __fmul_rn(lam, sig)), suf_at(e + 1));
            if (c < best) {
                best = c;
                best_e = e;
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o);
        const int oe = __shfl_xor(best_e, o);
        const uint32_t om = (uint32_t)__shfl_xor((int)best_m, o);
        if (ob < best || (ob == best && oe < best_e)) {
            best = ob;
            best_e = oe;
            best_m = om;
        }
    }
    // the decision, the same on every lane of the group, then its writes
    int keep = m;  // positions [keep, m) of the scan become zero
    int set_e = -1;
    if (act && ctxv) {
        if (dc_m != 0) {
            const uint32_t tok = dc_m < 3 ? dc_m : 3;
            float c0 = __fadd_rn(eob_bin[0], eob_base[0][tok - 1]);
            if (tok == 3) c0 = __fadd_rn(c0, hi_cost(dc_m, 0));
            c0 = __fadd_rn(c0, dc_sign[cf[scan[0]] < 0 ? 1 : 0]);
            const float total0 = __fadd_rn(__fadd_rn(__fmul_rn(lam, c0), dcur0), suf_at(1));
            if (total0 < best) {
                best = total0;
                best_e = 0;
                best_m = dc_m;
            }
        }
        const float skip = __fadd_rn(__fadd_rn(suf_at(1), dzero0), __fmul_rn(lam, 1.0f));
        if (best_e == 0x7fffffff || skip < best) {
            keep = 0;
        } else {
            keep = best_e + 1;
            set_e = best_e;
        }
    } else if (act) {
        const float skip = __fadd_rn(suf_at(0), __fmul_rn(lam, 1.0f));
        keep = (best_e == 0x7fffffff || skip < best) ? 0 : best_e + 1;
    }
    wave_lds_sync();  // every lane has read the levels the decision depends on
    if (act) {
        if (set_e >= 0 && gl == 0) {
            const uint32_t rc = scan[set_e];
            if (best_m != uabs32(cf[rc])) {
                const bool neg = cf[rc] != 0 ? cf[rc] < 0 : tf[rc] < 0.0f;
                cf[rc] = neg ? -(int32_t)best_m : (int32_t)best_m;
            }
        }
        for (int k = keep + gl; k < m; k += G) cf[scan[k]] = 0;
    }
    wave_lds_sync();
    return status;
}

// trellis_group_impl without the band-tilt division where no block of the wave tilts (the
// reference's encoder passes a zero tilt): the division is then never evaluated.
template <int G>
__device__ int trellis_group(const mt_trellisblk& d, bool have, int32_t* cf, const float* tf, float* sa, float* sb,
                             uint8_t* lb, const uint16_t* scan, const mtlib::TrellisTabs& T) {
    if (__ballot(have && d.band_tilt != 0.0f) != 0ull)
        return trellis_group_impl<G, true>(d, have, cf, tf, sa, sb, lb, scan, T);
    return trellis_group_impl<G, false>(d, have, cf, tf, sa, sb, lb, scan, T);
}

}  // namespace
