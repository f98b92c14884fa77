// CPU oracle: coefficient rate model (the decision-side mirror of the coefficient coder).
//
// TEST INFRASTRUCTURE ONLY -- the checker for the HIP library, never part of the product.
// Restates src/cost.rs:159-262 (cost table, CDF cost, Golomb, base-range ladder),
// src/rate.rs:50-301 (CoefCostTables, real_block_bits_bounded) and src/coeffs.rs:46
// (get_lo_ctx_2d), with Cdfs::new(qctx)'s coefficient part (src/coder/cdfs.rs:400-629).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "oracle.h"
#include "../maroontree_amd/csrc/av1_cdf_tables.h"

namespace {

inline uint32_t uabs(int32_t v) { return v < 0 ? 0u - (uint32_t)v : (uint32_t)v; }

// cost_q_table (src/cost.rs:159): Q22 -log2(p / 32768) from the C library's log2f.
const uint32_t* cost_q_table() {
    static uint32_t t[32769];
    static std::once_flag once;
    std::call_once(once, [] {
        t[0] = 0;
        for (int p = 1; p <= 32768; p++) {
            float bits = -log2f((float)p * (1.0f / 32768.0f));
            t[p] = (uint32_t)roundf(bits * (float)(1u << 22));
        }
    });
    return t;
}

// cdf_cost_with_table (src/cost.rs:185) on an inverse CDF (icdf form: 32768 - cdf values).
float cdf_cost(const uint16_t* icdf, int s) {
    int fl = s > 0 ? icdf[s - 1] : 32768;
    int fh = icdf[s];
    int p = fl - fh < 1 ? 1 : fl - fh;
    return (float)cost_q_table()[p] * (1.0f / (float)(1u << 22));
}

// golomb_cost (src/cost.rs:194)
float golomb_cost(uint32_t v) {
    uint32_t len = 32 - __builtin_clz(v + 1);
    return (float)(2 * len - 1);
}

// hi_tok_cost_with_table (src/cost.rs:235)
float hi_tok_cost(uint32_t m, const uint16_t* br_icdf) {
    int total_br = std::min((int)m - 3, 12);
    int coded = 0;
    float bits = 0.0f;
    for (int k = 0; k < 4; k++) {
        int s = std::min(total_br - coded, 3);
        bits += cdf_cost(br_icdf, s);
        coded += s;
        if (s < 3) break;
    }
    if (m >= 15) bits += golomb_cost(m - 15);
    return bits;
}

// br_cum_row_with_table (src/cost.rs:204)
void br_cum_row(const uint16_t* br_icdf, float* row) {
    float c[4];
    for (int s = 0; s < 4; s++) c[s] = cdf_cost(br_icdf, s);
    float c3x2 = c[3] + c[3];
    float c3x3 = c3x2 + c[3];
    float c3x4 = c3x3 + c[3];
    const float r[13] = {c[0], c[1], c[2], c[3] + c[0], c[3] + c[1], c[3] + c[2], c3x2 + c[0],
                         c3x2 + c[1], c3x2 + c[2], c3x3 + c[0], c3x3 + c[1], c3x3 + c[2], c3x4};
    memcpy(row, r, sizeof(r));
}

// level_byte (src/tables.rs:365)
uint8_t level_byte(uint32_t m) {
    if (m == 0) return 0;
    if (m <= 2) return (uint8_t)(m * 0x41);
    return (uint8_t)(std::min(m, 15u) + (3u << 6));
}

// The coefficient CDFs of Cdfs::new(qctx) in icdf form (src/coder/cdfs.rs:400-629,
// icdf src/tables.rs:31): every symbol row stored with its terminal 0.
struct CoefCdfs {
    uint16_t txb_skip[5][13][2];
    uint16_t base_tok[5][2][42][4];
    uint16_t br_tok[4][2][21][4];
    uint16_t eob_base[5][2][4][3];
    uint16_t eob_hi[5][2][11][2];
    uint16_t dc_sign[2][3][2];
    uint16_t eob_bin[17][11];  // see the eob-bin numbering in oracle.h
    uint16_t txtp4[13][7], txtp8[13][7], txtp16[13][5];
    // CoefCostTables (src/rate.rs:50-80)
    float bt[5][2][42][4];
    float br[4][2][21][13];
};

template <int K> void icdf_row(const uint16_t* raw, uint16_t* out) {
    for (int i = 0; i < K; i++) out[i] = (uint16_t)(32768 - raw[i]);
    out[K] = 0;
}

constexpr int kEobK[17] = {4, 4, 4, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 9, 9, 10, 10};

// The raw coefficient CDFs of Cdfs::new(qctx) (src/coder/cdfs.rs:400-629).
void raw_defaults(int qctx, orc_coef_cdfs& r) {
    memset(&r, 0, sizeof(r));
    const uint16_t* skip[5] = {MT_SKIP_TX4[qctx], MT_SKIP_TX8[qctx], MT_SKIP_TX16[qctx], MT_SKIP_TX32[qctx],
                               MT_SKIP_TX64[qctx]};
    for (int cls = 0; cls < 5; cls++)
        for (int ctx = 0; ctx < 13; ctx++) r.skip[cls][ctx] = skip[cls][ctx];
    const uint16_t(*bt_src[5][2])[3] = {
        {MT_BASE_TOK_TX4_LUMA_Q[qctx], MT_BASE_TOK_TX4_CHROMA_Q[qctx]},
        {MT_BASE_TOK_TX8_LUMA_Q[qctx], MT_BASE_TOK_TX8_CHROMA_Q[qctx]},
        {MT_BASE_TOK_TX16_LUMA_Q[qctx], MT_BASE_TOK_TX16_CHROMA_Q[qctx]},
        {MT_BASE_TOK_TX32_LUMA_Q[qctx], MT_BASE_TOK_TX32_CHROMA_Q[qctx]},
        {MT_BASE_TOK_TX64_LUMA_Q[qctx], nullptr}};
    const int bt_n[5] = {41, 41, 42, 42, 42};
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++)
            if (bt_src[cls][pl])
                for (int ctx = 0; ctx < bt_n[cls]; ctx++)
                    for (int i = 0; i < 3; i++) r.base_tok[cls][pl][ctx][i] = bt_src[cls][pl][ctx][i];
    const uint16_t(*br_src[4][2])[3] = {
        {MT_BR_TOK_TX4_LUMA_Q[qctx], MT_BR_TOK_TX4_CHROMA_Q[qctx]},
        {MT_BR_TOK_TX8_LUMA_Q[qctx], MT_BR_TOK_TX8_CHROMA_Q[qctx]},
        {MT_BR_TOK_TX16_LUMA_Q[qctx], MT_BR_TOK_TX16_CHROMA_Q[qctx]},
        {MT_BR_TOK_TX32_LUMA_Q[qctx], MT_BR_TOK_TX32_CHROMA_Q[qctx]}};
    for (int cls = 0; cls < 4; cls++)
        for (int pl = 0; pl < 2; pl++)
            for (int ctx = 0; ctx < 21; ctx++)
                for (int i = 0; i < 3; i++) r.br_tok[cls][pl][ctx][i] = br_src[cls][pl][ctx][i];
    const uint16_t(*eb_src[5][2])[2] = {
        {MT_EOB_BASE_TX4_LUMA_Q[qctx], MT_EOB_BASE_TX4_CHROMA_Q[qctx]},
        {MT_EOB_BASE_TX8_LUMA_Q[qctx], MT_EOB_BASE_TX8_CHROMA_Q[qctx]},
        {MT_EOB_BASE_TX16_LUMA_Q[qctx], MT_EOB_BASE_TX16_CHROMA_Q[qctx]},
        {MT_EOB_BASE_TX32_LUMA_Q[qctx], MT_EOB_BASE_TX32_CHROMA_Q[qctx]},
        {MT_EOB_BASE_TX64_LUMA_Q[qctx], nullptr}};
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++)
            if (eb_src[cls][pl])
                for (int ctx = 0; ctx < 4; ctx++)
                    for (int i = 0; i < 2; i++) r.eob_base[cls][pl][ctx][i] = eb_src[cls][pl][ctx][i];
    const uint16_t* eh_src[5][2] = {
        {MT_EOB_HI_TX4_LUMA[qctx], MT_EOB_HI_TX4_CHROMA[qctx]},
        {MT_EOB_HI_TX8_LUMA[qctx], MT_EOB_HI_TX8_CHROMA[qctx]},
        {MT_EOB_HI_TX16_LUMA[qctx], MT_EOB_HI_TX16_CHROMA[qctx]},
        {MT_EOB_HI_TX32_LUMA[qctx], MT_EOB_HI_TX32_CHROMA[qctx]},
        {MT_EOB_HI_TX64_LUMA[qctx], nullptr}};
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++)
            if (eh_src[cls][pl])
                for (int b = 0; b < 11; b++) r.eob_hi[cls][pl][b] = eh_src[cls][pl][b];
    for (int pl = 0; pl < 2; pl++)
        for (int ctx = 0; ctx < 3; ctx++) r.dc_sign[pl][ctx] = MT_DC_SIGN_Q[qctx][pl][ctx];
    const uint16_t* bins[17] = {MT_EOB_BIN_16_LUMA[qctx],    MT_EOB_BIN_16_CHROMA[qctx],   MT_EOB_BIN_16_LUMA_1D[qctx],
                                MT_EOB_BIN_32_LUMA[qctx],    MT_EOB_BIN_32_CHROMA[qctx],   MT_EOB_BIN_64_LUMA[qctx],
                                MT_EOB_BIN_64_CHROMA[qctx],  MT_EOB_BIN_64_LUMA_1D[qctx],  MT_EOB_BIN_128_LUMA[qctx],
                                MT_EOB_BIN_128_CHROMA[qctx], MT_EOB_BIN_128_LUMA_1D[qctx], MT_EOB_BIN_256_LUMA[qctx],
                                MT_EOB_BIN_256_CHROMA[qctx], MT_EOB_BIN_512_LUMA[qctx],    MT_EOB_BIN_512_CHROMA[qctx],
                                MT_EOB_BIN_1024_LUMA[qctx],  MT_EOB_BIN_1024_CHROMA[qctx]};
    for (int e = 0; e < 17; e++)
        for (int i = 0; i < kEobK[e]; i++) r.eob_bin[e][i] = bins[e][i];
    for (int m = 0; m < 13; m++) {
        for (int i = 0; i < 6; i++) r.txtp4[m][i] = MT_TXTP_INTRA1_TX4[m][i];
        for (int i = 0; i < 6; i++) r.txtp8[m][i] = MT_TXTP_INTRA1_TX8[m][i];
        for (int i = 0; i < 4; i++) r.txtp16[m][i] = MT_TXTP_INTRA2_TX16[m][i];
    }
}

// The icdf rows and CoefCostTables (src/rate.rs:50-80) of a raw snapshot.
void build_coef(const orc_coef_cdfs& r, CoefCdfs& d) {
    memset(&d, 0, sizeof(d));
    for (int cls = 0; cls < 5; cls++)
        for (int ctx = 0; ctx < 13; ctx++) icdf_row<1>(&r.skip[cls][ctx], d.txb_skip[cls][ctx]);
    const int bt_n[5] = {41, 41, 42, 42, 42};
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++)
            if (!(cls == 4 && pl == 1))
                for (int ctx = 0; ctx < bt_n[cls]; ctx++) icdf_row<3>(r.base_tok[cls][pl][ctx], d.base_tok[cls][pl][ctx]);
    for (int cls = 0; cls < 4; cls++)
        for (int pl = 0; pl < 2; pl++)
            for (int ctx = 0; ctx < 21; ctx++) icdf_row<3>(r.br_tok[cls][pl][ctx], d.br_tok[cls][pl][ctx]);
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++)
            if (!(cls == 4 && pl == 1))
                for (int ctx = 0; ctx < 4; ctx++) icdf_row<2>(r.eob_base[cls][pl][ctx], d.eob_base[cls][pl][ctx]);
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++)
            if (!(cls == 4 && pl == 1))
                for (int b = 0; b < 11; b++) icdf_row<1>(&r.eob_hi[cls][pl][b], d.eob_hi[cls][pl][b]);
    for (int pl = 0; pl < 2; pl++)
        for (int ctx = 0; ctx < 3; ctx++) icdf_row<1>(&r.dc_sign[pl][ctx], d.dc_sign[pl][ctx]);
    for (int e = 0; e < 17; e++) {
        for (int i = 0; i < kEobK[e]; i++) d.eob_bin[e][i] = (uint16_t)(32768 - r.eob_bin[e][i]);
        d.eob_bin[e][kEobK[e]] = 0;
    }
    for (int m = 0; m < 13; m++) {
        icdf_row<6>(r.txtp4[m], d.txtp4[m]);
        icdf_row<6>(r.txtp8[m], d.txtp8[m]);
        icdf_row<4>(r.txtp16[m], d.txtp16[m]);
    }
    // CoefCostTables::build (src/rate.rs:58)
    for (int cls = 0; cls < 5; cls++)
        for (int pl = 0; pl < 2; pl++) {
            for (int ctx = 0; ctx < 42; ctx++)
                for (int tok = 0; tok < 4; tok++) d.bt[cls][pl][ctx][tok] = cdf_cost(d.base_tok[cls][pl][ctx], tok);
            if (cls == 4) continue;
            for (int bc = 0; bc < 21; bc++) br_cum_row(d.br_tok[cls][pl][bc], d.br[cls][pl][bc]);
        }
}

// Snapshot slots: 0..3 Cdfs::new(qctx); 4..7 loaded by orc_rate_load_cdfs (Cdfs::new(0) until then).
CoefCdfs g_slots[8];
std::once_flag g_slots_once;

const CoefCdfs& coef_cdfs(int slot) {
    std::call_once(g_slots_once, [] {
        for (int q = 0; q < 8; q++) {
            orc_coef_cdfs r;
            raw_defaults(q < 4 ? q : 0, r);
            build_coef(r, g_slots[q]);
        }
    });
    return g_slots[slot];
}

// get_lo_ctx_2d (src/coeffs.rs:46)
inline void lo_ctx(const uint8_t* levels, int x, int y, const uint32_t (*off)[5], int stride,
                   int* ctx, uint32_t* hi_mag) {
    auto g = [&](int dx, int dy) { return (uint32_t)levels[(x + dx) * stride + (y + dy)]; };
    uint32_t hm = g(0, 1) + g(1, 0) + g(1, 1);
    uint32_t mag = hm + g(0, 2) + g(2, 0);
    uint32_t offset = off[std::min(y, 4)][std::min(x, 4)];
    *ctx = (int)(offset + (mag > 512 ? 4 : (mag + 64) >> 7));
    *hi_mag = hm;
}

}  // namespace

extern "C" uint32_t orc_cost_q(int p) { return cost_q_table()[p]; }

extern "C" float orc_cdf_cost_raw(const uint16_t* raw_cdf, int nsym, int s) {
    // cost of symbol s against a raw (AV1 spec) CDF of nsym-1 values
    uint16_t ic[17];
    for (int i = 0; i < nsym - 1; i++) ic[i] = (uint16_t)(32768 - raw_cdf[i]);
    ic[nsym - 1] = 0;
    return cdf_cost(ic, s);
}

// real_block_bits_bounded (src/rate.rs:149)
extern "C" float orc_block_rate(const int32_t* cf, const uint32_t* scan, const orc_rate_ctx* c,
                                float bound) {
    const CoefCdfs& d = coef_cdfs(c->qctx);
    const int cls = c->cls, pl = c->plane, w = c->w, h = c->h, n = w * h;
    const uint16_t* skip_cdf = d.txb_skip[cls][c->skip_ctx];
    int eob = -1;
    for (int i = n - 1; i >= 0; i--)
        if (cf[scan[i]] != 0) {
            eob = i;
            break;
        }
    if (eob < 0) return c->use_skip_bits ? c->skip_one : cdf_cost(skip_cdf, 1);

    const uint16_t(*br_tok)[4] = d.br_tok[std::min(cls, 3)][pl];
    const uint16_t(*eob_base)[3] = d.eob_base[cls][pl];
    const uint16_t(*eob_hi)[2] = d.eob_hi[cls][pl];
    const uint16_t* dc_sign = d.dc_sign[pl][c->dcs_ctx];
    const uint16_t* eob_bin_cdf = d.eob_bin[c->eob_bin];

    float bits = c->use_skip_bits ? c->skip_zero : cdf_cost(skip_cdf, 0);
    if (c->txtp_set > 0) {
        const uint16_t* tc = c->txtp_set == 1 ? d.txtp4[c->txtp_mode]
                           : c->txtp_set == 2 ? d.txtp8[c->txtp_mode] : d.txtp16[c->txtp_mode];
        bits += cdf_cost(tc, c->txtp);
    }
    if (eob == 0) {
        uint32_t m = uabs(cf[0]);
        bits += cdf_cost(eob_bin_cdf, 0);
        bits += cdf_cost(eob_base[0], (int)std::min(m, 3u) - 1);
        if (std::min(m, 3u) == 3) bits += hi_tok_cost(m, br_tok[0]);
        bits += cdf_cost(dc_sign, cf[0] < 0 ? 1 : 0);
        return bits;
    }
    const int eob_bin = eob < 2 ? eob : 32 - __builtin_clz((uint32_t)eob);
    bits += cdf_cost(eob_bin_cdf, eob_bin);
    if (eob_bin > 1) {
        const int nbits = eob_bin - 2;
        bits += cdf_cost(eob_hi[eob_bin], (eob >> nbits) & 1);
        bits += (float)nbits;
    }
    if (bits > bound) return INFINITY;

    const int log2h = __builtin_ctz(h);
    const int stride = h;
    const uint32_t(*offsets)[5] = w < h ? MT_LO_CTX_OFF_WLH : w > h ? MT_LO_CTX_OFF_WGH : MT_LO_CTX_OFF;
    const float(*bt_c)[4] = d.bt[cls][pl];
    const float(*br_c)[13] = d.br[std::min(cls, 3)][pl];
    std::vector<uint8_t> levels((size_t)(w + 2) * (h + 2), 0);

    const int ctx_e = 1 + (eob > (n >> 3) ? 1 : 0) + (eob > (n >> 2) ? 1 : 0);
    {
        const int rc = (int)scan[eob];
        const int ex = rc >> log2h, ey = rc & (h - 1);
        const uint32_t m = uabs(cf[rc]);
        const uint32_t eob_tok = std::min(m, 3u) - 1;
        bits += cdf_cost(eob_base[ctx_e], (int)eob_tok);
        if (eob_tok == 2) {
            const int bc = (ex | ey) > 1 ? 14 : 7;
            bits += hi_tok_cost(m, br_tok[bc]);
        }
        levels[ex * stride + ey] = level_byte(m);
    }
    uint32_t ac_nonzero = 1;
    for (int i = eob - 1; i >= 1; i--) {
        if ((i & 15) == 0 && bits > bound) return INFINITY;
        const int rc = (int)scan[i];
        const int x = rc >> log2h, y = rc & (h - 1);
        int ctx;
        uint32_t hi_mag;
        lo_ctx(levels.data(), x, y, offsets, stride, &ctx, &hi_mag);
        const uint32_t m = uabs(cf[rc]);
        const uint32_t tok = std::min(m, 3u);
        bits += bt_c[ctx][tok];
        if (tok == 3) {
            const uint32_t mag = hi_mag & 63;
            const int bc = ((y | x) > 1 ? 14 : 7) + (mag > 12 ? 6 : (int)((mag + 1) >> 1));
            const int total_br = std::min((int)m - 3, 12);
            bits += br_c[bc][total_br];
            if (m >= 15) bits += golomb_cost(m - 15);
        }
        if (m != 0) {
            levels[x * stride + y] = level_byte(m);
            ac_nonzero++;
        }
    }
    const uint32_t dm = uabs(cf[0]);
    const uint32_t dc_tok = std::min(dm, 3u);
    bits += cdf_cost(d.base_tok[cls][pl][0], (int)dc_tok);
    if (dc_tok == 3) {
        const uint32_t mag = ((uint32_t)levels[1] + levels[stride] + levels[stride + 1]) & 63;
        const int bc = mag > 12 ? 6 : (int)((mag + 1) >> 1);
        bits += hi_tok_cost(dm, br_tok[bc]);
    }
    if (cf[0] != 0) bits += cdf_cost(dc_sign, cf[0] < 0 ? 1 : 0);
    for (uint32_t k = 0; k < ac_nonzero; k++) bits += 1.0f;
    return bits;
}

extern "C" const uint32_t* orc_scan(int tx_size) {
    switch (tx_size) {
        case 0: return MT_SCAN_4X4;
        case 1: return MT_SCAN_8X8;
        case 2: return MT_SCAN_16X16;
        case 3: case 4: return MT_SCAN_32X32;
        case 5: return MT_SCAN_4X8;
        case 6: return MT_SCAN_8X4;
        case 7: return MT_SCAN_8X16;
        case 8: return MT_SCAN_16X8;
        case 9: return MT_SCAN_16X32;
        case 10: return MT_SCAN_32X16;
        case 13: return MT_SCAN_4X16;
        case 14: return MT_SCAN_16X4;
        default: return nullptr;
    }
}

extern "C" float orc_skip_cost(int qctx, int cls, int ctx, int s) {
    return cdf_cost(coef_cdfs(qctx).txb_skip[cls][ctx], s);
}

extern "C" void orc_rate_default_cdfs(int qctx, orc_coef_cdfs* out) { raw_defaults(qctx, *out); }

extern "C" int orc_rate_load_cdfs(int slot, const orc_coef_cdfs* cdfs) {
    if (slot < 4 || slot > 7 || !cdfs) return -1;
    coef_cdfs(0);
    build_coef(*cdfs, g_slots[slot]);
    return 0;
}

// ---- trellis quantization (src/trellis.rs, src/trellis_dist.rs, src/avx/trellis.rs) ------
namespace {

// COEF_RATE_LUT (src/cost.rs:73) and coef_rate_bits / coef_rate_bits_slow (:52-87): the
// table is the closed form below 64; f_fmlaf is unfused on the default x86-64 build
// (src/aq_common.rs:66-88).
const float kCoefRateLut[64] = {
    0.9f, 2.7f, 3.6f, 5.3f, 5.3f, 5.3f, 6.6f, 6.6f, 6.6f, 7.9f, 7.9f, 7.9f, 9.2f, 9.2f, 9.2f, 11.5f,
    13.5f, 13.5f, 15.5f, 15.5f, 15.5f, 15.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f, 17.5f,
    19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f, 19.5f,
    19.5f, 19.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f,
    21.5f, 21.5f, 21.5f, 21.5f, 21.5f, 21.5f};

float coef_rate_bits(uint32_t l) {
    if (l < 64) return kCoefRateLut[l];
    float b = 4.0f;
    const int total_br = std::min((int)l - 3, 12);
    b += (float)(total_br / 3 + 1) * 1.3f;
    const uint32_t v = l - 15;
    b += 2.0f * (float)(32 - __builtin_clz(v + 1)) + -1.0f;
    return b;
}

// scaled_trellis_lambda (src/trellis.rs:41-52): the aom KF law on dc_q, TRELLIS_AOM_CALIB
// (0.045025873597302174 * 0.5525, an f32 product) times the shipped trellis_calib 1.15
// (src/tuning.rs:953), scaled by lambda0 / TRELLIS_LAMBDA0 (0.05, src/cost.rs:93).
float trellis_lambda(float dc_q, float lambda0) {
    const float calib = 0.045025873597302174f * 0.5525f;
    const float aom = calib * 1.15f * dc_q * dc_q * (3.3f + 0.0015f * dc_q);
    return aom * (lambda0 / 0.05f);
}

int bitlen(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

// trellis_round_down_scan_avx2 (src/avx/trellis.rs:196): chunks of 8 scan positions, then
// of 4, then one at a time; a chunk with any |level| >= 64 runs the scalar form on all of its
// positions; otherwise cost = fma(q2, e*e, lambda * LUT[l]) for l and l - 1 (dc_q2 on the
// chunk's first lane at scan index 0).
void round_down_one(int32_t* cf, const float* tf, uint32_t rc, float dq2, float lambda) {
    const int32_t c = cf[rc];
    if (c == 0) return;
    const uint32_t l = uabs(c);
    const float at = fabsf(tf[rc]);
    const float el = at - (float)l, edn = at - (float)(l - 1);
    const float cost_l = dq2 * (el * el) + lambda * coef_rate_bits(l);
    const float cost_dn = dq2 * (edn * edn) + lambda * coef_rate_bits(l - 1);
    if (cost_dn < cost_l) cf[rc] = c < 0 ? -(int32_t)(l - 1) : (int32_t)(l - 1);
}

void round_down_chunk(int32_t* cf, const float* tf, const uint32_t* scan, int i0, int len, float dq2_dc,
                      float dq2_ac, float lambda) {
    bool big = false;
    for (int k = 0; k < len; k++) big |= uabs(cf[scan[i0 + k]]) > 63;
    if (big) {
        for (int k = 0; k < len; k++) {
            const uint32_t rc = scan[i0 + k];
            round_down_one(cf, tf, rc, rc == 0 ? dq2_dc : dq2_ac, lambda);
        }
        return;
    }
    bool lower[8];
    for (int k = 0; k < len; k++) {
        const uint32_t rc = scan[i0 + k];
        const uint32_t l = uabs(cf[rc]), ldn = l ? l - 1 : 0;
        const float at = fabsf(tf[rc]);
        const float q2 = (i0 == 0 && k == 0) ? dq2_dc : dq2_ac;
        const float el = at - (float)l, edn = at - (float)ldn;
        const float cost_l = fmaf(q2, el * el, lambda * kCoefRateLut[l]);
        const float cost_dn = fmaf(q2, edn * edn, lambda * kCoefRateLut[ldn]);
        lower[k] = cost_dn < cost_l;
    }
    for (int k = 0; k < len; k++) {
        const uint32_t rc = scan[i0 + k];
        const int32_t c = cf[rc];
        if (lower[k] && c != 0) cf[rc] = c < 0 ? c + 1 : c - 1;
    }
}

void round_down_scan(int32_t* cf, const float* tf, const uint32_t* scan, int n, float dq2_dc, float dq2_ac,
                     float lambda) {
    int i = 0;
    for (; i + 8 <= n; i += 8) round_down_chunk(cf, tf, scan, i, 8, dq2_dc, dq2_ac, lambda);
    for (; i + 4 <= n; i += 4) round_down_chunk(cf, tf, scan, i, 4, dq2_dc, dq2_ac, lambda);
    for (; i < n; i++) round_down_one(cf, tf, scan[i], scan[i] == 0 ? dq2_dc : dq2_ac, lambda);
}

int last_nonzero(const int32_t* cf, const uint32_t* scan, int n) {
    for (int i = n - 1; i >= 0; i--)
        if (cf[scan[i]] != 0) return i;
    return -1;
}

}  // namespace

// hbd_trellis_scale (src/coder/lossy_state.rs:158) on the shipped tuning (src/tuning.rs:954-959):
// above 8 bits at 4:2:0 (top 0.174, hi 36, lo 20) and 4:2:2 (1.0, 24, 10) the factor ramps
// from 1 down to top as qindex falls from hi to lo; f_fmlaf unfused.
extern "C" float orc_trellis_lambda0(int qidx, int bd, int sub) {
    float scale = 1.0f;
    if (bd > 8 && (sub == 1 || sub == 2)) {
        const float top = sub == 2 ? 0.174f : 1.0f, hi = sub == 2 ? 36.0f : 24.0f, lo = sub == 2 ? 20.0f : 10.0f;
        float w = (hi - (float)qidx) / std::max(hi - lo, 1.0f);
        w = std::min(std::max(w, 0.0f), 1.0f);
        scale = (top - 1.0f) * w + 1.0f;
    }
    return 0.05f * scale;
}

// trellis_optimize (src/trellis.rs:944): round-down (Step A), then the EOB position against
// the coef_rate_bits model with eob_sig(e) = bin * 0.9 + (extra + 2) (f_fmlaf, unfused), or
// all-zero at 1 bit.
extern "C" void orc_trellis(int32_t* cf, const float* tf, float dc_q, float ac_q, const uint32_t* scan, int n,
                            float lambda0) {
    if (lambda0 <= 0.0f) return;
    const float lambda = trellis_lambda(dc_q, lambda0);
    const float dq2_dc = dc_q * dc_q, dq2_ac = ac_q * ac_q;
    const int eob = last_nonzero(cf, scan, n);
    if (eob < 0) return;
    round_down_scan(cf, tf, scan, eob + 1, dq2_dc, dq2_ac, lambda);
    const int m = eob + 1;
    std::vector<float> dcur(m), dzero(m), suf(m + 1);
    for (int i = 0; i < m; i++) {
        const uint32_t rc = scan[i];
        const float q2 = rc == 0 ? dq2_dc : dq2_ac;
        const float at = fabsf(tf[rc]);
        const float e = at - (float)uabs(cf[rc]);
        dcur[i] = q2 * (e * e);
        dzero[i] = q2 * (at * at);
    }
    suf[m] = 0.0f;
    for (int i = m - 1; i >= 0; i--) suf[i] = suf[i + 1] + dzero[i];
    int best_e = -1;
    float best = INFINITY, pre = 0.0f;
    for (int e = 0; e < m; e++) {
        const uint32_t rc = scan[e];
        if (cf[rc] != 0) {
            const int bin = e < 2 ? e : bitlen((uint32_t)e);
            const int extra = bin > 1 ? bin - 2 : 0;
            const float sig = (float)bin * 0.9f + ((float)extra + 2.0f);
            const float c = pre + dcur[e] + lambda * sig + suf[e + 1];
            if (c < best) {
                best = c;
                best_e = e;
            }
        }
        pre = pre + dcur[e] + lambda * coef_rate_bits(uabs(cf[rc]));
    }
    const float skip = suf[0] + lambda * 1.0f;
    const int keep = (best_e < 0 || skip < best) ? 0 : best_e + 1;
    for (int i = keep; i < m; i++) cf[scan[i]] = 0;
}

// trellis_optimize_ctx (src/trellis.rs:55) on the shipped switches (src/tuning.rs:945-953:
// trellis_aom, trellis_eob_reprice, trellis_eob_zero_term, trellis_pad_stride and
// trellis_dc_only_scan off, trellis_dc_ctx0 on, trellis_zero_bits 1, trellis_dz 0, the
// tilt magnitude cap 2): Step A prices each interior level (reverse scan) and then DC in the
// contexts of the current level map; Step B picks the EOB position (each nonzero at its
// level or one below) against the prefix of interior costs, the DC-only terminal and the
// all-zero block.  Rates come from the coefficient CDFs of snapshot c->qctx.
extern "C" void orc_trellis_ctx(int32_t* cf, const float* tf, float dc_q, float ac_q, const uint32_t* scan,
                                float lambda0, const orc_rate_ctx* c, int qm_level, float band_tilt) {
    if (lambda0 <= 0.0f) return;
    const int w = c->w, h = c->h, n = w * h, cls = c->cls, pl = c->plane;
    const CoefCdfs& d = coef_cdfs(c->qctx);
    const bool qm_on = qm_level < 15, qm_full = qm_level >= 5;
    auto qm_at = [&](uint32_t rc) {
        const float wq = (float)orc_qm_weight(qm_level, pl != 0, (int)rc, w, h) / 32.0f;
        return qm_full ? wq * wq : wq;
    };
    const float lambda = trellis_lambda(dc_q, lambda0);
    const int log2h = __builtin_ctz(h), stride = h;
    const uint32_t(*lo_off)[5] = w < h ? MT_LO_CTX_OFF_WLH : w > h ? MT_LO_CTX_OFF_WGH : MT_LO_CTX_OFF;
    const uint16_t(*base_tok)[4] = d.base_tok[cls][pl];
    const uint16_t(*br_tok)[4] = d.br_tok[std::min(cls, 3)][pl];
    const uint16_t(*eob_hi)[2] = d.eob_hi[cls][pl];
    const uint16_t(*eob_base)[3] = d.eob_base[cls][pl];
    const uint16_t* dc_sign = d.dc_sign[pl][c->dcs_ctx];
    const uint16_t* eob_bin_cdf = d.eob_bin[c->eob_bin];
    const float dq2_dc = dc_q * dc_q, dq2_ac = ac_q * ac_q;
    const float n_inv = 1.0f / (float)n;
    auto distw = [&](int idx, uint32_t rc, uint32_t lev) {
        const float q2 = rc == 0 ? dq2_dc : dq2_ac;
        const float e = fabsf(tf[rc]) - (float)lev;
        float dd = q2 * (e * e);
        if (qm_on) dd *= qm_at(rc);
        if (band_tilt != 0.0f && fabsf(tf[rc]) < 2.0f) dd /= 1.0f + band_tilt * ((float)idx * n_inv);
        return dd;
    };
    const int eob = last_nonzero(cf, scan, n);
    if (eob < 0) return;
    // base-range tail: the cumulative ladder from n >= 64 up, the direct walk below
    // (the same values; src/trellis.rs:150-170)
    float br_cum[21][13];
    if (n >= 64)
        for (int b = 0; b < 21; b++) br_cum_row(br_tok[b], br_cum[b]);
    auto hi_cost = [&](uint32_t m, int bc) {
        if (n >= 64) {
            float bits = br_cum[bc][std::min((int)m - 3, 12)];
            if (m >= 15) bits += golomb_cost(m - 15);
            return bits;
        }
        return hi_tok_cost(m, br_tok[bc]);
    };
    std::vector<uint8_t> levels((size_t)(w + 2) * (h + 2), 0);
    auto at_map = [&](uint32_t rc) -> uint8_t& { return levels[(rc >> log2h) * stride + (rc & (h - 1))]; };
    for (int i = 0; i <= eob; i++)
        if (cf[scan[i]] != 0) at_map(scan[i]) = level_byte(uabs(cf[scan[i]]));
    auto interior_ctx = [&](uint32_t rc, int* ctx, int* bc) {
        const int x = (int)(rc >> log2h), y = (int)(rc & (h - 1));
        uint32_t hm;
        lo_ctx(levels.data(), x, y, lo_off, stride, ctx, &hm);
        const uint32_t mag = hm & 63;
        *bc = ((y | x) > 1 ? 14 : 7) + (mag > 12 ? 6 : (int)((mag + 1) >> 1));
    };
    auto dc_brc = [&]() {
        const uint32_t mag = ((uint32_t)levels[1] + levels[stride] + levels[stride + 1]) & 63;
        return mag > 12 ? 6 : (int)((mag + 1) >> 1);
    };
    auto interior_rate = [&](int ctx, int bc, uint32_t k) {
        if (k == 0) return cdf_cost(base_tok[ctx], 0);
        const uint32_t tok = std::min(k, 3u);
        float b = cdf_cost(base_tok[ctx], (int)tok);
        if (tok == 3) b += hi_cost(k, bc);
        return b + 1.0f;
    };
    auto eob_pt_cost = [&](int e) {
        const int bin = e < 2 ? e : bitlen((uint32_t)e);
        float cc = cdf_cost(eob_bin_cdf, bin);
        if (bin > 1) {
            const int nbits = bin - 2;
            cc += cdf_cost(eob_hi[bin], (e >> nbits) & 1);
            cc += (float)nbits;
        }
        return cc;
    };
    auto set_sign = [&](uint32_t rc, uint32_t k) { cf[rc] = cf[rc] < 0 ? -(int32_t)k : (int32_t)k; };

    // Step A: interior positions eob-1 .. 1, then DC
    for (int i = eob - 1; i >= 1; i--) {
        const uint32_t rc = scan[i], l = uabs(cf[rc]);
        if (l == 0) continue;
        int ctx, bc;
        interior_ctx(rc, &ctx, &bc);
        const float bt0 = cdf_cost(base_tok[ctx], 0), bt1 = cdf_cost(base_tok[ctx], 1);
        const float bt2 = cdf_cost(base_tok[ctx], 2), bt3 = cdf_cost(base_tok[ctx], 3);
        auto rate_k = [&](uint32_t k) {
            return k == 0 ? bt0 : k == 1 ? bt1 + 1.0f : k == 2 ? bt2 + 1.0f : (bt3 + hi_cost(k, bc)) + 1.0f;
        };
        uint32_t best_k = l;
        float best_c = distw(i, rc, l) + lambda * rate_k(l);
        for (int k = (int)l - 1; k >= 0; k--) {
            const float dk = distw(i, rc, (uint32_t)k);
            if (dk >= best_c) break;
            const float cc = dk + lambda * rate_k((uint32_t)k);
            if (cc < best_c) {
                best_c = cc;
                best_k = (uint32_t)k;
            }
        }
        if (best_k != l) {
            set_sign(rc, best_k);
            at_map(rc) = level_byte(best_k);
        }
    }
    {
        const uint32_t rc = scan[0], l = uabs(cf[rc]);
        if (l != 0) {
            const int bc = dc_brc();
            const int sgn = cf[rc] < 0 ? 1 : 0;
            auto dc_rate = [&](uint32_t k) {
                if (k == 0) return cdf_cost(base_tok[0], 0);
                const uint32_t tok = std::min(k, 3u);
                float b = cdf_cost(base_tok[0], (int)tok);
                if (tok == 3) b += hi_cost(k, bc);
                return b + cdf_cost(dc_sign, sgn);
            };
            uint32_t best_k = l;
            float best_c = distw(0, rc, l) + lambda * dc_rate(l);
            for (int k = (int)l - 1; k >= 0; k--) {
                const float dk = distw(0, rc, (uint32_t)k);
                if (dk >= best_c) break;
                const float cc = dk + lambda * dc_rate((uint32_t)k);
                if (cc < best_c) {
                    best_c = cc;
                    best_k = (uint32_t)k;
                }
            }
            if (best_k != l) {
                set_sign(rc, best_k);
                at_map(rc) = level_byte(best_k);
            }
        }
    }

    // Step B
    const int m = eob + 1;
    std::vector<float> dcur(m), dzero(m), suf(m + 1, 0.0f);
    for (int i = 0; i < m; i++) {
        const uint32_t rc = scan[i];
        const float q2 = rc == 0 ? dq2_dc : dq2_ac;
        const float at = fabsf(tf[rc]);
        const float e = at - (float)uabs(cf[rc]);
        dcur[i] = q2 * (e * e);
        dzero[i] = q2 * (at * at);
        if (qm_on) {
            const float w2 = qm_at(rc);
            dcur[i] *= w2;
            dzero[i] *= w2;
        }
        if (band_tilt != 0.0f && at < 2.0f) {
            const float wgt = 1.0f / (1.0f + band_tilt * ((float)i * n_inv));
            dzero[i] *= wgt;
            dcur[i] *= wgt;
        }
    }
    float sacc = 0.0f;
    for (int i = m - 1; i >= 1; i--) {
        sacc += dzero[i];
        suf[i] = sacc;
    }
    const uint32_t dc_rc = scan[0], dc_m = uabs(cf[dc_rc]);
    auto eob_cand = [&](int e, uint32_t mm) {
        const int ctx_e = 1 + (e > n / 8 ? 1 : 0) + (e > n / 4 ? 1 : 0);
        const uint32_t tok = std::min(mm, 3u);
        float cc = cdf_cost(eob_base[ctx_e], (int)tok - 1);
        if (tok == 3) {
            const uint32_t rc = scan[e];
            const int ex = (int)(rc >> log2h), ey = (int)(rc & (h - 1));
            cc += hi_cost(mm, (ex | ey) > 1 ? 14 : 7);
        }
        return lambda * (cc + 1.0f) + distw(e, scan[e], mm);
    };
    float dc_cost;
    if (dc_m == 0) {
        dc_cost = lambda * cdf_cost(base_tok[0], 0);
    } else {
        const uint32_t tok = std::min(dc_m, 3u);
        float b = cdf_cost(base_tok[0], (int)tok);
        if (tok == 3) b += hi_cost(dc_m, dc_brc());
        b += cdf_cost(dc_sign, cf[dc_rc] < 0 ? 1 : 0);
        dc_cost = lambda * b;
    }
    dc_cost = dc_cost + dcur[0];
    int best_e = -1;
    uint32_t best_m = 0;
    float best = INFINITY, pre = 0.0f;  // pre: interior cost of scan positions 1 .. e-1
    for (int e = 1; e < m; e++) {
        const uint32_t rc = scan[e], mm = uabs(cf[rc]);
        if (mm != 0) {
            const float base = dc_cost + pre + lambda * eob_pt_cost(e) + suf[e + 1];
            uint32_t mb = mm;
            float cc = base + eob_cand(e, mm);
            if (mm >= 2) {
                const float lo = base + eob_cand(e, mm - 1);
                if (lo < cc) {
                    cc = lo;
                    mb = mm - 1;
                }
            }
            if (cc < best) {
                best = cc;
                best_e = e;
                best_m = mb;
            }
        }
        int ctx, bc;
        interior_ctx(rc, &ctx, &bc);
        pre += lambda * interior_rate(ctx, bc, mm) + dcur[e];
    }
    if (dc_m != 0) {  // the DC-only terminal: eob_base[0], br context 0
        const uint32_t tok = std::min(dc_m, 3u);
        float c0 = cdf_cost(eob_bin_cdf, 0) + cdf_cost(eob_base[0], (int)tok - 1);
        if (tok == 3) c0 += hi_cost(dc_m, 0);
        c0 = c0 + cdf_cost(dc_sign, cf[dc_rc] < 0 ? 1 : 0);
        const float total0 = lambda * c0 + dcur[0] + suf[1];
        if (total0 < best) {
            best = total0;
            best_e = 0;
            best_m = dc_m;
        }
    }
    const float skip = suf[1] + dzero[0] + lambda * 1.0f;
    if (best_e < 0 || skip < best) {
        for (int i = 0; i < m; i++) cf[scan[i]] = 0;
        return;
    }
    const uint32_t rc = scan[best_e];
    if (best_m != uabs(cf[rc])) {
        const bool neg = cf[rc] != 0 ? cf[rc] < 0 : tf[rc] < 0.0f;
        cf[rc] = neg ? -(int32_t)best_m : (int32_t)best_m;
    }
    for (int i = best_e + 1; i < m; i++) cf[scan[i]] = 0;
}
