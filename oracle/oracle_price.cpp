// CPU oracle (TEST INFRASTRUCTURE ONLY) of the decision pricing helpers: the SATD partition
// price, the proxy units and the perceptual RD scale with the activity measures it reads.
//
// f_fmlaf (src/aq_common.rs:66) is a true fused multiply-add only when `fma` is a
// compile-time target feature; the reference's default x86-64 build has it off, so every
// f_fmlaf here is a product and a sum, each rounded (the library is built with
// -ffp-contract=off).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle.h"

namespace {
inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float fmlaf(float a, float b, float c) { return a * b + c; }
inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// 2^((i - 32) / 64) rounded to f32 (the table of dirty_exp2f, src/coder/lossy_state.rs:4270):
// computed here from its definition in long double and rounded once.
struct Exp2Tab {
    float v[64];
    Exp2Tab() {
        for (int i = 0; i < 64; i++) v[i] = (float)std::exp2l((long double)(i - 32) / 64.0L);
    }
};
const Exp2Tab& exp2tab() { static Exp2Tab t; return t; }

constexpr float kPartLambda = 2.775f;  // K_PART_LAMBDA (src/cost.rs:278)
}  // namespace

extern "C" {

// dirty_log1pf (src/aq_common.rs:90)
float orc_dirty_log1pf(float d) {
    const uint32_t ix = f_bits(1.0f + d);
    const int n = (int)((ix & 0x7f800000u) >> 23) - 0x7f;
    const float mantissa = bits_f((ix & 0x007fffffu) | 0x3f800000u);
    const float t = n == 0 ? d : mantissa - 1.0f;
    float p = 0.014539075084030628f;
    p = fmlaf(p, t, -0.0675969123840332f);
    p = fmlaf(p, t, 0.15056970715522766f);
    p = fmlaf(p, t, -0.23573730885982513f);
    p = fmlaf(p, t, 0.33125850558280945f);
    p = fmlaf(p, t, -0.4998837411403656f);
    p = fmlaf(p, t, 0.999998927116394f);
    return fmlaf((float)n, 0.693147180559945309f, t * p);
}

// dirty_exp2f (src/coder/lossy_state.rs:4284)
float orc_dirty_exp2f(float d) {
    const float redux = bits_f(0x4b400000u) / 64.0f;
    const uint32_t ui = f_bits(d + redux);
    uint32_t i0 = ui + 32u;
    const uint32_t k = i0 / 64u;
    i0 &= 63u;
    const float uf = bits_f(ui) - redux;
    const float z0 = exp2tab().v[i0];
    const float f = d - uf;
    float u = 0.24022668600082397f;
    u = fmlaf(u, f, 0.693149745464325f);
    u *= f;
    const float i2 = bits_f((k + 0x7fu) << 23);
    return fmlaf(u, z0, z0) * i2;
}

float orc_exp2_table(int i) { return exp2tab().v[i & 63]; }

// block_moments_i32 (src/coder/lossy_state.rs:91)
void orc_block_moments(const uint16_t* plane, int stride, int px, int py, int w, int h,
                       int64_t* sum, int64_t* sum_sq) {
    int64_t s = 0, q = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int64_t v = plane[(size_t)(py + y) * stride + px + x];
            s += v;
            q += v * v;
        }
    *sum = s;
    *sum_sq = q;
}

// sb_activity (src/coder/aq.rs:140)
float orc_sb_activity(const uint16_t* yp, int pw, int sb_y, int sb_x, int width, int height, int bd) {
    const int h = height > sb_y ? (height - sb_y < 64 ? height - sb_y : 64) : 0;
    const int w = width > sb_x ? (width - sb_x < 64 ? width - sb_x : 64) : 0;
    if (h == 0 || w == 0) return 0.0f;
    int64_t sum, sq;
    orc_block_moments(yp, pw, sb_x, sb_y, w, h, &sum, &sq);
    const float n = (float)(h * w);
    const float mean = (float)sum / n;
    float var = (float)sq / n - mean * mean;
    if (!(var >= 0.0f)) var = 0.0f;  // f32::max(0.0)
    const int b = bd < 8 ? 8 : bd;
    return orc_dirty_log1pf(var / (float)(1u << (2 * (b - 8))));
}

// tile_ref_activity (src/coder/aq.rs:163): raster-order f32 mean of sb_activity
float orc_tile_ref_activity(const uint16_t* yp, int pw, int w, int h, int bd) {
    float sum = 0.0f, cnt = 0.0f;
    for (int sb_y = 0; sb_y < h; sb_y += 64)
        for (int sb_x = 0; sb_x < w; sb_x += 64) {
            sum += orc_sb_activity(yp, pw, sb_y, sb_x, w, h, bd);
            cnt += 1.0f;
        }
    return cnt > 0.0f ? sum / cnt : 5.0f;
}

// luma_satd_scale (src/partition_rd.rs:30) at the shipped tuning (src/tuning.rs:854-857)
float orc_luma_satd_scale(int base_q_idx, int subsampled) {
    const float base = 1.25f, top = 2.921556f, lo = 20.0f, hi = 48.0f;
    if (subsampled) return base;
    const float qi = (float)base_q_idx;
    if (qi <= lo) return top;
    if (qi >= hi) return base;
    const float d = hi - lo;
    return top + (base - top) * (qi - lo) / (d > 1.0f ? d : 1.0f);
}

// k_satd (src/partition_rd.rs:54)
float orc_k_satd(int base_q_idx, int subsampled) {
    const float QI[3] = {28.0f, 81.0f, 140.0f};
    const float k444[3] = {35.0f, 21.2f, 18.1f}, ksub[3] = {24.5f, 21.3f, 18.2f};
    const float* k = subsampled ? ksub : k444;
    const float qi = (float)base_q_idx;
    if (qi <= QI[0]) return k[0];
    if (qi >= QI[2]) return k[2];
    if (qi <= QI[1]) return k[0] + (k[1] - k[0]) * (qi - QI[0]) / (QI[1] - QI[0]);
    return k[1] + (k[2] - k[1]) * (qi - QI[1]) / (QI[2] - QI[1]);
}

// luma_partition_distortion (src/coder/lossy_state.rs:2385) from the block's luma_satd, with
// the shipped proxy_luma_bridge (on) and proxy_luma_scale 7.130643 (src/tuning.rs:783-784)
float orc_luma_partition_distortion(uint64_t satd, float qstep, int base_q_idx, int subsampled) {
    float price = (float)satd * (qstep >= 1.0f ? qstep : 1.0f) * 0.25f *
                  orc_luma_satd_scale(base_q_idx, subsampled);
    price *= kPartLambda / orc_k_satd(base_q_idx, subsampled) * 7.130643f;
    return price;
}

// proxy_wsse (src/coder/lossy_state.rs:1829)
float orc_proxy_wsse(float d) { return d * kPartLambda; }

// perceptual_rd_scale (src/coder/lossy_state.rs:4189) at the shipped tuning: prdo_k 0.4,
// prdo_k_444 0.5 (src/tuning.rs:827-828), top_ease 1 / knee 20 / width 12 (:750-752),
// local_ref_blend 0.5 (:863), clamp [1/2, 1.5] (src/coder.rs:382-388).  fmt: MT_ENC_*
// numbering (0 4:2:0, 1 4:2:2, 2 4:4:4, 3 monochrome); ref_act: the AQ state's tile activity
// (0 when AQ is off); the source plane is the tile's luma (stride pw, w x h).
float orc_perceptual_rd_scale(const uint16_t* src, int pw, int w, int h, int bd, int base_q_idx,
                              int fmt, int mono_large, float ref_act, int px, int py, int dim) {
    const bool is444 = fmt == 2, is422 = fmt == 1, is420 = fmt == 0, mono = fmt == 3;
    float k = is444 ? 0.5f : 0.4f;
    if (k == 0.0f) return 1.0f;
    if (is422 || mono_large) {
        const float te = clampf((20.0f - (float)base_q_idx) / 12.0f, 0.0f, 1.0f);
        const float g = fmlaf(-1.0f, te, 1.0f);
        k *= g;
        if (k == 0.0f) return 1.0f;
    }
    if (!is422 && !mono) {
        const float fade = 1.0f / (100.0f - 50.0f);
        float t = clampf(((float)base_q_idx - 50.0f) * fade, 0.0f, 1.0f);
        const float floor = is420 ? 0.4f : 0.0f;
        t = fmlaf(1.0f - floor, t, floor);
        k *= t;
        if (k == 0.0f) return 1.0f;
    }
    const float a = 0.5f;
    const float sa = orc_sb_activity(src, pw, py & ~63, px & ~63, w, h, bd);
    const float refa = ref_act * (1.0f - a) + sa * a;
    if (refa <= 0.0f) return 1.0f;
    const int bw = dim < (w > px ? w - px : 0) ? dim : (w > px ? w - px : 0);
    const int bh = dim < (h > py ? h - py : 0) ? dim : (h > py ? h - py : 0);
    if (bw == 0 || bh == 0) return 1.0f;
    int64_t sum, sq;
    orc_block_moments(src, pw, px, py, bw, bh, &sum, &sq);
    const float n = (float)(bw * bh);
    const float mean = (float)sum / n;
    float var = (float)sq / n - mean * mean;
    if (!(var >= 0.0f)) var = 0.0f;
    const float act = orc_dirty_log1pf(var / (float)(1u << (2 * (bd - 8))));
    const float c = 2.0f;
    const float ln2 = 0.693147180559945309f, log2e = 1.44269504088896341f;
    const float exponent = clampf(k * (act - refa), -ln2, ln2);
    return clampf(orc_dirty_exp2f(exponent * log2e), 1.0f / c, 1.5f);
}

}  // extern "C"
