// CPU oracle: intra prediction, CfL, residual and distortion kernels.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Scalar restatement of the reference encoder's
// src/intrapred.rs, src/rd_sse.rs, src/partition_rd.rs and src/avx/rd.rs, written from
// the Rust source text.  Every function cites the reference function it follows.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>

#include "oracle.h"
#include "../maroontree_amd/csrc/av1_cdf_tables.h"

namespace {

constexpr int kOrigin = 2;     // IntraEdge::EDGE_ORIGIN (src/intrapred.rs:422)
constexpr int kEdgeCap = 132;  // EDGE_CAPACITY (src/intrapred.rs:423)

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline int ctz(unsigned v) { return __builtin_ctz(v); }

// One reference edge: samples at logical indices -2 .. 129.
struct Edge {
    int32_t s[kEdgeCap];
    int32_t& at(int i) { return s[i + kOrigin]; }
    int32_t at(int i) const { return s[i + kOrigin]; }
};

// intra_edge_filter_strength (src/intrapred.rs:917)
int edge_strength(int w, int h, bool smooth_nb, int delta) {
    const int d = std::abs(delta);
    const int wh = w + h;
    if (!smooth_nb) {
        if (wh <= 8) return d >= 56 ? 1 : 0;
        if (wh <= 16) return d >= 40 ? 1 : 0;
        if (wh <= 24) return d >= 32 ? 3 : d >= 16 ? 2 : d >= 8 ? 1 : 0;
        if (wh <= 32) return d >= 32 ? 3 : d >= 4 ? 2 : d >= 1 ? 1 : 0;
        return d >= 1 ? 3 : 0;
    }
    if (wh <= 8) return d >= 64 ? 2 : d >= 40 ? 1 : 0;
    if (wh <= 16) return d >= 48 ? 2 : d >= 20 ? 1 : 0;
    if (wh <= 24) return d >= 4 ? 3 : 0;
    return d >= 1 ? 3 : 0;
}

// use_intra_edge_upsample (src/intrapred.rs:1021)
bool edge_upsample(int w, int h, bool smooth_nb, int delta) {
    const int d = std::abs(delta);
    if (d == 0 || d >= 40) return false;
    return smooth_nb ? (w + h <= 8) : (w + h <= 16);
}

// filter_intra_edge (src/intrapred.rs:977): the `size` samples from index -1 go through a
// 5-tap kernel with clamped reads; outputs land on indices 0 .. size-2.
void filter_edge(Edge& e, int size, int strength) {
    if (strength == 0) return;
    static const int kTaps[3][5] = {{0, 4, 8, 4, 0}, {0, 5, 6, 5, 0}, {2, 4, 4, 4, 2}};
    const int* k = kTaps[strength - 1];
    int32_t src[kEdgeCap];
    for (int i = 0; i < size; ++i) src[i] = e.at(i - 1);
    for (int o = 0; o + 1 < size; ++o) {
        int sum = 0;
        for (int j = 0; j < 5; ++j) sum += k[j] * src[clampi(o - 1 + j, 0, size - 1)];
        e.at(o) = (sum + 8) >> 4;
    }
}

// upsample_intra_edge (src/intrapred.rs:1026)
void upsample_edge(Edge& e, int num_px, int bd) {
    int32_t in[kEdgeCap];
    in[0] = e.at(-1);
    in[1] = e.at(-1);
    for (int i = 0; i < num_px; ++i) in[i + 2] = e.at(i);
    in[num_px + 2] = e.at(num_px - 1);
    e.at(-2) = in[0];
    const int maxv = (1 << bd) - 1;
    for (int i = 0; i < num_px; ++i) {
        const int v = -in[i] + 9 * in[i + 1] + 9 * in[i + 2] - in[i + 3];
        e.at(2 * i - 1) = clampi((v + 8) >> 4, 0, maxv);
        e.at(2 * i) = in[i + 2];
    }
}

// dr_interp_scalar (src/intrapred.rs:617): `count` samples from e[base + i*step].
void dr_interp(const Edge& e, int base, int step, int shift, int32_t* out, int count) {
    for (int i = 0; i < count; ++i) {
        const int b = base + i * step;
        out[i] = (e.at(b) * (32 - shift) + e.at(b + 1) * shift + 16) >> 5;
    }
}

// dr_predict_scalar (src/intrapred.rs:629)
void dr_predict(int zone, int bw, int bh, int edge_len, int dx, int dy, int upa, int upl,
                const Edge& above, const Edge& left, int32_t* out) {
    if (zone == 1) {
        const int max_base_x = (edge_len - 1) << upa;
        const int frac_bits = 6 - upa;
        const int inc = 1 << upa;
        for (int y = 0; y < bh; ++y) {
            int32_t* row = out + y * bw;
            const int xpos = dx * (y + 1);
            const int shift = ((xpos << upa) & 0x3f) >> 1;
            const int base = xpos >> frac_bits;
            const int n_fit =
                base >= max_base_x ? 0 : std::min((max_base_x - base + inc - 1) / inc, bw);
            dr_interp(above, base, inc, shift, row, n_fit);
            for (int x = n_fit; x < bw; ++x) row[x] = above.at(max_base_x);
        }
    } else if (zone == 3) {
        const int max_base_y = (edge_len - 1) << upl;
        const int frac_bits = 6 - upl;
        const int inc = 1 << upl;
        for (int y = 0; y < bh; ++y)
            for (int x = 0; x < bw; ++x) {
                const int ypos = dy * (x + 1);
                const int shift = ((ypos << upl) & 0x3f) >> 1;
                const int base = (ypos >> frac_bits) + y * inc;
                out[y * bw + x] =
                    base >= max_base_y
                        ? left.at(max_base_y)
                        : (left.at(base) * (32 - shift) + left.at(base + 1) * shift + 16) >> 5;
            }
    } else {
        const int frac_y = 6 - upl;
        for (int y = 0; y < bh; ++y) {
            int32_t* row = out + y * bw;
            const int t = (y + 1) * dx - 64;
            const int x0 = t <= 0 ? 0 : std::min((t + 63) >> 6, bw);
            for (int x = 0; x < x0; ++x) {
                const int ypos = (y << 6) - (x + 1) * dy;
                const int base = ypos >> frac_y;
                const int shift = ((ypos * (1 << upl)) & 0x3f) >> 1;
                row[x] = (left.at(base) * (32 - shift) + left.at(base + 1) * shift + 16) >> 5;
            }
            if (x0 < bw) {
                const int xpos = (x0 << 6) - (y + 1) * dx;
                const int base = xpos >> (6 - upa);
                const int shift = ((xpos * (1 << upa)) & 0x3f) >> 1;
                dr_interp(above, base, 1 << upa, shift, row + x0, bw - x0);
            }
        }
    }
}

const int32_t* sm_w(int n) { return &MT_SM_WEIGHTS[n]; }  // sm_weights (src/intrapred.rs:1091)

// Edge assembly of intra_predict_nd_ad_impl (src/intrapred.rs:1376-1459): above/left rows
// with the availability fills, the corner, and the top-right / bottom-left extension by
// min(bw, bh, remaining frame) real samples followed by replication.
void build_edges(const uint16_t* recon, int stride, int ox, int oy, int bw, int bh, bool have_tr,
                 bool have_bl, int fw, int fh, int bd, Edge& above, Edge& left, int32_t& corner) {
    const bool top = oy > 0, lft = ox > 0;
    const int base = 1 << (bd - 1);
    const int n = bw + bh;
    auto R = [&](int x, int y) -> int32_t { return recon[(size_t)y * stride + x]; };
    if (top) {
        for (int i = 0; i < bw; ++i) above.at(i) = R(ox + i, oy - 1);
    } else {
        const int32_t f = lft ? R(ox - 1, oy) : base - 1;
        for (int i = 0; i < n; ++i) above.at(i) = f;
    }
    if (lft) {
        for (int j = 0; j < bh; ++j) left.at(j) = R(ox - 1, oy + j);
    } else {
        const int32_t f = top ? R(ox, oy - 1) : base + 1;
        for (int i = 0; i < n; ++i) left.at(i) = f;
    }
    corner = lft ? (top ? R(ox - 1, oy - 1) : R(ox - 1, oy)) : (top ? R(ox, oy - 1) : base);
    above.at(-1) = corner;
    left.at(-1) = corner;
    if (top) {
        const int have = have_tr ? std::min(std::min(bw, bh), std::max(fw - (ox + bw), 0)) : 0;
        for (int i = 0; i < have; ++i) above.at(bw + i) = R(ox + bw + i, oy - 1);
        const int32_t f = above.at(std::max(bw + have - 1, 0));
        for (int i = bw + have; i < n; ++i) above.at(i) = f;
    }
    if (lft) {
        const int have = have_bl ? std::min(std::min(bw, bh), std::max(fh - (oy + bh), 0)) : 0;
        for (int i = 0; i < have; ++i) left.at(bh + i) = R(ox - 1, oy + bh + i);
        const int32_t f = left.at(std::max(bh + have - 1, 0));
        for (int i = bh + have; i < n; ++i) left.at(i) = f;
    }
}

bool valid_dim(int v) { return v == 4 || v == 8 || v == 16 || v == 32 || v == 64; }

}  // namespace

extern "C" {

int orc_edge_filter_strength(int w, int h, int filter_type, int delta) {
    return edge_strength(w, h, filter_type != 0, delta);
}

int orc_use_edge_upsample(int w, int h, int filter_type, int delta) {
    return edge_upsample(w, h, filter_type != 0, delta) ? 1 : 0;
}

void orc_filter_edge(int32_t* edge_origin2, int size, int strength) {
    Edge e;
    std::memcpy(e.s, edge_origin2, sizeof(e.s));
    filter_edge(e, size, strength);
    std::memcpy(edge_origin2, e.s, sizeof(e.s));
}

void orc_upsample_edge(int32_t* edge_origin2, int num_px, int bd) {
    Edge e;
    std::memcpy(e.s, edge_origin2, sizeof(e.s));
    upsample_edge(e, num_px, bd);
    std::memcpy(edge_origin2, e.s, sizeof(e.s));
}

void orc_dr_predict(int zone, int bw, int bh, int edge_len, int dx, int dy, int up_above,
                    int up_left, const int32_t* above_origin2, const int32_t* left_origin2,
                    int32_t* out) {
    Edge a, l;
    std::memcpy(a.s, above_origin2, sizeof(a.s));
    std::memcpy(l.s, left_origin2, sizeof(l.s));
    dr_predict(zone, bw, bh, edge_len, dx, dy, up_above, up_left, a, l, out);
}

// dc_pred_from_sum (src/intrapred.rs:1772)
int32_t orc_dc_pred_from_sum(int32_t edge_sum, int w, int h, int have_top, int have_left, int bd) {
    if (have_top && have_left) {
        int32_t s = ((w + h) >> 1) + edge_sum;
        s >>= ctz((unsigned)(w + h));
        if (w != h) {
            const uint32_t mult = (w > 2 * h || h > 2 * w) ? 0x3334u : 0x5556u;
            s = (int32_t)(((uint32_t)s * mult) >> 16);
        }
        return s;
    }
    if (have_top) return ((w >> 1) + edge_sum) >> ctz((unsigned)w);
    if (have_left) return ((h >> 1) + edge_sum) >> ctz((unsigned)h);
    return 1 << (bd - 1);
}

// dc_pred (src/intrapred.rs:1740)
int32_t orc_dc_pred(const uint16_t* recon, int stride, int ox, int oy, int w, int h, int bd) {
    int32_t sum = 0;
    if (oy > 0)
        for (int i = 0; i < w; ++i) sum += recon[(size_t)(oy - 1) * stride + ox + i];
    if (ox > 0)
        for (int j = 0; j < h; ++j) sum += recon[(size_t)(oy + j) * stride + ox - 1];
    return orc_dc_pred_from_sum(sum, w, h, oy > 0, ox > 0, bd);
}

// filter_intra_predict_impl + scalar_filter_intra_cells (src/intrapred.rs:473, :533)
int orc_filter_intra(int fmode, const uint16_t* recon, int stride, int ox, int oy, int w, int h,
                     int bd, int32_t* out) {
    if (fmode < 0 || fmode > 4 || (w & 3) || (h & 1) || w > 32 || h > 32 || w < 4 || h < 2)
        return -1;
    const int base = 1 << (bd - 1);
    const bool top = oy > 0, lft = ox > 0;
    auto R = [&](int x, int y) -> int32_t { return recon[(size_t)y * stride + x]; };
    int32_t buf[33][33];
    buf[0][0] = top ? (lft ? R(ox - 1, oy - 1) : R(ox, oy - 1)) : (lft ? R(ox - 1, oy) : base);
    for (int x = 0; x < w; ++x)
        buf[0][x + 1] = top ? R(ox + x, oy - 1) : (lft ? R(ox - 1, oy) : base - 1);
    for (int y = 0; y < h; ++y)
        buf[y + 1][0] = lft ? R(ox - 1, oy + y) : (top ? R(ox, oy - 1) : base + 1);
    const int maxv = (1 << bd) - 1;
    for (int r = 1; r <= h; r += 2)
        for (int c = 1; c <= w; c += 4) {
            const int32_t p[7] = {buf[r - 1][c - 1], buf[r - 1][c],     buf[r - 1][c + 1],
                                  buf[r - 1][c + 2], buf[r - 1][c + 3], buf[r][c - 1],
                                  buf[r + 1][c - 1]};
            for (int k = 0; k < 8; ++k) {
                int32_t sum = 0;
                for (int t = 0; t < 7; ++t) sum += (int32_t)MT_INTRA_FILTER_TAPS[fmode][k][t] * p[t];
                buf[r + (k >> 2)][c + (k & 3)] = clampi((sum + 8) >> 4, 0, maxv);
            }
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) out[y * w + x] = buf[y + 1][x + 1];
    return 0;
}

// IntraPredDispatch::predict_nd_ad -> intra_predict_nd_ad_impl (src/intrapred.rs:359,
// :1357-1595) for V..PAETH (modes 1..12), with the edge filter on.  Mode 0 (DC) fills the
// block with dc_pred (the callers' use of IntraPredDispatch::dc_pred, :232); modes 16..20
// are filter intra mode 0..4 (filter_predict, :399).  Returns 0, or -1 for an invalid call.
int orc_intra_pred(int mode, int angle_delta, const uint16_t* recon, int stride, int ox, int oy,
                   int bw, int bh, int have_tr, int have_bl, int fw, int fh, int filter_type,
                   int bd, int32_t* out) {
    if (!valid_dim(bw) || !valid_dim(bh) || angle_delta < -3 || angle_delta > 3) return -1;
    if (mode >= 16) return orc_filter_intra(mode - 16, recon, stride, ox, oy, bw, bh, bd, out);
    if (mode == 0) {
        const int32_t dc = orc_dc_pred(recon, stride, ox, oy, bw, bh, bd);
        for (int i = 0; i < bw * bh; ++i) out[i] = dc;
        return 0;
    }
    if (mode > 12) return -1;
    Edge above, left;
    std::memset(&above, 0, sizeof(above));
    std::memset(&left, 0, sizeof(left));
    int32_t corner = 0;
    build_edges(recon, stride, ox, oy, bw, bh, have_tr != 0, have_bl != 0, fw, fh, bd, above,
                left, corner);
    const int n = bw + bh;
    const bool smooth_nb = filter_type != 0;
    static const int kNominal[9] = {0, 90, 180, 45, 135, 113, 157, 203, 67};
    const bool directional = mode >= 1 && mode <= 8;
    const int angle = directional ? kNominal[mode] + 3 * angle_delta : 0;
    int upa = 0, upl = 0;
    if (directional && angle != 90 && angle != 180) {
        if (angle > 90 && angle < 180 && n >= 24) {  // filter_intra_corner (:1014)
            const int32_t f = (5 * left.at(0) + 6 * above.at(-1) + 5 * above.at(0) + 8) >> 4;
            above.at(-1) = f;
            left.at(-1) = f;
        }
        filter_edge(above, bw + 1 + (angle < 90 ? bh : 0),
                    edge_strength(bw, bh, smooth_nb, angle - 90));
        filter_edge(left, bh + 1 + (angle > 180 ? bw : 0),
                    edge_strength(bh, bw, smooth_nb, angle - 180));
        upa = edge_upsample(bw, bh, smooth_nb, angle - 90) ? 1 : 0;
        upl = edge_upsample(bh, bw, smooth_nb, angle - 180) ? 1 : 0;
        if (upa) upsample_edge(above, bw + (angle < 90 ? bh : 0), bd);
        if (upl) upsample_edge(left, bh + (angle > 180 ? bw : 0), bd);
    }
    if (directional) {
        if (angle == 90) {  // vertical_scalar (:605)
            for (int y = 0; y < bh; ++y)
                for (int x = 0; x < bw; ++x) out[y * bw + x] = above.at(x);
        } else if (angle == 180) {  // horizontal_scalar (:611)
            for (int y = 0; y < bh; ++y)
                for (int x = 0; x < bw; ++x) out[y * bw + x] = left.at(y);
        } else if (angle < 90) {
            dr_predict(1, bw, bh, n, MT_DR_INTRA_DERIVATIVE[angle >> 1], 0, upa, 0, above, left,
                       out);
        } else if (angle > 180) {
            dr_predict(3, bw, bh, n, 0, MT_DR_INTRA_DERIVATIVE[(270 - angle) >> 1], 0, upl, above,
                       left, out);
        } else {
            dr_predict(2, bw, bh, n, MT_DR_INTRA_DERIVATIVE[(180 - angle) >> 1],
                       MT_DR_INTRA_DERIVATIVE[(angle - 90) >> 1], upa, upl, above, left, out);
        }
        return 0;
    }
    const int32_t* wv = sm_w(bh);
    const int32_t* wh = sm_w(bw);
    const int32_t right = above.at(bw - 1), bottom = left.at(bh - 1);
    for (int y = 0; y < bh; ++y)
        for (int x = 0; x < bw; ++x) {
            const int32_t t = above.at(x), l = left.at(y);
            int32_t v;
            switch (mode) {
                case 9:  // smooth_scalar (:1679)
                    v = (wv[y] * t + (256 - wv[y]) * bottom + wh[x] * l + (256 - wh[x]) * right +
                         256) >> 9;
                    break;
                case 10:  // smooth_v_scalar (:1691)
                    v = (wv[y] * t + (256 - wv[y]) * bottom + 128) >> 8;
                    break;
                case 11:  // smooth_h_scalar (:1702)
                    v = (wh[x] * l + (256 - wh[x]) * right + 128) >> 8;
                    break;
                default: {  // paeth_scalar (:1597)
                    const int32_t b = l + t - corner;
                    const int32_t ld = std::abs(l - b), td = std::abs(t - b),
                                  cd = std::abs(corner - b);
                    v = (ld <= td && ld <= cd) ? l : (td <= cd ? t : corner);
                }
            }
            out[y * bw + x] = v;
        }
    return 0;
}

// ---- CfL (src/intrapred.rs:1112-1260) --------------------------------------------------

// cfl_ac_444 (:1112)
void orc_cfl_ac_444(const uint16_t* luma, int w, int h, int32_t* ac) {
    orc_cfl_ac_sub(luma, w, w, h, 0, 0, ac);
}

// cfl_ac_sub (:1133); with no subsampling it is cfl_ac_444 (both scale by 8)
void orc_cfl_ac_sub(const uint16_t* luma, int lstride, int cw, int ch, int ss_hor, int ss_ver,
                    int32_t* ac) {
    const int shift = 1 + (ss_ver ? 0 : 1) + (ss_hor ? 0 : 1);
    for (int y = 0; y < ch; ++y)
        for (int x = 0; x < cw; ++x) {
            const int ly = y << (ss_ver ? 1 : 0), lx = x << (ss_hor ? 1 : 0);
            int32_t s = luma[(size_t)ly * lstride + lx];
            if (ss_hor) s += luma[(size_t)ly * lstride + lx + 1];
            if (ss_ver) {
                s += luma[(size_t)(ly + 1) * lstride + lx];
                if (ss_hor) s += luma[(size_t)(ly + 1) * lstride + lx + 1];
            }
            ac[y * cw + x] = s << shift;
        }
    const int n = cw * ch;
    const int log2sz = ctz((unsigned)cw) + ctz((unsigned)ch);
    int64_t sum = (int64_t(1) << log2sz) >> 1;
    for (int i = 0; i < n; ++i) sum += ac[i];
    const int32_t mean = (int32_t)(sum >> log2sz);
    for (int i = 0; i < n; ++i) ac[i] -= mean;
}

// cfl_pred_pixel (:1194)
int32_t orc_cfl_pred_pixel(int32_t dc, int32_t ac, int32_t alpha, int bd) {
    const int32_t diff = alpha * ac;
    const int32_t mag = (std::abs(diff) + 32) >> 6;
    return clampi(dc + (diff < 0 ? -mag : mag), 0, (1 << bd) - 1);
}

// cfl_best_alpha_strided (:1218)
int32_t orc_cfl_best_alpha(const int32_t* ac, int ac_stride, const uint16_t* src, int src_stride,
                           int w, int h, int32_t dc, int bd) {
    if (w == 0 || h == 0) return 0;
    int64_t num = 0, den = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int64_t a = ac[(size_t)y * ac_stride + x];
            num += (int64_t)((int32_t)src[(size_t)y * src_stride + x] - dc) * a;
            den += a * a;
        }
    if (den == 0) return 0;
    const int64_t sg = num > 0 ? 1 : (num < 0 ? -1 : 0);
    int64_t a0 = (64 * num + (den >> 1) * sg) / den;
    a0 = std::max<int64_t>(-16, std::min<int64_t>(16, a0));
    int32_t best_a = 0;
    int64_t best_e = INT64_MAX;
    for (int c = (int)a0 - 3; c <= (int)a0 + 3; ++c) {
        if (c < -16 || c > 16) continue;
        int64_t e = 0;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const int64_t d = (int32_t)src[(size_t)y * src_stride + x] -
                                  orc_cfl_pred_pixel(dc, ac[(size_t)y * ac_stride + x], c, bd);
                e += d * d;
            }
        if (e < best_e) {
            best_e = e;
            best_a = c;
        }
    }
    return best_a;
}

// The alpha s steps from a0 in the direction of a0's sign (s < 0: toward zero), as
// cfl_alpha_rd_refine's walk visits them (src/cost.rs:332-372): clipped to -16..16 and never
// past zero; a0 = 0 stays 0.
int32_t orc_cfl_alpha_step(int32_t a0, int32_t s) {
    const int32_t dir = a0 > 0 ? 1 : (a0 < 0 ? -1 : 0);
    int32_t a = clampi(a0 + dir * s, -16, 16);
    if (a * dir < 0) a = 0;
    return a;
}

// cfl_alpha_rd_refine (src/cost.rs:332): from the energy optimum a0 of walk cost c0, steps
// toward zero while the cost keeps falling; if the first such step does not lower it, steps
// away from zero the same way; at most `steps` each way.  cost[a + 16] is the walk cost of
// alpha a (read only for the alphas the walk visits).
int32_t orc_cfl_alpha_rd_refine(int32_t a0, float c0, int steps, const float* cost) {
    if (steps <= 0 || a0 == 0) return a0;
    const int32_t dir = a0 > 0 ? 1 : -1;
    int32_t best_a = a0;
    float best_c = c0;
    for (int s = 1; s <= steps; ++s) {
        const int32_t a = a0 - dir * s;
        if (a * dir < 0) break;
        const float c = cost[a + 16];
        if (!(c < best_c)) break;
        best_a = a;
        best_c = c;
    }
    if (best_a == a0)
        for (int s = 1; s <= steps; ++s) {
            const int32_t a = a0 + dir * s;
            if (std::abs(a) > 16) break;
            const float c = cost[a + 16];
            if (!(c < best_c)) break;
            best_a = a;
            best_c = c;
        }
    return best_a;
}

// ---- residual and distortion (src/rd_sse.rs, src/partition_rd.rs) ---------------------

// residual_pred_scalar (src/rd_sse.rs:1179)
void orc_residual_pred(const uint16_t* src, int stride, const int32_t* pred, int w, int h,
                       int32_t* dst) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            dst[y * w + x] = (int32_t)src[(size_t)y * stride + x] - pred[y * w + x];
}

// RdDispatch::sse_recon without the weighted branch (src/rd_sse.rs:454): plain SSE of
// clamp(pred + resid) over the visible vw x vh corner (sse_recon_scalar :1262 and
// sse_recon_strided_scalar :1290 give the same sum).
int64_t orc_sse_recon_vis(const int32_t* pred, const int32_t* resid, const uint16_t* src,
                          int stride, int w, int vw, int vh, int bd) {
    const int maxv = (1 << bd) - 1;
    int64_t sse = 0;
    for (int y = 0; y < vh; ++y)
        for (int x = 0; x < vw; ++x) {
            const int i = y * w + x;
            const int64_t d = (int32_t)src[(size_t)y * stride + x] -
                              clampi(pred[i] + (resid ? resid[i] : 0), 0, maxv);
            sse += d * d;
        }
    return sse;
}

namespace {
// WSSE_DCT4 and WSSE_WEIGHTS (src/rd_sse.rs:1402-1428): f32 literals, and the hpvca table
// divided by its mean in f32.
const float kD[4][4] = {{0.5f, 0.5f, 0.5f, 0.5f},
                        {0.6532815f, 0.27059805f, -0.27059805f, -0.6532815f},
                        {0.5f, -0.5f, -0.5f, 0.5f},
                        {0.27059805f, -0.6532815f, 0.6532815f, -0.27059805f}};
struct WsseW {
    float w[4][4];
    WsseW() {
        volatile float mean = 0.4625f;
        const float dc = 1.8f / mean, diag = 0.8f / mean, axis = 0.4f / mean, hf = 0.2f / mean,
                    ac1 = 1.0f / mean;
        const float t[4][4] = {
            {dc, ac1, axis, hf}, {ac1, diag, hf, hf}, {axis, hf, hf, hf}, {hf, hf, hf, hf}};
        std::memcpy(w, t, sizeof(w));
    }
};
const WsseW& wsse_w() {
    static const WsseW w;
    return w;
}
}  // namespace

void orc_wsse_tables(float* dct16, float* weights16) {
    std::memcpy(dct16, kD, sizeof(kD));
    std::memcpy(weights16, wsse_w().w, sizeof(kD));
}

// weighted_sse_scalar (src/rd_sse.rs:1433): separate f32 multiplies and adds.
float orc_weighted_sse_scalar(const int32_t* err, int w, int h) {
    const auto& W = wsse_w().w;
    float total = 0.0f;
    for (int r0 = 0; r0 < h; r0 += 4)
        for (int c0 = 0; c0 < w; c0 += 4) {
            float e[4][4];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) e[i][j] = (float)err[(r0 + i) * w + c0 + j];
            for (int k = 0; k < 4; ++k) {
                float t[4];
                for (int j = 0; j < 4; ++j)
                    t[j] = e[0][j] * kD[k][0] + e[1][j] * kD[k][1] + e[2][j] * kD[k][2] +
                           e[3][j] * kD[k][3];
                for (int l = 0; l < 4; ++l) {
                    const float c =
                        kD[l][0] * t[0] + kD[l][1] * t[1] + kD[l][2] * t[2] + kD[l][3] * t[3];
                    total += W[k][l] * c * c;
                }
            }
        }
    return total;
}

// weighted_sse_avx2 (src/avx/rd.rs:767-849), the kernel the reference selects on x86 with
// AVX2 and FMA: 32 accumulators acc[k][half][l] (row k of the coefficient tile, the tile's
// position in an 8-wide pair, column frequency l), each a sequential fused chain over the
// tile pairs in (r0, c0) order; then the fixed reduction tree, the 4-wide tail column, and
// the final horizontal sum.
float orc_weighted_sse_avx2(const int32_t* err, int w, int h) {
    const auto& W = wsse_w().w;
    float acc[4][2][4] = {};
    const int pairs = w / 8 * 8;
    auto tile_coef = [&](int r0, int c0, int k, float* c) {
        float t[4];
        for (int j = 0; j < 4; ++j) {
            float v = (float)err[(r0 + 0) * w + c0 + j] * kD[k][0];
            v = std::fmaf((float)err[(r0 + 1) * w + c0 + j], kD[k][1], v);
            v = std::fmaf((float)err[(r0 + 2) * w + c0 + j], kD[k][2], v);
            v = std::fmaf((float)err[(r0 + 3) * w + c0 + j], kD[k][3], v);
            t[j] = v;
        }
        for (int l = 0; l < 4; ++l) {
            float v = kD[l][0] * t[0];
            v = std::fmaf(kD[l][1], t[1], v);
            v = std::fmaf(kD[l][2], t[2], v);
            v = std::fmaf(kD[l][3], t[3], v);
            c[l] = v;
        }
    };
    for (int r0 = 0; r0 < h; r0 += 4)
        for (int c0 = 0; c0 < pairs; c0 += 8)
            for (int k = 0; k < 4; ++k)
                for (int half = 0; half < 2; ++half) {
                    float c[4];
                    tile_coef(r0, c0 + 4 * half, k, c);
                    for (int l = 0; l < 4; ++l)
                        acc[k][half][l] = std::fmaf(W[k][l] * c[l], c[l], acc[k][half][l]);
                }
    float sum[4];
    for (int l = 0; l < 4; ++l) {
        const float s01 = (acc[0][0][l] + acc[0][1][l]) + (acc[1][0][l] + acc[1][1][l]);
        const float s23 = (acc[2][0][l] + acc[2][1][l]) + (acc[3][0][l] + acc[3][1][l]);
        sum[l] = s01 + s23;
    }
    if (pairs < w)
        for (int r0 = 0; r0 < h; r0 += 4)
            for (int k = 0; k < 4; ++k) {
                float c[4];
                tile_coef(r0, pairs, k, c);
                for (int l = 0; l < 4; ++l) sum[l] = std::fmaf(W[k][l] * c[l], c[l], sum[l]);
            }
    const float p0 = sum[0] + sum[2], p1 = sum[1] + sum[3];
    return p0 + p1;
}

// RdDispatch::weighted_dist (src/rd_sse.rs:244): the visible vw x vh error zero-padded to
// whole 4x4 tiles, weighted SSE, times the scale, rounded half away from zero.
int64_t orc_weighted_dist(const int32_t* err_vis, int vw, int vh, float scale, int avx2) {
    const int pw = (vw + 3) & ~3, ph = (vh + 3) & ~3;
    if (pw == 0 || ph == 0) return 0;
    int32_t* buf = (int32_t*)std::calloc((size_t)pw * ph, sizeof(int32_t));
    for (int y = 0; y < vh; ++y)
        for (int x = 0; x < vw; ++x) buf[y * pw + x] = err_vis[y * vw + x];
    const float d = avx2 ? orc_weighted_sse_avx2(buf, pw, ph) : orc_weighted_sse_scalar(buf, pw, ph);
    std::free(buf);
    return (int64_t)std::round(d * scale);
}

namespace {
inline void had4(int32_t a, int32_t b, int32_t c, int32_t d, int32_t* o) {
    const int32_t e = a + c, f = a - c, g = b + d, h = b - d;
    o[0] = e + g;
    o[1] = f + h;
    o[2] = f - h;
    o[3] = e - g;
}
uint64_t satd_tiles(const int32_t* e, int w, int h, uint64_t* sad_out) {
    uint64_t satd = 0, sad = 0;
    for (int ty = 0; ty < h; ty += 4)
        for (int tx = 0; tx < w; tx += 4) {
            int32_t rows[4][4];
            for (int r = 0; r < 4; ++r) {
                const int32_t* d = e + (ty + r) * w + tx;
                for (int x = 0; x < 4; ++x) sad += (uint64_t)std::llabs(d[x]);
                had4(d[0], d[1], d[2], d[3], rows[r]);
            }
            for (int x = 0; x < 4; ++x) {
                int32_t col[4];
                had4(rows[0][x], rows[1][x], rows[2][x], rows[3][x], col);
                for (int i = 0; i < 4; ++i) satd += (uint64_t)std::llabs(col[i]);
            }
        }
    if (sad_out) *sad_out = sad;
    return satd;
}
}  // namespace

// RdDispatch::luma_satd (src/rd_sse.rs:635) -> luma_satd_scalar (src/partition_rd.rs:86):
// 4x4 Hadamard L1 of src - clamp(pred + resid); invisible samples of an edge block are set to
// the reconstruction (zero error).  resid may be NULL.
uint64_t orc_luma_satd(const uint16_t* src, int stride, int w, int h, int vw, int vh, int bd,
                       const int32_t* pred, const int32_t* resid) {
    const int maxv = (1 << bd) - 1;
    int32_t* e = (int32_t*)std::malloc((size_t)w * h * sizeof(int32_t));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int i = y * w + x;
            const int32_t rec = clampi(pred[i] + (resid ? resid[i] : 0), 0, maxv);
            e[i] = (x < vw && y < vh) ? (int32_t)src[(size_t)y * stride + x] - rec : 0;
        }
    const uint64_t s = satd_tiles(e, w, h, nullptr);
    std::free(e);
    return s;
}

// RdDispatch::satd_sad_proxy (src/rd_sse.rs:552) -> satd_sad_proxy_scalar (:1143):
// SAD + (SATD >> 2) of src - pred, invisible samples set to the prediction.
uint64_t orc_satd_sad_proxy(const uint16_t* src, int stride, int w, int h, int vw, int vh,
                            const int32_t* pred) {
    int32_t* e = (int32_t*)std::malloc((size_t)w * h * sizeof(int32_t));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int i = y * w + x;
            e[i] = (x < vw && y < vh) ? (int32_t)src[(size_t)y * stride + x] - pred[i] : 0;
        }
    uint64_t sad = 0;
    const uint64_t satd = satd_tiles(e, w, h, &sad);
    std::free(e);
    return sad + (satd >> 2);
}

}  // extern "C"
