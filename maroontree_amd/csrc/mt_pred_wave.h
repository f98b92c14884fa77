// Wave-level intra prediction of one block: the reference encoder's predict_nd_ad /
// dc_pred / filter_predict (src/intrapred.rs:232-415, :473-563, :1357-1711).
//
// One wave (64 lanes) predicts one block.  The two reference edges are assembled in LDS
// with the availability rules of the reference, run through the corner filter, the 5-tap
// edge filter and the 2x upsampler (all lanes, one sample each), then every lane predicts
// pixels lane, lane+64, ... straight from LDS.  Filter intra runs its 4x2 cells along
// anti-diagonals (a cell needs its left, upper and upper-left neighbours), 8 lanes per cell.
//
// Shared by the batched predictor and the superblock reconstruction kernel; each
// translation unit that includes it holds its own copy of the three small tables and
// uploads them with upload_pred_tables() at device initialisation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mtgpu.h"
#include "av1_cdf_tables.h"

namespace {

constexpr int kOrg = 2;    // edge arrays hold logical indices -2 .. 129 (IntraEdge, :422)
constexpr int kEdge = 132;

__constant__ int8_t c_taps[5][8][7];  // INTRA_FILTER_TAPS (src/tables.rs:50)
__constant__ int32_t c_deriv[44];     // DR_INTRA_DERIVATIVE (src/intrapred.rs:1081)
__constant__ int32_t c_smw[128];      // sm_weights(n) at [n, 2n) (src/intrapred.rs:1091)

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// intra_edge_filter_strength (src/intrapred.rs:917)
__device__ int edge_strength(int w, int h, bool smooth_nb, int delta) {
    const int d = abs(delta);
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
__device__ bool edge_upsample(int w, int h, bool smooth_nb, int delta) {
    const int d = abs(delta);
    if (d == 0 || d >= 40) return false;
    return smooth_nb ? (w + h <= 8) : (w + h <= 16);
}

// One filtered sample of filter_intra_edge (src/intrapred.rs:977): output index o reads the
// pre-filter samples o-2 .. o+2, clamped to the `size` samples that start at index -1.
__device__ __forceinline__ int edge_tap(const int32_t* e, int o, int size, int strength) {
    int w0, w1, w2;  // symmetric 5-tap kernels {w0, w1, w2, w1, w0}
    if (strength == 1) { w0 = 0; w1 = 4; w2 = 8; }
    else if (strength == 2) { w0 = 0; w1 = 5; w2 = 6; }
    else { w0 = 2; w1 = 4; w2 = 4; }
    auto s = [&](int j) { return e[kOrg - 1 + clampi(o - 1 + j, 0, size - 1)]; };
    const int sum = w0 * s(0) + w1 * s(1) + w2 * s(2) + w1 * s(3) + w0 * s(4);
    return (sum + 8) >> 4;
}

// Upsampled value at logical index 2i-1 (odd) or the copied sample at 2i
// (upsample_intra_edge, src/intrapred.rs:1026).  `e` holds the pre-upsample edge.
__device__ __forceinline__ int up_in(const int32_t* e, int k, int num) {
    // the reference's `input` array: [e(-1), e(-1), e(0) .. e(num-1), e(num-1)]
    const int idx = k < 2 ? -1 : (k - 2 < num ? k - 2 : num - 1);
    return e[kOrg + idx];
}

// V 90, H 180, D45 45, D135 135, D113 113, D157 157, D203 203, D67 67
__device__ __forceinline__ int nominal_angle(int mode) {
    switch (mode) {
        case MT_PRED_V: return 90;
        case MT_PRED_H: return 180;
        case MT_PRED_D45: return 45;
        case MT_PRED_D135: return 135;
        case MT_PRED_D113: return 113;
        case MT_PRED_D157: return 157;
        case MT_PRED_D203: return 203;
        default: return 67;
    }
}


// edge_tap through an accessor E(i) = logical edge sample i.
template <typename E>
__device__ __forceinline__ int edge_tap_fn(E e, int o, int size, int strength) {
    int w0, w1, w2;
    if (strength == 1) { w0 = 0; w1 = 4; w2 = 8; }
    else if (strength == 2) { w0 = 0; w1 = 5; w2 = 6; }
    else { w0 = 2; w1 = 4; w2 = 4; }
    auto sm = [&](int j) { return e(clampi(o - 1 + j, 0, size - 1) - 1); };
    return (w0 * sm(0) + w1 * sm(1) + w2 * sm(2) + w1 * sm(3) + w0 * sm(4) + 8) >> 4;
}

__device__ __forceinline__ int interp(const int32_t* e, int b, int shift) {
    return (e[kOrg + b] * (32 - shift) + e[kOrg + b + 1] * shift + 16) >> 5;
}

// Predicts block d (its x, y, bw, bh, mode, angle_delta, flags, bd) into dst[bw * bh]
// (row-major; global or LDS).  rc(x, y) reads reconstruction sample (x, y) of a plane of
// frame size pw x ph.  s_a, s_l: LDS edges of kEdge ints; s_fi: LDS of 33 * 33 ints.  Must
// be called by all 64 lanes of the wave; returns with dst written by this wave's lanes
// (the caller synchronises before other lanes read it).
template <class Src>
__device__ void wave_predict(const mt_predblk& d, int pw, int ph, Src rc, int32_t* dst,
                             int32_t* s_a, int32_t* s_l, int32_t* s_fi) {
    const int lane = threadIdx.x & 63;
    const int bw = d.bw, bh = d.bh, bd = d.bd, mode = d.mode, delta = d.angle_delta;
    const int ox = d.x, oy = d.y;
    const bool top = oy > 0, lft = ox > 0;
    const int base = 1 << (bd - 1);
    const int ne = bw + bh;
    // ---- edge assembly (src/intrapred.rs:1376-1459)
    const int have_tr = (top && (d.flags & MT_PRED_HAVE_TR))
                            ? min(min(bw, bh), max(pw - (ox + bw), 0)) : 0;
    const int have_bl = (lft && (d.flags & MT_PRED_HAVE_BL))
                            ? min(min(bw, bh), max(ph - (oy + bh), 0)) : 0;
    for (int i = lane; i < ne; i += 64) {
        s_a[kOrg + i] = top ? (int)rc(ox + min(i, bw + have_tr - 1), oy - 1)
                            : (lft ? (int)rc(ox - 1, oy) : base - 1);
        s_l[kOrg + i] = lft ? (int)rc(ox - 1, oy + min(i, bh + have_bl - 1))
                            : (top ? (int)rc(ox, oy - 1) : base + 1);
    }
    int corner;
    if (lft) corner = top ? rc(ox - 1, oy - 1) : rc(ox - 1, oy);
    else corner = top ? (int)rc(ox, oy - 1) : base;
    if (lane == 0) {
        s_a[kOrg - 1] = corner;
        s_l[kOrg - 1] = corner;
    }
    __syncthreads();

    const bool directional = mode >= MT_PRED_V && mode <= MT_PRED_D67;
    const int angle = directional ? nominal_angle(mode) + 3 * delta : 0;
    int upa = 0, upl = 0;
    if (directional && angle != 90 && angle != 180) {
        const bool snb = (d.flags & MT_PRED_SMOOTH_NB) != 0;
        if (angle > 90 && angle < 180 && ne >= 24) {  // filter_intra_corner (:1014)
            const int f = (5 * s_l[kOrg] + 6 * s_a[kOrg - 1] + 5 * s_a[kOrg] + 8) >> 4;
            __syncthreads();
            if (lane == 0) {
                s_a[kOrg - 1] = f;
                s_l[kOrg - 1] = f;
            }
            __syncthreads();
        }
        const int sa = bw + 1 + (angle < 90 ? bh : 0), sl = bh + 1 + (angle > 180 ? bw : 0);
        const int fa = edge_strength(bw, bh, snb, angle - 90);
        const int fl = edge_strength(bh, bw, snb, angle - 180);
        // every lane computes its outputs from the unfiltered edges, then all write
        int va[2], vl[2];
        for (int k = 0; k < 2; ++k) {
            const int o = lane + 64 * k;
            va[k] = (fa && o + 1 < sa) ? edge_tap(s_a, o, sa, fa) : 0;
            vl[k] = (fl && o + 1 < sl) ? edge_tap(s_l, o, sl, fl) : 0;
        }
        __syncthreads();
        for (int k = 0; k < 2; ++k) {
            const int o = lane + 64 * k;
            if (fa && o + 1 < sa) s_a[kOrg + o] = va[k];
            if (fl && o + 1 < sl) s_l[kOrg + o] = vl[k];
        }
        __syncthreads();
        upa = edge_upsample(bw, bh, snb, angle - 90) ? 1 : 0;
        upl = edge_upsample(bh, bw, snb, angle - 180) ? 1 : 0;
        if (upa | upl) {
            const int na = bw + (angle < 90 ? bh : 0), nl = bh + (angle > 180 ? bw : 0);
            const int mx = (1 << bd) - 1;
            int ua_odd = 0, ua_even = 0, ul_odd = 0, ul_even = 0, ua_m2 = 0, ul_m2 = 0;
            const int i = lane;  // num_px <= 16 whenever upsampling is selected
            if (upa && i < na) {
                const int v = -up_in(s_a, i, na) + 9 * up_in(s_a, i + 1, na) +
                              9 * up_in(s_a, i + 2, na) - up_in(s_a, i + 3, na);
                ua_odd = clampi((v + 8) >> 4, 0, mx);
                ua_even = up_in(s_a, i + 2, na);
            }
            if (upl && i < nl) {
                const int v = -up_in(s_l, i, nl) + 9 * up_in(s_l, i + 1, nl) +
                              9 * up_in(s_l, i + 2, nl) - up_in(s_l, i + 3, nl);
                ul_odd = clampi((v + 8) >> 4, 0, mx);
                ul_even = up_in(s_l, i + 2, nl);
            }
            ua_m2 = s_a[kOrg - 1];
            ul_m2 = s_l[kOrg - 1];
            __syncthreads();
            if (upa && i < na) {
                s_a[kOrg + 2 * i - 1] = ua_odd;
                s_a[kOrg + 2 * i] = ua_even;
            }
            if (upl && i < nl) {
                s_l[kOrg + 2 * i - 1] = ul_odd;
                s_l[kOrg + 2 * i] = ul_even;
            }
            if (lane == 0) {
                if (upa) s_a[kOrg - 2] = ua_m2;
                if (upl) s_l[kOrg - 2] = ul_m2;
            }
            __syncthreads();
        }
    }

    const int lw = __builtin_ctz(bw);
    const int npx = bw * bh;
    if (mode >= MT_PRED_FILTER0) {
        // ---- filter intra (src/intrapred.rs:473-563)
        const int fm = mode - MT_PRED_FILTER0;
        for (int i = lane; i <= bw; i += 64) s_fi[i] = i == 0 ? corner : s_a[kOrg + i - 1];
        for (int i = lane; i < bh; i += 64) s_fi[(i + 1) * 33] = s_l[kOrg + i];
        __syncthreads();
        const int rows = bh >> 1, cols = bw >> 2;
        const int mx = (1 << bd) - 1;
        const int k = lane & 7;
        for (int diag = 0; diag < rows + cols - 1; ++diag) {
            const int c_lo = max(0, diag - rows + 1);
            const int cj = c_lo + (lane >> 3);  // cell column of this lane
            const int ci = diag - cj;           // cell row
            const bool act = cj < cols && ci >= 0 && cj <= diag;
            int v = 0, r = 0, c = 0;
            if (act) {
                r = 2 * ci + 1;
                c = 4 * cj + 1;
                const int p[7] = {s_fi[(r - 1) * 33 + c - 1], s_fi[(r - 1) * 33 + c],
                                  s_fi[(r - 1) * 33 + c + 1], s_fi[(r - 1) * 33 + c + 2],
                                  s_fi[(r - 1) * 33 + c + 3], s_fi[r * 33 + c - 1],
                                  s_fi[(r + 1) * 33 + c - 1]};
                int sum = 0;
                for (int t = 0; t < 7; ++t) sum += (int)c_taps[fm][k][t] * p[t];
                v = clampi((sum + 8) >> 4, 0, mx);
            }
            __syncthreads();
            if (act) s_fi[(r + (k >> 2)) * 33 + c + (k & 3)] = v;
            __syncthreads();
        }
        for (int p = lane; p < npx; p += 64)
            dst[p] = s_fi[((p >> lw) + 1) * 33 + (p & (bw - 1)) + 1];
    } else if (mode == MT_PRED_DC) {
        // ---- dc_pred / dc_pred_from_sum (src/intrapred.rs:1740, :1772)
        int part = 0;
        for (int i = lane; i < 128; i += 64) {
            if (top && i < bw) part += s_a[kOrg + i];
            if (lft && i < bh) part += s_l[kOrg + i];
        }
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
        int dc;
        if (top && lft) {
            dc = (((bw + bh) >> 1) + part) >> __builtin_ctz(bw + bh);
            if (bw != bh) {
                const uint32_t mult = (bw > 2 * bh || bh > 2 * bw) ? 0x3334u : 0x5556u;
                dc = (int)(((uint32_t)dc * mult) >> 16);
            }
        } else if (top) {
            dc = ((bw >> 1) + part) >> lw;
        } else if (lft) {
            dc = ((bh >> 1) + part) >> __builtin_ctz(bh);
        } else {
            dc = base;
        }
        for (int p = lane; p < npx; p += 64) dst[p] = dc;
    } else if (directional) {
        // ---- vertical / horizontal / dr_predict_scalar (src/intrapred.rs:605-705)
        const int zone = angle == 90 ? 4 : angle == 180 ? 5 : angle < 90 ? 1 : angle > 180 ? 3 : 2;
        const int dx = zone == 1 ? c_deriv[angle >> 1] : zone == 2 ? c_deriv[(180 - angle) >> 1] : 0;
        const int dy = zone == 3 ? c_deriv[(270 - angle) >> 1] : zone == 2 ? c_deriv[(angle - 90) >> 1] : 0;
        const int max_base_x = (ne - 1) << upa, max_base_y = (ne - 1) << upl;
        for (int p = lane; p < npx; p += 64) {
            const int y = p >> lw, x = p & (bw - 1);
            int v;
            if (zone == 4) {
                v = s_a[kOrg + x];
            } else if (zone == 5) {
                v = s_l[kOrg + y];
            } else if (zone == 1) {
                const int xpos = dx * (y + 1);
                const int shift = ((xpos << upa) & 0x3f) >> 1;
                const int bx = (xpos >> (6 - upa)) + (x << upa);
                v = bx < max_base_x ? interp(s_a, bx, shift) : s_a[kOrg + max_base_x];
            } else if (zone == 3) {
                const int ypos = dy * (x + 1);
                const int shift = ((ypos << upl) & 0x3f) >> 1;
                const int by = (ypos >> (6 - upl)) + (y << upl);
                v = by < max_base_y ? interp(s_l, by, shift) : s_l[kOrg + max_base_y];
            } else {
                const int t = (y + 1) * dx - 64;
                const int x0 = t <= 0 ? 0 : min((t + 63) >> 6, bw);
                if (x < x0) {
                    const int ypos = (y << 6) - (x + 1) * dy;
                    v = interp(s_l, ypos >> (6 - upl), ((ypos * (1 << upl)) & 0x3f) >> 1);
                } else {
                    const int xpos = (x << 6) - (y + 1) * dx;
                    v = interp(s_a, xpos >> (6 - upa), ((xpos * (1 << upa)) & 0x3f) >> 1);
                }
            }
            dst[p] = v;
        }
    } else {
        // ---- smooth / smooth_v / smooth_h / paeth (src/intrapred.rs:1597-1711)
        const int right = s_a[kOrg + bw - 1], bottom = s_l[kOrg + bh - 1];
        for (int p = lane; p < npx; p += 64) {
            const int y = p >> lw, x = p & (bw - 1);
            const int t = s_a[kOrg + x], l = s_l[kOrg + y];
            const int wy = c_smw[bh + y], wx = c_smw[bw + x];
            int v;
            if (mode == MT_PRED_SMOOTH) {
                v = (wy * t + (256 - wy) * bottom + wx * l + (256 - wx) * right + 256) >> 9;
            } else if (mode == MT_PRED_SMOOTH_V) {
                v = (wy * t + (256 - wy) * bottom + 128) >> 8;
            } else if (mode == MT_PRED_SMOOTH_H) {
                v = (wx * l + (256 - wx) * right + 128) >> 8;
            } else {
                const int bs = l + t - corner;
                const int ld = abs(l - bs), td = abs(t - bs), cd = abs(corner - bs);
                v = (ld <= td && ld <= cd) ? l : (td <= cd ? t : corner);
            }
            dst[p] = v;
        }
    }
}

// Uploads the prediction tables of the including translation unit.
inline hipError_t upload_pred_tables() {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_taps), MT_INTRA_FILTER_TAPS, sizeof(c_taps));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_deriv), MT_DR_INTRA_DERIVATIVE, sizeof(c_deriv));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_smw), MT_SM_WEIGHTS, sizeof(c_smw));
    return e;
}

// ---- chroma from luma (src/intrapred.rs:1112-1260) -----------------------------------------

__device__ __forceinline__ int cfl_pixel(int dc, int ac, int alpha, int maxv) {
    const int diff = alpha * ac;
    const int mag = (abs(diff) + 32) >> 6;
    return clampi(dc + (diff < 0 ? -mag : mag), 0, maxv);
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One CfL block by one wave: cfl_ac_sub's subsampled, scaled luma minus its rounded mean
// (cfl_ac_444 without subsampling) from the luma plane at (lx, ly); unless alpha_given,
// cfl_best_alpha_strided's analytic alpha and its +/-3 energy search over the displayed part
// of the source block at (x, y), then alpha_step steps from that alpha along its sign
// (< 0 toward zero; clipped to -16..16 and never past zero, as cfl_alpha_rd_refine's walk
// visits them, src/cost.rs:332); cfl_pred_pixel over the whole block into dst (and the AC
// into ac_dst, src - prediction into resid_dst, when given).  s_ac: LDS of cw * ch ints.
// Returns the alpha.  The caller has checked that both blocks lie inside their planes.
__device__ int cfl_block(const mt_plane& lp, int lx, int ly, bool sh, bool sv, const mt_plane& sp, int x, int y,
                         int cw, int ch, int bd, int dc, bool alpha_given, int alpha_in, int32_t* s_ac,
                         int32_t* dst, int32_t* ac_dst, int16_t* resid_dst, int alpha_step = 0) {
    const int lane = threadIdx.x & 63;
    const int lcw = __builtin_ctz(cw), npx = cw * ch;
    const int shift = 1 + (sv ? 0 : 1) + (sh ? 0 : 1);
    const uint16_t* L = lp.data + (size_t)ly * lp.stride + lx;
    int64_t part = 0;
    for (int p = lane; p < npx; p += 64) {
        const int yy = p >> lcw, xx = p & (cw - 1);
        const uint16_t* r0 = L + (size_t)(yy << (sv ? 1 : 0)) * lp.stride + (xx << (sh ? 1 : 0));
        int s = r0[0];
        if (sh) s += r0[1];
        if (sv) {
            s += r0[lp.stride];
            if (sh) s += r0[lp.stride + 1];
        }
        s <<= shift;
        s_ac[p] = s;
        part += s;
    }
    const int log2sz = lcw + __builtin_ctz(ch);
    const int64_t sum = wave_sum(part) + ((int64_t(1) << log2sz) >> 1);
    const int mean = (int)(sum >> log2sz);
    __syncthreads();
    for (int p = lane; p < npx; p += 64) s_ac[p] -= mean;
    __syncthreads();
    const int maxv = (1 << bd) - 1;
    const int vw = min(max(sp.vis_w - x, 0), cw), vh = min(max(sp.vis_h - y, 0), ch);
    const uint16_t* S = sp.data + (size_t)y * sp.stride + x;
    int alpha = alpha_in;
    if (!alpha_given) {
        int64_t num = 0, den = 0;
        for (int p = lane; p < npx; p += 64) {
            const int yy = p >> lcw, xx = p & (cw - 1);
            if (xx < vw && yy < vh) {
                const int64_t a = s_ac[p];
                num += (int64_t)((int)S[(size_t)yy * sp.stride + xx] - dc) * a;
                den += a * a;
            }
        }
        num = wave_sum(num);
        den = wave_sum(den);
        if (den == 0 || vw == 0 || vh == 0) {
            alpha = 0;
        } else {
            const int64_t sg = num > 0 ? 1 : (num < 0 ? -1 : 0);
            int64_t a0 = (64 * num + (den >> 1) * sg) / den;
            a0 = a0 < -16 ? -16 : (a0 > 16 ? 16 : a0);
            int64_t e[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int p = lane; p < npx; p += 64) {
                const int yy = p >> lcw, xx = p & (cw - 1);
                if (xx < vw && yy < vh) {
                    const int s = S[(size_t)yy * sp.stride + xx];
                    for (int k = 0; k < 7; ++k) {
                        const int64_t dd = s - cfl_pixel(dc, s_ac[p], (int)a0 - 3 + k, maxv);
                        e[k] += dd * dd;
                    }
                }
            }
            int64_t best_e = INT64_MAX;
            alpha = 0;
            for (int k = 0; k < 7; ++k) {
                const int c = (int)a0 - 3 + k;
                const int64_t ek = wave_sum(e[k]);
                if (c < -16 || c > 16) continue;
                if (ek < best_e) {
                    best_e = ek;
                    alpha = c;
                }
            }
        }
        if (alpha_step != 0) {
            const int dir = alpha > 0 ? 1 : (alpha < 0 ? -1 : 0);
            int a = min(max(alpha + dir * alpha_step, -16), 16);
            if (a * dir < 0) a = 0;
            alpha = a;
        }
    }
    for (int p = lane; p < npx; p += 64) {
        const int v = cfl_pixel(dc, s_ac[p], alpha, maxv);
        dst[p] = v;
        if (ac_dst) ac_dst[p] = s_ac[p];
        if (resid_dst) resid_dst[p] = (int16_t)((int)S[(size_t)(p >> lcw) * sp.stride + (p & (cw - 1))] - v);
    }
    __syncthreads();
    return alpha;
}

}  // namespace
