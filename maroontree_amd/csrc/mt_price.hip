// Decision pricing helpers on the GPU: the SATD partition price and the proxy units
// (src/partition_rd.rs:30-74, src/coder/lossy_state.rs:1829, :2385) and the perceptual
// RD scale with the activity measures it reads (src/coder/lossy_state.rs:4189,
// src/coder/aq.rs:140-174).
//
// All of it is f32 arithmetic in the reference's order.  f_fmlaf (src/aq_common.rs:66) is
// unfused on the reference's default x86-64 build (no `fma` target feature), so it is a
// product and a sum here, each rounded (-ffp-contract=off keeps the compiler from fusing).
// The moments are exact integer sums, so the lanes may add them in any order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "mt_common.h"

namespace {

constexpr float kPartLambda = 2.775f;  // K_PART_LAMBDA (src/cost.rs:278)

// dirty_exp2f's table: 2^((i - 32) / 64) rounded to f32, built on the host from its
// definition and passed by value.
struct Exp2Tab {
    float v[64];
};

__host__ __device__ inline float fmlaf(float a, float b, float c) { return a * b + c; }
__host__ __device__ inline float clampf(float x, float lo, float hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

// dirty_log1pf (src/aq_common.rs:90)
__device__ inline float dirty_log1pf(float d) {
    const uint32_t ix = __float_as_uint(1.0f + d);
    const int n = (int)((ix & 0x7f800000u) >> 23) - 0x7f;
    const float mantissa = __uint_as_float((ix & 0x007fffffu) | 0x3f800000u);
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

// dirty_exp2f (src/coder/lossy_state.rs:4284): table value times a quadratic of the
// remainder, the exponent put in with integer arithmetic.
__device__ inline float dirty_exp2f(float d, const Exp2Tab& tab) {
    const float redux = __uint_as_float(0x4b400000u) / 64.0f;
    const uint32_t ui = __float_as_uint(d + redux);
    uint32_t i0 = ui + 32u;
    const uint32_t k = i0 / 64u;
    i0 &= 63u;
    const float uf = __uint_as_float(ui) - redux;
    const float z0 = tab.v[i0];
    const float f = d - uf;
    float u = 0.24022668600082397f;
    u = fmlaf(u, f, 0.693149745464325f);
    u *= f;
    const float i2 = __uint_as_float((k + 0x7fu) << 23);
    return fmlaf(u, z0, z0) * i2;
}

// the variance's activity: log1p of the 8-bit-scaled variance, from exact moments
__device__ inline float moments_activity(int64_t sum, int64_t sq, int count, int bd) {
    const float n = (float)count;
    const float mean = (float)sum / n;
    float var = (float)sq / n - mean * mean;
    if (!(var >= 0.0f)) var = 0.0f;  // f32::max(0.0)
    const int b = bd < 8 ? 8 : bd;
    return dirty_log1pf(var / (float)(1u << (2 * (b - 8))));
}

// Sum and sum of squares of a w x h window, one wave: each lane walks columns.
__device__ inline void wave_moments(const mt_plane& p, int px, int py, int w, int h, int64_t& sum,
                                    int64_t& sq) {
    const int lane = threadIdx.x & 63;
    int64_t s = 0, q = 0;
    for (int y = 0; y < h; y++) {
        const uint16_t* row = p.data + (size_t)(py + y) * p.stride + px;
        for (int x = lane; x < w; x += 64) {
            const int64_t v = row[x];
            s += v;
            q += v * v;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        q += __shfl_xor(q, o);
    }
    sum = s;
    sq = q;
}

// sb_activity (src/coder/aq.rs:140) of every superblock, one wave each, raster order.
__global__ void __launch_bounds__(64) sb_act_kernel(mt_plane p, int bd, int sbw, int nsb, float* act) {
    const int sb = blockIdx.x;
    if (sb >= nsb) return;
    const int sb_x = (sb % sbw) * 64, sb_y = (sb / sbw) * 64;
    const int h = min(64, p.height - sb_y), w = min(64, p.width - sb_x);
    int64_t sum, sq;
    wave_moments(p, sb_x, sb_y, w, h, sum, sq);
    if (threadIdx.x == 0) act[sb] = moments_activity(sum, sq, h * w, bd);
}

// tile_ref_activity (src/coder/aq.rs:163): the f32 sum in raster order, one lane.
__global__ void ref_act_kernel(const float* act, int nsb, float* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float sum = 0.0f, cnt = 0.0f;
    for (int i = 0; i < nsb; i++) {
        sum += act[i];
        cnt += 1.0f;
    }
    *out = cnt > 0.0f ? sum / cnt : 5.0f;
}

struct PrdoArgs {
    int bd, base_q, fmt, mono_large, sbw, n;
};

// perceptual_rd_scale (src/coder/lossy_state.rs:4189) at the shipped tuning: prdo_k 0.4,
// prdo_k_444 0.5 (src/tuning.rs:827-828), top_ease 1 with knee 20 and width 12 (:750-752),
// local_ref_blend 0.5 (:863), the clamp [1/2, 1.5] (src/coder.rs:382-388).  One wave per
// block; the superblock's activity comes from sb_act_kernel.
__global__ void __launch_bounds__(64) prdo_kernel(const mt_prdoblk* blocks, mt_plane p, PrdoArgs a,
                                                  const float* sb_act, const float* ref_act_d,
                                                  Exp2Tab tab, float* out) {
    const int i = blockIdx.x;
    if (i >= a.n) return;
    const mt_prdoblk b = blocks[i];
    const int px = b.x, py = b.y, dim = b.dim;
    const bool is444 = a.fmt == MT_ENC_YUV444, is422 = a.fmt == MT_ENC_YUV422,
               is420 = a.fmt == MT_ENC_YUV420, mono = a.fmt == MT_ENC_MONOCHROME;
    float k = is444 ? 0.5f : 0.4f;
    bool one = false;
    if (is422 || a.mono_large) {
        const float te = clampf((20.0f - (float)a.base_q) / 12.0f, 0.0f, 1.0f);
        k *= fmlaf(-1.0f, te, 1.0f);
        one = one || k == 0.0f;
    }
    if (!one && !is422 && !mono) {
        const float fade = 1.0f / (100.0f - 50.0f);
        float t = clampf(((float)a.base_q - 50.0f) * fade, 0.0f, 1.0f);
        const float floor = is420 ? 0.4f : 0.0f;
        t = fmlaf(1.0f - floor, t, floor);
        k *= t;
        one = k == 0.0f;
    }
    // a block outside the plane has an empty superblock window (activity 0) and no samples:
    // the reference returns 1 from the one test or the other
    const int bw = min(dim, max(p.width - px, 0)), bh = min(dim, max(p.height - py, 0));
    one = one || bw == 0 || bh == 0;
    float refa = 0.0f;
    if (!one) {
        const float g = ref_act_d ? *ref_act_d : 0.0f;
        const float sa = sb_act[(py >> 6) * a.sbw + (px >> 6)];
        refa = g * (1.0f - 0.5f) + sa * 0.5f;
        one = refa <= 0.0f;
    }
    if (one) {  // uniform over the wave: every lane leaves together
        if (threadIdx.x == 0) out[i] = 1.0f;
        return;
    }
    int64_t sum, sq;
    wave_moments(p, px, py, bw, bh, sum, sq);
    if (threadIdx.x != 0) return;
    const float act = moments_activity(sum, sq, bw * bh, a.bd);
    const float ln2 = 0.693147180559945309f, log2e = 1.44269504088896341f;
    const float exponent = clampf(k * (act - refa), -ln2, ln2);
    out[i] = clampf(dirty_exp2f(exponent * log2e, tab), 1.0f / 2.0f, 1.5f);
}

// luma_satd_scale (src/partition_rd.rs:30) at the shipped satd_base 1.25, satd_top 2.921556,
// knees 20 / 48 (src/tuning.rs:854-857)
__host__ __device__ inline float satd_scale(int base_q, int subsampled) {
    const float base = 1.25f, top = 2.921556f, lo = 20.0f, hi = 48.0f;
    if (subsampled) return base;
    const float qi = (float)base_q;
    if (qi <= lo) return top;
    if (qi >= hi) return base;
    const float d = hi - lo;
    return top + (base - top) * (qi - lo) / (d > 1.0f ? d : 1.0f);
}

// k_satd (src/partition_rd.rs:54)
__host__ __device__ inline float k_satd(int base_q, int subsampled) {
    const float q0 = 28.0f, q1 = 81.0f, q2 = 140.0f;
    const float k0 = subsampled ? 24.5f : 35.0f, k1 = subsampled ? 21.3f : 21.2f,
                k2 = subsampled ? 18.2f : 18.1f;
    const float qi = (float)base_q;
    if (qi <= q0) return k0;
    if (qi >= q2) return k2;
    if (qi <= q1) return k0 + (k1 - k0) * (qi - q0) / (q1 - q0);
    return k1 + (k2 - k1) * (qi - q1) / (q2 - q1);
}

// luma_partition_distortion (src/coder/lossy_state.rs:2385) with the shipped
// proxy_luma_bridge on and proxy_luma_scale 7.130643 (src/tuning.rs:783-784)
__host__ __device__ inline float satd_price(uint64_t satd, float qstep, int base_q, int subsampled) {
    float price = (float)satd * (qstep >= 1.0f ? qstep : 1.0f) * 0.25f * satd_scale(base_q, subsampled);
    price *= kPartLambda / k_satd(base_q, subsampled) * 7.130643f;
    return price;
}

__global__ void satd_price_kernel(const uint64_t* satd, const float* qstep_d, float qstep, int n, int base_q,
                                  int subsampled, float* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = satd_price(satd[i], qstep_d ? qstep_d[i] : qstep, base_q, subsampled);
}

const Exp2Tab& exp2_tab() {
    static Exp2Tab t = [] {
        Exp2Tab r;
        for (int i = 0; i < 64; i++) r.v[i] = (float)std::exp2l((long double)(i - 32) / 64.0L);
        return r;
    }();
    return t;
}

int check_plane(const mt_plane& p, const char* who) {
    if (!p.data || p.width <= 0 || p.height <= 0 || p.stride < p.width || p.width > 65535 ||
        p.height > 65535)
        return mtlib::set_error(MT_ERR_ARG, std::string(who) + ": bad source plane");
    return MT_OK;
}

}  // namespace

extern "C" float mt_luma_satd_scale(int base_q_idx, int subsampled) {
    return satd_scale(base_q_idx, subsampled);
}
extern "C" float mt_k_satd(int base_q_idx, int subsampled) { return k_satd(base_q_idx, subsampled); }
extern "C" float mt_proxy_wsse(float d) { return d * kPartLambda; }

extern "C" int mt_luma_partition_distortion(const uint64_t* satd_d, const float* qstep_d, float qstep, int n,
                                            int base_q_idx, int subsampled, float* price_d, void* stream) {
    int dev, cus;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!satd_d || !price_d)) || base_q_idx < 0 || base_q_idx > 255)
        return mtlib::set_error(MT_ERR_ARG, "mt_luma_partition_distortion: bad argument");
    if (n == 0) return MT_OK;
    const int grid = std::min((n + 255) / 256, cus * 8);
    hipLaunchKernelGGL(satd_price_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, satd_d, qstep_d,
                       qstep, n, base_q_idx, subsampled ? 1 : 0, price_d);
    if (hipGetLastError() != hipSuccess)
        return mtlib::set_error(MT_ERR_DEVICE, "mt_luma_partition_distortion: launch failed");
    return MT_OK;
}

extern "C" int mt_sb_activity(mt_plane src, int bd, float* act_d, float* ref_act_d, void* stream) {
    int dev, cus;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc) return rc;
    if ((rc = check_plane(src, "mt_sb_activity"))) return rc;
    if (!act_d || (bd != 8 && bd != 10 && bd != 12))
        return mtlib::set_error(MT_ERR_ARG, "mt_sb_activity: bad output or bit depth");
    const int sbw = (src.width + 63) / 64, nsb = sbw * ((src.height + 63) / 64);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(sb_act_kernel, dim3(nsb), dim3(64), 0, s, src, bd, sbw, nsb, act_d);
    if (ref_act_d) hipLaunchKernelGGL(ref_act_kernel, dim3(1), dim3(64), 0, s, act_d, nsb, ref_act_d);
    if (hipGetLastError() != hipSuccess) return mtlib::set_error(MT_ERR_DEVICE, "mt_sb_activity: launch failed");
    return MT_OK;
}

extern "C" int mt_perceptual_rd_scale(const mt_prdoblk* blocks_d, int n, mt_plane src, int bd, int base_q_idx,
                                      int format, int mono_large, const float* ref_act_d, const float* sb_act_d,
                                      float* scale_d, void* stream) {
    int dev, cus;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc) return rc;
    if ((rc = check_plane(src, "mt_perceptual_rd_scale"))) return rc;
    if (n < 0 || (n > 0 && (!blocks_d || !sb_act_d || !scale_d)) || (bd != 8 && bd != 10 && bd != 12) ||
        base_q_idx < 0 || base_q_idx > 255 || format < 0 || format > 3)
        return mtlib::set_error(MT_ERR_ARG, "mt_perceptual_rd_scale: bad argument");
    if (n == 0) return MT_OK;
    PrdoArgs a{bd, base_q_idx, format, mono_large ? 1 : 0, (src.width + 63) / 64, n};
    hipLaunchKernelGGL(prdo_kernel, dim3(n), dim3(64), 0, (hipStream_t)stream, blocks_d, src, a, sb_act_d,
                       ref_act_d, exp2_tab(), scale_d);
    if (hipGetLastError() != hipSuccess)
        return mtlib::set_error(MT_ERR_DEVICE, "mt_perceptual_rd_scale: launch failed");
    return MT_OK;
}
