// Intra prediction on the GPU: the reference encoder's predict_nd_ad / dc_pred /
// filter_predict (src/intrapred.rs:232-415, :473-563, :1357-1711) over a batch of blocks.
//
// Blocks larger than 64 samples: one wave (64 lanes) per block, persistent over the batch,
// running wave_predict (mt_pred_wave.h) and writing rows with coalesced stores.  Blocks of
// at most 64 samples: one thread per block (pred_small_kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/mtgpu.h"
#include "av1_cdf_tables.h"
#include "mt_common.h"
#include "mt_lists.h"
#include "mt_pred_wave.h"

namespace {

using mtlib::PlaneSet;

// Whether a descriptor is one the predictors accept (every sample it reads lies in its plane).
__device__ bool pred_valid(const mt_predblk& d, const PlaneSet& planes) {
    const int bw = d.bw, bh = d.bh, bd = d.bd, mode = d.mode, delta = d.angle_delta;
    const bool pow_ok = (bw == 4 || bw == 8 || bw == 16 || bw == 32 || bw == 64) &&
                        (bh == 4 || bh == 8 || bh == 16 || bh == 32 || bh == 64);
    bool ok = pow_ok && (bd == 8 || bd == 10 || bd == 12) && d.plane < planes.n &&
              delta >= -3 && delta <= 3 &&
              (mode <= MT_PRED_PAETH ||
               (mode >= MT_PRED_FILTER0 && mode <= MT_PRED_FILTER4 && bw <= 32 && bh <= 32));
    const mt_plane& pl = planes.p[ok ? d.plane : 0];
    return ok && (int)d.x + bw <= pl.width && (int)d.y + bh <= pl.height;
}

// Writes status, and sorts the valid blocks into one list per shape class.
__global__ void __launch_bounds__(256) pred_classify_kernel(const mt_predblk* __restrict__ blocks, int n,
                                                            PlaneSet planes, uint32_t* __restrict__ lists,
                                                            size_t cap, uint32_t* __restrict__ cnt,
                                                            int32_t* __restrict__ status) {
    auto cls = [&](int b, bool first) -> int {
        const mt_predblk d = blocks[b];
        const bool ok = pred_valid(d, planes);
        if (first && status) status[b] = ok ? MT_OK : MT_ERR_ARG;
        return ok ? mtlib::small_shape_class(d.bw, d.bh) : -1;
    };
    mtlib::classify_chunk<mtlib::kSmallShapes + 1>(n, cls, lists, cap, cnt);
}

// resid (optional): src - pred as int16, the source plane index in the descriptor's pad byte.
__global__ __launch_bounds__(64) void intra_pred_kernel(const mt_predblk* __restrict__ blocks,
                                                        const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count,
                                                        PlaneSet planes,
                                                        int32_t* __restrict__ out,
                                                        int16_t* __restrict__ resid) {
    __shared__ int32_t s_a[kEdge], s_l[kEdge];
    __shared__ int32_t s_fi[33 * 33];
    const int lane = threadIdx.x;
    const int nlist = (int)*count;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const uint32_t b = list[li];
        const mt_predblk d = blocks[b];
        const mt_plane pl = planes.p[d.plane];
        const uint16_t* rcp = pl.data;
        const int stride = pl.stride;
        int32_t* dst = out + d.out_offset;
        wave_predict(d, pl.width, pl.height,
                     [=](int x, int y) { return (int)rcp[(size_t)y * stride + x]; }, dst, s_a, s_l, s_fi);
        if (resid) {
            // residual_pred (src/rd_sse.rs:1179) of the prediction this wave just wrote
            const int bw = d.bw, npx = d.bw * d.bh, lw = __builtin_ctz(d.bw);
            const int ox = d.x, oy = d.y, lane = threadIdx.x;
            __threadfence_block();
            const mt_plane& sp = planes.p[d.pad];
            const uint16_t* src = sp.data + (size_t)oy * sp.stride + ox;
            for (int p = lane; p < npx; p += 64)
                resid[d.out_offset + p] =
                    (int16_t)((int)src[(size_t)(p >> lw) * sp.stride + (p & (bw - 1))] - dst[p]);
        }
        __syncthreads();  // the next block reuses the LDS edges
    }
}


// ---- blocks of at most 64 samples: one thread per block ------------------------------------
// Each thread keeps its block's two edges in its own column of LDS (sample i of thread t at
// [i][t]: a wave touching one sample index hits 64 consecutive halfwords), filters and
// upsamples them through static-size register copies, and predicts the block sample by
// sample, storing rows four samples at a time.

template <int W, int H>
__global__ void __launch_bounds__(256) pred_small_kernel(const mt_predblk* __restrict__ blocks,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ count,
                                                         PlaneSet planes, int32_t* __restrict__ out,
                                                         int16_t* __restrict__ resid) {
    constexpr int NE = W + H;           // edge samples 0 .. NE-1
    constexpr int CAP = 2 * NE + 2;     // logical -2 .. 2*NE-1 (an upsampled edge doubles)
    __shared__ uint16_t s_a[CAP][256], s_l[CAP][256];
    const int t = threadIdx.x;
    auto A = [&](int i) -> int { return s_a[i + 2][t]; };
    auto L = [&](int i) -> int { return s_l[i + 2][t]; };
    auto setA = [&](int i, int v) { s_a[i + 2][t] = (uint16_t)v; };
    auto setL = [&](int i, int v) { s_l[i + 2][t] = (uint16_t)v; };
    const int n = (int)*count;
    for (int li = blockIdx.x * 256 + t; li < n; li += gridDim.x * 256) {
        const uint32_t b = list[li];
        const mt_predblk d = blocks[b];
        const mt_plane pl = planes.p[d.plane];
        const uint16_t* rc = pl.data;
        const int stride = pl.stride, ox = d.x, oy = d.y, bd = d.bd, mode = d.mode;
        const bool top = oy > 0, lft = ox > 0;
        const int base = 1 << (bd - 1), mx = (1 << bd) - 1;
        // ---- edge assembly (src/intrapred.rs:1376-1459)
        const int have_tr = (top && (d.flags & MT_PRED_HAVE_TR)) ? min(min(W, H), max(pl.width - (ox + W), 0)) : 0;
        const int have_bl = (lft && (d.flags & MT_PRED_HAVE_BL)) ? min(min(W, H), max(pl.height - (oy + H), 0)) : 0;
        const int afill = lft ? (int)rc[(size_t)oy * stride + ox - 1] : base - 1;
        const int lfill = top ? (int)rc[(size_t)(oy - 1) * stride + ox] : base + 1;
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            setA(i, top ? (int)rc[(size_t)(oy - 1) * stride + ox + min(i, W + have_tr - 1)] : afill);
            setL(i, lft ? (int)rc[(size_t)(oy + min(i, H + have_bl - 1)) * stride + ox - 1] : lfill);
        }
        const int corner = lft ? (top ? (int)rc[(size_t)(oy - 1) * stride + ox - 1] : afill)
                               : (top ? lfill : base);
        setA(-1, corner);
        setL(-1, corner);
        const bool directional = mode >= MT_PRED_V && mode <= MT_PRED_D67;
        const int angle = directional ? nominal_angle(mode) + 3 * d.angle_delta : 0;
        int upa = 0, upl = 0;
        if (directional && angle != 90 && angle != 180) {
            // (no corner filter: it needs w + h >= 24)
            const bool snb = (d.flags & MT_PRED_SMOOTH_NB) != 0;
            const int sa = W + 1 + (angle < 90 ? H : 0), sl = H + 1 + (angle > 180 ? W : 0);
            const int fa = edge_strength(W, H, snb, angle - 90);
            const int fl = edge_strength(H, W, snb, angle - 180);
            if (fa | fl) {
                int na[NE], nl[NE];
#pragma unroll
                for (int o = 0; o < NE; ++o) {
                    na[o] = (fa && o + 1 < sa) ? edge_tap_fn(A, o, sa, fa) : 0;
                    nl[o] = (fl && o + 1 < sl) ? edge_tap_fn(L, o, sl, fl) : 0;
                }
#pragma unroll
                for (int o = 0; o < NE; ++o) {
                    if (fa && o + 1 < sa) setA(o, na[o]);
                    if (fl && o + 1 < sl) setL(o, nl[o]);
                }
            }
            upa = edge_upsample(W, H, snb, angle - 90) ? 1 : 0;
            upl = edge_upsample(H, W, snb, angle - 180) ? 1 : 0;
            if (NE <= 16 && (upa | upl)) {
                constexpr int U = NE <= 16 ? NE : 1;
                const int numa = W + (angle < 90 ? H : 0), numl = H + (angle > 180 ? W : 0);
                int ia[U + 3], il[U + 3];  // the upsampler's input arrays
#pragma unroll
                for (int k = 0; k < U + 3; ++k) {
                    ia[k] = A(k < 2 ? -1 : (k - 2 < numa ? k - 2 : numa - 1));
                    il[k] = L(k < 2 ? -1 : (k - 2 < numl ? k - 2 : numl - 1));
                }
                if (upa) {
                    setA(-2, ia[0]);
#pragma unroll
                    for (int i = 0; i < U; ++i)
                        if (i < numa) {
                            const int v = -ia[i] + 9 * ia[i + 1] + 9 * ia[i + 2] - ia[i + 3];
                            setA(2 * i - 1, clampi((v + 8) >> 4, 0, mx));
                            setA(2 * i, ia[i + 2]);
                        }
                }
                if (upl) {
                    setL(-2, il[0]);
#pragma unroll
                    for (int i = 0; i < U; ++i)
                        if (i < numl) {
                            const int v = -il[i] + 9 * il[i + 1] + 9 * il[i + 2] - il[i + 3];
                            setL(2 * i - 1, clampi((v + 8) >> 4, 0, mx));
                            setL(2 * i, il[i + 2]);
                        }
                }
            }
        }
        int32_t* dst = out + d.out_offset;
        const bool aligned = (d.out_offset & 3) == 0;
        int16_t* rdst = resid ? resid + d.out_offset : nullptr;
        const uint16_t* src = nullptr;
        int sstride = 0;
        if (resid) {
            const mt_plane& sp = planes.p[d.pad];
            src = sp.data + (size_t)oy * sp.stride + ox;
            sstride = sp.stride;
        }
        auto store4 = [&](int y, int x0, int v0, int v1, int v2, int v3) {
            int32_t* o = dst + y * W + x0;
            if (aligned) {
                *reinterpret_cast<int4*>(o) = make_int4(v0, v1, v2, v3);
            } else {
                o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v3;
            }
            if (rdst) {
                const uint16_t* sr = src + (size_t)y * sstride + x0;
                int16_t* r = rdst + y * W + x0;
                r[0] = (int16_t)(sr[0] - v0);
                r[1] = (int16_t)(sr[1] - v1);
                r[2] = (int16_t)(sr[2] - v2);
                r[3] = (int16_t)(sr[3] - v3);
            }
        };
        if (mode >= MT_PRED_FILTER0) {
            // ---- filter intra (src/intrapred.rs:473-563), cell rows top to bottom
            const int fm = mode - MT_PRED_FILTER0;
            int prow[W + 1];
            prow[0] = corner;
#pragma unroll
            for (int x = 0; x < W; ++x) prow[x + 1] = A(x);
#pragma unroll
            for (int r = 0; r < H / 2; ++r) {
                int ra[W + 1], rb[W + 1];
                ra[0] = L(2 * r);
                rb[0] = L(2 * r + 1);
#pragma unroll
                for (int c = 0; c < W / 4; ++c) {
                    const int p[7] = {prow[4 * c], prow[4 * c + 1], prow[4 * c + 2], prow[4 * c + 3],
                                      prow[4 * c + 4], ra[4 * c], rb[4 * c]};
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        int sum = 0;
#pragma unroll
                        for (int q = 0; q < 7; ++q) sum += (int)c_taps[fm][k][q] * p[q];
                        const int v = clampi((sum + 8) >> 4, 0, mx);
                        if (k < 4) ra[4 * c + 1 + k] = v;
                        else rb[4 * c + 1 + (k - 4)] = v;
                    }
                }
#pragma unroll
                for (int x = 0; x < W; x += 4) {
                    store4(2 * r, x, ra[x + 1], ra[x + 2], ra[x + 3], ra[x + 4]);
                    store4(2 * r + 1, x, rb[x + 1], rb[x + 2], rb[x + 3], rb[x + 4]);
                }
#pragma unroll
                for (int x = 0; x <= W; ++x) prow[x] = rb[x];
            }
        } else if (mode == MT_PRED_DC) {
            int sum = 0;
            if (top)
#pragma unroll
                for (int i = 0; i < W; ++i) sum += A(i);
            if (lft)
#pragma unroll
                for (int i = 0; i < H; ++i) sum += L(i);
            int dc;
            if (top && lft) {
                dc = (((W + H) >> 1) + sum) >> __builtin_ctz(W + H);
                if (W != H) {
                    const uint32_t mult = (W > 2 * H || H > 2 * W) ? 0x3334u : 0x5556u;
                    dc = (int)(((uint32_t)dc * mult) >> 16);
                }
            } else if (top) {
                dc = ((W >> 1) + sum) >> __builtin_ctz(W);
            } else if (lft) {
                dc = ((H >> 1) + sum) >> __builtin_ctz(H);
            } else {
                dc = base;
            }
#pragma unroll
            for (int y = 0; y < H; ++y)
#pragma unroll
                for (int x = 0; x < W; x += 4) store4(y, x, dc, dc, dc, dc);
        } else if (directional) {
            const int zone = angle == 90 ? 4 : angle == 180 ? 5 : angle < 90 ? 1 : angle > 180 ? 3 : 2;
            const int dx = zone == 1 ? c_deriv[angle >> 1] : zone == 2 ? c_deriv[(180 - angle) >> 1] : 0;
            const int dy = zone == 3 ? c_deriv[(270 - angle) >> 1] : zone == 2 ? c_deriv[(angle - 90) >> 1] : 0;
            const int mbx = (NE - 1) << upa, mby = (NE - 1) << upl;
            auto ip = [&](auto E, int bb, int sh) { return (E(bb) * (32 - sh) + E(bb + 1) * sh + 16) >> 5; };
            auto px = [&](int y, int x) -> int {
                if (zone == 4) return A(x);
                if (zone == 5) return L(y);
                if (zone == 1) {
                    const int xpos = dx * (y + 1);
                    const int bx = (xpos >> (6 - upa)) + (x << upa);
                    return bx < mbx ? ip(A, bx, ((xpos << upa) & 0x3f) >> 1) : A(mbx);
                }
                if (zone == 3) {
                    const int ypos = dy * (x + 1);
                    const int by = (ypos >> (6 - upl)) + (y << upl);
                    return by < mby ? ip(L, by, ((ypos << upl) & 0x3f) >> 1) : L(mby);
                }
                const int tt = (y + 1) * dx - 64;
                const int x0 = tt <= 0 ? 0 : min((tt + 63) >> 6, W);
                if (x < x0) {
                    const int ypos = (y << 6) - (x + 1) * dy;
                    return ip(L, ypos >> (6 - upl), ((ypos * (1 << upl)) & 0x3f) >> 1);
                }
                const int xpos = (x << 6) - (y + 1) * dx;
                return ip(A, xpos >> (6 - upa), ((xpos * (1 << upa)) & 0x3f) >> 1);
            };
#pragma unroll
            for (int y = 0; y < H; ++y)
#pragma unroll
                for (int x = 0; x < W; x += 4) store4(y, x, px(y, x), px(y, x + 1), px(y, x + 2), px(y, x + 3));
        } else {
            const int right = A(W - 1), bottom = L(H - 1);
            auto px = [&](int y, int x) -> int {
                const int tv = A(x), lv = L(y);
                const int wy = c_smw[H + y], wx = c_smw[W + x];
                if (mode == MT_PRED_SMOOTH)
                    return (wy * tv + (256 - wy) * bottom + wx * lv + (256 - wx) * right + 256) >> 9;
                if (mode == MT_PRED_SMOOTH_V) return (wy * tv + (256 - wy) * bottom + 128) >> 8;
                if (mode == MT_PRED_SMOOTH_H) return (wx * lv + (256 - wx) * right + 128) >> 8;
                const int bs = lv + tv - corner;
                const int ld = abs(lv - bs), td = abs(tv - bs), cd = abs(corner - bs);
                return (ld <= td && ld <= cd) ? lv : (td <= cd ? tv : corner);
            };
#pragma unroll
            for (int y = 0; y < H; ++y)
#pragma unroll
                for (int x = 0; x < W; x += 4) store4(y, x, px(y, x), px(y, x + 1), px(y, x + 2), px(y, x + 3));
        }
    }
}

// One wave per block of mt_cfl.
__global__ __launch_bounds__(64) void cfl_kernel(const mt_cflblk* __restrict__ blocks, int n,
                                                 PlaneSet planes, int32_t* __restrict__ pred,
                                                 int32_t* __restrict__ ac_out,
                                                 int32_t* __restrict__ alpha_out) {
    __shared__ int32_t s_ac[32 * 32];
    const int lane = threadIdx.x;
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const mt_cflblk d = blocks[b];
        const int cw = d.cw, ch = d.ch, bd = d.bd;
        const bool sh = d.flags & MT_CFL_SS_HOR, sv = d.flags & MT_CFL_SS_VER;
        bool ok = (cw == 4 || cw == 8 || cw == 16 || cw == 32) &&
                  (ch == 4 || ch == 8 || ch == 16 || ch == 32) && (bd == 8 || bd == 10 || bd == 12) &&
                  d.luma_plane < planes.n && d.src_plane < planes.n;
        const mt_plane lp = planes.p[ok ? d.luma_plane : 0];
        const mt_plane sp = planes.p[ok ? d.src_plane : 0];
        ok = ok && (int)d.lx + (cw << (sh ? 1 : 0)) <= lp.width && (int)d.ly + (ch << (sv ? 1 : 0)) <= lp.height &&
             (int)d.x + cw <= sp.width && (int)d.y + ch <= sp.height;
        if (!ok) {
            if (lane == 0) alpha_out[b] = -128;
            continue;
        }
        const int alpha = cfl_block(lp, d.lx, d.ly, sh, sv, sp, d.x, d.y, cw, ch, bd, d.dc,
                                    (d.flags & MT_CFL_ALPHA_GIVEN) != 0, d.alpha, s_ac, pred + d.out_offset,
                                    ac_out ? ac_out + d.out_offset : nullptr, nullptr);
        if (lane == 0) alpha_out[b] = alpha;
    }
}

// The CfL candidates of a candidate evaluation (indices list[0 .. *count)): their DC
// prediction is already in the prediction pool (its first sample is the DC value); the CfL
// prediction and src - prediction replace the pool's prediction and residual.
__global__ __launch_bounds__(64) void cfl_eval_kernel(const mt_cand* __restrict__ cands,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count, PlaneSet planes,
                                                      int32_t* __restrict__ pred, int16_t* __restrict__ resid,
                                                      int32_t* __restrict__ alpha_out) {
    __shared__ int32_t s_ac[32 * 32];
    const uint32_t n = *count;
    for (uint32_t b = blockIdx.x; b < n; b += gridDim.x) {
        const uint32_t i = list[b];
        const mt_cand c = cands[i];
        const int w = 4 << c.tx_size;  // square 4..32 (checked by the setup)
        const int si = (c.planes & 3) | ((c.planes_hi & 15) << 2);
        const bool sh = c.pflags & MT_CAND_CFL_SS_HOR, sv = c.pflags & MT_CAND_CFL_SS_VER;
        const int bdi = (c.bd_qm >> 4) & 3;
        const int bd = bdi == 0 ? 8 : bdi == 1 ? 10 : 12;
        const int dc = pred[c.offset];
        const int alpha = cfl_block(planes.p[c.txtp_mode], (int)c.x << (sh ? 1 : 0), (int)c.y << (sv ? 1 : 0), sh,
                                    sv, planes.p[si], c.x, c.y, w, w, bd, dc,
                                    (c.pflags & MT_CAND_CFL_ALPHA_GIVEN) != 0, c.angle_delta, s_ac,
                                    pred + c.offset, nullptr, resid + c.offset,
                                    (c.pflags & MT_CAND_CFL_ALPHA_STEP) ? c.angle_delta : 0);
        if ((threadIdx.x & 63) == 0) alpha_out[i] = alpha;
    }
}

}  // namespace

namespace mtlib {
int pred_init_device(int /*device*/) {
    MT_HIP_TRY(upload_pred_tables());
    return MT_OK;
}

int planes_from_host(const mt_plane* planes, int nplanes, const char* who, PlaneSet* ps) {
    if (nplanes < 1 || nplanes > MT_MAX_PLANES || !planes)
        return set_error(MT_ERR_ARG, std::string(who) + ": 1.." + std::to_string(MT_MAX_PLANES) +
                                         " planes required");
    for (int i = 0; i < nplanes; ++i) {
        const mt_plane& p = planes[i];
        if (!p.data || p.stride <= 0 || p.width <= 0 || p.height <= 0 || p.width > p.stride ||
            p.vis_w < 0 || p.vis_h < 0)
            return set_error(MT_ERR_ARG, std::string(who) + ": bad plane " + std::to_string(i));
        ps->p[i] = p;
    }
    ps->n = nplanes;
    return MT_OK;
}
}  // namespace mtlib

namespace {
// Library-owned class lists of one device, grown on demand: calls on one device share them.
struct PredWork {
    size_t cap = 0;
    uint32_t* lists = nullptr;  // kSmallShapes + 1 lists of cap entries
    uint32_t* cnt = nullptr;
};
std::mutex g_pred_mu;
std::map<mtlib::StreamKey, PredWork> g_pred_work;
}  // namespace

namespace mtlib {
// The class lists of (device, stream) for n blocks (grown, never shrunk).
int pred_grow(PredWork& w, size_t n) {
    constexpr int K = kSmallShapes + 1;
    if (w.cap < n) {
        if (w.lists) MT_HIP_TRY(hipFree(w.lists));
        w.lists = nullptr;
        const size_t cap = n < 65536 ? 65536 : n;
        MT_HIP_TRY(hipMalloc(&w.lists, K * cap * sizeof(uint32_t)));
        w.cap = cap;
    }
    if (!w.cnt) MT_HIP_TRY(hipMalloc(&w.cnt, K * sizeof(uint32_t)));
    return MT_OK;
}
int pred_reserve(size_t n, hipStream_t s) {
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_pred_mu);
    return pred_grow(g_pred_work[mtlib::stream_key(dev, s)], n);
}

// mt_intra_pred, optionally also writing src - pred (int16) into resid_d at each block's
// out_offset, the source plane index taken from each descriptor's pad byte.
int intra_pred_launch(const mt_predblk* blocks_d, int n, const PlaneSet& ps, int32_t* pred_d,
                      int32_t* status_d, int16_t* resid_d, hipStream_t s) {
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_pred_mu);
    PredWork& w = g_pred_work[mtlib::stream_key(dev, s)];
    constexpr int K = kSmallShapes + 1;
    if ((rc = pred_grow(w, (size_t)n))) return rc;
    ProfScope prof(kProfPred, s);
    hipLaunchKernelGGL(zero_counts_kernel<K>, dim3(1), dim3(64), 0, s, w.cnt);
    hipLaunchKernelGGL(pred_classify_kernel, dim3(std::max(1, std::min((n + 255) / 256, cus * 4))),
                       dim3(256), 0, s, blocks_d, n, ps, w.lists, w.cap, w.cnt, status_d);
    const uint32_t* L = w.lists;
    const size_t cap = w.cap;
    const int gs = std::max(1, std::min((n + 255) / 256, cus * 4));
    hipLaunchKernelGGL((pred_small_kernel<4, 4>), dim3(gs), dim3(256), 0, s, blocks_d, L, w.cnt, ps, pred_d, resid_d);
    hipLaunchKernelGGL((pred_small_kernel<4, 8>), dim3(gs), dim3(256), 0, s, blocks_d, L + cap, w.cnt + 1, ps, pred_d, resid_d);
    hipLaunchKernelGGL((pred_small_kernel<8, 4>), dim3(gs), dim3(256), 0, s, blocks_d, L + 2 * cap, w.cnt + 2, ps, pred_d, resid_d);
    hipLaunchKernelGGL((pred_small_kernel<8, 8>), dim3(gs), dim3(256), 0, s, blocks_d, L + 3 * cap, w.cnt + 3, ps, pred_d, resid_d);
    hipLaunchKernelGGL((pred_small_kernel<4, 16>), dim3(gs), dim3(256), 0, s, blocks_d, L + 4 * cap, w.cnt + 4, ps, pred_d, resid_d);
    hipLaunchKernelGGL((pred_small_kernel<16, 4>), dim3(gs), dim3(256), 0, s, blocks_d, L + 5 * cap, w.cnt + 5, ps, pred_d, resid_d);
    const int gw = n < cus * 32 ? n : cus * 32;
    hipLaunchKernelGGL(intra_pred_kernel, dim3(gw), dim3(64), 0, s, blocks_d, L + 6 * cap, w.cnt + 6, ps,
                       pred_d, resid_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib

namespace mtlib {
int cfl_eval_launch(const mt_cand* cands_d, const uint32_t* list_d, const uint32_t* count_d, int n_max,
                    const PlaneSet& ps, int32_t* pred_d, int16_t* resid_d, int32_t* alpha_d, hipStream_t s) {
    if (n_max <= 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    const int grid = n_max < cus * 16 ? n_max : cus * 16;
    hipLaunchKernelGGL(cfl_eval_kernel, dim3(grid), dim3(64), 0, s, cands_d, list_d, count_d, ps, pred_d, resid_d,
                       alpha_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib

extern "C" int mt_intra_pred(const mt_predblk* blocks_d, int n, const mt_plane* planes,
                             int nplanes, int32_t* pred_d, int32_t* status_d, void* stream) {
    if (n < 0 || (n > 0 && (!blocks_d || !pred_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_intra_pred: bad pointer or count");
    PlaneSet ps{};
    const int rc = mtlib::planes_from_host(planes, nplanes, "mt_intra_pred", &ps);
    if (rc != MT_OK) return rc;
    return mtlib::intra_pred_launch(blocks_d, n, ps, pred_d, status_d, nullptr, (hipStream_t)stream);
}

extern "C" int mt_cfl(const mt_cflblk* blocks_d, int n, const mt_plane* planes, int nplanes,
                      int32_t* pred_d, int32_t* ac_d, int32_t* alpha_d, void* stream) {
    if (n < 0 || (n > 0 && (!blocks_d || !pred_d || !alpha_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_cfl: bad pointer or count");
    PlaneSet ps{};
    int rc = mtlib::planes_from_host(planes, nplanes, "mt_cfl", &ps);
    if (rc != MT_OK) return rc;
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    const int grid = n < cus * 32 ? n : cus * 32;
    mtlib::ProfScope prof(mtlib::kProfPred, (hipStream_t)stream);
    hipLaunchKernelGGL(cfl_kernel, dim3(grid), dim3(64), 0, (hipStream_t)stream, blocks_d, n, ps,
                       pred_d, ac_d, alpha_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
