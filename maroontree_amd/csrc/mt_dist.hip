// Distortion of reconstructed candidates on the GPU: the reference encoder's RdDispatch
// measures (src/rd_sse.rs:244-271, :454, :552, :635, :682) over a batch of blocks.
//
// One wave per block, persistent over the batch.  Phase 1 builds the block's error
// src - clamp(pred + resid) in LDS, zero outside the displayed part, and reduces the plain
// SSE.  Phase 2 takes one 4x4 tile per lane for the Hadamard measures.  Phase 3 is the
// weighted SSE in the op order of the reference's x86 kernel weighted_sse_avx2
// (src/avx/rd.rs:767-849): lane a < 32 owns accumulator (k, half, l) = (a >> 3, (a >> 2) & 1,
// a & 3) and runs its fused chain over the tile pairs in order; lanes 0..3 then apply the
// fixed reduction tree and the 4-wide tail column.  Every product is rounded as the
// reference rounds it (__fmul_rn / __fmaf_rn / __fadd_rn).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/mtgpu.h"
#include "mt_common.h"
#include "mt_lists.h"

namespace {

using mtlib::PlaneSet;

__constant__ float c_wd[4][4];  // WSSE_DCT4 (src/rd_sse.rs:1420)
__constant__ float c_ww[4][4];  // WSSE_WEIGHTS (src/rd_sse.rs:1402)

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

__device__ __forceinline__ void had4(int a, int b, int c, int d, int* o) {
    const int e = a + c, f = a - c, g = b + d, h = b - d;
    o[0] = e + g;
    o[1] = f + h;
    o[2] = f - h;
    o[3] = e - g;
}

// Sum |H e H^T| of one 4x4 tile (had4 on rows, then on columns; luma_satd_scalar,
// src/partition_rd.rs:86).
template <typename T> __device__ uint32_t tile_satd(const T* e, int stride) {
    int rows[4][4];
    for (int r = 0; r < 4; ++r) had4(e[r * stride], e[r * stride + 1], e[r * stride + 2], e[r * stride + 3], rows[r]);
    uint32_t s = 0;
    for (int x = 0; x < 4; ++x) {
        int col[4];
        had4(rows[0][x], rows[1][x], rows[2][x], rows[3][x], col);
        for (int i = 0; i < 4; ++i) s += (uint32_t)abs(col[i]);
    }
    return s;
}

// Coefficient l of row frequency k of the 4x4 tile at (r0, c0) of the error block, in the
// vector program's order: t[j] = e0j*D[k][0] then three fmas down the column; then
// c = D[l][0]*t0 and three fmas across.
__device__ __forceinline__ float wsse_coef(const int16_t* e, int stride, int r0, int c0, int k, int l) {
    float t[4];
    for (int j = 0; j < 4; ++j) {
        const int16_t* col = e + r0 * stride + c0 + j;
        float v = __fmul_rn((float)col[0], c_wd[k][0]);
        v = __fmaf_rn((float)col[stride], c_wd[k][1], v);
        v = __fmaf_rn((float)col[2 * stride], c_wd[k][2], v);
        v = __fmaf_rn((float)col[3 * stride], c_wd[k][3], v);
        t[j] = v;
    }
    float c = __fmul_rn(c_wd[l][0], t[0]);
    c = __fmaf_rn(c_wd[l][1], t[1], c);
    c = __fmaf_rn(c_wd[l][2], t[2], c);
    c = __fmaf_rn(c_wd[l][3], t[3], c);
    return c;
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

constexpr int kES = 65;  // LDS error row stride (64 + 1)

__global__ __launch_bounds__(64) void dist_kernel(const mt_distblk* __restrict__ blocks,
                                                  const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ count,
                                                  PlaneSet planes, const int32_t* __restrict__ pred,
                                                  const int32_t* __restrict__ resid,
                                                  mt_distres* __restrict__ out) {
    __shared__ int16_t s_e[64 * kES];  // error src - recon (|e| < 2^12), zero where invisible
    __shared__ float s_acc[32];
    __shared__ float s_sum[4];
    __shared__ float s_t[16 * 16];  // column sums t[k][j] of one tile row: [tile][k * 4 + j]
    const int lane = threadIdx.x;
    const int n = (int)*count;
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const uint32_t b = list[li];
        const mt_distblk d = blocks[b];
        const int bw = d.bw, bh = d.bh, bd = d.bd;
        const mt_plane pl = planes.p[d.plane];  // validated by dist_classify_kernel
        const int vw = min(max(pl.vis_w - (int)d.x, 0), bw);
        const int vh = min(max(pl.vis_h - (int)d.y, 0), bh);
        const int maxv = (1 << bd) - 1;
        const bool has_res = !(d.flags & MT_DIST_NO_RESID);
        const int32_t* pr = pred + d.pred_offset;
        const int32_t* rs = resid + d.resid_offset;
        const uint16_t* src = pl.data + (size_t)d.y * pl.stride + d.x;
        const int lw = __builtin_ctz(bw);
        // ---- phase 1: error blocks and the plain SSE (sse_recon / chroma_sse)
        int64_t sse = 0;
#pragma unroll 8
        for (int p = lane; p < bw * bh; p += 64) {
            const int y = p >> lw, x = p & (bw - 1);
            int e = 0;
            if (x < vw && y < vh)
                e = (int)src[(size_t)y * pl.stride + x] - clampi(pr[p] + (has_res ? rs[p] : 0), 0, maxv);
            s_e[y * kES + x] = (int16_t)e;
            sse += (int64_t)e * e;
        }
        sse = wave_sum(sse);
        __syncthreads();
        // ---- phase 2: SATD of the reconstruction error, SAD + SATD/4 of the prediction error
        uint64_t satd = 0, psatd = 0, sad = 0;
        const int tiles = (bw >> 2) * (bh >> 2);
#pragma unroll 4
        for (int t = lane; t < tiles; t += 64) {
            const int ty = (t / (bw >> 2)) * 4, tx = (t % (bw >> 2)) * 4;
            satd += tile_satd(s_e + ty * kES + tx, kES);
            int ep[16];  // src - pred of the tile, zero where invisible
            for (int r = 0; r < 4; ++r)
                for (int x = 0; x < 4; ++x) {
                    const int yy = ty + r, xx = tx + x;
                    ep[r * 4 + x] = (xx < vw && yy < vh)
                                        ? (int)src[(size_t)yy * pl.stride + xx] - pr[yy * bw + xx] : 0;
                    sad += (uint64_t)abs(ep[r * 4 + x]);
                }
            psatd += tile_satd(ep, 4);
        }
        satd = wave_sum(satd);
        psatd = wave_sum(psatd);
        sad = wave_sum(sad);
        // ---- phase 3: weighted SSE of the visible error, zero-padded to whole 4x4 tiles
        int64_t dist = sse;
        if (d.wsse_scale > 0.0f) {
            const int pw = (vw + 3) & ~3, ph = (vh + 3) & ~3;
            float total = 0.0f;
            if (pw > 0 && ph > 0) {
                const int pairs = pw / 8 * 8;
                // Per row of tiles: all lanes first build the column sums t[k][j] of its
                // paired tiles (wsse_coef's first stage, same operations), then lane a < 32
                // finishes its coefficient (k, l) of tile c0 / 4 + half from them and runs its
                // chain over the tiles in order, as wsse_coef would.
                const int k = lane >> 3, half = (lane >> 2) & 1, l = lane & 3;
                float acc = 0.0f;
                const float wk = lane < 32 ? c_ww[k][l] : 0.0f;
                for (int r0 = 0; r0 < ph; r0 += 4) {
                    for (int e = lane; e < pairs * 4; e += 64) {  // pairs / 4 tiles x 16
                        const int q = e >> 4, kk = (e >> 2) & 3, j = e & 3;
                        const int16_t* col = s_e + r0 * kES + q * 4 + j;
                        float v = __fmul_rn((float)col[0], c_wd[kk][0]);
                        v = __fmaf_rn((float)col[kES], c_wd[kk][1], v);
                        v = __fmaf_rn((float)col[2 * kES], c_wd[kk][2], v);
                        v = __fmaf_rn((float)col[3 * kES], c_wd[kk][3], v);
                        s_t[e] = v;
                    }
                    __syncthreads();
                    if (lane < 32) {
                        for (int c0 = 0; c0 < pairs; c0 += 8) {
                            const float* t = s_t + ((c0 >> 2) + half) * 16 + k * 4;
                            float c = __fmul_rn(c_wd[l][0], t[0]);
                            c = __fmaf_rn(c_wd[l][1], t[1], c);
                            c = __fmaf_rn(c_wd[l][2], t[2], c);
                            c = __fmaf_rn(c_wd[l][3], t[3], c);
                            acc = __fmaf_rn(__fmul_rn(wk, c), c, acc);
                        }
                    }
                    __syncthreads();  // s_t is rewritten for the next row
                }
                if (lane < 32) s_acc[lane] = acc;
                __syncthreads();
                if (lane < 4) {
                    const int l = lane;
                    auto A = [&](int k, int half) { return s_acc[k * 8 + half * 4 + l]; };
                    const float s01 = __fadd_rn(__fadd_rn(A(0, 0), A(0, 1)), __fadd_rn(A(1, 0), A(1, 1)));
                    const float s23 = __fadd_rn(__fadd_rn(A(2, 0), A(2, 1)), __fadd_rn(A(3, 0), A(3, 1)));
                    float sum = __fadd_rn(s01, s23);
                    if (pairs < pw)
                        for (int r0 = 0; r0 < ph; r0 += 4)
                            for (int k = 0; k < 4; ++k) {
                                const float c = wsse_coef(s_e, kES, r0, pairs, k, l);
                                sum = __fmaf_rn(__fmul_rn(c_ww[k][l], c), c, sum);
                            }
                    s_sum[l] = sum;
                }
                __syncthreads();
                total = __fadd_rn(__fadd_rn(s_sum[0], s_sum[2]), __fadd_rn(s_sum[1], s_sum[3]));
            }
            dist = (int64_t)roundf(__fmul_rn(total, d.wsse_scale));
        }
        if (lane == 0) out[b] = mt_distres{sse, dist, satd, sad + (psatd >> 2)};
        __syncthreads();  // LDS reuse by the next block
    }
}

// ---- blocks of at most 64 samples: one thread per block, everything in registers ----------

// t[k][j] of the 4x4 tile at (r0, c0): column j through row frequency k, in the vector
// program's order (e0j*D[k][0], then three fmas down the column).
__device__ __forceinline__ void wsse_rows_reg(const int* e, int stride, int r0, int c0, float (&t)[4][4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int* col = e + r0 * stride + c0 + j;
            float v = __fmul_rn((float)col[0], c_wd[k][0]);
            v = __fmaf_rn((float)col[stride], c_wd[k][1], v);
            v = __fmaf_rn((float)col[2 * stride], c_wd[k][2], v);
            v = __fmaf_rn((float)col[3 * stride], c_wd[k][3], v);
            t[k][j] = v;
        }
}

// Coefficient (k, l) from t[k]: D[l][0]*t0, then three fmas across.
__device__ __forceinline__ float wsse_col(const float (&t)[4][4], int k, int l) {
    float c = __fmul_rn(c_wd[l][0], t[k][0]);
    c = __fmaf_rn(c_wd[l][1], t[k][1], c);
    c = __fmaf_rn(c_wd[l][2], t[k][2], c);
    c = __fmaf_rn(c_wd[l][3], t[k][3], c);
    return c;
}

// The same four measures as dist_kernel for one W x H block, by one thread.  Loops run over
// the block's static tile grid (so the error stays in registers); the displayed extent only
// predicates them.  The weighted SSE keeps weighted_sse_avx2's accumulators and order.
// err(i, y, x, ev, epv) gives sample i = y * W + x's reconstruction error ev and prediction
// error epv (both 0 outside the displayed extent).
template <int W, int H, class Err>
__device__ mt_distres small_measures(const Err& err, int vw, int vh, float scale) {
    int e[W * H];
    int64_t sse = 0;
    uint64_t satd = 0, psatd = 0, sad = 0;
#pragma unroll
    for (int ty = 0; ty < H; ty += 4)
#pragma unroll
        for (int tx = 0; tx < W; tx += 4) {
            int ep[16];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int y = ty + r, x = tx + c, i = y * W + x;
                    int ev, epv;
                    err(i, y, x, ev, epv);
                    e[i] = ev;
                    ep[r * 4 + c] = epv;
                    sse += (int64_t)ev * ev;
                    sad += (uint64_t)abs(epv);
                }
            satd += tile_satd(e + ty * W + tx, W);
            psatd += tile_satd(ep, 4);
        }
    int64_t dist = sse;
    if (scale > 0.0f) {
        const int pw = (vw + 3) & ~3, ph = (vh + 3) & ~3, pairs = pw / 8 * 8;
        float total = 0.0f;
        if (pw > 0 && ph > 0) {
            float acc[4][2][4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int hf = 0; hf < 2; ++hf)
#pragma unroll
                    for (int l = 0; l < 4; ++l) acc[k][hf][l] = 0.0f;
#pragma unroll
            for (int r0 = 0; r0 < H; r0 += 4)
#pragma unroll
                for (int c0 = 0; c0 + 8 <= W; c0 += 8)
                    if (r0 < ph && c0 < pairs)
#pragma unroll
                        for (int hf = 0; hf < 2; ++hf) {
                            float t[4][4];
                            wsse_rows_reg(e, W, r0, c0 + 4 * hf, t);
#pragma unroll
                            for (int k = 0; k < 4; ++k)
#pragma unroll
                                for (int l = 0; l < 4; ++l) {
                                    const float c = wsse_col(t, k, l);
                                    acc[k][hf][l] = __fmaf_rn(__fmul_rn(c_ww[k][l], c), c, acc[k][hf][l]);
                                }
                        }
            float sum[4];
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float s01 = __fadd_rn(__fadd_rn(acc[0][0][l], acc[0][1][l]), __fadd_rn(acc[1][0][l], acc[1][1][l]));
                const float s23 = __fadd_rn(__fadd_rn(acc[2][0][l], acc[2][1][l]), __fadd_rn(acc[3][0][l], acc[3][1][l]));
                sum[l] = __fadd_rn(s01, s23);
            }
            if (pairs < pw) {
#pragma unroll
                for (int r0 = 0; r0 < H; r0 += 4)
#pragma unroll
                    for (int tc = 0; tc < W; tc += 4)
                        if (r0 < ph && tc == pairs) {
                            float t[4][4];
                            wsse_rows_reg(e, W, r0, tc, t);
#pragma unroll
                            for (int k = 0; k < 4; ++k)
#pragma unroll
                                for (int l = 0; l < 4; ++l) {
                                    const float c = wsse_col(t, k, l);
                                    sum[l] = __fmaf_rn(__fmul_rn(c_ww[k][l], c), c, sum[l]);
                                }
                        }
            }
            total = __fadd_rn(__fadd_rn(sum[0], sum[2]), __fadd_rn(sum[1], sum[3]));
        }
        dist = (int64_t)roundf(__fmul_rn(total, scale));
    }
    return mt_distres{sse, dist, satd, sad + (psatd >> 2)};
}

using mtlib::kSmallShapes;

// The error samples of one block (residual_pred then sse_recon's difference): the
// reconstruction error and the prediction error, 0 outside the displayed extent.
struct SmallBlk {
    const uint16_t* src;
    const int32_t* pr;
    const int32_t* rs;  // null: no residual (prediction-only block)
    int sstride, vw, vh, maxv;
};
__device__ __forceinline__ void small_err(const SmallBlk& B, int i, int y, int x, int& ev, int& epv) {
    ev = 0;
    epv = 0;
    if (x < B.vw && y < B.vh) {
        const int sv = B.src[(size_t)y * B.sstride + x];
        const int p = B.pr[i];
        ev = sv - clampi(p + (B.rs ? B.rs[i] : 0), 0, B.maxv);
        epv = sv - p;
    }
}

// One wave per 64 list entries, one block per lane.  The wave first stages every block's
// error samples through LDS with loads spread over the lanes (a block's predictions and
// residuals are contiguous, so each load instruction reads whole lines) as packed 16-bit
// pairs; each lane then measures its block from LDS.  A wave whose errors do not fit 16 bits
// (samples or predictions outside the bit depth) measures from memory instead.
template <int W, int H>
__global__ void __launch_bounds__(64) dist_small_kernel(const mt_distblk* __restrict__ blocks,
                                                        const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count,
                                                        PlaneSet planes, const int32_t* __restrict__ pred,
                                                        const int32_t* __restrict__ resid,
                                                        mt_distres* __restrict__ out) {
    constexpr int E = W * H, BPI = 64 / E, RS = E + 1;
    __shared__ int32_t s_err[64 * RS];
    __shared__ SmallBlk s_blk[64];
    const int t = threadIdx.x;
    const int n = (int)*count;
    for (int base = blockIdx.x * 64; base < n; base += gridDim.x * 64) {
        const int nb = min(64, n - base);
        uint32_t b = 0;
        SmallBlk mine{};
        float scale = 0.0f;
        if (t < nb) {
            b = list[base + t];
            const mt_distblk d = blocks[b];
            const mt_plane pl = planes.p[d.plane];
            mine.src = pl.data + (size_t)d.y * pl.stride + d.x;
            mine.pr = pred + d.pred_offset;
            mine.rs = (d.flags & MT_DIST_NO_RESID) ? nullptr : resid + d.resid_offset;
            mine.sstride = pl.stride;
            mine.vw = min(max(pl.vis_w - (int)d.x, 0), W);
            mine.vh = min(max(pl.vis_h - (int)d.y, 0), H);
            mine.maxv = (1 << d.bd) - 1;
            scale = d.wsse_scale;
            s_blk[t] = mine;
        }
        __syncthreads();
        bool wide = false;
        for (int j0 = 0; j0 < nb; j0 += BPI) {
            const int j = j0 + t / E, i = t % E;
            if (j < nb) {
                int ev, epv;
                small_err(s_blk[j], i, i / W, i % W, ev, epv);
                wide |= ev != (int)(int16_t)ev || epv != (int)(int16_t)epv;
                s_err[j * RS + i] = (int)(((uint32_t)ev & 0xffffu) | ((uint32_t)epv << 16));
            }
        }
        const bool packed = __ballot(wide) == 0ull;
        __syncthreads();
        if (t < nb) {
            if (packed) {
                const int32_t* row = s_err + t * RS;
                out[b] = small_measures<W, H>(
                    [&](int i, int, int, int& ev, int& epv) {
                        const int v = row[i];
                        ev = (int)(int16_t)(v & 0xffff);
                        epv = v >> 16;
                    },
                    mine.vw, mine.vh, scale);
            } else {
                out[b] = small_measures<W, H>(
                    [&](int i, int y, int x, int& ev, int& epv) { small_err(mine, i, y, x, ev, epv); }, mine.vw,
                    mine.vh, scale);
            }
        }
        __syncthreads();  // the next entries reuse the LDS
    }
}

// Validates every descriptor (invalid ones are answered here) and sorts the valid ones into
// one list per shape class (list c at lists + c * cap).  cnt[0..6] must be zero on entry.
__global__ void __launch_bounds__(256) dist_classify_kernel(const mt_distblk* __restrict__ blocks, int n,
                                                            PlaneSet planes, uint32_t* __restrict__ lists,
                                                            size_t cap, uint32_t* __restrict__ cnt,
                                                            mt_distres* __restrict__ out) {
    auto cls = [&](int b, bool first) -> int {
        const mt_distblk d = blocks[b];
        const int bw = d.bw, bh = d.bh, bd = d.bd;
        bool ok = (bw == 4 || bw == 8 || bw == 16 || bw == 32 || bw == 64) &&
                  (bh == 4 || bh == 8 || bh == 16 || bh == 32 || bh == 64) &&
                  (bd == 8 || bd == 10 || bd == 12) && d.plane < planes.n;
        const mt_plane pl = planes.p[ok ? d.plane : 0];
        ok = ok && (int)d.x + bw <= pl.width && (int)d.y + bh <= pl.height;
        if (!ok) {
            if (first) out[b] = mt_distres{-1, -1, 0, 0};
            return -1;
        }
        return mtlib::small_shape_class(bw, bh);
    };
    mtlib::classify_chunk<kSmallShapes + 1>(n, cls, lists, cap, cnt);
}


}  // namespace

namespace mtlib {
void wsse_tables(float wd[4][4], float ww[4][4]) {
    // WSSE_DCT4: the f32 literals of the reference; WSSE_WEIGHTS: the raw table divided by
    // its mean (0.4625) in f32, as the reference's const evaluation rounds it.
    const float a = 0.5f, bb = 0.6532815f, c = 0.27059805f;
    const float d[4][4] = {{a, a, a, a}, {bb, c, -c, -bb}, {a, -a, -a, a}, {c, -bb, bb, -c}};
    volatile float mean = 0.4625f;
    const float dc = 1.8f / mean, diag = 0.8f / mean, axis = 0.4f / mean, hf = 0.2f / mean,
                ac1 = 1.0f / mean;
    const float w[4][4] = {{dc, ac1, axis, hf}, {ac1, diag, hf, hf}, {axis, hf, hf, hf}, {hf, hf, hf, hf}};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            wd[i][j] = d[i][j];
            ww[i][j] = w[i][j];
        }
}

int dist_init_device(int /*device*/) {
    float wd[4][4], ww[4][4];
    wsse_tables(wd, ww);
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_wd), wd, sizeof(wd)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_ww), ww, sizeof(ww)));
    return MT_OK;
}
}  // namespace mtlib

namespace {
// Library-owned lists of one device, grown on demand: calls on one device share them.
struct DistWork {
    size_t cap = 0;
    uint32_t* lists = nullptr;  // kSmallShapes + 1 lists of cap entries
    uint32_t* cnt = nullptr;
};
std::mutex g_dist_mu;
std::map<mtlib::StreamKey, DistWork> g_dist_work;
}  // namespace

namespace mtlib {
int dist_grow(DistWork& w, size_t n) {
    if (w.cap < n) {
        if (w.lists) MT_HIP_TRY(hipFree(w.lists));
        w.lists = nullptr;
        const size_t cap = n < 65536 ? 65536 : n;
        MT_HIP_TRY(hipMalloc(&w.lists, (kSmallShapes + 1) * cap * sizeof(uint32_t)));
        w.cap = cap;
    }
    if (!w.cnt) MT_HIP_TRY(hipMalloc(&w.cnt, (kSmallShapes + 1) * sizeof(uint32_t)));
    return MT_OK;
}
int dist_reserve(size_t n, hipStream_t s) {
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_dist_mu);
    return dist_grow(g_dist_work[mtlib::stream_key(dev, s)], n);
}

int distortion_launch(const mt_distblk* blocks_d, int n, const PlaneSet& ps, const int32_t* pred_d,
                      const int32_t* resid_d, mt_distres* out_d, hipStream_t s) {
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_dist_mu);
    DistWork& w = g_dist_work[mtlib::stream_key(dev, s)];
    if ((rc = dist_grow(w, (size_t)n))) return rc;
    const int32_t* resid = resid_d ? resid_d : pred_d;
    mtlib::ProfScope prof(mtlib::kProfDist, s);
    hipLaunchKernelGGL(mtlib::zero_counts_kernel<kSmallShapes + 1>, dim3(1), dim3(64), 0, s, w.cnt);
    hipLaunchKernelGGL(dist_classify_kernel, dim3(std::max(1, std::min((n + 255) / 256, cus * 4))),
                       dim3(256), 0, s, blocks_d, n, ps, w.lists, w.cap, w.cnt, out_d);
    const int gs = std::max(1, std::min((n + 63) / 64, cus * 16));
    const uint32_t* L = w.lists;
    const size_t cap = w.cap;
    hipLaunchKernelGGL((dist_small_kernel<4, 4>), dim3(gs), dim3(64), 0, s, blocks_d, L, w.cnt, ps, pred_d, resid, out_d);
    hipLaunchKernelGGL((dist_small_kernel<4, 8>), dim3(gs), dim3(64), 0, s, blocks_d, L + cap, w.cnt + 1, ps, pred_d, resid, out_d);
    hipLaunchKernelGGL((dist_small_kernel<8, 4>), dim3(gs), dim3(64), 0, s, blocks_d, L + 2 * cap, w.cnt + 2, ps, pred_d, resid, out_d);
    hipLaunchKernelGGL((dist_small_kernel<8, 8>), dim3(gs), dim3(64), 0, s, blocks_d, L + 3 * cap, w.cnt + 3, ps, pred_d, resid, out_d);
    hipLaunchKernelGGL((dist_small_kernel<4, 16>), dim3(gs), dim3(64), 0, s, blocks_d, L + 4 * cap, w.cnt + 4, ps, pred_d, resid, out_d);
    hipLaunchKernelGGL((dist_small_kernel<16, 4>), dim3(gs), dim3(64), 0, s, blocks_d, L + 5 * cap, w.cnt + 5, ps, pred_d, resid, out_d);
    const int gw = n < cus * 32 ? n : cus * 32;
    hipLaunchKernelGGL(dist_kernel, dim3(gw), dim3(64), 0, s, blocks_d, L + 6 * cap, w.cnt + 6, ps,
                       pred_d, resid, out_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib

extern "C" int mt_distortion(const mt_distblk* blocks_d, int n, const mt_plane* planes, int nplanes,
                             const int32_t* pred_d, const int32_t* resid_d, mt_distres* out_d,
                             void* stream) {
    if (n < 0 || (n > 0 && (!blocks_d || !pred_d || !out_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_distortion: bad pointer or count");
    PlaneSet ps{};
    int rc = mtlib::planes_from_host(planes, nplanes, "mt_distortion", &ps);
    if (rc != MT_OK) return rc;
    return mtlib::distortion_launch(blocks_d, n, ps, pred_d, resid_d, out_d, (hipStream_t)stream);
}
