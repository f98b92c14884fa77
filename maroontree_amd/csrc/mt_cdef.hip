// CDEF on the GPU: the reference encoder's frame_cdef (src/coder/frame_cdef.rs:49) over a
// deblocked reconstruction.
//
//  - cdef_dir_kernel: one thread per luma 8x8, the direction search of cdef_find_dir
//    (src/cdef.rs:103-179: the eight line-projection partial sums, the costs, the argmax and
//    the directional variance);
//  - cdef_tables_kernel: one wave per non-skip luma 8x8 and its chroma sub-block; each lane
//    filters one sample at all twelve signalled strength pairs at once (the reference's
//    cdef_block_candidates, :651: the constrained tap sums per strength shared across the
//    pairs), the errors against the source go to LDS, and one lane per (plane, pair)
//    measures the block (weighted SSE in the x86 kernel's op order, or the plain SSE) and
//    adds it to the 64x64 unit's table with an integer atomic (exact in any order);
//  - the decision on the host over the tables (units x 12 x 2 int64): the directional-
//    variance gate, the Slow / Fast candidate pairs, the greedy set construction, the
//    mildest-entry assignment with the margin, the cdef_bits choice on rate
//    (:140-558), at the shipped tuning (margin 40 per mille, gate 15000, wSSE, lambda_d);
//  - cdef_apply_kernel: one thread per sample, every tap from the pre-CDEF planes, out of
//    place (apply_cdef_plane / apply_cdef_chroma, :995-1125).
// HBM-bound: the tables pass reads each reconstructed sample's 5x5 neighbourhood from L2
// and the source once; the apply pass reads and writes each sample once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/mtgpu.h"
#include "mt_common.h"

namespace {

constexpr int kVeryLarge = 0x4000;
__constant__ int8_t c_dir[8][2][2] = {{{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}},
                                      {{1, 1}, {2, 2}},   {{1, 0}, {2, 1}},  {{1, 0}, {2, 0}}, {{1, 0}, {2, -1}}};

struct PlaneV {
    const uint16_t* p;
    int stride, w, h;
};
struct Wsse {
    float d[4][4], w[4][4];
};

__device__ __forceinline__ int log2_floor(int v) { return v <= 0 ? 0 : 31 - __clz(v); }
__device__ __forceinline__ int str_shift(int s, int damp) { return s ? max(0, damp - log2_floor(s)) : 0; }
__device__ __forceinline__ int constrain(int diff, int s, int shift) {
    const int a = abs(diff);
    const int c = min(max(s - (a >> shift), 0), a);
    return diff < 0 ? -c : c;
}
__device__ __forceinline__ int tap(const PlaneV& P, int x, int y) {
    return (x < 0 || y < 0 || x >= P.w || y >= P.h) ? kVeryLarge : (int)P.p[(size_t)y * P.stride + x];
}
// adjust_pri (src/cdef.rs:181)
__host__ __device__ __forceinline__ int adjust_pri(int pri, int var) {
    if (var == 0) return 0;
    const int vs = var >> 6;
    int l = 0;
    if (vs) {
        l = 31 - __builtin_clz((unsigned)vs);
        l = l < 12 ? l : 12;
    }
    return (pri * (4 + l) + 8) >> 4;
}

// One sample at one strength pair (cdef_filter_block_generic's body, src/cdef.rs:514):
// effective strengths pri / sec (bit-depth scaled, luma pri adjusted).
__device__ int filter_px(const PlaneV& P, int x, int y, int pri, int sec, int dir, int damping, int bd) {
    const int c = P.p[(size_t)y * P.stride + x];
    const int cs = bd - 8, damp = max(damping, 1);
    const bool ep = pri != 0, es = sec != 0, clip = ep && es;
    const int pt0 = ((pri >> cs) & 1) ? 3 : 4, pt1 = ((pri >> cs) & 1) ? 3 : 2;
    const int psh = str_shift(pri, damp), ssh = str_shift(sec, damp);
    int sum = 0, mn = c, mx = c;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (ep)
#pragma unroll
            for (int sg = 1; sg >= -1; sg -= 2) {
                const int v = tap(P, x + sg * c_dir[dir][k][1], y + sg * c_dir[dir][k][0]);
                if (v == kVeryLarge) continue;
                sum += (k ? pt1 : pt0) * constrain(v - c, pri, psh);
                if (clip) {
                    mn = min(mn, v);
                    mx = max(mx, v);
                }
            }
        if (es)
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                const int sd = (dir + (o ? 6 : 2)) & 7;
#pragma unroll
                for (int sg = 1; sg >= -1; sg -= 2) {
                    const int v = tap(P, x + sg * c_dir[sd][k][1], y + sg * c_dir[sd][k][0]);
                    if (v == kVeryLarge) continue;
                    sum += (k ? 1 : 2) * constrain(v - c, sec, ssh);
                    if (clip) {
                        mn = min(mn, v);
                        mx = max(mx, v);
                    }
                }
            }
    }
    int out = c + ((8 + sum - (sum < 0 ? 1 : 0)) >> 4);
    if (clip) out = min(max(out, mn), mx);
    return min(max(out, 0), (1 << bd) - 1);
}

// The twelve signalled pairs of one sample (cdef_block_candidates, src/cdef.rs:651):
// out[si * 4 + pi] = (CAND_PRI[pi] = 0, 1, 2, 4; CAND_SEC[si] = 0, 1, 2); var < 0: chroma.
// A zero signalled primary filters the secondary taps along direction 0.
// tapf(dx, dy): the sample at that offset from the centre, kVeryLarge outside the plane.
template <class Tap> __device__ void px12(Tap tapf, int dir, int var, int damping, int bd, int* out) {
    const int c = tapf(0, 0);
    const int cs = bd - 8, damp = max(damping, 1), maxv = (1 << bd) - 1;
    int pe[3], psh[3], pw0[3], pw1[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int sp = (q == 2 ? 4 : q + 1) << cs;
        pe[q] = var >= 0 ? adjust_pri(sp, var) : sp;
        psh[q] = str_shift(pe[q], damp);
        const bool odd = (pe[q] >> cs) & 1;
        pw0[q] = odd ? 3 : 4;
        pw1[q] = odd ? 3 : 2;
    }
    const int se[2] = {1 << cs, 2 << cs};
    const int ssh[2] = {str_shift(se[0], damp), str_shift(se[1], damp)};
    int psum[3] = {0, 0, 0}, ssum[2] = {0, 0}, ssum0[2] = {0, 0};
    int mn = c, mx = c;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
#pragma unroll
        for (int sg = 1; sg >= -1; sg -= 2) {
            const int v = tapf(sg * c_dir[dir][k][1], sg * c_dir[dir][k][0]);
            if (v != kVeryLarge) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    if (pe[q]) psum[q] += (k ? pw1[q] : pw0[q]) * constrain(v - c, pe[q], psh[q]);
                mn = min(mn, v);
                mx = max(mx, v);
            }
        }
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {  // secondary taps along dir (min/max), then dir 0
            if (pass == 1 && dir == 0) continue;
            const int d0 = pass ? 0 : dir;
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                const int sd = (d0 + (o ? 6 : 2)) & 7;
#pragma unroll
                for (int sg = 1; sg >= -1; sg -= 2) {
                    const int v = tapf(sg * c_dir[sd][k][1], sg * c_dir[sd][k][0]);
                    if (v == kVeryLarge) continue;
#pragma unroll
                    for (int q = 0; q < 2; ++q) (pass ? ssum0 : ssum)[q] += (k ? 1 : 2) * constrain(v - c, se[q], ssh[q]);
                    if (!pass) {
                        mn = min(mn, v);
                        mx = max(mx, v);
                    }
                }
            }
        }
    }
    if (dir == 0) {
        ssum0[0] = ssum[0];
        ssum0[1] = ssum[1];
    }
#pragma unroll
    for (int si = 0; si < 3; ++si)
#pragma unroll
        for (int pi = 0; pi < 4; ++pi) {
            const bool enp = pi != 0 && pe[pi ? pi - 1 : 0] != 0, ens = si != 0;
            int sum;
            if (pi == 0) sum = ens ? ssum0[si - 1] : 0;
            else sum = (enp ? psum[pi - 1] : 0) + (ens ? ssum[si - 1] : 0);
            int v = c + ((8 + sum - (sum < 0 ? 1 : 0)) >> 4);
            if (enp && ens) v = min(max(v, mn), mx);
            out[si * 4 + pi] = min(max(v, 0), maxv);
        }
}

// Weighted SSE of a w x h (4 or 8) error block in the op order of weighted_sse_avx2
// (src/avx/rd.rs:767-849): per 4-row band, the 8-wide tile pairs accumulate (k, half, l)
// with rounded weight products fused into the sums, a fixed tree over k and halves, the
// 4-wide tail column fused into the per-l sums, then (l0 + l2) + (l1 + l3).
__device__ float wsse_block(const int* err, int w, int h, const Wsse& T) {
    float acc[4][2][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int hh = 0; hh < 2; ++hh)
#pragma unroll
            for (int l = 0; l < 4; ++l) acc[k][hh][l] = 0.0f;
    const int pairs = w / 8 * 8;
    auto tile = [&](int r0, int c0, int k, float* cf) {
        float t[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v = __fmul_rn((float)err[(r0 + 0) * w + c0 + j], T.d[k][0]);
            v = __fmaf_rn((float)err[(r0 + 1) * w + c0 + j], T.d[k][1], v);
            v = __fmaf_rn((float)err[(r0 + 2) * w + c0 + j], T.d[k][2], v);
            v = __fmaf_rn((float)err[(r0 + 3) * w + c0 + j], T.d[k][3], v);
            t[j] = v;
        }
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            float v = __fmul_rn(T.d[l][0], t[0]);
            v = __fmaf_rn(T.d[l][1], t[1], v);
            v = __fmaf_rn(T.d[l][2], t[2], v);
            v = __fmaf_rn(T.d[l][3], t[3], v);
            cf[l] = v;
        }
    };
    for (int r0 = 0; r0 < h; r0 += 4)
        for (int c0 = 0; c0 < pairs; c0 += 8)
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    float cf[4];
                    tile(r0, c0 + 4 * hh, k, cf);
#pragma unroll
                    for (int l = 0; l < 4; ++l) acc[k][hh][l] = __fmaf_rn(__fmul_rn(T.w[k][l], cf[l]), cf[l], acc[k][hh][l]);
                }
    float sum[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const float s01 = __fadd_rn(__fadd_rn(acc[0][0][l], acc[0][1][l]), __fadd_rn(acc[1][0][l], acc[1][1][l]));
        const float s23 = __fadd_rn(__fadd_rn(acc[2][0][l], acc[2][1][l]), __fadd_rn(acc[3][0][l], acc[3][1][l]));
        sum[l] = __fadd_rn(s01, s23);
    }
    if (pairs < w)
        for (int r0 = 0; r0 < h; r0 += 4)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float cf[4];
                tile(r0, pairs, k, cf);
#pragma unroll
                for (int l = 0; l < 4; ++l) sum[l] = __fmaf_rn(__fmul_rn(T.w[k][l], cf[l]), cf[l], sum[l]);
            }
    return __fadd_rn(__fadd_rn(sum[0], sum[2]), __fadd_rn(sum[1], sum[3]));
}

// RdDispatch::wsse_err (src/rd_sse.rs:277): the scaled weighted SSE rounded half away from
// zero, or the plain SSE when the scale is 0.
__device__ long long block_dist(const int* err, int w, int h, float scale, const Wsse& T) {
    if (scale > 0.0f) return (long long)roundf(__fmul_rn(wsse_block(err, w, h, T), scale));
    long long s = 0;
    for (int i = 0; i < w * h; ++i) s += (long long)err[i] * err[i];
    return s;
}

__global__ void __launch_bounds__(256) cdef_dir_kernel(PlaneV Y, int nbx, int nby, int bd, int* __restrict__ dirs,
                                                       int* __restrict__ vars) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbx * nby) return;
    const int x = (b % nbx) * 8, y = (b / nbx) * 8, cs = bd - 8;
    int part[8][15];
#pragma unroll
    for (int d = 0; d < 8; ++d)
#pragma unroll
        for (int i = 0; i < 15; ++i) part[d][i] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int v = ((int)Y.p[(size_t)min(y + i, Y.h - 1) * Y.stride + min(x + j, Y.w - 1)] >> cs) - 128;
            part[0][i + j] += v;
            part[1][i + j / 2] += v;
            part[2][i] += v;
            part[3][3 + i - j / 2] += v;
            part[4][7 + i - j] += v;
            part[5][3 - i / 2 + j] += v;
            part[6][j] += v;
            part[7][i / 2 + j] += v;
        }
    const long long div[8] = {0, 840, 420, 280, 210, 168, 140, 120};
    long long cost[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto sq = [](int v) { return (long long)v * v; };
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        cost[2] += sq(part[2][i]);
        cost[6] += sq(part[6][i]);
    }
    cost[2] *= 105;
    cost[6] *= 105;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        cost[0] += (sq(part[0][i]) + sq(part[0][14 - i])) * div[i + 1];
        cost[4] += (sq(part[4][i]) + sq(part[4][14 - i])) * div[i + 1];
    }
    cost[0] += sq(part[0][7]) * 105;
    cost[4] += sq(part[4][7]) * 105;
#pragma unroll
    for (int s = 1; s < 8; s += 2) {
#pragma unroll
        for (int i = 0; i < 5; ++i) cost[s] += sq(part[s][3 + i]) * 105;
#pragma unroll
        for (int i = 0; i < 3; ++i) cost[s] += (sq(part[s][i]) + sq(part[s][10 - i])) * div[2 * i + 2];
    }
    int best = 0;
#pragma unroll
    for (int d = 1; d < 8; ++d)
        if (cost[d] > cost[best]) best = d;
    dirs[b] = best;
    vars[b] = (int)((cost[best] - cost[(best + 4) & 7]) >> 10);
}

struct TabArgs {
    PlaneV rec[3], src[3];
    const uint8_t* skip8;
    const int *dirs, *vars;
    int nbx, nby, uc, sub_x, sub_y, mono, damping, bd;
    float scale;
    Wsse T;
    long long *ltab, *ctab;  // [units][12]
};

// One wave per luma 8x8 (grid-stride): twelve pairs per sample, errors in LDS, one lane per
// (plane, pair) measures and adds to the unit's table.
__global__ void __launch_bounds__(64) cdef_tables_kernel(TabArgs a) {
    __shared__ int s_err[3][12][64];
    __shared__ int s_tile[3][12 * 12];  // each plane's block with a 2-sample border
    const int lane = threadIdx.x;
    const int nb = a.nbx * a.nby;
    const int cbw = 8 >> a.sub_x, cbh = 8 >> a.sub_y, cn = cbw * cbh;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        if (a.skip8[b]) continue;
        const int bx = b % a.nbx, by = b / a.nbx, x = bx * 8, y = by * 8;
        const int u = (by / 8) * a.uc + bx / 8, dir = a.dirs[b];
        __syncthreads();  // the previous block's tiles are read
        for (int pl = 0; pl < (a.mono ? 1 : 3); ++pl) {
            const int tw = (pl ? cbw : 8) + 4, th = (pl ? cbh : 8) + 4;
            const int x0 = (pl ? x >> a.sub_x : x) - 2, y0 = (pl ? y >> a.sub_y : y) - 2;
            for (int i = lane; i < tw * th; i += 64) s_tile[pl][i] = tap(a.rec[pl], x0 + i % tw, y0 + i / tw);
        }
        __syncthreads();
        int o[12];
        if (a.scale <= 0.0f) {  // plain SSE: integer sums, reduced across the wave in any order
            unsigned sq[12], csq[12];
            {
                const int px = x + (lane & 7), py = y + (lane >> 3);
                const int* T0 = s_tile[0] + ((lane >> 3) + 2) * 12 + (lane & 7) + 2;
                px12([&](int dx, int dy) { return T0[dy * 12 + dx]; }, dir, a.vars[b], a.damping, a.bd, o);
                const int s = a.src[0].p[(size_t)py * a.src[0].stride + px];
#pragma unroll
                for (int c = 0; c < 12; ++c) {
                    const int e = o[c] - s;
                    sq[c] = (unsigned)(e * e);
                    csq[c] = 0u;
                }
            }
            if (!a.mono)
                for (int pl = 1; pl < 3; ++pl)
                    if (lane < cn) {
                        const int px = (x >> a.sub_x) + lane % cbw, py = (y >> a.sub_y) + lane / cbw;
                        const int tw = cbw + 4;
                        const int* Tc = s_tile[pl] + (lane / cbw + 2) * tw + lane % cbw + 2;
                        px12([&](int dx, int dy) { return Tc[dy * tw + dx]; }, dir, -1, a.damping - 1, a.bd, o);
                        const int s = a.src[pl].p[(size_t)py * a.src[pl].stride + px];
#pragma unroll
                        for (int c = 0; c < 12; ++c) csq[c] += (unsigned)((o[c] - s) * (o[c] - s));
                    }
            // 64 x 4095^2 < 2^32 per plane; U + V summed in 64-bit below
#pragma unroll
            for (int c = 0; c < 12; ++c) {
                unsigned long long l = sq[c], cc = csq[c];
                for (int off = 32; off > 0; off >>= 1) {
                    l += __shfl_xor(l, off);
                    cc += __shfl_xor(cc, off);
                }
                sq[c] = 0u;
                if (lane == c) atomicAdd((unsigned long long*)&a.ltab[(size_t)u * 12 + c], l);
                if (!a.mono && lane == 16 + c) atomicAdd((unsigned long long*)&a.ctab[(size_t)u * 12 + c], cc);
            }
            continue;
        }
        {
            const int px = x + (lane & 7), py = y + (lane >> 3);
            const int* T0 = s_tile[0] + ((lane >> 3) + 2) * 12 + (lane & 7) + 2;
            px12([&](int dx, int dy) { return T0[dy * 12 + dx]; }, dir, a.vars[b], a.damping, a.bd, o);
            const int s = a.src[0].p[(size_t)py * a.src[0].stride + px];
#pragma unroll
            for (int c = 0; c < 12; ++c) s_err[0][c][lane] = o[c] - s;
        }
        if (!a.mono)
            for (int pl = 1; pl < 3; ++pl)
                if (lane < cn) {
                    const int px = (x >> a.sub_x) + lane % cbw, py = (y >> a.sub_y) + lane / cbw;
                    const int tw = cbw + 4;
                    const int* Tc = s_tile[pl] + (lane / cbw + 2) * tw + lane % cbw + 2;
                    px12([&](int dx, int dy) { return Tc[dy * tw + dx]; }, dir, -1, a.damping - 1, a.bd, o);
                    const int s = a.src[pl].p[(size_t)py * a.src[pl].stride + px];
#pragma unroll
                    for (int c = 0; c < 12; ++c) s_err[pl][c][lane] = o[c] - s;
                }
        __syncthreads();
        if (lane < 12) {
            atomicAdd((unsigned long long*)&a.ltab[(size_t)u * 12 + lane],
                      (unsigned long long)block_dist(s_err[0][lane], 8, 8, a.scale, a.T));
        } else if (!a.mono && lane >= 16 && lane < 40) {
            const int pl = 1 + (lane - 16) / 12, c = (lane - 16) % 12;
            atomicAdd((unsigned long long*)&a.ctab[(size_t)u * 12 + c],
                      (unsigned long long)block_dist(s_err[pl][c], cbw, cbh, a.scale, a.T));
        }
        __syncthreads();
    }
}

struct ApplyArgs {
    PlaneV in;
    uint16_t* out;
    int ostride;
    const uint8_t* skip8;
    const int *dirs, *vars;
    const uint8_t* unit_str;  // [units][4]: y pri, y sec, uv pri, uv sec
    int nbx, uc, sub_x, sub_y, luma, damping, bd;
};

__global__ void __launch_bounds__(256) cdef_apply_kernel(ApplyArgs a) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.in.w || y >= a.in.h) return;
    const int lx = x << a.sub_x, ly = y << a.sub_y, bx = lx >> 3, by = ly >> 3, b = by * a.nbx + bx;
    const uint8_t* st = a.unit_str + 4 * ((by >> 3) * a.uc + (bx >> 3));
    const int pri = st[a.luma ? 0 : 2], sec = st[a.luma ? 1 : 3], cs = a.bd - 8;
    int v = a.in.p[(size_t)y * a.in.stride + x];
    if (!a.skip8[b] && (pri || sec)) {
        const int pe = a.luma ? adjust_pri(pri << cs, a.vars[b]) : pri << cs;
        v = filter_px(a.in, x, y, pe, sec << cs, pri ? a.dirs[b] : 0, a.damping, a.bd);
    }
    a.out[(size_t)y * a.ostride + x] = (uint16_t)v;
}

// Per 64x64 unit: signalled (some non-skip 8x8) and trusted by the directional-variance gate
// (mean variance of its non-skip 8x8s >= 15000, src/coder/frame_cdef.rs:163-177).
__global__ void cdef_units_kernel(const uint8_t* __restrict__ skip8, const int* __restrict__ vars, int nbx, int nby,
                                  int uc, int nu, uint8_t* __restrict__ flags) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int ux = (u % uc) * 8, uy = (u / uc) * 8;
    long long s = 0, n = 0;
    for (int by = uy; by < min(uy + 8, nby); ++by)
        for (int bx = ux; bx < min(ux + 8, nbx); ++bx)
            if (!skip8[by * nbx + bx]) {
                s += vars[by * nbx + bx];
                ++n;
            }
    flags[u] = (uint8_t)((n > 0 ? 1 : 0) | ((s / (n > 0 ? n : 1) >= 15000) ? 2 : 0));
}

__global__ void zero_i64_kernel(long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Per (device, stream): directions, variances, tables, unit strengths.
struct CdefWork {
    size_t cap_b = 0, cap_u = 0;
    int *dirs = nullptr, *vars = nullptr;
    long long* tab = nullptr;
    uint8_t* ustr = nullptr;
};
std::mutex g_cdef_mu;
std::map<mtlib::StreamKey, CdefWork> g_cdef;

int cdef_work(int dev, hipStream_t s, size_t nb, size_t nu, CdefWork** out) {
    std::lock_guard<std::mutex> lk(g_cdef_mu);
    CdefWork& w = g_cdef[mtlib::stream_key(dev, s)];
    if (w.cap_b < nb) {
        if (w.dirs) MT_HIP_TRY(hipFree(w.dirs));
        if (w.vars) MT_HIP_TRY(hipFree(w.vars));
        w.dirs = w.vars = nullptr;
        w.cap_b = 0;
        MT_HIP_TRY(hipMalloc(&w.dirs, nb * sizeof(int)));
        MT_HIP_TRY(hipMalloc(&w.vars, nb * sizeof(int)));
        w.cap_b = nb;
    }
    if (w.cap_u < nu) {
        if (w.tab) MT_HIP_TRY(hipFree(w.tab));
        if (w.ustr) MT_HIP_TRY(hipFree(w.ustr));
        w.tab = nullptr;
        w.ustr = nullptr;
        w.cap_u = 0;
        MT_HIP_TRY(hipMalloc(&w.tab, nu * 24 * sizeof(long long)));
        MT_HIP_TRY(hipMalloc(&w.ustr, nu * 4));
        w.cap_u = nu;
    }
    *out = &w;
    return MT_OK;
}

const int kCandPri[4] = {0, 1, 2, 4};
const int kCandSec[3] = {0, 1, 2};
int cand_index(int pri, int sec) {
    int pi = 0, si = 0;
    while (kCandPri[pi] != pri) ++pi;
    while (kCandSec[si] != sec) ++si;
    return si * 4 + pi;
}

// The frame decision over the unit tables (src/coder/frame_cdef.rs:140-500).  Returns false
// when CDEF stays off.
bool decide(const std::vector<long long>& lt, const std::vector<long long>& ct, const std::vector<char>& sig,
            const std::vector<char>& trusted, int nu, bool mono, bool slow, float lambda, mt_cdef_params* prm,
            std::vector<uint8_t>& grid) {
    const long long margin = 40;
    int n_sig = 0, n_trusted = 0;
    for (int u = 0; u < nu; ++u) {
        n_sig += sig[u];
        n_trusted += trusted[u];
    }
    if (n_sig == 0) return false;
    auto clears = [&](long long off, long long on) -> long long {
        const long long thr = off - off * margin / 1000;
        return on < thr ? off - on : 0;
    };
    std::vector<std::pair<int, int>> cands{{0, 0}};
    if (slow) {
        const int pl[3] = {1, 2, 4};
        std::vector<std::pair<long long, int>> ranked;
        for (int i = 0; i < 3; ++i) {
            const int c = cand_index(pl[i], 0);
            long long g = 0;
            for (int u = 0; u < nu; ++u) g += clears(lt[(size_t)u * 12], lt[(size_t)u * 12 + c]);
            ranked.push_back({g, i});
            cands.push_back({pl[i], 0});
        }
        std::stable_sort(ranked.begin(), ranked.end(), [](const std::pair<long long, int>& p, const std::pair<long long, int>& q) {
            return p.first != q.first ? p.first > q.first : p.second < q.second;
        });
        for (int sec = 1; sec <= 2; ++sec) {
            cands.push_back({0, sec});
            for (int k = 0; k < 3; ++k)
                if (ranked[k].first > 0) cands.push_back({pl[ranked[k].second], sec});
        }
    } else {
        for (int pri : kCandPri)
            for (int sec : kCandSec)
                if (pri || sec) cands.push_back({pri, sec});
    }
    const int ny = (int)cands.size(), nc = mono ? 1 : ny, npairs = ny * nc;
    std::vector<int> ci(ny);
    for (int i = 0; i < ny; ++i) ci[i] = cand_index(cands[i].first, cands[i].second);
    // every pair's distortion per unit, pair-major: D[p * nu + u]
    std::vector<long long> D((size_t)npairs * nu);
    for (int p = 0; p < npairs; ++p)
        for (int u = 0; u < nu; ++u)
            D[(size_t)p * nu + u] = lt[(size_t)u * 12 + ci[p / nc]] + (mono ? 0 : ct[(size_t)u * 12 + ci[p % nc]]);
    auto dpair = [&](int p, int u) { return D[(size_t)p * nu + u]; };
    std::vector<long long> doff(nu), thr(nu);
    long long total_off = 0, off_unsig = 0;
    std::vector<int> su;  // the signalled units
    for (int u = 0; u < nu; ++u) {
        doff[u] = dpair(0, u);
        thr[u] = doff[u] - doff[u] * margin / 1000;
        total_off += doff[u];
        if (sig[u]) su.push_back(u);
        else off_unsig += doff[u];
    }
    // each pair's total over the signalled units (a few hundred thousand additions per
    // round: cheaper on one host thread than the cost of starting more); argmins take the
    // first minimum in pair order
    auto for_pairs = [&](const std::function<long long(int)>& total, std::vector<long long>& out) {
        out.assign(npairs, 0);
        for (int p = 0; p < npairs; ++p) out[p] = total(p);
    };
    auto argmin = [](const std::vector<long long>& v, long long* best) {
        int bi = 0;
        long long b = INT64_MAX;
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] < b) {
                b = v[i];
                bi = (int)i;
            }
        *best = b;
        return bi;
    };
    std::vector<long long> tot;
    for_pairs([&](int p) {
        long long t = off_unsig;
        const long long* d = &D[(size_t)p * nu];
        for (int u : su) t += d[u];
        return t;
    }, tot);
    long long prefix0;
    const int global_entry = argmin(tot, &prefix0);
    std::vector<long long> cur = doff;
    std::vector<int> entries{0};
    while (entries.size() < 8) {
        for_pairs([&](int p) {
            long long t = off_unsig;
            const long long* d = &D[(size_t)p * nu];
            for (int u : su) {
                const long long b = std::min(cur[u], d[u]);
                t += b >= thr[u] ? doff[u] : b;
            }
            return t;
        }, tot);
        long long best_t;
        const int best_p = argmin(tot, &best_t);
        entries.push_back(best_p);
        for (int u = 0; u < nu; ++u) cur[u] = std::min(cur[u], dpair(best_p, u));
    }
    auto mag = [&](int p) {
        return cands[p / nc].first + cands[p / nc].second + (mono ? 0 : cands[p % nc].first + cands[p % nc].second);
    };
    auto assign = [&](int nb, std::vector<uint8_t>& g) -> long long {
        int z = -1;
        for (int e = 0; e < nb && z < 0; ++e)
            if (entries[e] == 0) z = e;
        long long total = 0;
        g.assign(nu, 0);
        for (int u = 0; u < nu; ++u) {
            if (!sig[u]) {
                total += doff[u];
                continue;
            }
            int pm = -1, pe = 0, re = 0;
            long long pd = 0, rb = INT64_MAX;
            for (int e = 0; e < nb; ++e) {
                const long long d = dpair(entries[e], u);
                if (d < rb) {
                    rb = d;
                    re = e;
                }
                if (d < thr[u]) {
                    const int m = mag(entries[e]);
                    if (pm < 0 || m < pm || (m == pm && d < pd)) {
                        pm = m;
                        pd = d;
                        pe = e;
                    }
                }
            }
            if (pm >= 0) {
                total += pd;
                g[u] = (uint8_t)pe;
            } else if (z >= 0) {
                total += doff[u];
                g[u] = (uint8_t)z;
            } else {
                total += rb;
                g[u] = (uint8_t)re;
            }
        }
        return total;
    };
    const float per_entry = mono ? 6.0f : 12.0f;
    float best_cost = (float)total_off;
    int best_bits = -1;
    std::vector<uint8_t> tmp;
    for (int b = 0; b <= 3; ++b) {
        const int nb = 1 << b;
        const long long total = b == 0 ? prefix0 : assign(nb, tmp);
        if (total == INT64_MAX) continue;
        if (b == 0 && (global_entry == 0 || n_trusted * 4 < n_sig * 3 || total >= total_off - total_off * margin / 1000))
            continue;
        const float cost = (float)total + lambda * (per_entry * (float)(nb - 1) + (float)b * (float)n_sig);
        if (cost < best_cost) {
            best_cost = cost;
            best_bits = b;
        }
    }
    if (best_bits < 0) return false;
    std::vector<int> set;
    grid.assign(nu, 0);
    if (best_bits == 0) {
        set.push_back(global_entry);
    } else {
        set.assign(entries.begin(), entries.begin() + (1 << best_bits));
        assign(1 << best_bits, grid);
    }
    prm->bits = best_bits;
    prm->n = (int)set.size();
    for (size_t e = 0; e < set.size(); ++e) {
        const int p = set[e];
        prm->y_pri[e] = (uint8_t)cands[p / nc].first;
        prm->y_sec[e] = (uint8_t)cands[p / nc].second;
        prm->uv_pri[e] = (uint8_t)(mono ? 0 : cands[p % nc].first);
        prm->uv_sec[e] = (uint8_t)(mono ? 0 : cands[p % nc].second);
    }
    return true;
}

int check_cdef_plane(const uint16_t* p, int stride, int w, int h, const char* what) {
    if (!p || stride < w || w < 1 || h < 1) return mtlib::set_error(MT_ERR_ARG, std::string("mt_cdef_frame: bad ") + what);
    return MT_OK;
}

}  // namespace

extern "C" int mt_cdef_frame(const mt_cdef_frame_args* fa, mt_cdef_params* prm, uint8_t* grid, int64_t* tables,
                             void* stream) {
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    if (!fa || !prm) return mtlib::set_error(MT_ERR_ARG, "mt_cdef_frame: null arguments");
    const int w = fa->width, h = fa->height, bd = fa->bd;
    const bool mono = fa->chroma == MT_RDO_LUMA_ONLY;
    if (w < 8 || h < 8 || (w & 7) || (h & 7) || (bd != 8 && bd != 10 && bd != 12) || fa->chroma > MT_RDO_CHROMA_444 ||
        fa->qidx < 0 || fa->qidx > 255 || !fa->skip8 || !(fa->wsse_scale >= 0.0f))
        return mtlib::set_error(MT_ERR_ARG, "mt_cdef_frame: bad frame (sizes are multiples of 8)");
    const int sub = fa->chroma == MT_RDO_CHROMA_420 ? 1 : 0, cw = w >> sub, ch = h >> sub;
    const uint16_t* in[3] = {fa->rec_y, fa->rec_u, fa->rec_v};
    const uint16_t* src[3] = {fa->src_y, fa->src_u, fa->src_v};
    uint16_t* out[3] = {fa->dst_y, fa->dst_u, fa->dst_v};
    for (int pl = 0; pl < (mono ? 1 : 3); ++pl) {
        const int pw = pl ? cw : w, ph = pl ? ch : h, st = pl ? fa->uv_stride : fa->y_stride;
        const int sst = pl ? fa->src_uv_stride : fa->src_y_stride;
        if ((rc = check_cdef_plane(in[pl], st, pw, ph, "reconstruction")) ||
            (rc = check_cdef_plane(src[pl], sst, pw, ph, "source")) ||
            (rc = check_cdef_plane(out[pl], st, pw, ph, "output")))
            return rc;
        if (out[pl] == in[pl]) return mtlib::set_error(MT_ERR_ARG, "mt_cdef_frame: in place (the filter reads neighbours)");
    }
    hipStream_t s = (hipStream_t)stream;
    const int nbx = w / 8, nby = h / 8, nb = nbx * nby, uc = (w + 63) / 64, ur = (h + 63) / 64, nu = uc * ur;
    CdefWork* wk = nullptr;
    if ((rc = cdef_work(dev, s, (size_t)nb, (size_t)nu, &wk))) return rc;
    const int sig_damp = 3 + std::min(fa->qidx / 64, 3), damping = sig_damp + (bd - 8);
    PlaneV R[3], S[3];
    for (int pl = 0; pl < 3; ++pl) {
        const bool c = pl > 0;
        R[pl] = PlaneV{mono && c ? in[0] : in[pl], c ? fa->uv_stride : fa->y_stride, c ? cw : w, c ? ch : h};
        S[pl] = PlaneV{mono && c ? src[0] : src[pl], c ? fa->src_uv_stride : fa->src_y_stride, c ? cw : w, c ? ch : h};
    }
    hipLaunchKernelGGL(cdef_dir_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, R[0], nbx, nby, bd, wk->dirs, wk->vars);
    hipLaunchKernelGGL(zero_i64_kernel, dim3((nu * 24 + 255) / 256), dim3(256), 0, s, wk->tab, nu * 24);
    TabArgs ta{};
    for (int pl = 0; pl < 3; ++pl) {
        ta.rec[pl] = R[pl];
        ta.src[pl] = S[pl];
    }
    ta.skip8 = fa->skip8;
    ta.dirs = wk->dirs;
    ta.vars = wk->vars;
    ta.nbx = nbx;
    ta.nby = nby;
    ta.uc = uc;
    ta.sub_x = ta.sub_y = sub;
    ta.mono = mono ? 1 : 0;
    ta.damping = damping;
    ta.bd = bd;
    ta.scale = fa->wsse_scale;
    mtlib::wsse_tables(ta.T.d, ta.T.w);
    ta.ltab = wk->tab;
    ta.ctab = wk->tab + (size_t)nu * 12;
    hipLaunchKernelGGL(cdef_tables_kernel, dim3(std::min(nb, cus * 16)), dim3(64), 0, s, ta);
    MT_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(cdef_units_kernel, dim3((nu + 255) / 256), dim3(256), 0, s, fa->skip8, wk->vars, nbx, nby, uc,
                       nu, wk->ustr);
    MT_HIP_TRY(hipGetLastError());
    std::vector<long long> tab((size_t)nu * 24);
    std::vector<uint8_t> uflags(nu);
    MT_HIP_TRY(hipMemcpyAsync(tab.data(), wk->tab, tab.size() * sizeof(long long), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(uflags.data(), wk->ustr, (size_t)nu, hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipStreamSynchronize(s));
    if (tables) std::copy(tab.begin(), tab.end(), tables);
    // signalled units and the directional-variance gate (src/coder/frame_cdef.rs:120-177)
    std::vector<char> sig(nu, 0), trusted(nu, 0);
    for (int u = 0; u < nu; ++u) {
        sig[u] = (uflags[u] & 1) ? 1 : 0;
        trusted[u] = (uflags[u] & 2) ? 1 : 0;
    }
    std::vector<long long> lt(tab.begin(), tab.begin() + (size_t)nu * 12), ct(tab.begin() + (size_t)nu * 12, tab.end());
    *prm = mt_cdef_params{};
    prm->damping = sig_damp;
    prm->n = 1;
    std::vector<uint8_t> g;
    const bool on = decide(lt, ct, sig, trusted, nu, mono, fa->slow != 0, fa->lambda, prm, g);
    prm->on = on ? 1 : 0;
    if (!on) g.assign(nu, 0);
    if (grid) std::copy(g.begin(), g.end(), grid);
    std::vector<uint8_t> ustr((size_t)nu * 4);
    for (int u = 0; u < nu; ++u) {
        const int e = g[u];
        ustr[4 * u] = on ? prm->y_pri[e] : 0;
        ustr[4 * u + 1] = on ? prm->y_sec[e] : 0;
        ustr[4 * u + 2] = on ? prm->uv_pri[e] : 0;
        ustr[4 * u + 3] = on ? prm->uv_sec[e] : 0;
    }
    MT_HIP_TRY(hipMemcpyAsync(wk->ustr, ustr.data(), ustr.size(), hipMemcpyHostToDevice, s));
    for (int pl = 0; pl < (mono ? 1 : 3); ++pl) {
        ApplyArgs aa{};
        aa.in = R[pl];
        aa.out = out[pl];
        aa.ostride = pl ? fa->uv_stride : fa->y_stride;
        aa.skip8 = fa->skip8;
        aa.dirs = wk->dirs;
        aa.vars = wk->vars;
        aa.unit_str = wk->ustr;
        aa.nbx = nbx;
        aa.uc = uc;
        aa.sub_x = aa.sub_y = pl ? sub : 0;
        aa.luma = pl == 0;
        aa.damping = pl ? damping - 1 : damping;
        aa.bd = bd;
        hipLaunchKernelGGL(cdef_apply_kernel, dim3((R[pl].w + 63) / 64, (R[pl].h + 3) / 4), dim3(256), 0, s, aa);
    }
    MT_HIP_TRY(hipGetLastError());
    // the upload's host buffer must outlive the copy
    MT_HIP_TRY(hipStreamSynchronize(s));
    return MT_OK;
}
