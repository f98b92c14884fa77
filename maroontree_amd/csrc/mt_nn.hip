// Learned candidate pruning on the GPU: a small bf16 CNN on MFMA (fp32 accumulation) that
// reads a 64x64 luma superblock plus its qindex and predicts, per quadtree node 64..8, whether
// the exact RDO splits it, and per superblock which intra modes its leaves use.  The masks it
// produces feed mt_intra_rdo's prune argument, at the points where the reference prunes by
// hand (proxy_mode_beam src/coder/partition_search.rs:187, split_breakout_k :135-160,
// rank_luma_modes src/coder/superblock.rs:102).
//
// Network (60,031 parameters, fp32 master copy, bf16 in every product):
//   x[64x64] -> conv 3x3 s2 (1->16) ReLU -> conv 3x3 s2 (16->32) ReLU -> conv 3x3 s2 (32->64)
//   ReLU -> 4x4 average pool [2x2x64] ++ qindex/255 -> fc 257->32 ReLU -> fc 32->863
//   (outputs 10 n + k: logit of AV1 partition type k of node n = 0..84; 850 + j: the score of
//   intra mode j = 0..12).
// Every convolution and fully connected layer, forward and backward, is one call of the
// MFMA GEMM below (im2col / col2im around the convolutions); activations are NHWC.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/mtgpu.h"
#include "mt_common.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// fp32 -> bf16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------- GEMM
// C[m][n] (fp32, ldc) = sum_k A(m, k) B(k, n), A(m, k) = A[m * sam + k * sak],
// B(k, n) = B[k * sbk + n * sbn], both bf16; k in [k0, k1) of the split z = blockIdx.z,
// written to C + z * split_stride.  One 256-thread workgroup computes a 64x64 tile: four
// waves of v_mfma_f32_32x32x8bf16_1k tiles.  Operand layout of the instruction (per lane l,
// item t of its 4): A row l % 32, k = 4 (l / 32) + t; B column l % 32, same k; result
// register a holds row 8 (a / 4) + 4 (l / 32) + a % 4, column l % 32.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const uint16_t* __restrict__ A,
                                                        long sam, long sak, const uint16_t* __restrict__ B,
                                                        long sbk, long sbn, float* __restrict__ C, long ldc,
                                                        int kchunk, long split_stride) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.x * 64 + (wave & 1) * 32, n0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
    const int r = lane & 31, kq = (lane >> 5) * 4;
    const int am = m0 + r, bn = n0 + r;
    v16f acc;
#pragma unroll
    for (int a = 0; a < 16; ++a) acc[a] = 0.0f;
    for (int k = k0; k < k1; k += 8) {
        v4s av, bv;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int kk = k + kq + t;
            const bool ka = kk < k1;
            av[t] = (short)((ka && am < M) ? A[(long)am * sam + (long)kk * sak] : 0);
            bv[t] = (short)((ka && bn < N) ? B[(long)kk * sbk + (long)bn * sbn] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    float* Cz = C + (long)blockIdx.z * split_stride;
    const int cn = n0 + r;
    if (cn >= N) return;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
        const int cm = m0 + 8 * (a >> 2) + kq + (a & 3);
        if (cm < M) Cz[(long)cm * ldc + cn] = acc[a];
    }
}

// C[m][n] = sum over z of P[z][m][n] (in order), for split-K partials (ldc = N).
__global__ void __launch_bounds__(256) splitk_sum_kernel(const float* __restrict__ P, int splits, long mn,
                                                         float* __restrict__ C) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < mn; i += (long)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int z = 0; z < splits; ++z) s += P[(long)z * mn + i];
        C[i] = s;
    }
}

// ---------------------------------------------------------------------------- layers
constexpr int kC1 = 16, kC2 = 32, kC3 = 64, kF = 257, kH = 32, kOut = MT_PNET_OUTPUTS;
constexpr int kNodes = 85, kParts = 10, kModeOff = kNodes * kParts, kLabels = MT_PNET_LABELS;
// parameter offsets in the flat fp32 array (weights [K][N], then biases)
constexpr int kW1 = 0, kB1 = kW1 + 9 * kC1;
constexpr int kW2 = kB1 + kC1, kB2 = kW2 + 9 * kC1 * kC2;
constexpr int kW3 = kB2 + kC2, kB3 = kW3 + 9 * kC2 * kC3;
constexpr int kW4 = kB3 + kC3, kB4 = kW4 + kF * kH;
constexpr int kW5 = kB4 + kH, kB5 = kW5 + kH * kOut;
constexpr int kParams = kB5 + kOut;
static_assert(kParams == MT_PNET_PARAMS, "parameter count");

// Superblock pixels -> normalised bf16 input [B][64][64], and the qindex feature.
__global__ void __launch_bounds__(256) pnet_input_kernel(const uint16_t* __restrict__ pix, const int32_t* __restrict__ idx,
                                                         const uint8_t* __restrict__ qidx, int bd, int B,
                                                         uint16_t* __restrict__ x0, float* __restrict__ qf) {
    const float scale = 1.0f / (float)((1 << bd) - 1);
    const long n = (long)B * 4096;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long b = i >> 12;
        const long s = idx ? idx[b] : b;
        x0[i] = f2bf((float)pix[s * 4096 + (i & 4095)] * scale - 0.5f);
        if ((i & 4095) == 0) qf[b] = (float)qidx[s] * (1.0f / 255.0f);
    }
}

// im2col of a 3x3, stride 2, pad 1 convolution on NHWC input [B][H][H][C] (bf16):
// col[(b, oy, ox)][(ky * 3 + kx) * C + c].
__global__ void __launch_bounds__(256) im2col_kernel(const uint16_t* __restrict__ in, int B, int H, int C,
                                                     uint16_t* __restrict__ col) {
    const int OH = H / 2;
    const long n = (long)B * OH * OH * 9 * C;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int kc = (int)(i % (9 * C));
        const long row = i / (9 * C);
        const int tap = kc / C, c = kc % C;
        const int ox = (int)(row % OH), oy = (int)((row / OH) % OH);
        const long b = row / ((long)OH * OH);
        const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
        col[i] = (iy >= 0 && iy < H && ix >= 0 && ix < H) ? in[((b * H + iy) * H + ix) * C + c] : 0;
    }
}

// Gradient of the input of that convolution from the gradient of its columns (gather: each
// input sample sums the <= 4 column entries that read it), times the ReLU mask of the layer
// that produced the input: dz[b][y][x][c] = (a > 0) * sum dcol.  Output bf16 (next GEMM).
__global__ void __launch_bounds__(256) col2im_relu_kernel(const float* __restrict__ dcol, const uint16_t* __restrict__ act,
                                                          int B, int H, int C, uint16_t* __restrict__ dz_bf,
                                                          float* __restrict__ dz) {
    const int OH = H / 2;
    const long n = (long)B * H * H * C;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int x = (int)((i / C) % H), y = (int)((i / ((long)C * H)) % H);
        const long b = i / ((long)C * H * H);
        float s = 0.0f;
        for (int ky = 0; ky < 3; ++ky) {
            const int t = y + 1 - ky;
            if (t < 0 || (t & 1) || (t >> 1) >= OH) continue;
            for (int kx = 0; kx < 3; ++kx) {
                const int u = x + 1 - kx;
                if (u < 0 || (u & 1) || (u >> 1) >= OH) continue;
                s += dcol[((b * OH + (t >> 1)) * OH + (u >> 1)) * (9L * C) + (ky * 3 + kx) * C + c];
            }
        }
        const float g = bf2f(act[i]) > 0.0f ? s : 0.0f;
        dz[i] = g;
        dz_bf[i] = f2bf(g);
    }
}

// a = ReLU(z + bias) (z [M][N] fp32 in place gets the biased value), as bf16.
__global__ void __launch_bounds__(256) bias_relu_kernel(float* __restrict__ z, const float* __restrict__ bias, long M,
                                                        int N, int relu, uint16_t* __restrict__ a) {
    const long n = M * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = z[i] + bias[i % N];
        z[i] = v;
        if (a) a[i] = f2bf(relu ? fmaxf(v, 0.0f) : v);
    }
}

// 4x4 average pool of the 8x8x64 map -> features [B][257] (bf16): f[(py * 2 + px) * 64 + c],
// then the qindex feature.
__global__ void __launch_bounds__(256) pool_kernel(const uint16_t* __restrict__ a3, const float* __restrict__ qf, int B,
                                                   uint16_t* __restrict__ feat) {
    const long n = (long)B * kF;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long b = i / kF;
        const int f = (int)(i % kF);
        if (f == 256) {
            feat[i] = f2bf(qf[b]);
            continue;
        }
        const int c = f % 64, q = f / 64, py = q >> 1, px = q & 1;
        float s = 0.0f;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) s += bf2f(a3[((b * 8 + py * 4 + y) * 8 + px * 4 + x) * 64 + c]);
        feat[i] = f2bf(s * (1.0f / 16.0f));
    }
}

// dz3[b][y][x][c] = (a3 > 0) * dfeat[b][(y / 4 * 2 + x / 4) * 64 + c] / 16
__global__ void __launch_bounds__(256) unpool_relu_kernel(const float* __restrict__ dfeat, const uint16_t* __restrict__ a3,
                                                          int B, uint16_t* __restrict__ dz_bf, float* __restrict__ dz) {
    const long n = (long)B * 64 * 64;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % 64), x = (int)((i / 64) % 8), y = (int)((i / 512) % 8);
        const long b = i / 4096;
        const float g = bf2f(a3[i]) > 0.0f ? dfeat[b * kF + ((y >> 2) * 2 + (x >> 2)) * 64 + c] * (1.0f / 16.0f) : 0.0f;
        dz[i] = g;
        dz_bf[i] = f2bf(g);
    }
}

// dz = (a > 0) * dA for the fully connected hidden layer (dA [M][N] fp32).
__global__ void __launch_bounds__(256) relu_back_kernel(const float* __restrict__ da, const uint16_t* __restrict__ a, long n,
                                                        uint16_t* __restrict__ dz_bf, float* __restrict__ dz) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float g = bf2f(a[i]) > 0.0f ? da[i] : 0.0f;
        dz[i] = g;
        dz_bf[i] = f2bf(g);
    }
}

// The loss of one label column, one thread each: column n < 85 the masked softmax
// cross-entropy of node n's 10 partition logits with its type t (dlogit_k = w (p_k - [k = t])),
// column 85 + j the masked binary cross-entropy of mode score j (dlogit = w (sigmoid(z) - t)).
// w = mask * norm of the column's head; the weighted loss terms go to loss_terms (summed on
// the host).
__global__ void __launch_bounds__(256) pnet_loss_kernel(const float* __restrict__ z, const float* __restrict__ t,
                                                        const float* __restrict__ m, int B, float norm_split,
                                                        float norm_mode, uint16_t* __restrict__ dz_bf,
                                                        float* __restrict__ dz, float* __restrict__ loss_terms) {
    const long n = (long)B * kLabels;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long b = i / kLabels;
        const int c = (int)(i % kLabels);
        if (c < kNodes) {
            const float w = m[i] * norm_split;
            const long o = b * kOut + c * kParts;
            const int tc = (int)t[i];
            float mx = z[o];
            for (int k = 1; k < kParts; ++k) mx = fmaxf(mx, z[o + k]);
            float se = 0.0f;
            for (int k = 0; k < kParts; ++k) se += expf(z[o + k] - mx);
            for (int k = 0; k < kParts; ++k) {
                const float g = w * (expf(z[o + k] - mx) / se - (k == tc ? 1.0f : 0.0f));
                dz[o + k] = g;
                dz_bf[o + k] = f2bf(g);
            }
            loss_terms[i] = w * (logf(se) + mx - z[o + (tc >= 0 && tc < kParts ? tc : 0)]);
        } else {
            const float w = m[i] * norm_mode;
            const long o = b * kOut + kModeOff + (c - kNodes);
            const float zz = z[o], tt = t[i];
            const float g = w * (1.0f / (1.0f + expf(-zz)) - tt);
            dz[o] = g;
            dz_bf[o] = f2bf(g);
            // log(1 + e^z) - t z, stable
            loss_terms[i] = w * (fmaxf(zz, 0.0f) - zz * tt + log1pf(expf(-fabsf(zz))));
        }
    }
}

// db[n] = sum over rows of dz[M][N], one workgroup per column, fixed order.
__global__ void __launch_bounds__(256) colsum_kernel(const float* __restrict__ dz, long M, int N, float* __restrict__ db) {
    __shared__ float s[256];
    const int n = blockIdx.x;
    float a = 0.0f;
    for (long i = threadIdx.x; i < M; i += 256) a += dz[i * N + n];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = s[0];
}

__global__ void __launch_bounds__(256) to_bf16_kernel(const float* __restrict__ in, long n, uint16_t* __restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) out[i] = f2bf(in[i]);
}

// Adam (decoupled from nothing: plain Adam with bias correction) on the fp32 master weights.
__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                                   float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                                                   float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Extracts the 64x64 superblocks of a frame (raster order; samples beyond the frame edge
// replicate the last row / column) into pix[nsb][4096].
__global__ void __launch_bounds__(256) sb_extract_kernel(mt_plane p, int C, int nsb, uint16_t* __restrict__ pix) {
    const long n = (long)nsb * 4096;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int sb = (int)(i >> 12), k = (int)(i & 4095);
        const int x = min((sb % C) * 64 + (k & 63), p.width - 1), y = min((sb / C) * 64 + (k >> 6), p.height - 1);
        pix[i] = p.data[(size_t)y * p.stride + x];
    }
}

// Per-level policy of the prune masks: split thresholds of levels 0..3 (64x64 .. 8x8
// nodes), the number of intra modes kept at levels 0..4 (64x64 .. 4x4) and the confidence
// below which a level falls back to all modes.
struct SplitThresholds {
    float lo[4], hi[4];
    int topk[5];
    float cover;
};

// Pruning masks from the network's logits, one thread per (SB, node): the SB's top-k intra
// modes by score, k = topk[level of the node] (ties to the lower mode; DC always; filter
// intra with DC) and, for nodes 0..84 of level L, NO_SPLIT when P(split) < lo[L],
// FORCE_SPLIT when P(split) > hi[L].  Nodes 85..340 (4x4) take the mode mask only.
// Confidence fallback of the modes: the mode scores are independent probabilities that a
// leaf of the SB uses the mode, so prod over the dropped modes of (1 - p_j) (ascending j)
// estimates the probability that the kept modes cover every mode the SB uses; below
// `cover` the node keeps all 13 modes.
__global__ void __launch_bounds__(256) prune_mask_kernel(const float* __restrict__ logits, int B,
                                                         SplitThresholds th, uint16_t* __restrict__ prune) {
    const long n = (long)B * MT_RDO_NODES;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long b = i / MT_RDO_NODES;
        const int node = (int)(i % MT_RDO_NODES);
        const float* z = logits + b * kOut;
        const float* zm = z + kModeOff;
        const int lv = node >= 85 ? 4 : node >= 21 ? 3 : node >= 5 ? 2 : node >= 1 ? 1 : 0;
        const int topk = th.topk[lv];
        uint16_t m = (1u << 0) | (1u << 13);
        float covered = 1.0f;
        for (int j = 0; j < 13; ++j) {
            int rank = 0;
            for (int q = 0; q < 13; ++q) rank += (zm[q] > zm[j] || (zm[q] == zm[j] && q < j)) ? 1 : 0;
            if (rank < topk) m |= (uint16_t)(1u << j);
            else if (j != 0) covered *= 1.0f - 1.0f / (1.0f + expf(-zm[j]));
        }
        if (covered < th.cover) m |= 0x1fffu;
        if (node < 85) {
            const int L = node >= 21 ? 3 : node >= 5 ? 2 : node >= 1 ? 1 : 0;
            // the probability that the node is partitioned at all: 1 - softmax(...)[NONE]
            const float* zp = z + node * kParts;
            float mx = zp[0];
            for (int k = 1; k < kParts; ++k) mx = fmaxf(mx, zp[k]);
            float se = 0.0f;
            for (int k = 0; k < kParts; ++k) se += expf(zp[k] - mx);
            const float ps = 1.0f - expf(zp[0] - mx) / se;
            if (ps < th.lo[L]) m |= MT_RDO_NO_SPLIT;
            else if (ps > th.hi[L]) m |= MT_RDO_FORCE_SPLIT;
        }
        prune[i] = m;
    }
}

int grid_for(long n, int cus) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, (long)cus * 8)); }

int gemm(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn,
         float* C, float* split_ws, long split_ws_elems, hipStream_t s) {
    const int gm = (M + 63) / 64, gn = (N + 63) / 64;
    int splits = 1;
    // split a long K when the output tile grid is small
    while (splits < 256 && (long)gm * gn * splits < 512 && K / (splits * 2) >= 256 &&
           (long)M * N * splits * 2 <= split_ws_elems)
        splits *= 2;
    const int kchunk = ((K + splits - 1) / splits + 7) / 8 * 8;
    if (splits == 1) {
        hipLaunchKernelGGL(gemm_bf16_kernel, dim3(gm, gn, 1), dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                           C, (long)N, kchunk, 0L);
    } else {
        hipLaunchKernelGGL(gemm_bf16_kernel, dim3(gm, gn, splits), dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk,
                           sbn, split_ws, (long)N, kchunk, (long)M * N);
        hipLaunchKernelGGL(splitk_sum_kernel, dim3(grid_for((long)M * N, 64)), dim3(256), 0, s, split_ws, splits,
                           (long)M * N, C);
    }
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

// Workspace layout for a batch of B (byte offsets, 256-aligned).
struct Ws {
    size_t wbf, x0, qf, col1, z1, a1, col2, z2, a2, col3, z3, a3, feat, z4, a4, z5;
    size_t dz5f, dz5b, loss, da4, dz4f, dz4b, dfeat, dz3f, dz3b, dcol3, dz2f, dz2b, dcol2, dz1f, dz1b, split, total;
    long split_elems;
};
Ws layout(int B) {
    Ws w{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += (bytes + 255) / 256 * 256;
        return at;
    };
    const long M1 = (long)B * 1024, M2 = (long)B * 256, M3 = (long)B * 64;
    w.wbf = take(kParams * 2);
    w.x0 = take(B * 4096L * 2);
    w.qf = take(B * 4L);
    w.col1 = take(M1 * 9 * 2);
    w.z1 = take(M1 * kC1 * 4);
    w.a1 = take(M1 * kC1 * 2);
    w.col2 = take(M2 * 9 * kC1 * 2);
    w.z2 = take(M2 * kC2 * 4);
    w.a2 = take(M2 * kC2 * 2);
    w.col3 = take(M3 * 9 * kC2 * 2);
    w.z3 = take(M3 * kC3 * 4);
    w.a3 = take(M3 * kC3 * 2);
    w.feat = take((long)B * kF * 2);
    w.z4 = take((long)B * kH * 4);
    w.a4 = take((long)B * kH * 2);
    w.z5 = take((long)B * kOut * 4);
    w.dz5f = take((long)B * kOut * 4);
    w.dz5b = take((long)B * kOut * 2);
    w.loss = take((long)B * kOut * 4);
    w.da4 = take((long)B * kH * 4);
    w.dz4f = take((long)B * kH * 4);
    w.dz4b = take((long)B * kH * 2);
    w.dfeat = take((long)B * kF * 4);
    w.dz3f = take(M3 * kC3 * 4);
    w.dz3b = take(M3 * kC3 * 2);
    w.dcol3 = take(M3 * 9 * kC2 * 4);
    w.dz2f = take(M2 * kC2 * 4);
    w.dz2b = take(M2 * kC2 * 2);
    w.dcol2 = take(M2 * 9 * kC1 * 4);
    w.dz1f = take(M1 * kC1 * 4);
    w.dz1b = take(M1 * kC1 * 2);
    w.split_elems = 1L << 20;
    w.split = take(w.split_elems * 4);
    w.total = o;
    return w;
}

}  // namespace

extern "C" int mt_gemm_bf16(int M, int N, int K, const uint16_t* A_d, long sam, long sak, const uint16_t* B_d,
                            long sbk, long sbn, float* C_d, void* stream) {
    if (M <= 0 || N <= 0 || K <= 0 || !A_d || !B_d || !C_d)
        return mtlib::set_error(MT_ERR_ARG, "mt_gemm_bf16: bad arguments");
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    const int gm = (M + 63) / 64, gn = (N + 63) / 64;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3(gm, gn, 1), dim3(256), 0, (hipStream_t)stream, M, N, K, A_d, sam, sak,
                       B_d, sbk, sbn, C_d, (long)N, (K + 7) / 8 * 8, 0L);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

extern "C" size_t mt_pnet_workspace_bytes(int batch) { return batch > 0 ? layout(batch).total : 0; }

extern "C" int mt_pnet_extract(const mt_plane* plane, uint16_t* pix_d, void* stream) {
    if (!plane || !plane->data || !pix_d || plane->width <= 0 || plane->height <= 0 || plane->stride < plane->width)
        return mtlib::set_error(MT_ERR_ARG, "mt_pnet_extract: bad plane");
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    const int C = (plane->width + 63) / 64, R = (plane->height + 63) / 64;
    hipLaunchKernelGGL(sb_extract_kernel, dim3(grid_for((long)C * R * 4096, cus)), dim3(256), 0, (hipStream_t)stream,
                       *plane, C, C * R, pix_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

extern "C" int mt_pnet_forward(const float* params_d, const uint16_t* pix_d, const int32_t* idx_d,
                               const uint8_t* qidx_d, int bd, int batch, void* work_d, float* logits_d,
                               void* stream) {
    if (!params_d || !pix_d || !qidx_d || !work_d || batch <= 0 || (bd != 8 && bd != 10 && bd != 12))
        return mtlib::set_error(MT_ERR_ARG, "mt_pnet_forward: bad arguments");
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int B = batch;
    const Ws w = layout(B);
    char* base = (char*)work_d;
    auto bf = [&](size_t o) { return (uint16_t*)(base + o); };
    auto f32 = [&](size_t o) { return (float*)(base + o); };
    uint16_t* wbf = bf(w.wbf);
    float* split = f32(w.split);
    const long M1 = (long)B * 1024, M2 = (long)B * 256, M3 = (long)B * 64;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(kParams, cus)), dim3(256), 0, s, params_d, (long)kParams, wbf);
    hipLaunchKernelGGL(pnet_input_kernel, dim3(grid_for((long)B * 4096, cus)), dim3(256), 0, s, pix_d, idx_d, qidx_d,
                       bd, B, bf(w.x0), f32(w.qf));
    // conv1: [M1 x 9] x [9 x 16]
    hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(M1 * 9, cus)), dim3(256), 0, s, bf(w.x0), B, 64, 1, bf(w.col1));
    if ((rc = gemm((int)M1, kC1, 9, bf(w.col1), 9, 1, wbf + kW1, kC1, 1, f32(w.z1), split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(bias_relu_kernel, dim3(grid_for(M1 * kC1, cus)), dim3(256), 0, s, f32(w.z1), params_d + kB1, M1,
                       kC1, 1, bf(w.a1));
    // conv2
    hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(M2 * 9 * kC1, cus)), dim3(256), 0, s, bf(w.a1), B, 32, kC1,
                       bf(w.col2));
    if ((rc = gemm((int)M2, kC2, 9 * kC1, bf(w.col2), 9 * kC1, 1, wbf + kW2, kC2, 1, f32(w.z2), split, w.split_elems, s)))
        return rc;
    hipLaunchKernelGGL(bias_relu_kernel, dim3(grid_for(M2 * kC2, cus)), dim3(256), 0, s, f32(w.z2), params_d + kB2, M2,
                       kC2, 1, bf(w.a2));
    // conv3
    hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(M3 * 9 * kC2, cus)), dim3(256), 0, s, bf(w.a2), B, 16, kC2,
                       bf(w.col3));
    if ((rc = gemm((int)M3, kC3, 9 * kC2, bf(w.col3), 9 * kC2, 1, wbf + kW3, kC3, 1, f32(w.z3), split, w.split_elems, s)))
        return rc;
    hipLaunchKernelGGL(bias_relu_kernel, dim3(grid_for(M3 * kC3, cus)), dim3(256), 0, s, f32(w.z3), params_d + kB3, M3,
                       kC3, 1, bf(w.a3));
    // pool + qindex, fc1, fc2
    hipLaunchKernelGGL(pool_kernel, dim3(grid_for((long)B * kF, cus)), dim3(256), 0, s, bf(w.a3), f32(w.qf), B,
                       bf(w.feat));
    if ((rc = gemm(B, kH, kF, bf(w.feat), kF, 1, wbf + kW4, kH, 1, f32(w.z4), split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(bias_relu_kernel, dim3(grid_for((long)B * kH, cus)), dim3(256), 0, s, f32(w.z4), params_d + kB4,
                       (long)B, kH, 1, bf(w.a4));
    if ((rc = gemm(B, kOut, kH, bf(w.a4), kH, 1, wbf + kW5, kOut, 1, f32(w.z5), split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(bias_relu_kernel, dim3(grid_for((long)B * kOut, cus)), dim3(256), 0, s, f32(w.z5),
                       params_d + kB5, (long)B, kOut, 0, (uint16_t*)nullptr);
    if (logits_d)
        MT_HIP_TRY(hipMemcpyAsync(logits_d, f32(w.z5), (size_t)B * kOut * 4, hipMemcpyDeviceToDevice, s));
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

extern "C" int mt_pnet_backward(const float* params_d, const float* targets_d, const float* mask_d, int batch,
                                float norm_split, float norm_mode, void* work_d, float* grads_d,
                                float* loss_terms_d, void* stream) {
    if (!params_d || !targets_d || !mask_d || !work_d || !grads_d || batch <= 0)
        return mtlib::set_error(MT_ERR_ARG, "mt_pnet_backward: bad arguments");
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int B = batch;
    const Ws w = layout(B);
    char* base = (char*)work_d;
    auto bf = [&](size_t o) { return (uint16_t*)(base + o); };
    auto f32 = [&](size_t o) { return (float*)(base + o); };
    uint16_t* wbf = bf(w.wbf);
    float* split = f32(w.split);
    float* g = grads_d;
    const long M1 = (long)B * 1024, M2 = (long)B * 256, M3 = (long)B * 64;
    hipLaunchKernelGGL(pnet_loss_kernel, dim3(grid_for((long)B * kLabels, cus)), dim3(256), 0, s, f32(w.z5), targets_d, mask_d, B,
                       norm_split, norm_mode, bf(w.dz5b), f32(w.dz5f), loss_terms_d ? loss_terms_d : f32(w.loss));
    // fc2: dW5 = a4^T dz5, db5, da4 = dz5 W5^T
    if ((rc = gemm(kH, kOut, B, bf(w.a4), 1, kH, bf(w.dz5b), kOut, 1, g + kW5, split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(colsum_kernel, dim3(kOut), dim3(256), 0, s, f32(w.dz5f), (long)B, kOut, g + kB5);
    if ((rc = gemm(B, kH, kOut, bf(w.dz5b), kOut, 1, wbf + kW5, 1, kOut, f32(w.da4), split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(relu_back_kernel, dim3(grid_for((long)B * kH, cus)), dim3(256), 0, s, f32(w.da4), bf(w.a4),
                       (long)B * kH, bf(w.dz4b), f32(w.dz4f));
    // fc1
    if ((rc = gemm(kF, kH, B, bf(w.feat), 1, kF, bf(w.dz4b), kH, 1, g + kW4, split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(colsum_kernel, dim3(kH), dim3(256), 0, s, f32(w.dz4f), (long)B, kH, g + kB4);
    if ((rc = gemm(B, kF, kH, bf(w.dz4b), kH, 1, wbf + kW4, 1, kH, f32(w.dfeat), split, w.split_elems, s))) return rc;
    // pool -> conv3
    hipLaunchKernelGGL(unpool_relu_kernel, dim3(grid_for(M3 * kC3, cus)), dim3(256), 0, s, f32(w.dfeat), bf(w.a3), B,
                       bf(w.dz3b), f32(w.dz3f));
    if ((rc = gemm(9 * kC2, kC3, (int)M3, bf(w.col3), 1, 9 * kC2, bf(w.dz3b), kC3, 1, g + kW3, split, w.split_elems, s)))
        return rc;
    hipLaunchKernelGGL(colsum_kernel, dim3(kC3), dim3(256), 0, s, f32(w.dz3f), M3, kC3, g + kB3);
    if ((rc = gemm((int)M3, 9 * kC2, kC3, bf(w.dz3b), kC3, 1, wbf + kW3, 1, kC3, f32(w.dcol3), split, w.split_elems, s)))
        return rc;
    hipLaunchKernelGGL(col2im_relu_kernel, dim3(grid_for(M2 * kC2, cus)), dim3(256), 0, s, f32(w.dcol3), bf(w.a2), B, 16,
                       kC2, bf(w.dz2b), f32(w.dz2f));
    // conv2
    if ((rc = gemm(9 * kC1, kC2, (int)M2, bf(w.col2), 1, 9 * kC1, bf(w.dz2b), kC2, 1, g + kW2, split, w.split_elems, s)))
        return rc;
    hipLaunchKernelGGL(colsum_kernel, dim3(kC2), dim3(256), 0, s, f32(w.dz2f), M2, kC2, g + kB2);
    if ((rc = gemm((int)M2, 9 * kC1, kC2, bf(w.dz2b), kC2, 1, wbf + kW2, 1, kC2, f32(w.dcol2), split, w.split_elems, s)))
        return rc;
    hipLaunchKernelGGL(col2im_relu_kernel, dim3(grid_for(M1 * kC1, cus)), dim3(256), 0, s, f32(w.dcol2), bf(w.a1), B, 32,
                       kC1, bf(w.dz1b), f32(w.dz1f));
    // conv1
    if ((rc = gemm(9, kC1, (int)M1, bf(w.col1), 1, 9, bf(w.dz1b), kC1, 1, g + kW1, split, w.split_elems, s))) return rc;
    hipLaunchKernelGGL(colsum_kernel, dim3(kC1), dim3(256), 0, s, f32(w.dz1f), M1, kC1, g + kB1);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

extern "C" int mt_pnet_adam(float* params_d, const float* grads_d, float* m_d, float* v_d, int step, float lr,
                            float beta1, float beta2, float eps, void* stream) {
    if (!params_d || !grads_d || !m_d || !v_d || step < 1)
        return mtlib::set_error(MT_ERR_ARG, "mt_pnet_adam: bad arguments");
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, params_d, grads_d,
                       m_d, v_d, kParams, lr, beta1, beta2, eps, c1, c2);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

extern "C" int mt_pnet_prune_masks(const float* logits_d, int batch, const int* topk5, const float* lo4,
                                   const float* hi4, float mode_cover, uint16_t* prune_d, void* stream) {
    if (!logits_d || !prune_d || !lo4 || !hi4 || !topk5 || batch <= 0)
        return mtlib::set_error(MT_ERR_ARG, "mt_pnet_prune_masks: bad arguments");
    if (!(mode_cover >= 0.0f && mode_cover <= 1.0f))
        return mtlib::set_error(MT_ERR_ARG, "mt_pnet_prune_masks: mode_cover outside [0, 1]");
    SplitThresholds th;
    th.cover = mode_cover;
    for (int L = 0; L < 5; ++L) {
        if (topk5[L] < 0 || topk5[L] > 13) return mtlib::set_error(MT_ERR_ARG, "mt_pnet_prune_masks: topk");
        th.topk[L] = topk5[L];
    }
    for (int L = 0; L < 4; ++L) {
        if (!(lo4[L] <= hi4[L])) return mtlib::set_error(MT_ERR_ARG, "mt_pnet_prune_masks: lo > hi");
        th.lo[L] = lo4[L];
        th.hi[L] = hi4[L];
    }
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    hipLaunchKernelGGL(prune_mask_kernel, dim3(grid_for((long)batch * MT_RDO_NODES, cus)), dim3(256), 0,
                       (hipStream_t)stream, logits_d, batch, th, prune_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
