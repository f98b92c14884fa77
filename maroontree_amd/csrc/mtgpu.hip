// mtgpu: batched intra candidate transform chain on MI355X (gfx950).
//
// Per (TX size) one kernel template walks tiles of B same-type blocks through LDS:
//   load int16 residual -> forward pass 1 -> forward pass 2 + quant (+QM) + dequant
//   -> inverse row pass -> inverse column pass -> SSE / nnz / checksum.
// Each 1-D transform is one lane's register array; LDS (rows padded to W+1 words) is the
// transpose between passes.  Integer VALU only: the stage-wise rounding of the reference
// encoder (src/dct.rs, src/idct.rs) must stay bit-exact, so nothing here uses MFMA.
// 64x64 has its own kernel: its forward transform is the reference's f64 separable basis.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mtgpu.h"
#include "av1_tables.h"
#include "mt_common.h"
#include "mt_tx1d.h"

using namespace mt;

// ------------------------------------------------------------------------- error state
static thread_local std::string g_err;
static int set_err(int code, const char* what) {
    g_err = what;
    return code;
}
int mtlib::set_error(int code, const std::string& what) {
    g_err = what;
    return code;
}
#define HIP_TRY(expr)                                                               \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            g_err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
            return MT_ERR_DEVICE;                                                   \
        }                                                                           \
    } while (0)

// ------------------------------------------------------------------------- device tables
__constant__ uint16_t c_dcq[3][256];
__constant__ uint16_t c_acq[3][256];
__constant__ int32_t c_fadst4[16];
__constant__ int32_t c_fadst8[64];
__constant__ int32_t c_fadst16[256];
__device__ uint8_t g_iqm[MT_QM_LEVELS][2][MT_QM_USED_SIZE];

enum TxKind { KD = 0, KA = 1, KI = 2 };  // DCT, ADST, identity

// Dense class index of the 13 reference sizes (-1: none) and of the 7 reference types.
__host__ __device__ inline int size_slot(int s) {
    switch (s) {
        case MT_TX_4X4: return 0;
        case MT_TX_4X8: return 1;
        case MT_TX_8X4: return 2;
        case MT_TX_4X16: return 3;
        case MT_TX_16X4: return 4;
        case MT_TX_8X8: return 5;
        case MT_TX_8X16: return 6;
        case MT_TX_16X8: return 7;
        case MT_TX_16X16: return 8;
        case MT_TX_16X32: return 9;
        case MT_TX_32X16: return 10;
        case MT_TX_32X32: return 11;
        case MT_TX_64X64: return 12;
        default: return -1;
    }
}
static const int kSlotSize[13] = {MT_TX_4X4,   MT_TX_4X8,   MT_TX_8X4,   MT_TX_4X16, MT_TX_16X4,
                                  MT_TX_8X8,   MT_TX_8X16,  MT_TX_16X8,  MT_TX_16X16,
                                  MT_TX_16X32, MT_TX_32X16, MT_TX_32X32, MT_TX_64X64};
__host__ __device__ inline int type_slot(int t) {
    switch (t) {
        case MT_TXTP_DCT_DCT: return 0;
        case MT_TXTP_ADST_DCT: return 1;
        case MT_TXTP_DCT_ADST: return 2;
        case MT_TXTP_ADST_ADST: return 3;
        case MT_TXTP_IDTX: return 4;
        case MT_TXTP_V_DCT: return 5;
        case MT_TXTP_H_DCT: return 6;
        default: return -1;
    }
}
__host__ __device__ inline int type_of_slot(int t) {
    return t == 0 ? MT_TXTP_DCT_DCT : t == 1 ? MT_TXTP_ADST_DCT : t == 2 ? MT_TXTP_DCT_ADST
         : t == 3 ? MT_TXTP_ADST_ADST : t == 4 ? MT_TXTP_IDTX : t == 5 ? MT_TXTP_V_DCT
         : MT_TXTP_H_DCT;
}
static const int kSlotType[7] = {MT_TXTP_DCT_DCT, MT_TXTP_ADST_DCT, MT_TXTP_DCT_ADST,
                                 MT_TXTP_ADST_ADST, MT_TXTP_IDTX, MT_TXTP_V_DCT, MT_TXTP_H_DCT};

// The reference's forward coverage (src/dct.rs:2416-2470, :2624-2950).
__host__ __device__ inline bool tx_supported(int s, int t) {
    const bool hyb = t == MT_TXTP_DCT_DCT || t == MT_TXTP_ADST_ADST || t == MT_TXTP_ADST_DCT ||
                     t == MT_TXTP_DCT_ADST;
    const bool all7 = hyb || t == MT_TXTP_V_DCT || t == MT_TXTP_H_DCT || t == MT_TXTP_IDTX;
    switch (s) {
        case MT_TX_4X4: case MT_TX_8X8: case MT_TX_8X16: case MT_TX_16X8: return all7;
        case MT_TX_4X8: case MT_TX_8X4: return hyb;
        case MT_TX_16X16: return hyb || t == MT_TXTP_IDTX;
        case MT_TX_4X16: case MT_TX_16X4: case MT_TX_16X32: case MT_TX_32X16:
        case MT_TX_32X32: case MT_TX_64X64:
            return t == MT_TXTP_DCT_DCT;
        default: return false;
    }
}

static const int kTxW[19] = {4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64};
static const int kTxH[19] = {4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16};

// Offset of a size's inverse QM row inside one 3344-entry level table (src/quant.rs:59).
__host__ __device__ constexpr int qm_offset(int w, int h) {
    return (w == 4 && h == 4) ? 0 : (w == 8 && h == 8) ? 16 : (w == 16 && h == 16) ? 80
         : (w == 32 && h == 32) ? 336 : (w == 4 && h == 8) ? 1360 : (w == 8 && h == 4) ? 1392
         : (w == 8 && h == 16) ? 1424 : (w == 16 && h == 8) ? 1552 : (w == 16 && h == 32) ? 1680
         : (w == 32 && h == 16) ? 2192 : (w == 4 && h == 16) ? 2704 : (w == 16 && h == 4) ? 2768
         : (w == 8 && h == 32) ? 2832 : (w == 32 && h == 8) ? 3088 : -1;
}

// ------------------------------------------------------------------------- per-size rules
// Everything that differs between the reference's per-size forward/inverse functions,
// as compile-time properties of (W, H).
template <int W, int H> struct TxRules {
    static constexpr int N = W * H;
    // forward: multiplier ratio 8/sqrt(W*H) in Q16 for the 2-D types (src/dct.rs:3167-3169)
    static constexpr int32_t ratio = (W == 4 && H == 4) ? 131072
                                   : ((W == 4 && H == 8) || (W == 8 && H == 4)) ? 92682
                                   : ((W == 16 && H == 32) || (W == 32 && H == 16)) ? 23170 : 0;
    // forward: B = 6 headroom shift of the 32-point sizes (src/dct.rs:705, :3134, :3342)
    static constexpr int pre_shift = (W == 32 || H == 32) ? 6 : 0;
    // forward: row transform first (dct8x4 src/dct.rs:3253, dct32x16 src/dct.rs:3342)
    static constexpr bool row_first = (W == 8 && H == 4) || (W == 32 && H == 16);
    // forward post-scale: 0 none, 1 x 1/sqrt2, 2 x 1/2, 3 (x + 32) >> 6, 4 the 32x32 i64 form
    static constexpr int post = ((W == 8 && H == 16) || (W == 16 && H == 8)) ? 1
                              : (W == 16 && H == 16) ? 2
                              : ((W == 16 && H == 32) || (W == 32 && H == 16)) ? 3
                              : (W == 32 && H == 32) ? 4 : 0;
    // inverse: rect2 prescale for 2:1 shapes, inter-pass shift, dequant shift
    static constexpr bool rect2 = (W == 2 * H) || (H == 2 * W);
    static constexpr int mid = (N <= 32) ? 0 : ((W == 16 && H == 16) || (W == 32 && H == 32)) ? 2 : 1;
    static constexpr int dq_shift = (N >= 512) ? 1 : 0;
    static constexpr int qmoff = qm_offset(W, H);
    // tile geometry: B blocks per 256-thread workgroup
    static constexpr int MINWH = W < H ? W : H;
    static constexpr int B0 = 256 / MINWH;
    static constexpr int B1 = 4096 / N;
    static constexpr int B = B0 > B1 ? B0 : B1;
    static constexpr int RS = W + 1;      // padded LDS row stride
    static constexpr int BS = H * RS;     // LDS words per block
};

// ------------------------------------------------------------------------- per-block params
struct BlkParam {
    int32_t dc, ac;          // dequant steps
    int32_t qdc, qac;        // q multipliers, un-ratioed
    int32_t rmin, rmax, cmin, cmax, cf_max;
    int32_t qm_level, qm_chroma, bd;
};

// The quantizer of a chain descriptor: Quant::new_with_qm (src/quant.rs:354) for luma; with
// MT_TXBLK_CHROMA, Quant::new_chroma_with_delta_qm (:394) with the block's (dc, ac) qindex
// deltas from `qdelta` (two int8 per block; both 0 without the array).
__device__ inline void param_from_desc(const mt_txblk& d, BlkParam& p, const int8_t* qdelta,
                                       uint32_t g) {
    const int qm = d.bd_qm & 15;
    const int bdi = (d.bd_qm >> 4) & 3;
    const int bd = bdi == 0 ? 8 : bdi == 1 ? 10 : 12;
    const int q = d.qidx;
    const bool chroma = (d.bd_qm & MT_TXBLK_CHROMA) != 0;
    if (chroma) {
        const int ddc = qdelta ? qdelta[2 * (size_t)g] : 0, dac = qdelta ? qdelta[2 * (size_t)g + 1] : 0;
        p.dc = c_dcq[bdi][min(255, max(0, q + ddc))];
        p.ac = c_acq[bdi][min(255, max(0, q + dac))];
    } else {
        const int dci = q - 8 < 0 ? 0 : q - 8;  // luma DC delta -8 (src/quant.rs:350)
        p.dc = c_dcq[bdi][dci];
        p.ac = c_acq[bdi][q];
    }
    // round(65536 / step) of the reference's f64 (src/quant.rs:362): no step divides
    // 131072 oddly, so the integer round-half-up below is the same value.
    p.qdc = (131072 + p.dc) / (2 * p.dc);
    p.qac = (131072 + p.ac) / (2 * p.ac);
    if (bd <= 8) {
        p.rmin = p.cmin = -32768;
        p.rmax = p.cmax = 32767;
        p.cf_max = 32767;
    } else {
        const int32_t rmax = (1 << (bd + 7)) - 1, cmax = (1 << (bd + 5)) - 1;
        p.rmin = ~rmax;
        p.rmax = rmax;
        p.cmin = ~cmax;
        p.cmax = cmax;
        p.cf_max = rmax;
    }
    p.qm_level = qm;
    p.qm_chroma = chroma ? 1 : 0;
    p.bd = bd;
}

__device__ inline void param_from_quant(const mt_quant& q, BlkParam& p) {
    p.dc = q.dc;
    p.ac = q.ac;
    p.qdc = q.q_mult_dc;
    p.qac = q.q_mult_ac;
    p.rmin = q.rmin;
    p.rmax = q.rmax;
    p.cmin = q.cmin;
    p.cmax = q.cmax;
    p.cf_max = q.cf_max;
    p.qm_level = q.qm_level;
    p.qm_chroma = q.qm_chroma;
    p.bd = q.bd;
}

// ------------------------------------------------------------------------- scalar helpers
// quant_q16 (src/dct.rs:70): half-step dead zone, round to nearest.
__device__ inline int32_t quant_q16(int32_t c, int32_t m) {
    const int64_t prod = (int64_t)c * (int64_t)m;
    const uint64_t mag = prod < 0 ? (uint64_t)0 - (uint64_t)prod : (uint64_t)prod;
    if (mag < 32768) return 0;
    const int32_t l = (int32_t)((mag + 32768) >> 16);
    return prod >= 0 ? l : -l;
}
// quant_fullstep levels (src/dct.rs:373): full-step dead zone.
__device__ inline int32_t quant_full(int32_t c, int32_t m) {
    const int64_t prod = (int64_t)c * (int64_t)m;
    const uint64_t mag = prod < 0 ? (uint64_t)0 - (uint64_t)prod : (uint64_t)prod;
    if (mag < 65536) return 0;
    const int32_t l = (int32_t)((mag + 32768) >> 16);
    return prod >= 0 ? l : -l;
}
// Target of a Q16 level: f32(coeff) * f32(m) * 2^-16, each product rounded (src/dct.rs:1204;
// the x86 selection computes the 1-D family the same way, src/avx/dct.rs:184).
__device__ inline float q16_target(int32_t c, int32_t m) {
    return __fmul_rn(__fmul_rn((float)c, (float)m), 1.0f / 65536.0f);
}
// apply_qmatrix on one coefficient (src/avx/qmatrix.rs:34-66): scale the target by the
// forward weight (1024 + iwt/2) / iwt / 32, then round half away from zero.
__constant__ int32_t c_qmw[256];  // (1024 + iwt / 2) / iwt for every 8-bit iwt (0 at iwt 0)
__device__ inline int32_t qm_apply(float& t, int iwt) {
    const int w = c_qmw[iwt & 255];
    t = __fmul_rn(t, (float)w * (1.0f / 32.0f));
    const float tr = truncf(t);
    const float fr = fabsf(__fsub_rn(t, tr));
    const float step = fr >= 0.5f ? copysignf(1.0f, t) : 0.0f;
    return (int32_t)__fadd_rn(tr, step);
}
// Dequantize one level (src/idct.rs:1576-1590): 24-bit mask, dq shift, cf_max clamp.
__device__ inline int32_t dequant(int32_t lvl, int32_t step, int dq_shift, int32_t cf_max) {
    if (lvl == 0) return 0;
    const uint32_t a = lvl < 0 ? 0u - (uint32_t)lvl : (uint32_t)lvl;
    // the low 24 bits of a product depend only on the low 24 bits of its operands
    int32_t mag = (int32_t)((__umul24(a, (uint32_t)step) & 0xffffffu) >> dq_shift);
    const int32_t lim = cf_max + (lvl < 0 ? 1 : 0);
    mag = mag < lim ? mag : lim;
    return lvl < 0 ? -mag : mag;
}
// Order-independent block checksum: a sum over coefficients and samples of
// fold(value * odd 24-bit weight of its position), all full-rate 24-bit products.
__device__ inline uint32_t hw24(uint32_t p) { return (__umul24(p, 0x9E3779u) & 0xFFFFFFu) | 1u; }
__device__ inline uint32_t fold13(uint32_t t) { return t ^ (t >> 13); }
__device__ inline uint32_t coef_hash(int rc, int32_t lvl, float t) {
    const uint32_t tb = __float_as_uint(t);
    return fold13(__umul24((uint32_t)lvl, hw24((uint32_t)rc))) +
           fold13(__umul24(tb, hw24((uint32_t)rc + 4096u)) ^ (tb >> 24));
}
__device__ inline uint32_t pix_hash(int j, int32_t r) { return fold13(__umul24((uint32_t)r, hw24((uint32_t)j + 8192u))); }

// One forward 1-D transform of length L on a register array.
template <int L> __device__ inline void fwd_line(bool adst, int32_t* v) {
    if (adst) {
        if constexpr (L == 4) fadst<4>(c_fadst4, v);
        else if constexpr (L == 8) fadst<8>(c_fadst8, v);
        else if constexpr (L == 16) fadst<16>(c_fadst16, v);
    } else {
        FDct<L>::run(v);
    }
}

// One inverse 1-D transform of length L on a register array.
template <int L> __device__ inline void inv_line(int kind, int32_t* v, Clip c) {
    if (kind == KI) {
#pragma unroll
        for (int i = 0; i < L; i++) v[i] = L == 4 ? iid4(v[i]) : L == 8 ? iid8(v[i]) : iid16(v[i]);
    } else if (kind == KA) {
        if constexpr (L == 4) iadst4(v);
        else if constexpr (L == 8) iadst8(v, c);
        else if constexpr (L == 16) iadst16(v, c);
    } else {
        if constexpr (L == 4) idct4<1>(v, c);
        else if constexpr (L == 8) idct8<1>(v, c);
        else if constexpr (L == 16) idct16<1>(v, c);
        else if constexpr (L == 32) idct32<1>(v, c);
    }
}

// ------------------------------------------------------------------------- tile kernel
// MODE_CHAIN: the whole chain of a bucketed batch; MODE_FWD / MODE_INV: one class of
// contiguous blocks (the parity entry points); MODE_CFWD / MODE_CINV: the chain split around a
// pass that rewrites the levels (the trellis): the forward half writes levels and targets at
// each block's pool offset, the inverse half reads them back and measures like MODE_CHAIN.
enum { MODE_CHAIN = 0, MODE_FWD = 1, MODE_INV = 2, MODE_CFWD = 3, MODE_CINV = 4 };
template <int MODE> struct ModeTraits {
    static constexpr bool bucketed = MODE == MODE_CHAIN || MODE == MODE_CFWD || MODE == MODE_CINV;
    static constexpr bool fwd = MODE != MODE_INV && MODE != MODE_CINV;
    static constexpr bool inv = MODE != MODE_FWD && MODE != MODE_CFWD;
    static constexpr bool measure = MODE == MODE_CHAIN || MODE == MODE_CINV;
};

struct KArgs {
    // chain mode
    const mt_txblk* blocks;
    const uint32_t* order;       // bucketed block indices
    const uint32_t* cls_start;   // [13*7] first index of each class in `order`
    const uint32_t* cls_count;   // [13*7]
    const int16_t* pool16;
    mt_txres* res;
    // parity modes (contiguous per block) and optional chain outputs (pool layout)
    const int32_t* in32;
    const mt_quant* quants;
    int32_t* levels;
    float* targets;
    int32_t* recon;
    const int8_t* qdelta;  // chain mode, optional: chroma (dc, ac) qindex deltas per block
    int n;      // parity modes: block count
    int type;   // parity modes: tx type
};

// TS scales the tile: B << 1 blocks per workgroup (TS = 1), B (0) or B >> 1 (-1).  Every
// line loop strides over the tile's blocks with the 256 threads, so any B gives the same
// results; the launch schedule of each size picks it.
template <int W, int H, int MODE, int TS = 0>
__global__ void __launch_bounds__(256) tx_tile_kernel(KArgs a, int size_slot_idx) {
    using R = TxRules<W, H>;
    using M = ModeTraits<MODE>;
    constexpr int N = R::N, RS = R::RS, BS = R::BS;
    constexpr int B = TS > 0 ? R::B << 1 : TS < 0 ? R::B >> 1 : R::B;
    constexpr int NT = 256;

    __shared__ int32_t work[B * BS];
    __shared__ int16_t in16[M::measure ? B * N : 1];
    __shared__ BlkParam prm[B];
    __shared__ uint32_t gidx[B];      // global block index (chain) or block number (parity)
    __shared__ uint32_t goff[B];      // element offset of the block in the pool / arrays
    __shared__ unsigned long long s_sse[B];
    __shared__ uint32_t s_nnz[B], s_hash[B];
    __shared__ int s_nb, s_type;

    const int tid = threadIdx.x;

    // virtual tile space: chain mode walks the 7 type classes of this size
    __shared__ uint32_t s_cnt[7], s_st[7];
    if (M::bucketed && tid < 7) {
        s_cnt[tid] = a.cls_count[size_slot_idx * 7 + tid];
        s_st[tid] = a.cls_start[size_slot_idx * 7 + tid];
    }
    __syncthreads();
    int total_tiles = 0;
    if (M::bucketed) {
#pragma unroll
        for (int t = 0; t < 7; t++) total_tiles += (int)((s_cnt[t] + B - 1) / B);
    } else {
        total_tiles = (a.n + B - 1) / B;
    }

    for (int tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        // ---- tile setup: which blocks, their parameters
        if (tid == 0) {
            if (M::bucketed) {
                int rem = tile, t = 0;
                for (; t < 7; t++) {
                    const int nt = (int)((s_cnt[t] + B - 1) / B);
                    if (rem < nt) break;
                    rem -= nt;
                }
                const uint32_t first = (uint32_t)rem * B;
                s_type = type_of_slot(t);
                s_nb = (int)min((uint32_t)B, s_cnt[t] - first);
                gidx[0] = s_st[t] + first;  // position in `order` of the first block
            } else {
                s_type = a.type;
                s_nb = min(B, a.n - tile * B);
                gidx[0] = (uint32_t)tile * B;
            }
        }
        __syncthreads();
        const int nb = s_nb;
        const int type = s_type;
        const uint32_t base = gidx[0];
        __syncthreads();
        for (int b = tid; b < B; b += NT) {
            if (b < nb) {
                if (M::bucketed) {
                    const uint32_t g = a.order[base + b];
                    const mt_txblk d = a.blocks[g];
                    gidx[b] = g;
                    goff[b] = d.offset;
                    param_from_desc(d, prm[b], a.qdelta, g);
                } else {
                    const uint32_t g = base + b;
                    gidx[b] = g;
                    goff[b] = g;
                    param_from_quant(a.quants[g], prm[b]);
                }
            }
            s_sse[b] = 0;
            s_nnz[b] = 0;
            s_hash[b] = 0;
        }
        __syncthreads();

        const bool is_idtx = type == MT_TXTP_IDTX;
        const bool is_1d = type == MT_TXTP_V_DCT || type == MT_TXTP_H_DCT;
        const bool col_adst = type == MT_TXTP_ADST_ADST || type == MT_TXTP_ADST_DCT;
        const bool row_adst = type == MT_TXTP_ADST_ADST || type == MT_TXTP_DCT_ADST;
        // the identity family is quantized and dequantized flat (src/quant.rs:300)
        const bool use_qm_type = !(is_idtx || is_1d);

        // Quantize one coefficient (fx, fy) of block b, record it, and leave its dequantized
        // (and rect2-prescaled) value for the inverse in work[b][fy][fx].  The block's hash
        // and nonzero count gather in the caller's registers (hs, nz) and reach LDS once per
        // line (flush): both are sums, so the order does not matter.
        auto flush = [&](int b, uint32_t hs, uint32_t nz) {
            if (MODE == MODE_CHAIN) {
                atomicAdd(&s_hash[b], hs);
                if (nz) atomicAdd(&s_nnz[b], nz);
            }
        };
        auto emit_coef = [&](int b, int fx, int fy, int32_t lvl, float tgt, uint32_t& hs, uint32_t& nz) {
            const int rc = fx * H + fy;
            const BlkParam& p = prm[b];
            if (MODE == MODE_FWD || MODE == MODE_CFWD || (MODE == MODE_CHAIN && a.levels)) {
                const size_t o = (MODE == MODE_FWD ? (size_t)gidx[b] * N : (size_t)goff[b]) + rc;
                a.levels[o] = lvl;
                if (MODE == MODE_FWD || MODE == MODE_CFWD || a.targets) a.targets[o] = tgt;
            }
            if (MODE == MODE_CHAIN) {
                hs += coef_hash(rc, lvl, tgt);
                nz += lvl != 0;
            }
            if (M::inv) {
                int32_t step = rc == 0 ? p.dc : p.ac;
                if (use_qm_type && p.qm_level < 15)
                    step = (step * (int32_t)g_iqm[p.qm_level][p.qm_chroma][R::qmoff + rc] + 16) >> 5;
                int32_t c = dequant(lvl, step, R::dq_shift, p.cf_max);
                if (R::rect2) c = (c * 181 + 128) >> 8;
                work[b * BS + fy * RS + fx] = c;
            }
        };
        // Level and target of a 2-D-type coefficient (store_quant_target + apply_qmatrix).
        auto quant2d = [&](int b, int fx, int fy, int32_t coeff, uint32_t& hs, uint32_t& nz) {
            const BlkParam& p = prm[b];
            const int rc = fx * H + fy;
            int32_t m = rc == 0 ? p.qdc : p.qac;
            if (R::ratio) m = mulq16(m, R::ratio);
            int32_t lvl = quant_q16(coeff, m);
            float tgt = q16_target(coeff, m);
            if (p.qm_level < 15) lvl = qm_apply(tgt, g_iqm[p.qm_level][p.qm_chroma][R::qmoff + rc]);
            emit_coef(b, fx, fy, lvl, tgt, hs, nz);
        };

        if (M::fwd) {
            // ---- load the residual into LDS (row-major, padded rows)
            for (int i = tid; i < nb * N; i += NT) {
                const int b = i / N, p = i % N, y = p / W, x = p % W;
                int32_t v;
                if (M::bucketed) {
                    const int16_t s = a.pool16[(size_t)goff[b] + p];
                    if (M::measure) in16[b * N + p] = s;
                    v = s;
                } else {
                    v = a.in32[(size_t)gidx[b] * N + p];
                }
                work[b * BS + y * RS + x] = v;
            }
            __syncthreads();

            if (is_idtx) {
                // identity_quant_t_direct (src/dct.rs:1586): division by the dequant step.
                for (int i = tid; i < nb * N; i += NT) {
                    const int b = i / N, p = i % N, y = p / W, x = p % W;
                    const BlkParam& pp = prm[b];
                    const int rc = y + x * H;
                    const int32_t qd = rc == 0 ? pp.dc : pp.ac;
                    const int32_t num = work[b * BS + y * RS + x] * 8;
                    const float tgt = __fdiv_rn((float)num, (float)qd);
                    const int32_t am = num < 0 ? -num : num;
                    int32_t lvl = 0;
                    if (am >= qd) {
                        const int32_t l = (am + qd / 2) / qd;
                        lvl = num < 0 ? -l : l;
                    }
                    uint32_t hs = 0, nz = 0;
                    emit_coef(b, x, y, lvl, tgt, hs, nz);
                    flush(b, hs, nz);
                }
            } else if (is_1d) {
                // f1ddct*: DCT along one axis, identity along the other, per-length scale,
                // full-step quantizer, un-ratioed multipliers (src/dct.rs:328, :391, :1044).
                const bool vertical = type == MT_TXTP_V_DCT;
                if (vertical) {
                    for (int line = tid; line < nb * W; line += NT) {
                        const int b = line / W, x = line % W;
                        int32_t v[H];
#pragma unroll
                        for (int y = 0; y < H; y++) v[y] = work[b * BS + y * RS + x];
                        if constexpr (H <= 16) FDct<H>::run(v);
                        const BlkParam& p = prm[b];
                        uint32_t hs = 0, nz = 0;
#pragma unroll
                        for (int y = 0; y < H; y++) {
                            int32_t f;
                            if (W == 4 && H == 4) f = mulq16(v[y] * 2, kSqrt2Q16);
                            else if (H == 8) f = v[y] * 2;
                            else f = mulq16(v[y], kSqrt2Q16);
                            const int rc = y + x * H;
                            const int32_t m = rc == 0 ? p.qdc : p.qac;
                            emit_coef(b, x, y, quant_full(f, m), q16_target(f, m), hs, nz);
                        }
                        flush(b, hs, nz);
                    }
                } else {
                    for (int line = tid; line < nb * H; line += NT) {
                        const int b = line / H, y = line % H;
                        int32_t v[W];
#pragma unroll
                        for (int x = 0; x < W; x++) v[x] = work[b * BS + y * RS + x];
                        if constexpr (W <= 16) FDct<W>::run(v);
                        const BlkParam& p = prm[b];
                        uint32_t hs = 0, nz = 0;
#pragma unroll
                        for (int x = 0; x < W; x++) {
                            int32_t f;
                            if (W == 4 && H == 4) f = mulq16(v[x] * 2, kSqrt2Q16);
                            else if (W == 8) f = v[x] * 2;
                            else f = mulq16(v[x], kSqrt2Q16);
                            const int rc = y + x * H;
                            const int32_t m = rc == 0 ? p.qdc : p.qac;
                            emit_coef(b, x, y, quant_full(f, m), q16_target(f, m), hs, nz);
                        }
                        flush(b, hs, nz);
                    }
                }
            } else if (!R::row_first) {
                // column pass (length H), then row pass (length W) + quantization
                for (int line = tid; line < nb * W; line += NT) {
                    const int b = line / W, x = line % W;
                    int32_t v[H];
#pragma unroll
                    for (int y = 0; y < H; y++)
                        v[y] = (int32_t)((uint32_t)work[b * BS + y * RS + x] << R::pre_shift);
                    fwd_line<H>(col_adst, v);
#pragma unroll
                    for (int y = 0; y < H; y++) work[b * BS + y * RS + x] = v[y];
                }
                __syncthreads();
                for (int line = tid; line < nb * H; line += NT) {
                    const int b = line / H, fy = line % H;
                    int32_t v[W];
#pragma unroll
                    for (int x = 0; x < W; x++) v[x] = work[b * BS + fy * RS + x];
                    fwd_line<W>(row_adst, v);
                    uint32_t hs = 0, nz = 0;
#pragma unroll
                    for (int fx = 0; fx < W; fx++) {
                        int32_t c = v[fx];
                        if (R::post == 1) c = mulq16(c, 46341);
                        else if (R::post == 2) c = mulq16(c, 32768);
                        else if (R::post == 3) c = (c + 32) >> 6;
                        else if (R::post == 4) c = (int32_t)(((int64_t)c * 16384 + (1ll << 21)) >> 22);
                        quant2d(b, fx, fy, c, hs, nz);
                    }
                    flush(b, hs, nz);
                }
            } else {
                // row pass (length W), then column pass (length H) + quantization
                for (int line = tid; line < nb * H; line += NT) {
                    const int b = line / H, y = line % H;
                    int32_t v[W];
#pragma unroll
                    for (int x = 0; x < W; x++)
                        v[x] = (int32_t)((uint32_t)work[b * BS + y * RS + x] << R::pre_shift);
                    fwd_line<W>(row_adst, v);
#pragma unroll
                    for (int x = 0; x < W; x++) work[b * BS + y * RS + x] = v[x];
                }
                __syncthreads();
                for (int line = tid; line < nb * W; line += NT) {
                    const int b = line / W, fx = line % W;
                    int32_t v[H];
#pragma unroll
                    for (int y = 0; y < H; y++) v[y] = work[b * BS + y * RS + fx];
                    fwd_line<H>(col_adst, v);
                    uint32_t hs = 0, nz = 0;
#pragma unroll
                    for (int fy = 0; fy < H; fy++) {
                        int32_t c = v[fy];
                        if (R::post == 3) c = (c + 32) >> 6;
                        quant2d(b, fx, fy, c, hs, nz);
                    }
                    flush(b, hs, nz);
                }
            }
            __syncthreads();
        } else if (MODE == MODE_INV) {
            // ---- inverse-only: dequantize the given levels into LDS
            const bool use_qm = use_qm_type;
            for (int i = tid; i < nb * N; i += NT) {
                const int b = i / N, rc = i % N, fx = rc / H, fy = rc % H;
                const BlkParam& p = prm[b];
                const int32_t lvl = a.in32[(size_t)gidx[b] * N + rc];
                int32_t step = rc == 0 ? p.dc : p.ac;
                if (use_qm && p.qm_level < 15)
                    step = (step * (int32_t)g_iqm[p.qm_level][p.qm_chroma][R::qmoff + rc] + 16) >> 5;
                int32_t c = dequant(lvl, step, R::dq_shift, p.cf_max);
                if (R::rect2) c = (c * 181 + 128) >> 8;
                work[b * BS + fy * RS + fx] = c;
            }
            __syncthreads();
        } else {
            // ---- the split chain's inverse half: the residual for the distortion, and the
            // (rewritten) levels with their targets from the pool, one coefficient column
            // (fx) per line so the checksum and nonzero count reach LDS once per line
            for (int i = tid; i < nb * N; i += NT) {
                const int b = i / N, p = i % N;
                in16[b * N + p] = a.pool16[(size_t)goff[b] + p];
            }
            const bool use_qm = use_qm_type;
            for (int line = tid; line < nb * W; line += NT) {
                const int b = line / W, fx = line % W;
                const BlkParam& p = prm[b];
                uint32_t hs = 0, nz = 0;
                for (int fy = 0; fy < H; fy++) {
                    const int rc = fx * H + fy;
                    const size_t o = (size_t)goff[b] + rc;
                    const int32_t lvl = a.levels[o];
                    hs += coef_hash(rc, lvl, a.targets[o]);
                    nz += lvl != 0;
                    int32_t step = rc == 0 ? p.dc : p.ac;
                    if (use_qm && p.qm_level < 15)
                        step = (step * (int32_t)g_iqm[p.qm_level][p.qm_chroma][R::qmoff + rc] + 16) >> 5;
                    int32_t c = dequant(lvl, step, R::dq_shift, p.cf_max);
                    if (R::rect2) c = (c * 181 + 128) >> 8;
                    work[b * BS + fy * RS + fx] = c;
                }
                atomicAdd(&s_hash[b], hs);
                if (nz) atomicAdd(&s_nnz[b], nz);
            }
            __syncthreads();
        }

        if (M::inv) {
            // inverse kinds: ADST_DCT is vertical ADST (column pass)
            int rowk = KD, colk = KD;
            if (type == MT_TXTP_ADST_ADST) { rowk = KA; colk = KA; }
            else if (type == MT_TXTP_ADST_DCT) { rowk = KD; colk = KA; }
            else if (type == MT_TXTP_DCT_ADST) { rowk = KA; colk = KD; }
            else if (type == MT_TXTP_V_DCT) { rowk = KI; colk = KD; }
            else if (type == MT_TXTP_H_DCT) { rowk = KD; colk = KI; }
            else if (type == MT_TXTP_IDTX) { rowk = KI; colk = KI; }

            // ---- inverse row pass (length W) with the row clip, inter-pass shift + clip
            for (int line = tid; line < nb * H; line += NT) {
                const int b = line / H, y = line % H;
                const BlkParam& p = prm[b];
                int32_t v[W];
#pragma unroll
                for (int x = 0; x < W; x++) v[x] = work[b * BS + y * RS + x];
                inv_line<W>(rowk, v, Clip{p.rmin, p.rmax});
                const Clip cc{p.cmin, p.cmax};
#pragma unroll
                for (int x = 0; x < W; x++) {
                    int32_t t = v[x];
                    if (R::mid) t = (t + (1 << (R::mid - 1))) >> R::mid;
                    work[b * BS + y * RS + x] = cc(t);
                }
            }
            __syncthreads();
            // ---- inverse column pass (length H), final (t + 8) >> 4, distortion
            for (int line = tid; line < nb * W; line += NT) {
                const int b = line / W, x = line % W;
                const BlkParam& p = prm[b];
                int32_t v[H];
#pragma unroll
                for (int y = 0; y < H; y++) v[y] = work[b * BS + y * RS + x];
                inv_line<H>(colk, v, Clip{p.cmin, p.cmax});
                uint32_t sse = 0, hsh = 0;
                const int P = 1 << (p.bd - 1), maxv = (1 << p.bd) - 1;
#pragma unroll
                for (int y = 0; y < H; y++) {
                    const int32_t r = (v[y] + 8) >> 4;
                    const int j = y * W + x;
                    if (MODE == MODE_INV) {
                        a.recon[(size_t)gidx[b] * N + j] = r;
                    } else {
                        if (a.recon) a.recon[(size_t)goff[b] + j] = r;
                        const int32_t s = in16[b * N + j];
                        const int32_t src = min(max(P + s, 0), maxv);
                        const int32_t rec = min(max(P + r, 0), maxv);
                        const int32_t d = src - rec;
                        sse += (uint32_t)__mul24(d, d);  // |d| < 2^12
                        hsh += pix_hash(j, r);
                    }
                }
                if (M::measure) {
                    atomicAdd(&s_sse[b], (unsigned long long)sse);
                    atomicAdd(&s_hash[b], hsh);
                }
            }
        }
        __syncthreads();
        if (M::measure) {
            for (int b = tid; b < nb; b += NT) {
                mt_txres r;
                r.sse = (int64_t)s_sse[b];
                r.nnz = s_nnz[b];
                r.hash = s_hash[b];
                a.res[gidx[b]] = r;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------- 64x64
// TX_64X64: the reference codes the low 32x32 of a 64-point DCT.  Forward = f64 separable
// basis (dct64x64_coeffs, src/dct.rs:4155: vertical pass accumulated over y, then an
// ordered sum over x, round half away from zero of 8x the sum); quantizer with the 32x32
// QM; inverse = idct_dequant_64x64 (src/idct.rs:2677).
//
// One wave per block.  Lane x holds the 32 vertical chains of column x in f64 registers
// (basis rows come through the scalar unit from constant memory); the vertical output goes
// through LDS 16 rows at a time to the horizontal pass, where lanes (u, v parity) run 8
// chains each.  Every f64 product and sum is rounded separately in the reference's order.
// LDS stays at 12 KB per wave (the row-pass buffer and the column pass's output halves reuse
// the same 8 KB), so about three blocks share each SIMD.
__constant__ double c_dct64_basis[32 * 64];    // [k][i]
__constant__ double c_dct64_basis_t[64 * 32];  // [i][k], the vertical pass's scalar stream

template <int MODE>
__global__ void __launch_bounds__(64) tx64_kernel(KArgs a) {
    __shared__ double vt[16 * 64];          // half of the vertical output; then the row pass
    __shared__ int32_t coef[32 * 33];       // dequantized coded coefficients [fy][fx]
    __shared__ unsigned long long s_sse;
    __shared__ uint32_t s_nnz, s_hash;
    using M = ModeTraits<MODE>;
    int32_t* rowbuf = reinterpret_cast<int32_t*>(vt);  // 32 x 64 int32 after the forward
    const int lane = threadIdx.x;
    const int nblocks = M::bucketed ? (int)a.cls_count[12 * 7] : a.n;
    const uint32_t chain_start = M::bucketed ? a.cls_start[12 * 7] : 0;

    for (int blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        BlkParam prm;
        uint32_t g, off;
        if (M::bucketed) {
            g = a.order[chain_start + blk];
            const mt_txblk d = a.blocks[g];
            off = d.offset;
            param_from_desc(d, prm, a.qdelta, g);
        } else {
            g = blk;
            off = blk;
            param_from_quant(a.quants[blk], prm);
        }
        if (lane == 0) {
            s_sse = 0;
            s_nnz = 0;
            s_hash = 0;
        }
        uint32_t nnz = 0, hsh = 0;
        const bool qm = prm.qm_level < 15;
        const uint8_t* qmrow = &g_iqm[qm ? prm.qm_level : 0][prm.qm_chroma][336];
        if (M::fwd) {
            // Vertical pass: lane x runs the 32 chains of column x over y ascending; the
            // residual is staged in LDS as int16 (the chain's input type; the parity entry
            // point reads its int32 input directly), the basis arrives wave-uniform through
            // the scalar unit.  Its output goes through LDS in two halves of 16 rows, each
            // followed by the horizontal pass of those rows (lane (u, parity) runs 8 chains
            // over x ascending on LDS broadcasts).
            int16_t* in16 = reinterpret_cast<int16_t*>(vt);
            if (M::bucketed) {
#pragma unroll 16
                for (int y = 0; y < 64; y++) in16[y * 64 + lane] = a.pool16[(size_t)off + y * 64 + lane];
                __syncthreads();
            }
            double acc[32];
#pragma unroll
            for (int j = 0; j < 32; j++) acc[j] = 0.0;
#pragma unroll 1
            for (int y = 0; y < 64; y++) {
                const double c = M::bucketed ? (double)in16[y * 64 + lane]
                                                    : (double)a.in32[(size_t)g * 4096 + y * 64 + lane];
                // the row's 32 basis values in one batch of scalar loads (one wait per row)
                double bv[32];
#pragma unroll
                for (int j = 0; j < 32; j++) bv[j] = c_dct64_basis_t[y * 32 + j];
#pragma unroll
                for (int j = 0; j < 32; j++) acc[j] = __dadd_rn(acc[j], __dmul_rn(bv[j], c));
            }
            __syncthreads();  // the staged residual is dead
#pragma unroll
            for (int hf = 0; hf < 2; hf++) {
#pragma unroll
                for (int j = 0; j < 16; j++) vt[j * 64 + lane] = acc[hf * 16 + j];
                __syncthreads();
                {
                    const int u = lane & 31, vp = lane >> 5;
                    double hacc[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) hacc[k] = -0.0;
                    // the lane's basis values stream one pair of x ahead of their use (the
                    // last prefetch re-reads x = 62, 63 and goes unused)
                    double bn0 = c_dct64_basis_t[u], bn1 = c_dct64_basis_t[32 + u];
#pragma unroll 2
                    for (int x = 0; x < 64; x += 2) {
                        const double b0 = bn0, b1 = bn1;
                        const int xn = x + 2 < 64 ? x + 2 : 62;
                        bn0 = c_dct64_basis_t[xn * 32 + u];
                        bn1 = c_dct64_basis_t[(xn + 1) * 32 + u];
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            hacc[k] = __dadd_rn(hacc[k], __dmul_rn(vt[(2 * k + vp) * 64 + x], b0));
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            hacc[k] = __dadd_rn(hacc[k], __dmul_rn(vt[(2 * k + vp) * 64 + x + 1], b1));
                    }
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        coef[(hf * 16 + 2 * k + vp) * 33 + u] = (int32_t)round(__dmul_rn(hacc[k], 8.0));
                }
                __syncthreads();  // the next half overwrites vt
            }
            // quantize the lane's 16 coefficients, leave their dequantized values in place
            {
                const int u = lane & 31, vp = lane >> 5;
#pragma unroll 4
                for (int k = 0; k < 16; k++) {
                    const int v = 2 * k + vp;
                    const int32_t coeff = coef[v * 33 + u];
                    const int rc = u * 32 + v;
                    const int32_t m = rc == 0 ? prm.qdc : prm.qac;
                    int32_t lvl = quant_q16(coeff, m);
                    float tgt = q16_target(coeff, m);
                    const int iwt = qm ? qmrow[rc] : 32;
                    if (qm) lvl = qm_apply(tgt, iwt);
                    if (MODE == MODE_FWD || MODE == MODE_CFWD || (MODE == MODE_CHAIN && a.levels)) {
                        const size_t o = (MODE == MODE_FWD ? (size_t)g * 1024 : (size_t)off) + rc;
                        a.levels[o] = lvl;
                        if (MODE == MODE_FWD || MODE == MODE_CFWD || a.targets) a.targets[o] = tgt;
                    }
                    if (MODE == MODE_CHAIN) {
                        int32_t step = rc == 0 ? prm.dc : prm.ac;
                        if (qm) step = (step * iwt + 16) >> 5;
                        hsh += coef_hash(rc, lvl, tgt);
                        nnz += lvl != 0;
                        coef[v * 33 + u] = dequant(lvl, step, 2, prm.cf_max);
                    }
                }
            }
            __syncthreads();
        } else {
            for (int rc = lane; rc < 1024; rc += 64) {
                const int u = rc >> 5, v = rc & 31;
                int32_t step = rc == 0 ? prm.dc : prm.ac;
                if (qm) step = (step * (int32_t)qmrow[rc] + 16) >> 5;
                int32_t lvl;
                if (MODE == MODE_CINV) {  // the split chain's inverse half: the pool's levels
                    lvl = a.levels[(size_t)off + rc];
                    hsh += coef_hash(rc, lvl, a.targets[(size_t)off + rc]);
                    nnz += lvl != 0;
                } else {
                    lvl = a.in32[(size_t)g * 1024 + rc];
                }
                coef[v * 33 + u] = dequant(lvl, step, 2, prm.cf_max);
            }
            __syncthreads();
        }
        if (M::inv) {
            // row pass over the 32 coded rows, inter-pass (t + 2) >> 2 clip
            if (lane < 32) {
                const int y = lane;
                int32_t v[64];
#pragma unroll
                for (int x = 0; x < 64; x++) v[x] = x < 32 ? coef[y * 33 + x] : 0;
                idct64<1>(v, Clip{prm.rmin, prm.rmax});
                const Clip cc{prm.cmin, prm.cmax};
#pragma unroll
                for (int x = 0; x < 64; x++) rowbuf[y * 64 + x] = cc((v[x] + 2) >> 2);
            }
            __syncthreads();
            {
                const int x = lane;
                int32_t v[64];
#pragma unroll
                for (int y = 0; y < 64; y++) v[y] = y < 32 ? rowbuf[y * 64 + x] : 0;
                idct64<1>(v, Clip{prm.cmin, prm.cmax});
                if (MODE == MODE_INV) {
#pragma unroll
                    for (int y = 0; y < 64; y++) a.recon[(size_t)g * 4096 + y * 64 + x] = (v[y] + 8) >> 4;
                } else {
                    // column x is measured in two halves of 32 rows, each parked in the
                    // lane's own column of the row-pass buffer (which the lane alone read)
                    uint32_t sse = 0;
                    const int P = 1 << (prm.bd - 1), maxv = (1 << prm.bd) - 1;
#pragma unroll
                    for (int h2 = 0; h2 < 2; h2++) {
#pragma unroll
                        for (int y = 0; y < 32; y++) rowbuf[y * 64 + x] = (v[h2 * 32 + y] + 8) >> 4;
#pragma unroll 8
                        for (int y = 0; y < 32; y++) {
                            const int j = (h2 * 32 + y) * 64 + x;
                            const int32_t r = rowbuf[y * 64 + x];
                            if (a.recon) a.recon[(size_t)off + j] = r;
                            const int32_t src = min(max(P + (int32_t)a.pool16[(size_t)off + j], 0), maxv);
                            const int32_t rec = min(max(P + r, 0), maxv);
                            const int32_t dd = src - rec;
                            sse += (uint32_t)__mul24(dd, dd);  // |dd| < 2^12
                            hsh += pix_hash(j, r);
                        }
                    }
                    if (M::measure) {
                        atomicAdd(&s_sse, (unsigned long long)sse);
                        atomicAdd(&s_hash, hsh);
                        atomicAdd(&s_nnz, nnz);
                    }
                }
            }
        }
        __syncthreads();
        if (M::measure && lane == 0) {
            mt_txres r;
            r.sse = (int64_t)s_sse;
            r.nnz = s_nnz;
            r.hash = s_hash;
            a.res[g] = r;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------- RDO reconstruction
#include "mt_sbrecon.h"

// ------------------------------------------------------------------------- bucketing
// Counting sort of the block list by (size slot, type slot): 13 x 7 classes.  Blocks of an
// unsupported class are answered directly (sse = -1).
constexpr int kClasses = 13 * 7;

// Workspace of mt_tx_chain_noalloc, held by every device the library runs on until the
// process ends: the class counters, starts and cursors (kNoAllocSlots words each) and the
// block order for up to MT_NOALLOC_MAX_BLOCKS blocks (4 MiB + 1.5 KiB).  Host code never
// uses this name as an address: it uses the per-device pointer hipGetSymbolAddress returns.
constexpr int kNoAllocSlots = 128;
static __device__ __align__(256) uint32_t scratch[3 * kNoAllocSlots + MT_NOALLOC_MAX_BLOCKS];

__global__ void __launch_bounds__(256) bucket_count_kernel(const mt_txblk* blocks, int n,
                                                          uint32_t* counts, mt_txres* res) {
    __shared__ uint32_t h[kClasses];
    for (int i = threadIdx.x; i < kClasses; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const mt_txblk d = blocks[i];
        const int ss = size_slot(d.tx_size), ts = type_slot(d.tx_type);
        if (ss >= 0 && ts >= 0 && tx_supported(d.tx_size, d.tx_type)) {
            atomicAdd(&h[ss * 7 + ts], 1u);
        } else {
            mt_txres r;
            r.sse = -1;
            r.nnz = 0;
            r.hash = 0;
            res[i] = r;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kClasses; i += blockDim.x)
        if (h[i]) atomicAdd(&counts[i], h[i]);
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Exclusive scan of the class counts into starts and scatter cursors (one workgroup).
__global__ void bucket_scan_kernel(const uint32_t* counts, uint32_t* starts, uint32_t* cursor) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kClasses; i++) {
            starts[i] = acc;
            cursor[i] = acc;
            acc += counts[i];
        }
    }
}

__global__ void __launch_bounds__(256) bucket_scatter_kernel(const mt_txblk* blocks, int n,
                                                            uint32_t* cursor, uint32_t* order) {
    // Per-workgroup LDS histogram, one global reservation per class, then placement.
    __shared__ uint32_t h[kClasses], basepos[kClasses];
    const int chunk_begin = blockIdx.x * 4096;
    const int chunk_end = min(n, chunk_begin + 4096);
    for (int i = threadIdx.x; i < kClasses; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int i = chunk_begin + threadIdx.x; i < chunk_end; i += blockDim.x) {
        const mt_txblk d = blocks[i];
        const int ss = size_slot(d.tx_size), ts = type_slot(d.tx_type);
        if (ss >= 0 && ts >= 0 && tx_supported(d.tx_size, d.tx_type)) atomicAdd(&h[ss * 7 + ts], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kClasses; i += blockDim.x) {
        basepos[i] = h[i] ? atomicAdd(&cursor[i], h[i]) : 0;
        h[i] = 0;
    }
    __syncthreads();
    for (int i = chunk_begin + threadIdx.x; i < chunk_end; i += blockDim.x) {
        const mt_txblk d = blocks[i];
        const int ss = size_slot(d.tx_size), ts = type_slot(d.tx_type);
        if (ss >= 0 && ts >= 0 && tx_supported(d.tx_size, d.tx_type)) {
            const int c = ss * 7 + ts;
            order[basepos[c] + atomicAdd(&h[c], 1u)] = (uint32_t)i;
        }
    }
}

// ------------------------------------------------------------------------- host side
namespace {

// Launch schedule of one transform size in the chain: tile scale (tx_tile_kernel's TS) and the
// persistent grid's workgroups per CU (64x64: waves per CU).  The defaults are the hand-picked
// shapes; mt_tx_schedule_set installs the ones an autotuner measured.
struct TxSched {
    int8_t tile_shift = 0;
    uint8_t wg_per_cu = 8;
};
inline TxSched default_sched(int tx_size) {
    TxSched sc;
    if (tx_size == MT_TX_64X64) sc.wg_per_cu = 16;
    return sc;
}

struct DeviceState {
    bool init = false;
    int cus = 256;
    TxSched sched[19];  // the chain's launch schedule per transform size
    // this device's address of `scratch` (mt_tx_chain_noalloc), from hipGetSymbolAddress
    uint32_t* scratch_d = nullptr;
};
// the chain's workspace of one (device, stream)
struct ChainWork {
    uint32_t* counts = nullptr;  // [kClasses]
    uint32_t* starts = nullptr;  // [kClasses]
    uint32_t* cursor = nullptr;  // [kClasses]
    uint32_t* order = nullptr;   // [cap]
    size_t order_cap = 0;
};
std::mutex g_mu;
DeviceState g_dev[64];
std::map<mtlib::StreamKey, ChainWork> g_chain;

// Optional per-launch timing: while enabled, every launch of mt_tx_chain is bracketed by
// a pair of HIP events on the caller's stream; mt_profile_end sums them per kernel slot.
struct ProfState {
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> recs;  // (slot, events)
    std::vector<hipEvent_t> pool;
};
ProfState g_prof;

hipEvent_t prof_event() {
    if (!g_prof.pool.empty()) {
        hipEvent_t e = g_prof.pool.back();
        g_prof.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int current_device(int* dev) {
    HIP_TRY(hipGetDevice(dev));
    if (*dev < 0 || *dev >= 64) return set_err(MT_ERR_ARG, "device index out of range");
    if (!g_dev[*dev].init) return set_err(MT_ERR_NOT_INIT, "mt_init was not called for this device");
    return MT_OK;
}

}  // namespace

namespace mtlib {
ProfScope::ProfScope(int slot_, hipStream_t s_) : slot(slot_), s(s_) {
    if (!g_prof.on || slot < 0) return;  // a negative slot records nothing
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;  // in a graph
    a = prof_event();
    b = prof_event();
    if (a) (void)hipEventRecord(a, s);
}

ProfScope::~ProfScope() {
    if (!a || !b) return;
    (void)hipEventRecord(b, s);
    g_prof.recs.push_back({slot, {a, b}});
}

int device_ready(int* dev, int* cus) {
    int d = 0;
    MT_HIP_TRY(hipGetDevice(&d));
    if (d < 0 || d >= 64) return set_error(MT_ERR_ARG, "device index out of range");
    if (!g_dev[d].init) return set_error(MT_ERR_NOT_INIT, "mt_init was not called for this device");
    if (dev) *dev = d;
    if (cus) *cus = g_dev[d].cus;
    return MT_OK;
}
}  // namespace mtlib

namespace {
using mtlib::ProfScope;

int ensure_order(ChainWork& st, size_t n) {
    if (!st.counts) {
        HIP_TRY(hipMalloc(&st.counts, kClasses * sizeof(uint32_t)));
        HIP_TRY(hipMalloc(&st.starts, kClasses * sizeof(uint32_t)));
        HIP_TRY(hipMalloc(&st.cursor, kClasses * sizeof(uint32_t)));
    }
    if (st.order_cap >= n && st.order) return MT_OK;
    if (st.order) HIP_TRY(hipFree(st.order));
    st.order = nullptr;
    size_t cap = std::max<size_t>(n, 1 << 16);
    HIP_TRY(hipMalloc(&st.order, cap * sizeof(uint32_t)));
    st.order_cap = cap;
    return MT_OK;
}


template <int W, int H, int MODE, int TS> int launch_tiles_ts(const KArgs& a, int slot, int n,
                                                              int wg_per_cu, int cus, hipStream_t s) {
    using R = TxRules<W, H>;
    constexpr int B = TS > 0 ? R::B << 1 : TS < 0 ? R::B >> 1 : R::B;
    const int tiles = (n + B - 1) / B;
    const int grid = std::max(1, std::min(tiles, cus * wg_per_cu));
    hipLaunchKernelGGL((tx_tile_kernel<W, H, MODE, TS>), dim3(grid), dim3(256), 0, s, a, slot);
    HIP_TRY(hipGetLastError());
    return MT_OK;
}

// n: the batch's block count (an upper bound of the class's); the chain's whole-batch mode
// takes the schedule's tile scale, the other modes the default tile.
template <int W, int H, int MODE> int launch_tiles(const KArgs& a, int slot, int n, TxSched sc,
                                                  int cus, hipStream_t s) {
    if constexpr (MODE == MODE_CHAIN) {
        if (sc.tile_shift > 0) return launch_tiles_ts<W, H, MODE, 1>(a, slot, n, sc.wg_per_cu, cus, s);
        if (sc.tile_shift < 0) return launch_tiles_ts<W, H, MODE, -1>(a, slot, n, sc.wg_per_cu, cus, s);
    }
    return launch_tiles_ts<W, H, MODE, 0>(a, slot, n, sc.wg_per_cu, cus, s);
}

template <int MODE> int launch_size(int tx_size, const KArgs& a, int n, TxSched sc, int cus,
                                    hipStream_t s) {
    switch (tx_size) {
        case MT_TX_4X4: return launch_tiles<4, 4, MODE>(a, 0, n, sc, cus, s);
        case MT_TX_4X8: return launch_tiles<4, 8, MODE>(a, 1, n, sc, cus, s);
        case MT_TX_8X4: return launch_tiles<8, 4, MODE>(a, 2, n, sc, cus, s);
        case MT_TX_4X16: return launch_tiles<4, 16, MODE>(a, 3, n, sc, cus, s);
        case MT_TX_16X4: return launch_tiles<16, 4, MODE>(a, 4, n, sc, cus, s);
        case MT_TX_8X8: return launch_tiles<8, 8, MODE>(a, 5, n, sc, cus, s);
        case MT_TX_8X16: return launch_tiles<8, 16, MODE>(a, 6, n, sc, cus, s);
        case MT_TX_16X8: return launch_tiles<16, 8, MODE>(a, 7, n, sc, cus, s);
        case MT_TX_16X16: return launch_tiles<16, 16, MODE>(a, 8, n, sc, cus, s);
        case MT_TX_16X32: return launch_tiles<16, 32, MODE>(a, 9, n, sc, cus, s);
        case MT_TX_32X16: return launch_tiles<32, 16, MODE>(a, 10, n, sc, cus, s);
        case MT_TX_32X32: return launch_tiles<32, 32, MODE>(a, 11, n, sc, cus, s);
        case MT_TX_64X64: {
            const int grid = std::max(1, std::min(n, cus * (int)sc.wg_per_cu));
            hipLaunchKernelGGL((tx64_kernel<MODE>), dim3(grid), dim3(64), 0, s, a);
            HIP_TRY(hipGetLastError());
            return MT_OK;
        }
        default: return set_err(MT_ERR_UNSUPPORTED, "unsupported transform size");
    }
}

// round(65536 / step) as the reference computes it (f64, round half away from zero).
int32_t q_mult(int32_t step) { return (int32_t)std::round(65536.0 / (double)step); }

void itx_clips(int bd, mt_quant* q) {  // src/quant.rs:243
    if (bd <= 8) {
        q->rmin = q->cmin = -32768;
        q->rmax = q->cmax = 32767;
        q->cf_max = 32767;
    } else {
        const int32_t rmax = (1 << (bd + 7)) - 1, cmax = (1 << (bd + 5)) - 1;
        q->rmin = ~rmax;
        q->rmax = rmax;
        q->cmin = ~cmax;
        q->cmax = cmax;
        q->cf_max = rmax;
    }
}

int bd_index(int bd) { return bd == 10 ? 1 : bd == 12 ? 2 : 0; }

}  // namespace

extern "C" const char* mt_last_error(void) { return g_err.c_str(); }
extern "C" const char* mt_version(void) { return "mtgpu 0.1 (gfx950)"; }

extern "C" int mt_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= 64) return set_err(MT_ERR_ARG, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    DeviceState& st = g_dev[device];
    if (st.init) return MT_OK;
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_dcq), MT_DC_QLOOKUP, sizeof(MT_DC_QLOOKUP)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_acq), MT_AC_QLOOKUP, sizeof(MT_AC_QLOOKUP)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_fadst4), MT_FADST4_Q12, sizeof(MT_FADST4_Q12)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_fadst8), MT_FADST8_Q12, sizeof(MT_FADST8_Q12)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_fadst16), MT_FADST16_Q12, sizeof(MT_FADST16_Q12)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_iqm), MT_IQM, sizeof(MT_IQM)));
    {
        int32_t qmw[256];
        for (int i = 0; i < 256; i++) qmw[i] = i ? (1024 + i / 2) / i : 0;
        HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_qmw), qmw, sizeof(qmw)));
    }
    HIP_TRY(upload_pred_tables());  // this unit's copy, for the RDO reconstruction
    // dct64_basis (src/dct.rs:4132): built on the host with the C library's cos, as the
    // reference builds it, so every f64 weight is the same.
    std::vector<double> basis(32 * 64);
    for (int k = 0; k < 32; k++) {
        const double a = k == 0 ? std::sqrt(1.0 / 64.0) : std::sqrt(2.0 / 64.0);
        for (int i = 0; i < 64; i++)
            basis[k * 64 + i] = a * std::cos(M_PI * (double)(2 * i + 1) * (double)k / 128.0);
    }
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_dct64_basis), basis.data(), basis.size() * sizeof(double)));
    std::vector<double> basis_t(64 * 32);
    for (int k = 0; k < 32; k++)
        for (int i = 0; i < 64; i++) basis_t[i * 32 + k] = basis[k * 64 + i];
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_dct64_basis_t), basis_t.data(), basis_t.size() * sizeof(double)));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    st.cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    for (int t = 0; t < 19; t++) st.sched[t] = default_sched(t);
    {
        int rc = mtlib::rate_init_device(device);
        if (rc) return rc;
        rc = mtlib::pred_init_device(device);
        if (rc) return rc;
        rc = mtlib::dist_init_device(device);
        if (rc) return rc;
        rc = mtlib::rdo_init_device(device);
        if (rc) return rc;
    }
    // the allocation-free chain's workspace: this device's copy of `scratch`
    {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(MT_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        st.scratch_d = (uint32_t*)p;
    }
    HIP_TRY(hipDeviceSynchronize());
    st.init = true;
    return MT_OK;
}

extern "C" int mt_quant_new(int base_q_idx, int bd, int qm_level, mt_quant* out) {
    if (!out || base_q_idx < 0 || base_q_idx > 255 || (bd != 8 && bd != 10 && bd != 12))
        return set_err(MT_ERR_ARG, "mt_quant_new: bad argument");
    memset(out, 0, sizeof(*out));
    itx_clips(bd, out);
    const int dc_idx = std::min(255, std::max(0, base_q_idx - 8));  // luma DC delta -8
    out->dc = MT_DC_QLOOKUP[bd_index(bd)][dc_idx];
    out->ac = MT_AC_QLOOKUP[bd_index(bd)][base_q_idx];
    out->q_mult_dc = q_mult(out->dc);
    out->q_mult_ac = q_mult(out->ac);
    out->qm_level = std::min(std::max(qm_level, 0), 15);
    out->qm_chroma = 0;
    out->qidx = base_q_idx;
    out->bd = bd;
    return MT_OK;
}

extern "C" int mt_quant_new_chroma(int base_q_idx, int dc_delta, int ac_delta, int bd,
                                   int qm_level, mt_quant* out) {
    if (!out || base_q_idx < 0 || base_q_idx > 255 || (bd != 8 && bd != 10 && bd != 12))
        return set_err(MT_ERR_ARG, "mt_quant_new_chroma: bad argument");
    memset(out, 0, sizeof(*out));
    const int dc_idx = std::min(255, std::max(0, base_q_idx + dc_delta));
    const int ac_idx = std::min(255, std::max(0, base_q_idx + ac_delta));
    itx_clips(bd, out);
    out->dc = MT_DC_QLOOKUP[bd_index(bd)][dc_idx];
    out->ac = MT_AC_QLOOKUP[bd_index(bd)][ac_idx];
    out->q_mult_dc = q_mult(out->dc);
    out->q_mult_ac = q_mult(out->ac);
    out->qm_level = std::min(std::max(qm_level, 0), 15);
    out->qm_chroma = 1;
    out->qidx = ac_idx;
    out->bd = bd;
    return MT_OK;
}

// FastRound (src/util.rs:35-59): a one-ulp nudge, then the 1.5 * 2^23 magic round.
static float fast_round(float x) {
    const float magic = (float)((1u << 23) + (1u << 22));
    uint32_t bits;
    memcpy(&bits, &x, 4);
    bits += 1;
    float y;
    memcpy(&y, &bits, 4);
    volatile float s = y + magic;
    return s - magic;
}

extern "C" int mt_qm_level_law(int base_q_idx, int sub) {
    // qm_level_law (src/quant.rs:137-154)
    const bool subsampled = sub != 0;
    const float hi = 48.0f, lo = 128.0f;
    const float qi = (float)base_q_idx;
    if (qi <= 20.0f) return subsampled ? 8 : 15;
    if (qi <= hi) {
        const float top = subsampled ? 8.0f : 15.0f;
        volatile float num = (8.0f - top) * (qi - 20.0f);
        volatile float v = top + num / (hi - 20.0f);
        return (int)(uint8_t)fast_round(v);
    }
    if (qi >= lo) return 0;
    volatile float num = 8.0f * (lo - qi);
    volatile float v = num / (lo - hi);
    return (int)(uint8_t)fast_round(v);
}

// Frame chroma quantizer laws with the shipped tuning (src/tuning.rs:750-791: top_ease 1,
// qm_hq_knee 12, qm_hq_width 5, top_444_uac 6, uv420_tail_delta 28, uv420_mid_d 28).
static int uv420_delta_law(int qi) {
    // uv420_mid_delta (src/quant.rs:170) with tail = d = 28
    const int d = 28, tail = 28;
    if (qi <= 20) return 0;
    if (qi < 64) return -(d * (qi - 20)) / 44;
    if (qi <= 96) return -d;
    if (qi < 136) return -tail - ((d - tail) * (136 - qi)) / 40;
    return -tail;
}

extern "C" int mt_chroma_q_deltas(int base_q_idx, int sub, int* dc_delta, int* ac_delta) {
    if (base_q_idx < 0 || base_q_idx > 255 || (sub != 0 && sub != 2) || !dc_delta || !ac_delta)
        return mtlib::set_error(MT_ERR_ARG, "mt_chroma_q_deltas: qidx 0..255, sub 0 (4:4:4) or 2 (4:2:0)");
    const int q = base_q_idx;
    // chroma_ac_delta (src/quant.rs:204)
    *ac_delta = sub == 2 ? uv420_delta_law(q) : q <= 20 ? 6 : q >= 96 ? 24 : 24 * (q - 20) / 76;
    // chroma_dc_delta (src/quant.rs:223)
    const int deep = q <= 160 ? 0 : -std::min((q - 160) / 3, 32);
    *dc_delta = sub == 2 ? deep + uv420_delta_law(q) : deep;
    return MT_OK;
}

extern "C" int mt_qm_chroma_level_law(int base_q_idx, int sub) {
    // qm_chroma_level_law (src/quant.rs:155)
    if (sub != 0) return mt_qm_level_law(base_q_idx, sub);
    const int pinned = std::min(2, mt_qm_level_law(base_q_idx, 0));
    volatile float t = (12.0f - (float)base_q_idx) / 5.0f;  // qm_high_quality_t
    float tc = t < 0.0f ? 0.0f : t > 1.0f ? 1.0f : (float)t;
    tc = 1.0f * tc;  // top_ease
    if (tc > 0.0f) {
        volatile float prod = (15.0f - (float)pinned) * tc;
        volatile float v = (float)pinned + prod;
        return (int)(uint8_t)std::round((float)v);
    }
    return pinned;
}

extern "C" int mt_tx_width(int s) { return (s >= 0 && s < 19) ? kTxW[s] : 0; }
extern "C" int mt_tx_height(int s) { return (s >= 0 && s < 19) ? kTxH[s] : 0; }
extern "C" int mt_tx_ncoef(int s) {
    if (s < 0 || s >= 19) return 0;
    return std::min(kTxW[s], 32) * std::min(kTxH[s], 32);
}
extern "C" int mt_tx_supported(int s, int t) { return (s >= 0 && s < 19 && tx_supported(s, t)) ? 1 : 0; }

static int parity_launch(int mode, int tx_size, int tx_type, int n, const int32_t* in,
                         const mt_quant* quants, int32_t* levels, float* targets, int32_t* out,
                         void* stream) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!in || !quants))) return set_err(MT_ERR_ARG, "bad pointer or count");
    if (tx_size < 0 || tx_size >= 19 || !tx_supported(tx_size, tx_type))
        return set_err(MT_ERR_UNSUPPORTED, "no reference transform for this (size, type)");
    if (n == 0) return MT_OK;
    KArgs a{};
    a.in32 = in;
    a.quants = quants;
    a.levels = levels;
    a.targets = targets;
    a.recon = out;
    a.n = n;
    a.type = tx_type;
    hipStream_t s = (hipStream_t)stream;
    const TxSched sc = default_sched(tx_size);
    if (mode == MODE_FWD) return launch_size<MODE_FWD>(tx_size, a, n, sc, g_dev[dev].cus, s);
    return launch_size<MODE_INV>(tx_size, a, n, sc, g_dev[dev].cus, s);
}

extern "C" int mt_fwd_txfm(int tx_size, int tx_type, int n, const int32_t* resid_d,
                           const mt_quant* quants_d, int32_t* levels_d, float* targets_d,
                           void* stream) {
    if (n > 0 && (!levels_d || !targets_d)) return set_err(MT_ERR_ARG, "null output");
    return parity_launch(MODE_FWD, tx_size, tx_type, n, resid_d, quants_d, levels_d, targets_d,
                         nullptr, stream);
}

extern "C" int mt_inv_txfm(int tx_size, int tx_type, int n, const int32_t* levels_d,
                           const mt_quant* quants_d, int32_t* out_d, void* stream) {
    if (n > 0 && !out_d) return set_err(MT_ERR_ARG, "null output");
    return parity_launch(MODE_INV, tx_size, tx_type, n, levels_d, quants_d, nullptr, nullptr,
                         out_d, stream);
}

int mtlib::chain_reserve(size_t n, hipStream_t stream) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    return ensure_order(g_chain[mtlib::stream_key(dev, stream)], n);
}

extern "C" int mt_tx_chain(const mt_txblk* blocks_d, int n, const int16_t* pool_d, mt_txres* res_d,
                           int32_t* levels_d, float* targets_d, int32_t* recon_d, void* stream) {
    return mtlib::tx_chain_launch(blocks_d, n, pool_d, res_d, levels_d, targets_d, recon_d, nullptr,
                                  (hipStream_t)stream);
}

namespace {
// The chain's launches on a given workspace: counts / starts / cursor hold kClasses words,
// order n words.  Shared by the library-workspace and the allocation-free entry points.
// trellis_d (optional): the chain runs forward + quantization for every block, the trellis
// over the levels (its workspace: recon_d), then dequantization + inverse + measurement.
int enqueue_chain(uint32_t* counts, uint32_t* starts, uint32_t* cursor, uint32_t* order, int cus,
                  const TxSched* sched,
                  const mt_txblk* blocks_d, int n, const int16_t* pool_d, mt_txres* res_d,
                  int32_t* levels_d, float* targets_d, int32_t* recon_d, const int8_t* qdelta_d,
                  hipStream_t s, bool profile, const mt_trellisblk* trellis_d = nullptr) {
    int rc;
    {
    ProfScope ps(profile ? 0 : -1, s);
    // cleared by a kernel, not hipMemsetAsync, so the sequence stays graph-capturable
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(128), 0, s, counts, kClasses);
    HIP_TRY(hipGetLastError());
    const int cgrid = std::min((n + 255) / 256, cus * 4);
    hipLaunchKernelGGL(bucket_count_kernel, dim3(cgrid), dim3(256), 0, s, blocks_d, n, counts, res_d);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(bucket_scan_kernel, dim3(1), dim3(64), 0, s, counts, starts, cursor);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(bucket_scatter_kernel, dim3((n + 4095) / 4096), dim3(256), 0, s, blocks_d, n,
                       cursor, order);
    HIP_TRY(hipGetLastError());
    }
    KArgs a{};
    a.blocks = blocks_d;
    a.order = order;
    a.cls_start = starts;
    a.cls_count = counts;
    a.pool16 = pool_d;
    a.res = res_d;
    a.levels = levels_d;
    a.targets = targets_d;
    a.recon = recon_d;
    a.qdelta = qdelta_d;
    // The class sizes live on the device; each launch is persistent over its own tiles, so
    // the grid only needs an upper bound: n blocks can never need more tiles than this.
    for (int slot = 0; slot < 13; slot++) {
        const int sz = kSlotSize[slot];
        ProfScope ps(profile ? 1 + slot : -1, s);
        rc = trellis_d ? launch_size<MODE_CFWD>(sz, a, n, sched[sz], cus, s)
                       : launch_size<MODE_CHAIN>(sz, a, n, sched[sz], cus, s);
        if (rc) return rc;
    }
    if (!trellis_d) return MT_OK;
    if ((rc = mtlib::trellis_launch(trellis_d, n, levels_d, targets_d, nullptr, s)))
        return rc;
    for (int slot = 0; slot < 13; slot++) {
        const int sz = kSlotSize[slot];
        ProfScope ps(profile ? 1 + slot : -1, s);
        if ((rc = launch_size<MODE_CINV>(sz, a, n, sched[sz], cus, s))) return rc;
    }
    return MT_OK;
}
}  // namespace

int mtlib::tx_chain_launch(const mt_txblk* blocks_d, int n, const int16_t* pool_d, mt_txres* res_d,
                           int32_t* levels_d, float* targets_d, int32_t* recon_d,
                           const int8_t* qdelta_d, hipStream_t stream, const mt_trellisblk* trellis_d) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!blocks_d || !pool_d || !res_d)))
        return set_err(MT_ERR_ARG, "mt_tx_chain: bad pointer or count");
    if (targets_d && !levels_d)
        return set_err(MT_ERR_ARG, "mt_tx_chain: targets need levels");
    if (trellis_d && (!levels_d || !targets_d || !recon_d))
        return set_err(MT_ERR_ARG, "mt_tx_chain_trellis: levels, targets and recon are required");
    if (n == 0) return MT_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    ChainWork& cw = g_chain[mtlib::stream_key(dev, stream)];
    rc = ensure_order(cw, (size_t)n);
    if (rc) return rc;
    return enqueue_chain(cw.counts, cw.starts, cw.cursor, cw.order, g_dev[dev].cus, g_dev[dev].sched,
                         blocks_d, n, pool_d,
                         res_d, levels_d, targets_d, recon_d, qdelta_d, stream, true, trellis_d);
}

extern "C" int mt_tx_chain_trellis(const mt_txblk* blocks_d, const mt_trellisblk* trellis_d, int n,
                                   const int16_t* pool_d, mt_txres* res_d, int32_t* levels_d,
                                   float* targets_d, int32_t* recon_d, void* stream) {
    if (n > 0 && !trellis_d) return set_err(MT_ERR_ARG, "mt_tx_chain_trellis: no trellis descriptors");
    return mtlib::tx_chain_launch(blocks_d, n, pool_d, res_d, levels_d, targets_d, recon_d, nullptr,
                                  (hipStream_t)stream, trellis_d);
}

extern "C" int mt_tx_chain_noalloc(const mt_txblk* blocks_d, int n, const int16_t* pool_d,
                                   mt_txres* res_d, int32_t* levels_d, float* targets_d,
                                   int32_t* recon_d, void* stream) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!blocks_d || !pool_d || !res_d)))
        return set_err(MT_ERR_ARG, "mt_tx_chain_noalloc: bad pointer or count");
    if (n > MT_NOALLOC_MAX_BLOCKS)
        return set_err(MT_ERR_ARG, "mt_tx_chain_noalloc: more blocks than MT_NOALLOC_MAX_BLOCKS "
                                   "(1048576); split the batch or use mt_tx_chain");
    if (targets_d && !levels_d)
        return set_err(MT_ERR_ARG, "mt_tx_chain_noalloc: targets need levels");
    if (n == 0) return MT_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState& st = g_dev[dev];
    if (!st.scratch_d) return set_err(MT_ERR_NOT_INIT, "mt_tx_chain_noalloc: no scratch address");
    uint32_t* base = st.scratch_d;
    return enqueue_chain(base, base + kNoAllocSlots, base + 2 * kNoAllocSlots,
                         base + 3 * kNoAllocSlots, st.cus, st.sched, blocks_d, n, pool_d, res_d, levels_d,
                         targets_d, recon_d, nullptr, (hipStream_t)stream, false);
}

extern "C" int mt_tx_schedule_set(int device, int tx_size, int tile_shift, int wg_per_cu) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= 64 || !g_dev[device].init)
        return set_err(MT_ERR_NOT_INIT, "mt_tx_schedule_set: mt_init was not called for this device");
    if (tx_size < 0 || tx_size >= 19 || !tx_supported(tx_size, MT_TXTP_DCT_DCT))
        return set_err(MT_ERR_UNSUPPORTED, "mt_tx_schedule_set: no chain kernel for this size");
    if (tile_shift < -1 || tile_shift > 1 || (tx_size == MT_TX_64X64 && tile_shift != 0))
        return set_err(MT_ERR_ARG, "mt_tx_schedule_set: tile_shift must be -1, 0 or 1 (0 at 64x64)");
    if (wg_per_cu < 1 || wg_per_cu > 64)
        return set_err(MT_ERR_ARG, "mt_tx_schedule_set: wg_per_cu must be 1..64");
    g_dev[device].sched[tx_size].tile_shift = (int8_t)tile_shift;
    g_dev[device].sched[tx_size].wg_per_cu = (uint8_t)wg_per_cu;
    return MT_OK;
}

extern "C" int mt_tx_schedule_get(int device, int tx_size, int* tile_shift, int* wg_per_cu) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= 64 || !g_dev[device].init)
        return set_err(MT_ERR_NOT_INIT, "mt_tx_schedule_get: mt_init was not called for this device");
    if (tx_size < 0 || tx_size >= 19) return set_err(MT_ERR_ARG, "mt_tx_schedule_get: bad size");
    if (tile_shift) *tile_shift = g_dev[device].sched[tx_size].tile_shift;
    if (wg_per_cu) *wg_per_cu = g_dev[device].sched[tx_size].wg_per_cu;
    return MT_OK;
}

extern "C" int mt_profile_begin(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof.on = true;
    return MT_OK;
}

extern "C" int mt_profile_end(double* kernel_ms, int* launches, int max_slots) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof.on = false;
    for (int i = 0; i < max_slots; i++) {
        if (kernel_ms) kernel_ms[i] = 0.0;
        if (launches) launches[i] = 0;
    }
    int rc = MT_OK;
    for (auto& r : g_prof.recs) {
        hipEvent_t a = r.second.first, b = r.second.second;
        float ms = 0.0f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
            rc = set_err(MT_ERR_DEVICE, "profiling event failed");
        if (r.first < max_slots) {
            if (kernel_ms) kernel_ms[r.first] += ms;
            if (launches) launches[r.first] += 1;
        }
        g_prof.pool.push_back(a);
        g_prof.pool.push_back(b);
    }
    g_prof.recs.clear();
    return rc;
}

// ------------------------------------------------------------------------- generator
namespace {
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double unit(uint64_t& x) { return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }

const int kAllSizes[19] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18};

struct GenBlock {
    int size, type, qidx, bd, qm;
};

GenBlock gen_block_header(uint64_t seed, int i, int all_sizes) {
    uint64_t st = seed ^ (0xA0761D6478BD642Full * (uint64_t)(i + 1));
    splitmix64(st);
    GenBlock g;
    if (all_sizes) {
        g.size = kAllSizes[splitmix64(st) % 19];
    } else {
        g.size = kSlotSize[splitmix64(st) % 13];
    }
    int types[7], nt = 0;
    for (int t = 0; t < 7; t++)
        if (tx_supported(g.size, kSlotType[t])) types[nt++] = kSlotType[t];
    if (nt == 0) types[nt++] = MT_TXTP_DCT_DCT;  // non-reference sizes: DCT_DCT request
    g.type = types[splitmix64(st) % nt];
    g.qidx = (int)(splitmix64(st) & 255);
    g.bd = (splitmix64(st) & 1) ? 10 : 8;
    g.qm = (splitmix64(st) & 1) ? mt_qm_level_law(g.qidx, 0) : 15;
    return g;
}
}  // namespace

extern "C" uint64_t mt_gen_pool_size(uint64_t seed, int n, int all_sizes) {
    uint64_t total = 0;
    for (int i = 0; i < n; i++) {
        const GenBlock g = gen_block_header(seed, i, all_sizes);
        total += (uint64_t)kTxW[g.size] * kTxH[g.size];
    }
    return total;
}

extern "C" int mt_gen_tx_blocks(uint64_t seed, int n, int all_sizes, mt_txblk* blocks, int16_t* pool,
                                uint64_t pool_elems) {
    if (n < 0 || !blocks || !pool) return set_err(MT_ERR_ARG, "mt_gen_tx_blocks: bad argument");
    std::vector<GenBlock> hdr((size_t)n);
    std::vector<uint64_t> off((size_t)n);
    uint64_t total = 0;
    for (int i = 0; i < n; i++) {
        hdr[i] = gen_block_header(seed, i, all_sizes);
        off[i] = total;
        total += (uint64_t)kTxW[hdr[i].size] * kTxH[hdr[i].size];
    }
    if (total > pool_elems || total > 0xFFFFFFFFull) return set_err(MT_ERR_ARG, "pool too small");
    // discrete Laplace(b = 12): difference of two geometric(p = 1 - e^(-1/12)) draws,
    // floor(-12 ln U) each; clipped to the bit depth's residual range.
    const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int k = 0; k < nth; k++) {
        th.emplace_back([&, k]() {
            for (int i = k; i < n; i += nth) {
                const GenBlock& g = hdr[i];
                uint64_t st = seed * 0xD1B54A32D192ED03ull + (uint64_t)i * 0x9E3779B97F4A7C15ull + 7;
                const int npx = kTxW[g.size] * kTxH[g.size];
                const int lim = g.bd == 8 ? 255 : 1023;
                for (int p = 0; p < npx; p++) {
                    const double u1 = 1.0 - unit(st), u2 = 1.0 - unit(st);
                    const int g1 = (int)std::floor(-12.0 * std::log(u1));
                    const int g2 = (int)std::floor(-12.0 * std::log(u2));
                    int v = g1 - g2;
                    v = v < -lim ? -lim : (v > lim ? lim : v);
                    pool[off[i] + p] = (int16_t)v;
                }
                mt_txblk b;
                b.offset = (uint32_t)off[i];
                b.tx_size = (uint8_t)g.size;
                b.tx_type = (uint8_t)g.type;
                b.qidx = (uint8_t)g.qidx;
                b.bd_qm = (uint8_t)((g.qm & 15) | ((g.bd == 10 ? 1 : g.bd == 12 ? 2 : 0) << 4));
                blocks[i] = b;
            }
        });
    }
    for (auto& t : th) t.join();
    return MT_OK;
}
