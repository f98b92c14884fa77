// Coefficient rate of transform blocks on the GPU: the decision-side mirror of the AV1
// coefficient coder (reference src/rate.rs:149 real_block_bits_bounded, src/cost.rs:159-262,
// src/coeffs.rs:46 get_lo_ctx_2d) against the frame-initial CDFs Cdfs::new(qctx).
//
// The walk is sequential per block (every f32 addition in the reference's order).  Blocks of
// up to 128 coefficients take one lane each; larger ones one wave each, whose lanes price the
// scan positions in parallel before one lane adds the costs in order.  The neighbour "levels" plane
// the reference fills while walking the reverse scan needs no scratch here: at step i a
// position holds its level byte exactly when its scan index is above i, so every context
// is read straight from the block's levels through the inverse scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/mtgpu.h"
#include "av1_cdf_tables.h"
#include "av1_tables.h"
#include "mt_common.h"
#include "mt_lists.h"
#include "mt_trellis.h"

namespace {

using mtlib::RateTables;

// Scans (scan index -> x-major rc) and inverse scans, concatenated per scan size.
constexpr int kScanTotal = 16 + 32 + 32 + 64 + 64 + 64 + 128 + 128 + 256 + 512 + 512 + 1024;

}  // namespace

__device__ RateTables g_rate[MT_RATE_SLOTS];
__constant__ uint16_t c_scan[kScanTotal];
__constant__ uint16_t c_iscan[kScanTotal];
__constant__ uint8_t c_lo_off[3][5][5];  // square, w < h, w > h (src/tables.rs:354, :456, :466)
// the trellis's quantizer steps and inverse quantizer matrices (the chain's tables)
__constant__ uint16_t c_tr_dcq[3][256];
__constant__ uint16_t c_tr_acq[3][256];
__device__ uint8_t g_tr_iqm[MT_QM_LEVELS][2][MT_QM_USED_SIZE];

namespace {

// A decoded, validated rate descriptor.
struct RateBlk {
    int so, cls, cb, w, h, N, log2h, oi;
};

// Validation and the size-derived quantities of one descriptor (false: invalid, priced NaN).
__device__ inline bool rate_decode(const mt_rateblk& d, RateBlk& r) {
    const int s = d.tx_size;
    r.so = s < 19 ? scan_offset(s) : -1;
    if (r.so < 0 || d.plane > 1 || d.eob_bin > 16 || d.skip_ctx > 12 || d.dcs_ctx > 2 || d.qctx >= MT_RATE_SLOTS ||
        d.txtp_set > 3 || d.txtp_mode > 12 || d.txtp > 6 || (s == MT_TX_64X64 && d.plane))
        return false;
    const int wf = s == MT_TX_4X4 || s == MT_TX_4X8 || s == MT_TX_4X16 ? 4
                 : s == MT_TX_8X8 || s == MT_TX_8X4 || s == MT_TX_8X16 ? 8
                 : s == MT_TX_16X16 || s == MT_TX_16X8 || s == MT_TX_16X4 || s == MT_TX_16X32 ? 16
                 : s == MT_TX_64X64 ? 64 : 32;
    const int hf = s == MT_TX_4X4 || s == MT_TX_8X4 || s == MT_TX_16X4 ? 4
                 : s == MT_TX_8X8 || s == MT_TX_4X8 || s == MT_TX_16X8 ? 8
                 : s == MT_TX_16X16 || s == MT_TX_8X16 || s == MT_TX_4X16 || s == MT_TX_32X16 ? 16
                 : s == MT_TX_64X64 ? 64 : 32;
    // transform-size context of the coefficient CDFs: (log2(w/4) + log2(h/4) + 1) >> 1
    r.cls = ((31 - __clz(wf)) - 2 + (31 - __clz(hf)) - 2 + 1) >> 1;
    r.cb = r.cls < 3 ? r.cls : 3;
    r.w = wf > 32 ? 32 : wf;
    r.h = hf > 32 ? 32 : hf;
    r.N = r.w * r.h;
    r.log2h = 31 - __clz(r.h);
    r.oi = r.w < r.h ? 1 : (r.w > r.h ? 2 : 0);
    // the eob_bin CDF must be the one of this coefficient count (the reference indexes its
    // symbols by the eob's bin; a smaller table would run past its end)
    const int ebn = d.eob_bin < 3 ? 16 : d.eob_bin < 5 ? 32 : d.eob_bin < 8 ? 64 : d.eob_bin < 11 ? 128
                  : d.eob_bin < 13 ? 256 : d.eob_bin < 15 ? 512 : 1024;
    return ebn == r.N && !(d.txtp_set == 3 && d.txtp > 4);
}

// Blocks of at least this many coded coefficients take the wave path when split.
constexpr int kRateWaveN = 256;

// The 8-bit wrap of the reference's u8 level plane is irrelevant: entries stay below 256.
// nnz (optional): the transform chain's per-block results; their nonzero-level counts let
// the eob search stop after the last nonzero level instead of walking every scan position.
// skip_large: leave valid blocks of kRateWaveN or more coefficients to rate_wave_kernel.
__global__ void __launch_bounds__(256) rate_kernel(const mt_rateblk* __restrict__ blocks, int n,
                                                  const int32_t* __restrict__ levels,
                                                  float* __restrict__ out,
                                                  const mt_txres* __restrict__ nnz, int skip_large) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const mt_rateblk d = blocks[b];
    RateBlk R;
    if (!rate_decode(d, R)) {
        out[b] = __int_as_float(0x7fc00000);  // invalid descriptor: NaN
        return;
    }
    if (skip_large && R.N >= kRateWaveN) return;
    const int so = R.so, cls = R.cls, w = R.w, h = R.h, N = R.N;
    const int pl = d.plane;
    const RateTables& T = g_rate[d.qctx];
    const int32_t* cf = levels + d.offset;
    const uint16_t* scan = c_scan + so;
    const uint16_t* iscan = c_iscan + so;

    int eob = N - 1;
    if (nnz) {
        // eob = the largest scan index of a nonzero level: walk the x-major levels four at a
        // time until all of the block's nonzero levels have been seen
        const uint32_t want = nnz[b].nnz;
        uint32_t seen = 0;
        eob = -1;
        if ((d.offset & 3) == 0) {
            const int4* cf4 = reinterpret_cast<const int4*>(cf);
            for (int k4 = 0; k4 < (N >> 2) && seen < want; k4++) {
                const int4 v = cf4[k4];
                const int k = k4 << 2;
                if (v.x) { seen++; eob = max(eob, (int)iscan[k]); }
                if (v.y) { seen++; eob = max(eob, (int)iscan[k + 1]); }
                if (v.z) { seen++; eob = max(eob, (int)iscan[k + 2]); }
                if (v.w) { seen++; eob = max(eob, (int)iscan[k + 3]); }
            }
        } else {
            for (int k = 0; k < N && seen < want; k++)
                if (cf[k]) { seen++; eob = max(eob, (int)iscan[k]); }
        }
    } else {
        while (eob >= 0 && cf[scan[eob]] == 0) eob--;
    }
    const bool adaptive_skip = d.flags & 1;
    if (eob < 0) {
        out[b] = adaptive_skip ? d.skip_one : T.skip[cls][d.skip_ctx][1];
        return;
    }
    const int cb = R.cb;
    float bits = adaptive_skip ? d.skip_zero : T.skip[cls][d.skip_ctx][0];
    if (d.txtp_set) bits += T.txtp[d.txtp_set - 1][d.txtp_mode][d.txtp];
    if (eob == 0) {
        const uint32_t m = uabs32(cf[0]);
        const uint32_t t = m < 3 ? m : 3;
        bits += T.eob_bin[d.eob_bin][0];
        bits += T.eob_base[cls][pl][0][t - 1];
        if (t == 3) {
            float hi = T.br[cb][pl][0][(m - 3) < 12 ? (m - 3) : 12];
            if (m >= 15) hi += golomb_bits(m - 15);
            bits += hi;
        }
        bits += T.dc_sign[pl][d.dcs_ctx][cf[0] < 0 ? 1 : 0];
        out[b] = bits;
        return;
    }
    const int eob_bin = eob < 2 ? eob : 32 - __clz(eob);
    bits += T.eob_bin[d.eob_bin][eob_bin];
    if (eob_bin > 1) {
        const int nbits = eob_bin - 2;
        bits += T.eob_hi[cls][pl][eob_bin][(eob >> nbits) & 1];
        bits += (float)nbits;
    }
    if (bits > d.bound) {
        out[b] = INFINITY;
        return;
    }
    const int log2h = R.log2h;
    const int oi = R.oi;
    (void)w;
    // level byte of x-major index k as the reference's plane holds it at step i
    auto lvl_at = [&](int k, int i) -> uint32_t {
        return (k < N && (int)iscan[k] > i) ? level_byte(uabs32(cf[k])) : 0u;
    };
    {
        const int rc = scan[eob];
        const int ex = rc >> log2h, ey = rc & (h - 1);
        const uint32_t m = uabs32(cf[rc]);
        const uint32_t et = (m < 3 ? m : 3) - 1;
        const int ctx_e = 1 + (eob > (N >> 3) ? 1 : 0) + (eob > (N >> 2) ? 1 : 0);
        bits += T.eob_base[cls][pl][ctx_e][et];
        if (et == 2) {
            const int bc = (ex | ey) > 1 ? 14 : 7;
            float hi = T.br[cb][pl][bc][(m - 3) < 12 ? (m - 3) : 12];
            if (m >= 15) hi += golomb_bits(m - 15);
            bits += hi;
        }
    }
    uint32_t ac_nonzero = 1;
    for (int i = eob - 1; i >= 1; i--) {
        if ((i & 15) == 0 && bits > d.bound) {
            out[b] = INFINITY;
            return;
        }
        const int rc = scan[i];
        const int x = rc >> log2h, y = rc & (h - 1);
        const int base = x * h + y;
        const uint32_t hm = lvl_at(base + 1, i) + lvl_at(base + h, i) + lvl_at(base + h + 1, i);
        const uint32_t mag = hm + lvl_at(base + 2, i) + lvl_at(base + 2 * h, i);
        const int ctx = c_lo_off[oi][y < 4 ? y : 4][x < 4 ? x : 4] + (mag > 512 ? 4 : (int)((mag + 64) >> 7));
        const uint32_t m = uabs32(cf[rc]);
        const uint32_t tok = m < 3 ? m : 3;
        bits += T.bt[cls][pl][ctx][tok];
        if (tok == 3) {
            const uint32_t mg = hm & 63;
            const int bc = ((y | x) > 1 ? 14 : 7) + (mg > 12 ? 6 : (int)((mg + 1) >> 1));
            bits += T.br[cb][pl][bc][(m - 3) < 12 ? (m - 3) : 12];
            if (m >= 15) bits += golomb_bits(m - 15);
        }
        ac_nonzero += m != 0;
    }
    const uint32_t dm = uabs32(cf[0]);
    const uint32_t dt = dm < 3 ? dm : 3;
    bits += T.bt[cls][pl][0][dt];
    if (dt == 3) {
        const uint32_t mg = (lvl_at(1, 0) + lvl_at(h, 0) + lvl_at(h + 1, 0)) & 63;
        const int bc = mg > 12 ? 6 : (int)((mg + 1) >> 1);
        float hi = T.br[cb][pl][bc][(dm - 3) < 12 ? (dm - 3) : 12];
        if (dm >= 15) hi += golomb_bits(dm - 15);
        bits += hi;
    }
    if (cf[0] != 0) bits += T.dc_sign[pl][d.dcs_ctx][cf[0] < 0 ? 1 : 0];
    for (uint32_t k = 0; k < ac_nonzero; k++) bits += 1.0f;
    out[b] = bits;
}

// Blocks of kRateWaveN or more coefficients, one wave per block.  The lanes stage the level
// bytes and the inverse scan in LDS, find the eob, and price every scan position's symbols
// in parallel (the contexts only read levels coded before the position, which the staged
// plane already holds); lane 0 then adds the per-position costs in the reference's order,
// with the same bound checks, so the f32 sum is the sequential walk's.
__global__ void __launch_bounds__(64) rate_wave_kernel(const mt_rateblk* __restrict__ blocks,
                                                       const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ count,
                                                       const int32_t* __restrict__ levels,
                                                       float* __restrict__ out) {
    __shared__ uint8_t s_lb[1024];    // level byte of |cf[k]|, x-major
    __shared__ uint16_t s_is[1024];   // inverse scan
    // per scan position: base token and base range costs, and the Golomb bits; a symbol the
    // position does not code costs +0.0, whose addition leaves every finite sum unchanged
    __shared__ float2 s_t01[1024];
    __shared__ float s_t2[1024];
    const int lane = threadIdx.x;
    const int nlist = (int)*count;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const uint32_t b = list[li];
        const mt_rateblk d = blocks[b];
        RateBlk R;
        rate_decode(d, R);  // valid: only valid blocks are listed
        const int cls = R.cls, cb = R.cb, h = R.h, N = R.N, log2h = R.log2h, oi = R.oi;
        const int pl = d.plane;
        const RateTables& T = g_rate[d.qctx];
        const int32_t* cf = levels + d.offset;
        const uint16_t* scan = c_scan + R.so;
        int emax = -1, nzc = 0;
        for (int k = lane; k < N; k += 64) {
            const int32_t v = cf[k];
            const int is = c_iscan[R.so + k];
            s_lb[k] = (uint8_t)level_byte(uabs32(v));
            s_is[k] = (uint16_t)is;
            if (v) emax = max(emax, is);
        }
        for (int off = 32; off > 0; off >>= 1) emax = max(emax, __shfl_xor(emax, off));
        const int eob = emax;
        __syncthreads();
        const bool adaptive_skip = d.flags & 1;
        if (eob < 1) {
            // all zero, or DC only: the thread kernel's short paths
            if (lane == 0) {
                float bits;
                if (eob < 0) {
                    bits = adaptive_skip ? d.skip_one : T.skip[cls][d.skip_ctx][1];
                } else {
                    bits = adaptive_skip ? d.skip_zero : T.skip[cls][d.skip_ctx][0];
                    if (d.txtp_set) bits += T.txtp[d.txtp_set - 1][d.txtp_mode][d.txtp];
                    const uint32_t m = uabs32(cf[0]);
                    const uint32_t t = m < 3 ? m : 3;
                    bits += T.eob_bin[d.eob_bin][0];
                    bits += T.eob_base[cls][pl][0][t - 1];
                    if (t == 3) {
                        float hi = T.br[cb][pl][0][(m - 3) < 12 ? (m - 3) : 12];
                        if (m >= 15) hi += golomb_bits(m - 15);
                        bits += hi;
                    }
                    bits += T.dc_sign[pl][d.dcs_ctx][cf[0] < 0 ? 1 : 0];
                }
                out[b] = bits;
            }
            __syncthreads();
            continue;
        }
        auto lvl_at = [&](int k, int i) -> uint32_t {
            return (k < N && (int)s_is[k] > i) ? (uint32_t)s_lb[k] : 0u;
        };
        // every position 1 .. eob-1: its symbols' costs
        for (int i = 1 + lane; i < eob; i += 64) {
            const int rc = scan[i];
            const int x = rc >> log2h, y = rc & (h - 1);
            const int base = x * h + y;
            const uint32_t hm = lvl_at(base + 1, i) + lvl_at(base + h, i) + lvl_at(base + h + 1, i);
            const uint32_t mag = hm + lvl_at(base + 2, i) + lvl_at(base + 2 * h, i);
            const int ctx = c_lo_off[oi][y < 4 ? y : 4][x < 4 ? x : 4] + (mag > 512 ? 4 : (int)((mag + 64) >> 7));
            const uint32_t m = uabs32(cf[rc]);
            const uint32_t tok = m < 3 ? m : 3;
            float t1 = 0.0f, t2 = 0.0f;
            if (tok == 3) {
                const uint32_t mg = hm & 63;
                const int bc = ((y | x) > 1 ? 14 : 7) + (mg > 12 ? 6 : (int)((mg + 1) >> 1));
                t1 = T.br[cb][pl][bc][(m - 3) < 12 ? (m - 3) : 12];
                if (m >= 15) t2 = golomb_bits(m - 15);
            }
            s_t01[i] = make_float2(T.bt[cls][pl][ctx][tok], t1);
            s_t2[i] = t2;
            nzc += m != 0;
        }
        for (int off = 32; off > 0; off >>= 1) nzc += __shfl_xor(nzc, off);
        __syncthreads();
        if (lane == 0) {
            float bits = adaptive_skip ? d.skip_zero : T.skip[cls][d.skip_ctx][0];
            if (d.txtp_set) bits += T.txtp[d.txtp_set - 1][d.txtp_mode][d.txtp];
            const int eob_bin = eob < 2 ? eob : 32 - __clz(eob);
            bits += T.eob_bin[d.eob_bin][eob_bin];
            if (eob_bin > 1) {
                const int nbits = eob_bin - 2;
                bits += T.eob_hi[cls][pl][eob_bin][(eob >> nbits) & 1];
                bits += (float)nbits;
            }
            bool over = bits > d.bound;
            if (!over) {
                const int rc = scan[eob];
                const int ex = rc >> log2h, ey = rc & (h - 1);
                const uint32_t m = uabs32(cf[rc]);
                const uint32_t et = (m < 3 ? m : 3) - 1;
                const int ctx_e = 1 + (eob > (N >> 3) ? 1 : 0) + (eob > (N >> 2) ? 1 : 0);
                bits += T.eob_base[cls][pl][ctx_e][et];
                if (et == 2) {
                    const int bc = (ex | ey) > 1 ? 14 : 7;
                    float hi = T.br[cb][pl][bc][(m - 3) < 12 ? (m - 3) : 12];
                    if (m >= 15) hi += golomb_bits(m - 15);
                    bits += hi;
                }
                // positions i .. down to the next multiple of 16 form one run between the
                // reference's bound checks (made before pricing every 16th position)
                int i = eob - 1;
                while (i >= 1) {
                    if ((i & 15) == 0 && bits > d.bound) {
                        over = true;
                        break;
                    }
                    const int lo = max(1, (i & 15) == 0 ? i - 15 : (i & ~15) + 1);
                    for (; i >= lo; i--) {
                        const float2 t01 = s_t01[i];
                        bits += t01.x;
                        bits += t01.y;
                        bits += s_t2[i];
                    }
                }
            }
            if (!over) {
                const uint32_t dm = uabs32(cf[0]);
                const uint32_t dt = dm < 3 ? dm : 3;
                bits += T.bt[cls][pl][0][dt];
                if (dt == 3) {
                    const uint32_t mg = (lvl_at(1, 0) + lvl_at(h, 0) + lvl_at(h + 1, 0)) & 63;
                    const int bc = mg > 12 ? 6 : (int)((mg + 1) >> 1);
                    float hi = T.br[cb][pl][bc][(dm - 3) < 12 ? (dm - 3) : 12];
                    if (dm >= 15) hi += golomb_bits(dm - 15);
                    bits += hi;
                }
                if (cf[0] != 0) bits += T.dc_sign[pl][d.dcs_ctx][cf[0] < 0 ? 1 : 0];
                const uint32_t ac_nonzero = 1 + (uint32_t)nzc;
                for (uint32_t k = 0; k < ac_nonzero; k++) bits += 1.0f;
            }
            out[b] = over ? INFINITY : bits;
        }
        __syncthreads();
    }
}

// Lists the valid blocks of kRateWaveN or more coefficients.
__global__ void __launch_bounds__(256) rate_classify_kernel(const mt_rateblk* __restrict__ blocks, int n,
                                                            uint32_t* __restrict__ list,
                                                            uint32_t* __restrict__ cnt) {
    auto cls = [&](int b, bool) -> int {
        RateBlk R;
        return (rate_decode(blocks[b], R) && R.N >= kRateWaveN) ? 0 : -1;
    };
    mtlib::classify_chunk<1>(n, cls, list, 0, cnt);
}

// ---------------------------------------------------------------- host-side tables
// cost_q_table (src/cost.rs:159): Q22 -log2(p / 32768) of the C library's log2f, as the
// reference builds it at run time (so it is built on the host and uploaded, never
// recomputed on the device).
const uint32_t* cost_q() {
    static uint32_t t[32769];
    static std::once_flag once;
    std::call_once(once, [] {
        t[0] = 0;
        for (int p = 1; p <= 32768; p++) {
            const float bits = -log2f((float)p * (1.0f / 32768.0f));
            t[p] = (uint32_t)roundf(bits * (float)(1u << 22));
        }
    });
    return t;
}

// Cost of symbol s of a raw AV1 CDF with `k` stored values (cdf_cost_with_table on the
// icdf form, src/cost.rs:185 with icdf src/tables.rs:31).
float sym_cost(const uint16_t* raw, int k, int s) {
    const int fl = s > 0 ? 32768 - raw[s - 1] : 32768;
    const int fh = s < k ? 32768 - raw[s] : 0;
    const int p = fl - fh < 1 ? 1 : fl - fh;
    return (float)cost_q()[p] * (1.0f / (float)(1u << 22));
}

// Stored values of the eob_bin CDF of each mt_rateblk.eob_bin.
constexpr int kEobBinK[17] = {4, 4, 4, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 9, 9, 10, 10};

void ladder(const uint16_t* raw3, float* row) {  // br_cum_row_with_table (src/cost.rs:204)
    float c[4];
    for (int s = 0; s < 4; s++) c[s] = sym_cost(raw3, 3, s);
    const float c3x2 = c[3] + c[3];
    const float c3x3 = c3x2 + c[3];
    const float c3x4 = c3x3 + c[3];
    const float r[13] = {c[0], c[1], c[2], c[3] + c[0], c[3] + c[1], c[3] + c[2], c3x2 + c[0],
                         c3x2 + c[1], c3x2 + c[2], c3x3 + c[0], c3x3 + c[1], c3x3 + c[2], c3x4};
    memcpy(row, r, sizeof(r));
}

// The coefficient CDFs of Cdfs::new(q) (src/coder/cdfs.rs:400) in raw form.
void default_cdfs(int q, mt_coef_cdfs& c) {
    memset(&c, 0, sizeof(c));
    const uint16_t* skip[5] = {MT_SKIP_TX4[q], MT_SKIP_TX8[q], MT_SKIP_TX16[q], MT_SKIP_TX32[q], MT_SKIP_TX64[q]};
    for (int k = 0; k < 5; k++) memcpy(c.skip[k], skip[k], 13 * sizeof(uint16_t));
    const uint16_t(*bt[5][2])[3] = {{MT_BASE_TOK_TX4_LUMA_Q[q], MT_BASE_TOK_TX4_CHROMA_Q[q]},
                                    {MT_BASE_TOK_TX8_LUMA_Q[q], MT_BASE_TOK_TX8_CHROMA_Q[q]},
                                    {MT_BASE_TOK_TX16_LUMA_Q[q], MT_BASE_TOK_TX16_CHROMA_Q[q]},
                                    {MT_BASE_TOK_TX32_LUMA_Q[q], MT_BASE_TOK_TX32_CHROMA_Q[q]},
                                    {MT_BASE_TOK_TX64_LUMA_Q[q], nullptr}};
    const int nctx[5] = {41, 41, 42, 42, 42};
    for (int k = 0; k < 5; k++)
        for (int p = 0; p < 2; p++)
            if (bt[k][p]) memcpy(c.base_tok[k][p], bt[k][p], nctx[k] * 3 * sizeof(uint16_t));
    const uint16_t(*br[4][2])[3] = {{MT_BR_TOK_TX4_LUMA_Q[q], MT_BR_TOK_TX4_CHROMA_Q[q]},
                                    {MT_BR_TOK_TX8_LUMA_Q[q], MT_BR_TOK_TX8_CHROMA_Q[q]},
                                    {MT_BR_TOK_TX16_LUMA_Q[q], MT_BR_TOK_TX16_CHROMA_Q[q]},
                                    {MT_BR_TOK_TX32_LUMA_Q[q], MT_BR_TOK_TX32_CHROMA_Q[q]}};
    for (int k = 0; k < 4; k++)
        for (int p = 0; p < 2; p++) memcpy(c.br_tok[k][p], br[k][p], 21 * 3 * sizeof(uint16_t));
    const uint16_t(*eb[5][2])[2] = {{MT_EOB_BASE_TX4_LUMA_Q[q], MT_EOB_BASE_TX4_CHROMA_Q[q]},
                                    {MT_EOB_BASE_TX8_LUMA_Q[q], MT_EOB_BASE_TX8_CHROMA_Q[q]},
                                    {MT_EOB_BASE_TX16_LUMA_Q[q], MT_EOB_BASE_TX16_CHROMA_Q[q]},
                                    {MT_EOB_BASE_TX32_LUMA_Q[q], MT_EOB_BASE_TX32_CHROMA_Q[q]},
                                    {MT_EOB_BASE_TX64_LUMA_Q[q], nullptr}};
    for (int k = 0; k < 5; k++)
        for (int p = 0; p < 2; p++)
            if (eb[k][p]) memcpy(c.eob_base[k][p], eb[k][p], 4 * 2 * sizeof(uint16_t));
    const uint16_t* eh[5][2] = {{MT_EOB_HI_TX4_LUMA[q], MT_EOB_HI_TX4_CHROMA[q]},
                                {MT_EOB_HI_TX8_LUMA[q], MT_EOB_HI_TX8_CHROMA[q]},
                                {MT_EOB_HI_TX16_LUMA[q], MT_EOB_HI_TX16_CHROMA[q]},
                                {MT_EOB_HI_TX32_LUMA[q], MT_EOB_HI_TX32_CHROMA[q]},
                                {MT_EOB_HI_TX64_LUMA[q], nullptr}};
    for (int k = 0; k < 5; k++)
        for (int p = 0; p < 2; p++)
            if (eh[k][p]) memcpy(c.eob_hi[k][p], eh[k][p], 11 * sizeof(uint16_t));
    memcpy(c.dc_sign, MT_DC_SIGN_Q[q], sizeof(c.dc_sign));
    const uint16_t* bins[17] = {MT_EOB_BIN_16_LUMA[q],   MT_EOB_BIN_16_CHROMA[q],  MT_EOB_BIN_16_LUMA_1D[q],
                                MT_EOB_BIN_32_LUMA[q],   MT_EOB_BIN_32_CHROMA[q],  MT_EOB_BIN_64_LUMA[q],
                                MT_EOB_BIN_64_CHROMA[q], MT_EOB_BIN_64_LUMA_1D[q], MT_EOB_BIN_128_LUMA[q],
                                MT_EOB_BIN_128_CHROMA[q], MT_EOB_BIN_128_LUMA_1D[q], MT_EOB_BIN_256_LUMA[q],
                                MT_EOB_BIN_256_CHROMA[q], MT_EOB_BIN_512_LUMA[q],  MT_EOB_BIN_512_CHROMA[q],
                                MT_EOB_BIN_1024_LUMA[q], MT_EOB_BIN_1024_CHROMA[q]};
    for (int e = 0; e < 17; e++) memcpy(c.eob_bin[e], bins[e], kEobBinK[e] * sizeof(uint16_t));
    memcpy(c.txtp4, MT_TXTP_INTRA1_TX4, sizeof(c.txtp4));
    memcpy(c.txtp8, MT_TXTP_INTRA1_TX8, sizeof(c.txtp8));
    memcpy(c.txtp16, MT_TXTP_INTRA2_TX16, sizeof(c.txtp16));
}

// Symbol costs of a snapshot (CoefCostTables::build, src/rate.rs:58).
void build_tables(const mt_coef_cdfs& c, RateTables& t) {
    memset(&t, 0, sizeof(t));
    for (int k = 0; k < 5; k++)
        for (int x = 0; x < 13; x++)
            for (int s = 0; s < 2; s++) t.skip[k][x][s] = sym_cost(&c.skip[k][x], 1, s);
    // Contexts past a table's end stay 0.0, as in the reference's zero-initialised table
    // (the 2-D context walk never reaches them).
    const int nctx[5] = {41, 41, 42, 42, 42};
    for (int k = 0; k < 5; k++)
        for (int p = 0; p < 2; p++)
            if (!(k == 4 && p == 1))
                for (int x = 0; x < nctx[k]; x++)
                    for (int s = 0; s < 4; s++) t.bt[k][p][x][s] = sym_cost(c.base_tok[k][p][x], 3, s);
    for (int k = 0; k < 4; k++)
        for (int p = 0; p < 2; p++)
            for (int x = 0; x < 21; x++) ladder(c.br_tok[k][p][x], t.br[k][p][x]);
    for (int k = 0; k < 5; k++)
        for (int p = 0; p < 2; p++)
            if (!(k == 4 && p == 1))
                for (int x = 0; x < 4; x++)
                    for (int s = 0; s < 3; s++) t.eob_base[k][p][x][s] = sym_cost(c.eob_base[k][p][x], 2, s);
    for (int k = 0; k < 5; k++)
        for (int p = 0; p < 2; p++)
            if (!(k == 4 && p == 1))
                for (int x = 0; x < 11; x++)
                    for (int s = 0; s < 2; s++) t.eob_hi[k][p][x][s] = sym_cost(&c.eob_hi[k][p][x], 1, s);
    for (int p = 0; p < 2; p++)
        for (int x = 0; x < 3; x++)
            for (int s = 0; s < 2; s++) t.dc_sign[p][x][s] = sym_cost(&c.dc_sign[p][x], 1, s);
    for (int e = 0; e < 17; e++)
        for (int s = 0; s <= kEobBinK[e]; s++) t.eob_bin[e][s] = sym_cost(c.eob_bin[e], kEobBinK[e], s);
    for (int m = 0; m < 13; m++) {
        for (int s = 0; s <= 6; s++) t.txtp[0][m][s] = sym_cost(c.txtp4[m], 6, s);
        for (int s = 0; s <= 6; s++) t.txtp[1][m][s] = sym_cost(c.txtp8[m], 6, s);
        for (int s = 0; s <= 4; s++) t.txtp[2][m][s] = sym_cost(c.txtp16[m], 4, s);
    }
}

// The trellis's lane-group classes: the coded sizes by coefficient count and shorter side
// (the longest anti-diagonal); a block runs on as many lanes as its shorter side.
constexpr int kTrClasses = 8;

__device__ inline int trellis_class(const TrellisGeom& g) {
    switch (g.N) {
        case 16: return 0;
        case 32: return 1;
        case 64: return min(g.w, g.h) == 4 ? 2 : 3;
        case 128: return 4;
        case 256: return 5;
        case 512: return 6;
        default: return 7;
    }
}

// Sorts the blocks that need work into one list per class (class c's list at order + c * n,
// its length in counts[c]); writes the status of the others: MT_ERR_ARG for an invalid
// descriptor, MT_OK for lambda0 <= 0.  One reservation per class per workgroup.
__global__ void __launch_bounds__(256) trellis_sort_kernel(const mt_trellisblk* __restrict__ blocks, int n,
                                                          uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ order,
                                                          int32_t* __restrict__ status) {
    __shared__ uint32_t h[kTrClasses], basepos[kTrClasses];
    const int chunk_begin = blockIdx.x * 4096;
    const int chunk_end = min(n, chunk_begin + 4096);
    if (threadIdx.x < kTrClasses) h[threadIdx.x] = 0;
    __syncthreads();
    for (int i = chunk_begin + threadIdx.x; i < chunk_end; i += blockDim.x) {
        const mt_trellisblk d = blocks[i];
        TrellisGeom g;
        if (!trellis_valid(d, g)) {
            if (status) status[i] = MT_ERR_ARG;
        } else if (!(d.lambda0 > 0.0f)) {
            if (status) status[i] = MT_OK;
        } else {
            atomicAdd(&h[trellis_class(g)], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < kTrClasses) {
        const int c = threadIdx.x;
        basepos[c] = h[c] ? atomicAdd(&counts[c], h[c]) : 0;
        h[c] = 0;
    }
    __syncthreads();
    for (int i = chunk_begin + threadIdx.x; i < chunk_end; i += blockDim.x) {
        const mt_trellisblk d = blocks[i];
        TrellisGeom g;
        if (trellis_valid(d, g) && d.lambda0 > 0.0f) {
            const int c = trellis_class(g);
            order[(size_t)c * n + basepos[c] + atomicAdd(&h[c], 1u)] = (uint32_t)i;
        }
    }
}

// The blocks of one class (NE coefficients, G lanes each), 64 / G per one-wave workgroup,
// persistent over the class's list: each group stages its block's levels and targets in LDS
// (16-byte loads where aligned), runs trellis_group, and writes the levels back.
template <int G, int NE, int NS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
trellis_group_kernel(const mt_trellisblk* __restrict__ blocks, const uint32_t* __restrict__ list,
                     const uint32_t* __restrict__ count, int32_t* __restrict__ levels,
                     const float* __restrict__ targets, int32_t* __restrict__ status, mtlib::TrellisTabs tabs,
                     int so0) {
    constexpr int NG = 64 / G;
    __shared__ uint16_t s_scan[NS * NE];  // the class's scans (NS sizes, one after another)
    for (int k = threadIdx.x; k < NS * NE; k += 64) s_scan[k] = tabs.scan[so0 + k];
    __shared__ __align__(16) int32_t s_cf[NG * NE];
    __shared__ __align__(16) float s_tf[NG * NE];
    __shared__ __align__(16) float s_a[NG * NE];
    __shared__ __align__(16) float s_b[NG * NE];
    constexpr int LBN = NE + 2 * (NE / G) + 4;  // the level bytes and their zero pad
    __shared__ uint8_t s_lb[NG * LBN];
    const int g = threadIdx.x / G, gl = threadIdx.x & (G - 1);
    int32_t* cf = s_cf + g * NE;
    float* tf = s_tf + g * NE;
    const uint32_t cnt = *count;
    for (uint32_t base = blockIdx.x * NG; base < cnt; base += gridDim.x * NG) {
        const uint32_t j = base + g;
        const bool have = j < cnt;
        const uint32_t b = have ? list[j] : 0u;
        mt_trellisblk d{};
        if (have) d = blocks[b];
        TrellisGeom gm{};
        const int ne = have && trellis_geom(d.tx_size, gm) && gm.N == NE ? NE : 0;
        int32_t* lv = levels + d.offset;
        const float* tg = targets + d.offset;
        const bool vec = ((reinterpret_cast<uintptr_t>(lv) | reinterpret_cast<uintptr_t>(tg)) & 15) == 0;
        if (vec) {
            for (int k = gl; k < ne / 4; k += G) {
                reinterpret_cast<int4*>(cf)[k] = reinterpret_cast<const int4*>(lv)[k];
                reinterpret_cast<float4*>(tf)[k] = reinterpret_cast<const float4*>(tg)[k];
            }
        } else {
            for (int k = gl; k < ne; k += G) {
                cf[k] = lv[k];
                tf[k] = tg[k];
            }
        }
        const int rc = trellis_group<G>(d, ne > 0, cf, tf, s_a + g * NE, s_b + g * NE, s_lb + g * LBN,
                                        s_scan + (ne > 0 ? scan_offset(d.tx_size) - so0 : 0), tabs);
        if (rc == MT_OK) {
            if (vec) {
                for (int k = gl; k < ne / 4; k += G) reinterpret_cast<int4*>(lv)[k] = reinterpret_cast<const int4*>(cf)[k];
            } else {
                for (int k = gl; k < ne; k += G) lv[k] = cf[k];
            }
        }
        if (have && gl == 0 && status) status[b] = ne > 0 ? rc : MT_ERR_ARG;
        __syncthreads();  // the write-back has read the LDS the next block is staged into
    }
}

struct ScanTabs {
    uint16_t scan[kScanTotal], iscan[kScanTotal];
};

void build_scans(ScanTabs& st) {
    struct {
        int s;
        const uint32_t* tab;
        int n;
    } list[12] = {{MT_TX_4X4, MT_SCAN_4X4, 16},       {MT_TX_4X8, MT_SCAN_4X8, 32},
                  {MT_TX_8X4, MT_SCAN_8X4, 32},       {MT_TX_8X8, MT_SCAN_8X8, 64},
                  {MT_TX_4X16, MT_SCAN_4X16, 64},     {MT_TX_16X4, MT_SCAN_16X4, 64},
                  {MT_TX_8X16, MT_SCAN_8X16, 128},    {MT_TX_16X8, MT_SCAN_16X8, 128},
                  {MT_TX_16X16, MT_SCAN_16X16, 256},  {MT_TX_16X32, MT_SCAN_16X32, 512},
                  {MT_TX_32X16, MT_SCAN_32X16, 512},  {MT_TX_32X32, MT_SCAN_32X32, 1024}};
    for (auto& e : list) {
        const int o = scan_offset(e.s);
        for (int i = 0; i < e.n; i++) {
            st.scan[o + i] = (uint16_t)e.tab[i];
            st.iscan[o + e.tab[i]] = (uint16_t)i;
        }
    }
}

}  // namespace

namespace {
mtlib::TrellisTabs g_tabs[64];
}  // namespace

namespace mtlib {
TrellisTabs trellis_tabs(int device) { return g_tabs[device & 63]; }

int rate_init_device(int device) {
    static RateTables host[MT_RATE_SLOTS];
    static ScanTabs scans;
    static std::once_flag once;
    std::call_once(once, [] {
        for (int q = 0; q < MT_RATE_SLOTS; q++) {  // the loadable slots start as Cdfs::new(0)
            mt_coef_cdfs c;
            default_cdfs(q < MT_RATE_SLOTS_FIXED ? q : 0, c);
            build_tables(c, host[q]);
        }
        build_scans(scans);
    });
    uint8_t lo[3][5][5];
    for (int y = 0; y < 5; y++)
        for (int x = 0; x < 5; x++) {
            lo[0][y][x] = (uint8_t)MT_LO_CTX_OFF[y][x];
            lo[1][y][x] = (uint8_t)MT_LO_CTX_OFF_WLH[y][x];
            lo[2][y][x] = (uint8_t)MT_LO_CTX_OFF_WGH[y][x];
        }
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_rate), host, sizeof(host)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_scan), scans.scan, sizeof(scans.scan)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_iscan), scans.iscan, sizeof(scans.iscan)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_lo_off), lo, sizeof(lo)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tr_dcq), MT_DC_QLOOKUP, sizeof(MT_DC_QLOOKUP)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tr_acq), MT_AC_QLOOKUP, sizeof(MT_AC_QLOOKUP)));
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_tr_iqm), MT_IQM, sizeof(MT_IQM)));
    // the trellis's tables by address, for the kernels of every unit
    TrellisTabs t{};
    void* p = nullptr;
    MT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(g_rate)));
    t.rate = (const RateTables*)p;
    MT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(c_scan)));
    t.scan = (const uint16_t*)p;
    MT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(c_lo_off)));
    t.lo_off = (const uint8_t*)p;
    MT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(c_tr_dcq)));
    t.dcq = (const uint16_t*)p;
    MT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(c_tr_acq)));
    t.acq = (const uint16_t*)p;
    MT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(g_tr_iqm)));
    t.iqm = (const uint8_t*)p;
    if (!t.rate || !t.scan || !t.lo_off || !t.dcq || !t.acq || !t.iqm)
        return set_error(MT_ERR_DEVICE, "rate tables: no device address");
    g_tabs[device & 63] = t;
    return MT_OK;
}
}  // namespace mtlib

namespace mtlib {
int rate_launch_with_nnz(const mt_rateblk* blocks_d, int n, const int32_t* levels_d, float* bits_d,
                         const mt_txres* nnz_d, hipStream_t s);
}

extern "C" int mt_block_rate(const mt_rateblk* blocks_d, int n, const int32_t* levels_d, float* bits_d,
                             void* stream) {
    if (n < 0 || (n > 0 && (!blocks_d || !levels_d || !bits_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_block_rate: bad pointer or count");
    return mtlib::rate_launch_with_nnz(blocks_d, n, levels_d, bits_d, nullptr, (hipStream_t)stream);
}

namespace {
struct RateWork {
    size_t cap = 0;
    uint32_t* list = nullptr;
    uint32_t* cnt = nullptr;
};
std::mutex g_rate_mu;
std::map<mtlib::StreamKey, RateWork> g_rate_work;
}  // namespace

namespace mtlib {
int rate_grow(RateWork& w, size_t n) {
    if (w.cap < n) {
        if (w.list) MT_HIP_TRY(hipFree(w.list));
        w.list = nullptr;
        const size_t cap = n < 65536 ? 65536 : n;
        MT_HIP_TRY(hipMalloc(&w.list, cap * sizeof(uint32_t)));
        w.cap = cap;
    }
    if (!w.cnt) MT_HIP_TRY(hipMalloc(&w.cnt, sizeof(uint32_t)));
    return MT_OK;
}
int rate_reserve(size_t n, hipStream_t s) {
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_rate_mu);
    return rate_grow(g_rate_work[mtlib::stream_key(dev, s)], n);
}

// mt_block_rate for the candidate pipeline: the small blocks one lane each with the chain's
// nonzero counts steering the eob search, the large blocks one wave each.
int rate_launch_with_nnz(const mt_rateblk* blocks_d, int n, const int32_t* levels_d, float* bits_d,
                         const mt_txres* nnz_d, hipStream_t s) {  // levels of any alignment
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_rate_mu);
    RateWork& w = g_rate_work[mtlib::stream_key(dev, s)];
    if ((rc = rate_grow(w, (size_t)n))) return rc;
    ProfScope prof(kProfRate, s);
    hipLaunchKernelGGL(zero_counts_kernel<1>, dim3(1), dim3(64), 0, s, w.cnt);
    hipLaunchKernelGGL(rate_classify_kernel, dim3(std::max(1, std::min((n + 255) / 256, cus * 4))),
                       dim3(256), 0, s, blocks_d, n, w.list, w.cnt);
    hipLaunchKernelGGL(rate_kernel, dim3((n + 255) / 256), dim3(256), 0, s, blocks_d, n, levels_d,
                       bits_d, nnz_d, 1);
    const int gw = n < cus * 32 ? n : cus * 32;
    hipLaunchKernelGGL(rate_wave_kernel, dim3(gw), dim3(64), 0, s, blocks_d, w.list, w.cnt, levels_d,
                       bits_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib

namespace mtlib {
float cdf_cost_raw(const uint16_t* raw_cdf, int nsym, int s) { return sym_cost(raw_cdf, nsym - 1, s); }
}  // namespace mtlib

extern "C" int mt_rate_default_cdfs(int qctx, mt_coef_cdfs* out) {
    if (qctx < 0 || qctx >= MT_RATE_SLOTS_FIXED || !out)
        return mtlib::set_error(MT_ERR_ARG, "mt_rate_default_cdfs: bad argument");
    default_cdfs(qctx, *out);
    return MT_OK;
}

extern "C" int mt_rate_load_cdfs(int slot, const mt_coef_cdfs* cdfs) {
    if (slot < MT_RATE_SLOTS_FIXED || slot >= MT_RATE_SLOTS || !cdfs)
        return mtlib::set_error(MT_ERR_ARG, "mt_rate_load_cdfs: bad slot or pointer");
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    RateTables t;
    build_tables(*cdfs, t);
    MT_HIP_TRY(hipDeviceSynchronize());
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_rate), &t, sizeof(t), (size_t)slot * sizeof(t)));
    MT_HIP_TRY(hipDeviceSynchronize());
    return MT_OK;
}

namespace {
struct TrellisWork {
    size_t cap = 0;
    uint32_t* order = nullptr;   // kTrClasses lists of cap entries
    uint32_t* counts = nullptr;  // kTrClasses
};
std::mutex g_trellis_mu;
std::map<mtlib::StreamKey, TrellisWork> g_trellis_work;

// so0: the scan offset of the class's first size (its sizes' scans lie one after another)
template <int G, int NE, int NS>
void launch_trellis_class(const mt_trellisblk* blocks_d, const TrellisWork& w, int c, int so0, int n, int cus,
                          int32_t* levels_d, const float* targets_d, int32_t* status_d, hipStream_t s, int dev) {
    constexpr int NG = 64 / G;
    const int grid = std::max(1, std::min((n + NG - 1) / NG, cus * 8));
    hipLaunchKernelGGL((trellis_group_kernel<G, NE, NS>), dim3(grid), dim3(64), 0, s, blocks_d,
                       w.order + (size_t)c * n, w.counts + c, levels_d, targets_d, status_d, g_tabs[dev & 63], so0);
}
}  // namespace

namespace mtlib {
int trellis_grow(TrellisWork& w, size_t n) {
    if (w.cap < n) {
        if (w.order) MT_HIP_TRY(hipFree(w.order));
        w.order = nullptr;
        w.cap = 0;
        const size_t cap = n < 65536 ? 65536 : n;
        MT_HIP_TRY(hipMalloc(&w.order, kTrClasses * cap * sizeof(uint32_t)));
        w.cap = cap;
    }
    if (!w.counts) MT_HIP_TRY(hipMalloc(&w.counts, kTrClasses * sizeof(uint32_t)));
    return MT_OK;
}
int trellis_reserve(size_t n, hipStream_t s) {
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_trellis_mu);
    return trellis_grow(g_trellis_work[mtlib::stream_key(dev, s)], n);
}

// Sorts the blocks into lane-group classes (lists per device and stream, grown on demand),
// then one persistent launch per class.
int trellis_launch(const mt_trellisblk* blocks_d, int n, int32_t* levels_d, const float* targets_d,
                   int32_t* status_d, hipStream_t s) {
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_trellis_mu);
    TrellisWork& w = g_trellis_work[mtlib::stream_key(dev, s)];
    if ((rc = trellis_grow(w, (size_t)n))) return rc;
    ProfScope prof(kProfTrellis, s);
    hipLaunchKernelGGL(zero_counts_kernel<kTrClasses>, dim3(1), dim3(64), 0, s, w.counts);
    hipLaunchKernelGGL(trellis_sort_kernel, dim3((n + 4095) / 4096), dim3(256), 0, s, blocks_d, n, w.counts,
                       w.order, status_d);
    // the lists are n apart whatever the capacity, so their offsets match the sort's
    launch_trellis_class<4, 16, 1>(blocks_d, w, 0, scan_offset(MT_TX_4X4), n, cus, levels_d, targets_d, status_d, s, dev);
    launch_trellis_class<4, 32, 2>(blocks_d, w, 1, scan_offset(MT_TX_4X8), n, cus, levels_d, targets_d, status_d, s, dev);
    launch_trellis_class<4, 64, 2>(blocks_d, w, 2, scan_offset(MT_TX_4X16), n, cus, levels_d, targets_d, status_d, s,
                                   dev);
    launch_trellis_class<8, 64, 1>(blocks_d, w, 3, scan_offset(MT_TX_8X8), n, cus, levels_d, targets_d, status_d, s, dev);
    launch_trellis_class<8, 128, 2>(blocks_d, w, 4, scan_offset(MT_TX_8X16), n, cus, levels_d, targets_d, status_d, s,
                                    dev);
    launch_trellis_class<16, 256, 1>(blocks_d, w, 5, scan_offset(MT_TX_16X16), n, cus, levels_d, targets_d, status_d,
                                     s, dev);
    launch_trellis_class<16, 512, 2>(blocks_d, w, 6, scan_offset(MT_TX_16X32), n, cus, levels_d, targets_d, status_d,
                                     s, dev);
    launch_trellis_class<32, 1024, 1>(blocks_d, w, 7, scan_offset(MT_TX_32X32), n, cus, levels_d, targets_d, status_d,
                                      s, dev);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib

extern "C" int mt_trellis(const mt_trellisblk* blocks_d, int n, int32_t* levels_d, const float* targets_d,
                          float* work_d, int32_t* status_d, void* stream) {
    if (n < 0 || (n > 0 && (!blocks_d || !levels_d || !targets_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_trellis: bad pointer or count");
    (void)work_d;  // the trellis works in LDS
    return mtlib::trellis_launch(blocks_d, n, levels_d, targets_d, status_d, (hipStream_t)stream);
}

// trellis_lambda() * hbd_trellis_scale (src/cost.rs:110, src/coder/lossy_state.rs:158): above
// 8 bits at 4:2:0 / 4:2:2 the lambda ramps from 1 to `top` as qindex falls from `hi` to `lo`
// (shipped 0.174 / 36 / 20 at 4:2:0, 1.0 / 24 / 10 at 4:2:2, src/tuning.rs:954-959)
extern "C" float mt_trellis_lambda0(int qidx, int bd, int sub) {
    float scale = 1.0f;
    if (bd > 8 && (sub == 1 || sub == 2)) {
        const float top = sub == 2 ? 0.174f : 1.0f, hi = sub == 2 ? 36.0f : 24.0f, lo = sub == 2 ? 20.0f : 10.0f;
        float wgt = (hi - (float)qidx) / std::max(hi - lo, 1.0f);
        wgt = std::min(std::max(wgt, 0.0f), 1.0f);
        scale = (top - 1.0f) * wgt + 1.0f;
    }
    return 0.05f * scale;
}
