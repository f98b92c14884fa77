// Superblock reconstruction of the intra RDO (part of the mtgpu.hip translation unit: it
// uses that unit's transform tables).
//
// One wave per 64x64 superblock walks the superblock's chosen leaves in coding order.  Per
// leaf: the prediction (wave_predict, reading the reconstruction plane: the superblock's
// earlier leaves and the finished superblocks around it), the residual against the source,
// the forward DCT_DCT + quantization (+QM), dequantization and the inverse (the chain of
// the reference's DctDispatch / IdctDispatch for square sizes, src/dct.rs:2624-2950,
// src/idct.rs:168-728, :2677), and the reconstruction clamp(pred + residual) written back
// to the plane.  Every stage keeps the batched kernels' arithmetic, so the reconstruction
// equals the one the candidate evaluation measured for the same prediction.
#pragma once
#include "mt_pred_wave.h"
#include "mt_trellis.h"

namespace {

// The trellis between quantization and dequantization of a reconstruction chain: the
// block's descriptor (its offset unused), the tables, and LDS for the levels, the targets
// and the trellis's workspace (one element per coefficient each).
struct ReconTrellis {
    const mt_trellisblk* d;  // nullptr: no trellis
    const mtlib::TrellisTabs* tabs;
    int32_t* lv;
    float* tg;
    float* suf;   // the trellis's workspaces (16-byte aligned): pre is the block's residual
    float* pre;   // buffer and lb the chain's work buffer, both free from the forward's last
    uint8_t* lb;  // read of them to the dequantization
};

// Square chain of one N x N block (N <= 32) by one wave: lane i runs line i of each pass.
// type: DCT_DCT, or at N <= 16 ADST_DCT / DCT_ADST / ADST_ADST (vertical kind first, the
// 2-D types the chain kernels run with the same per-size rules).  res: the residual (LDS,
// row-major); out: the reconstructed residual (may alias res); work: LDS of N * (N + 1) ints.
// With tr.d: the levels and targets go to LDS, the wave runs the trellis over them, and the
// dequantization reads the trellis's levels.
// lv_out (optional, global): the levels the dequantization reads, frequency (fx, fy) at
// lv_out[fy * lv_stride + fx].
template <int N>
__device__ void wave_chain_sq(const int32_t* res, int32_t* out, int32_t* work, const BlkParam& p,
                              int type = MT_TXTP_DCT_DCT, ReconTrellis tr = ReconTrellis{},
                              int32_t* lv_out = nullptr, int lv_stride = 0) {
    using R = TxRules<N, N>;
    constexpr int RS = N + 1;
    const int lane = threadIdx.x & 63;
    const bool col_adst = N <= 16 && (type == MT_TXTP_ADST_ADST || type == MT_TXTP_ADST_DCT);
    const bool row_adst = N <= 16 && (type == MT_TXTP_ADST_ADST || type == MT_TXTP_DCT_ADST);
    // forward column pass (length N)
    if (lane < N) {
        int32_t v[N];
#pragma unroll
        for (int y = 0; y < N; y++) v[y] = (int32_t)((uint32_t)res[y * N + lane] << R::pre_shift);
        fwd_line<N>(col_adst, v);
#pragma unroll
        for (int y = 0; y < N; y++) work[y * RS + lane] = v[y];
    }
    __syncthreads();
    // forward row pass, post-scale, quant (+QM), dequant, in place
    if (lane < N) {
        const int fy = lane;
        int32_t v[N];
#pragma unroll
        for (int x = 0; x < N; x++) v[x] = work[fy * RS + x];
        fwd_line<N>(row_adst, v);
#pragma unroll
        for (int fx = 0; fx < N; fx++) {
            int32_t c = v[fx];
            if (R::post == 2) c = mulq16(c, 32768);
            else if (R::post == 4) c = (int32_t)(((int64_t)c * 16384 + (1ll << 21)) >> 22);
            const int rc = fx * N + fy;
            int32_t m = rc == 0 ? p.qdc : p.qac;
            if (R::ratio) m = mulq16(m, R::ratio);
            int32_t lvl = quant_q16(c, m);
            int32_t step = rc == 0 ? p.dc : p.ac;
            float tgt = q16_target(c, m);
            if (p.qm_level < 15) {
                const int iwt = g_iqm[p.qm_level][p.qm_chroma][R::qmoff + rc];
                lvl = qm_apply(tgt, iwt);
                step = (step * iwt + 16) >> 5;
            }
            if (tr.d) {
                tr.lv[rc] = lvl;
                tr.tg[rc] = tgt;
            } else {
                work[fy * RS + fx] = dequant(lvl, step, R::dq_shift, p.cf_max);
                if (lv_out) lv_out[fy * lv_stride + fx] = lvl;
            }
        }
    }
    __syncthreads();
    if (tr.d) {
        (void)trellis_group<64>(*tr.d, true, tr.lv, tr.tg, tr.suf, tr.pre, tr.lb,
                                tr.tabs->scan + scan_offset(tr.d->tx_size), *tr.tabs);
        __syncthreads();
        if (lane < N) {
            const int fy = lane;
#pragma unroll
            for (int fx = 0; fx < N; fx++) {
                const int rc = fx * N + fy;
                int32_t step = rc == 0 ? p.dc : p.ac;
                if (p.qm_level < 15) step = (step * (int32_t)g_iqm[p.qm_level][p.qm_chroma][R::qmoff + rc] + 16) >> 5;
                work[fy * RS + fx] = dequant(tr.lv[rc], step, R::dq_shift, p.cf_max);
                if (lv_out) lv_out[fy * lv_stride + fx] = tr.lv[rc];
            }
        }
        __syncthreads();
    }
    // inverse row pass (length N) with the row clip, inter-pass shift and column clip
    if (lane < N) {
        const int y = lane;
        int32_t v[N];
#pragma unroll
        for (int x = 0; x < N; x++) v[x] = work[y * RS + x];
        inv_line<N>(row_adst ? KA : KD, v, Clip{p.rmin, p.rmax});
        const Clip cc{p.cmin, p.cmax};
#pragma unroll
        for (int x = 0; x < N; x++) {
            int32_t t = v[x];
            if (R::mid) t = (t + (1 << (R::mid - 1))) >> R::mid;
            work[y * RS + x] = cc(t);
        }
    }
    __syncthreads();
    // inverse column pass (length N), (t + 8) >> 4
    if (lane < N) {
        const int x = lane;
        int32_t v[N];
#pragma unroll
        for (int y = 0; y < N; y++) v[y] = work[y * RS + x];
        inv_line<N>(col_adst ? KA : KD, v, Clip{p.cmin, p.cmax});
#pragma unroll
        for (int y = 0; y < N; y++) out[y * N + x] = (v[y] + 8) >> 4;
    }
    __syncthreads();
}

// TX_64X64 chain of one block by one wave: the f64 separable forward of dct64x64_coeffs
// (src/dct.rs:4155) with every product and sum rounded on its own in the reference's order,
// the 32x32-QM quantizer, and idct_dequant_64x64 (src/idct.rs:2677); the arithmetic of the
// batched 64x64 kernel.  vt: LDS of 16 * 64 doubles; coef: LDS of 32 * 33 ints.  With tr.d
// the trellis runs between quantization and dequantization (its levels and targets may live
// in vt, which the forward no longer needs by then).
__device__ void wave_chain64(const int32_t* res, int32_t* out, double* vt, int32_t* coef,
                             const BlkParam& p, ReconTrellis tr = ReconTrellis{}, int32_t* lv_out = nullptr,
                             int lv_stride = 0) {
    const int lane = threadIdx.x & 63;
    int32_t* rowbuf = reinterpret_cast<int32_t*>(vt);  // 32 x 64 int32 after the forward
    const bool qm = p.qm_level < 15;
    const uint8_t* qmrow = &g_iqm[qm ? p.qm_level : 0][p.qm_chroma][336];
    double acc[32];
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = 0.0;
#pragma unroll 1
    for (int y = 0; y < 64; y++) {
        const double c = (double)res[y * 64 + lane];
        double bv[32];  // the basis row in one batch of scalar loads (one wait per row)
#pragma unroll
        for (int j = 0; j < 32; j++) bv[j] = c_dct64_basis_t[y * 32 + j];
#pragma unroll
        for (int j = 0; j < 32; j++) acc[j] = __dadd_rn(acc[j], __dmul_rn(bv[j], c));
    }
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
            // the lane's basis values stream one pair of x ahead of their use
            double bn0 = c_dct64_basis_t[u], bn1 = c_dct64_basis_t[32 + u];
#pragma unroll 2
            for (int x = 0; x < 64; x += 2) {
                const double b0 = bn0, b1 = bn1;
                const int xn = x + 2 < 64 ? x + 2 : 62;
                bn0 = c_dct64_basis_t[xn * 32 + u];
                bn1 = c_dct64_basis_t[(xn + 1) * 32 + u];
#pragma unroll
                for (int k = 0; k < 8; k++) hacc[k] = __dadd_rn(hacc[k], __dmul_rn(vt[(2 * k + vp) * 64 + x], b0));
#pragma unroll
                for (int k = 0; k < 8; k++)
                    hacc[k] = __dadd_rn(hacc[k], __dmul_rn(vt[(2 * k + vp) * 64 + x + 1], b1));
            }
#pragma unroll
            for (int k = 0; k < 8; k++)
                coef[(hf * 16 + 2 * k + vp) * 33 + u] = (int32_t)round(__dmul_rn(hacc[k], 8.0));
        }
        __syncthreads();
    }
    {
        const int u = lane & 31, vp = lane >> 5;
#pragma unroll 1
        for (int k = 0; k < 16; k++) {
            const int v = 2 * k + vp;
            const int32_t coeff = coef[v * 33 + u];
            const int rc = u * 32 + v;
            const int32_t m = rc == 0 ? p.qdc : p.qac;
            int32_t lvl = quant_q16(coeff, m);
            float tgt = q16_target(coeff, m);
            const int iwt = qm ? qmrow[rc] : 32;
            if (qm) lvl = qm_apply(tgt, iwt);
            int32_t step = rc == 0 ? p.dc : p.ac;
            if (qm) step = (step * iwt + 16) >> 5;
            if (tr.d) {
                tr.lv[rc] = lvl;
                tr.tg[rc] = tgt;
            } else {
                coef[v * 33 + u] = dequant(lvl, step, 2, p.cf_max);
                if (lv_out) lv_out[v * lv_stride + u] = lvl;
            }
        }
    }
    __syncthreads();
    if (tr.d) {
        (void)trellis_group<64>(*tr.d, true, tr.lv, tr.tg, tr.suf, tr.pre, tr.lb,
                                tr.tabs->scan + scan_offset(tr.d->tx_size), *tr.tabs);
        __syncthreads();
        for (int rc = lane; rc < 1024; rc += 64) {
            int32_t step = rc == 0 ? p.dc : p.ac;
            if (qm) step = (step * (int32_t)qmrow[rc] + 16) >> 5;
            coef[(rc & 31) * 33 + (rc >> 5)] = dequant(tr.lv[rc], step, 2, p.cf_max);
            if (lv_out) lv_out[(rc & 31) * lv_stride + (rc >> 5)] = tr.lv[rc];
        }
        __syncthreads();
    }
    if (lane < 32) {
        const int y = lane;
        int32_t v[64];
#pragma unroll
        for (int x = 0; x < 64; x++) v[x] = x < 32 ? coef[y * 33 + x] : 0;
        idct64<1>(v, Clip{p.rmin, p.rmax});
        const Clip cc{p.cmin, p.cmax};
#pragma unroll
        for (int x = 0; x < 64; x++) rowbuf[y * 64 + x] = cc((v[x] + 2) >> 2);
    }
    __syncthreads();
    {
        const int x = lane;
        int32_t v[64];
#pragma unroll
        for (int y = 0; y < 64; y++) v[y] = y < 32 ? rowbuf[y * 64 + x] : 0;
        idct64<1>(v, Clip{p.cmin, p.cmax});
#pragma unroll
        for (int y = 0; y < 64; y++) out[y * 64 + x] = (v[y] + 8) >> 4;
    }
    __syncthreads();
}

// One workgroup (one wave) per superblock: leaves[first[b] .. first[b] + count[b]) in
// coding order, all of one frame.  Each leaf descriptor carries its prediction (x, y, size,
// mode, angle delta, availability and filter-type flags, bit depth) and, in `plane`, its
// frame's index in `frames` (source, reconstruction plane of the frame size, quantizer).
__global__ __launch_bounds__(64) void sb_recon_kernel(const mt_predblk* __restrict__ leaves,
                                                      const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ count,
                                                      const mtlib::ReconFrame* __restrict__ frames,
                                                      mtlib::TrellisTabs tabs) {
    __shared__ int32_t s_pred[4096];
    __shared__ __align__(16) int32_t s_res[4096];
    __shared__ int32_t s_a[kEdge], s_l[kEdge];
    __shared__ __align__(16) int32_t s_fi[33 * 33];
    __shared__ double s_vt[16 * 64];
    __shared__ int32_t s_work[32 * 33];
    const int lane = threadIdx.x;
    const uint32_t f0 = first[blockIdx.x], n = count[blockIdx.x];
    if (n == 0) return;
    const mtlib::ReconFrame fr = frames[leaves[f0].plane];
    const mt_plane src = fr.src;
    uint16_t* __restrict__ rec = fr.rec;
    const int rstride = fr.rstride;
    mt_txblk q{};
    q.qidx = (uint8_t)fr.qidx;
    q.bd_qm = (uint8_t)fr.bd_qm;
    BlkParam prm;
    param_from_desc(q, prm, nullptr, 0);
    const int maxv = (1 << prm.bd) - 1;
    // the trellis of the candidates (the levels in s_vt, its workspace in s_fi)
    mt_trellisblk td{};
    td.qidx = (uint8_t)fr.qidx;
    td.bd_qm = (uint8_t)fr.bd_qm;
    td.flags = MT_TRELLIS_CTX;
    td.qctx = (uint8_t)mtlib::rdo_qctx(fr.qidx);
    td.lambda0 = trellis_lambda0(fr.qidx, prm.bd, fr.tsub);
    const ReconTrellis rt{fr.trellis ? &td : nullptr, &tabs, reinterpret_cast<int32_t*>(s_vt),
                          reinterpret_cast<float*>(s_vt) + 1024, reinterpret_cast<float*>(s_fi),
                          reinterpret_cast<float*>(s_res), reinterpret_cast<uint8_t*>(s_work)};
    for (uint32_t i = 0; i < n; i++) {
        const mt_predblk d = leaves[f0 + i];
        const int s = d.bw, npx = s * s, lw = __builtin_ctz(s);
        td.tx_size = (uint8_t)(s == 4 ? MT_TX_4X4 : s == 8 ? MT_TX_8X8 : s == 16 ? MT_TX_16X16
                               : s == 32 ? MT_TX_32X32 : MT_TX_64X64);
        td.eob_bin = (uint8_t)(s == 4 ? 0 : s == 8 ? 5 : s == 16 ? 11 : 15);
        const int x0 = d.x, y0 = d.y;
        wave_predict(d, src.width, src.height,
                     [=](int x, int y) { return (int)rec[(size_t)y * rstride + x]; }, s_pred, s_a, s_l,
                     s_fi);
        __syncthreads();
        for (int p = lane; p < npx; p += 64)
            s_res[p] = (int)src.data[(size_t)(y0 + (p >> lw)) * src.stride + x0 + (p & (s - 1))] - s_pred[p];
        __syncthreads();
        int32_t* lvo = fr.coef ? fr.coef + (size_t)y0 * fr.coef_stride + x0 : nullptr;
        const int ls = fr.coef_stride;
        switch (s) {
            case 4: wave_chain_sq<4>(s_res, s_res, s_work, prm, MT_TXTP_DCT_DCT, rt, lvo, ls); break;
            case 8: wave_chain_sq<8>(s_res, s_res, s_work, prm, MT_TXTP_DCT_DCT, rt, lvo, ls); break;
            case 16: wave_chain_sq<16>(s_res, s_res, s_work, prm, MT_TXTP_DCT_DCT, rt, lvo, ls); break;
            case 32: wave_chain_sq<32>(s_res, s_res, s_work, prm, MT_TXTP_DCT_DCT, rt, lvo, ls); break;
            default: wave_chain64(s_res, s_res, s_vt, s_work, prm, rt, lvo, ls); break;
        }
        for (int p = lane; p < npx; p += 64) {
            const int v = s_pred[p] + s_res[p];
            rec[(size_t)(y0 + (p >> lw)) * rstride + x0 + (p & (s - 1))] =
                (uint16_t)(v < 0 ? 0 : v > maxv ? maxv : v);
        }
        __syncthreads();  // the next leaf predicts from these samples
    }
}

// Chroma of the RDO: one wave per (superblock, plane) walks the superblock's chroma
// transform blocks of that plane (blockIdx.y: 0 U, 1 V) in desc[first[b] .. + count[b]),
// in coding order (the planes do not depend on each other).  Per block: the prediction from the chroma
// reconstruction (CfL, mode 13: the DC prediction plus alpha (angle_delta) times the AC of
// the frame's luma reconstruction), the residual, the chain of the transform type in
// out_offset >> 1 with the chroma quantizer, and the clamped reconstruction; out_offset & 1
// selects V over U; `plane` is the frame's index in `frames`.
__global__ __launch_bounds__(64) void sb_recon_uv_kernel(const mt_predblk* __restrict__ desc,
                                                         const uint32_t* __restrict__ first,
                                                         const uint32_t* __restrict__ count,
                                                         const mtlib::ReconFrame* __restrict__ frames,
                                                         mtlib::TrellisTabs tabs) {
    __shared__ int32_t s_pred[1024];
    __shared__ __align__(16) int32_t s_res[1024];
    __shared__ int32_t s_a[kEdge], s_l[kEdge];
    __shared__ __align__(16) int32_t s_fi[33 * 33];
    __shared__ int32_t s_work[32 * 33];
    __shared__ int32_t s_tr[2048];  // the trellis's levels and targets
    const int lane = threadIdx.x;
    const uint32_t f0 = first[blockIdx.x], n = count[blockIdx.x];
    if (n == 0) return;
    const mtlib::ReconFrame fr = frames[desc[f0].plane];
    mt_txblk q{};
    q.qidx = (uint8_t)fr.qidx;
    q.bd_qm = (uint8_t)fr.uv_bd_qm;
    const int8_t qd[2] = {(int8_t)fr.uv_dc, (int8_t)fr.uv_ac};
    BlkParam prm;
    param_from_desc(q, prm, qd, 0);
    const int maxv = (1 << prm.bd) - 1;
    const int ustride = fr.uv_stride;
    mt_trellisblk td{};
    td.qidx = (uint8_t)fr.qidx;
    td.bd_qm = (uint8_t)fr.uv_bd_qm;
    td.dc_delta = (int8_t)fr.uv_dc;
    td.ac_delta = (int8_t)fr.uv_ac;
    td.flags = MT_TRELLIS_CTX;
    td.qctx = (uint8_t)mtlib::rdo_qctx(fr.qidx);
    td.plane = 1;
    td.lambda0 = trellis_lambda0(fr.qidx, prm.bd, fr.tsub);
    const ReconTrellis rt{fr.trellis ? &td : nullptr, &tabs, s_tr, reinterpret_cast<float*>(s_tr) + 1024,
                          reinterpret_cast<float*>(s_fi), reinterpret_cast<float*>(s_res),
                          reinterpret_cast<uint8_t*>(s_work)};
    const mt_plane luma{fr.rec, fr.rstride, fr.src.width, fr.src.height, fr.src.width, fr.src.height, 0};
    for (uint32_t i = 0; i < n; i++) {
        mt_predblk d = desc[f0 + i];
        const int pl = (int)(d.out_offset & 1), type = (int)(d.out_offset >> 1);
        if (pl != (int)blockIdx.y) continue;
        const mt_plane src = fr.src_uv[pl];
        uint16_t* __restrict__ rec = fr.rec_uv[pl];
        const int s = d.bw, npx = s * s, lw = __builtin_ctz(s);
        const int x0 = d.x, y0 = d.y;
        td.tx_size = (uint8_t)(s == 4 ? MT_TX_4X4 : s == 8 ? MT_TX_8X8 : s == 16 ? MT_TX_16X16 : MT_TX_32X32);
        td.eob_bin = (uint8_t)(s == 4 ? 1 : s == 8 ? 6 : s == 16 ? 12 : 16);
        const bool cfl = d.mode == MT_PRED_CFL;
        const int alpha = d.angle_delta;
        if (cfl) {
            d.mode = MT_PRED_DC;
            d.angle_delta = 0;
        }
        wave_predict(d, src.width, src.height,
                     [=](int x, int y) { return (int)rec[(size_t)y * ustride + x]; }, s_pred, s_a, s_l, s_fi);
        __syncthreads();
        if (cfl) {
            const int dc = s_pred[0];
            __syncthreads();
            cfl_block(luma, x0 << fr.ss, y0 << fr.ss, fr.ss != 0, fr.ss != 0, src, x0, y0, s, s, prm.bd, dc, true,
                      alpha, s_work, s_pred, nullptr, nullptr);
        }
        for (int p = lane; p < npx; p += 64)
            s_res[p] = (int)src.data[(size_t)(y0 + (p >> lw)) * src.stride + x0 + (p & (s - 1))] - s_pred[p];
        __syncthreads();
        int32_t* lvo = fr.coef_uv[pl] ? fr.coef_uv[pl] + (size_t)y0 * fr.coef_uv_stride + x0 : nullptr;
        const int ls = fr.coef_uv_stride;
        switch (s) {
            case 4: wave_chain_sq<4>(s_res, s_res, s_work, prm, type, rt, lvo, ls); break;
            case 8: wave_chain_sq<8>(s_res, s_res, s_work, prm, type, rt, lvo, ls); break;
            case 16: wave_chain_sq<16>(s_res, s_res, s_work, prm, type, rt, lvo, ls); break;
            default: wave_chain_sq<32>(s_res, s_res, s_work, prm, type, rt, lvo, ls); break;
        }
        for (int p = lane; p < npx; p += 64) {
            const int v = s_pred[p] + s_res[p];
            rec[(size_t)(y0 + (p >> lw)) * ustride + x0 + (p & (s - 1))] = (uint16_t)(v < 0 ? 0 : v > maxv ? maxv : v);
        }
        __syncthreads();  // the next block predicts from these samples
    }
}

}  // namespace

namespace mtlib {
int sb_recon_uv_launch(const mt_predblk* desc_d, const uint32_t* first_d, const uint32_t* count_d, int nsb,
                       const ReconFrame* frames_d, hipStream_t s) {
    if (nsb <= 0) return MT_OK;
    int dev = 0;
    MT_HIP_TRY(hipGetDevice(&dev));
    hipLaunchKernelGGL(sb_recon_uv_kernel, dim3(nsb, 2), dim3(64), 0, s, desc_d, first_d, count_d, frames_d,
                       trellis_tabs(dev));
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
int sb_recon_launch(const mt_predblk* leaves_d, const uint32_t* first_d, const uint32_t* count_d,
                    int nsb, const ReconFrame* frames_d, hipStream_t s) {
    if (nsb <= 0) return MT_OK;
    int dev = 0;
    MT_HIP_TRY(hipGetDevice(&dev));
    hipLaunchKernelGGL(sb_recon_kernel, dim3(nsb), dim3(64), 0, s, leaves_d, first_d, count_d, frames_d,
                       trellis_tabs(dev));
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib
