// CPU oracle: one intra candidate evaluated end to end, and a threaded batch of them.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Composes the oracle's restated kernels in the
// order of the reference's candidate loop body (src/coder/partition_search.rs:471-520):
// predict_nd_ad / dc_pred / filter_predict, residual_pred, DctDispatch forward + quant
// (+QM), IdctDispatch inverse, sse_recon / weighted distortion / luma_satd over the
// displayed part (src/rd_sse.rs:454, :244, :635), real_block_bits_bounded
// (src/rate.rs:149) plus the caller's extra bits, rd_cost (src/partition_rd.rs:140).
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

int tx_dim_log2(int v) { return __builtin_ctz((unsigned)v); }

// recon (optional): clamp(pred + inverse residual), w*h row-major; levels_out (optional):
// the levels the inverse reads (after the trellis), ncoef x-major.
void eval_one(const orc_cand& c, const orc_plane* planes, int nplanes, orc_candres& r,
              int32_t* recon = nullptr, int32_t* levels_out = nullptr) {
    std::memset(&r, 0, sizeof(r));
    r.sse = -1;
    r.dist = -1;
    r.cost = __builtin_inff();
    const int s = c.tx_size, t = c.tx_type;
    if (s > 18) {
        r.status = -1;
        return;
    }
    const int w = orc_tx_w(s), h = orc_tx_h(s);
    const int si = (c.planes & 3) | ((c.planes_hi & 15) << 2);
    const int ri = ((c.planes >> 2) & 3) | ((c.planes_hi >> 4) << 2);
    const int bdi = (c.bd_qm >> 4) & 3;
    if (si >= nplanes || ri >= nplanes || bdi > 2) {
        r.status = -1;
        return;
    }
    const int bd = bdi == 0 ? 8 : bdi == 1 ? 10 : 12, qm = c.bd_qm & 15;
    const orc_plane& sp = planes[si];
    const orc_plane& rp = planes[ri];
    if (c.x + w > sp.width || c.y + h > sp.height || c.x + w > rp.width || c.y + h > rp.height) {
        r.status = -1;
        return;
    }
    if (!orc_tx_supported(s, t)) {
        r.status = -2;
        return;
    }
    std::vector<int32_t> pred(w * h), resid(w * h), inv(w * h), levels(orc_tx_ncoef(s));
    std::vector<float> targets(orc_tx_ncoef(s));
    const bool cfl = c.mode == 13;
    const int ssx = (c.pflags >> 3) & 1, ssy = (c.pflags >> 4) & 1;
    if (cfl && (w > 32 || h > 32 || c.txtp_set != 0 || c.txtp_mode >= nplanes || c.angle_delta < -16 ||
                c.angle_delta > 16 || ((c.x + w) << ssx) > planes[c.txtp_mode].width ||
                ((c.y + h) << ssy) > planes[c.txtp_mode].height)) {
        r.status = -1;
        return;
    }
    if (orc_intra_pred(cfl ? 0 : c.mode, cfl ? 0 : c.angle_delta, rp.data, rp.stride, c.x, c.y, w, h,
                       cfl ? 0 : c.pflags & 1, cfl ? 0 : (c.pflags >> 1) & 1, rp.width, rp.height,
                       cfl ? 0 : (c.pflags >> 2) & 1, bd, pred.data()) != 0) {
        r.status = -1;
        return;
    }
    const uint16_t* src = sp.data + (size_t)c.y * sp.stride + c.x;
    if (cfl) {  // cfl_ac_sub / cfl_ac_444, cfl_best_alpha, cfl_pred_pixel (src/intrapred.rs:1112-1218)
        const orc_plane& lp = planes[c.txtp_mode];
        std::vector<int32_t> ac(w * h);
        orc_cfl_ac_sub(lp.data + (size_t)(c.y << ssy) * lp.stride + (c.x << ssx), lp.stride, w, h, ssx, ssy,
                       ac.data());
        const int dc = pred[0];
        const int vw = std::min(std::max(sp.vis_w - (int)c.x, 0), w);
        const int vh = std::min(std::max(sp.vis_h - (int)c.y, 0), h);
        int alpha = (c.pflags & 32) ? c.angle_delta : orc_cfl_best_alpha(ac.data(), w, src, sp.stride, vw, vh, dc, bd);
        if (!(c.pflags & 32) && (c.pflags & 64)) alpha = orc_cfl_alpha_step(alpha, c.angle_delta);
        for (int i = 0; i < w * h; ++i) pred[i] = orc_cfl_pred_pixel(dc, ac[i], alpha, bd);
        r.alpha = alpha;
    }
    orc_residual_pred(src, sp.stride, pred.data(), w, h, resid.data());
    orc_quant q;
    if (c.bd_qm & 0x40)  // Quant::new_chroma_with_delta_qm (src/quant.rs:394)
        orc_quant_new_chroma(c.qidx, c.dc_delta, c.ac_delta, bd, qm, &q);
    else  // Quant::new_with_qm (src/quant.rs:354)
        orc_quant_new(c.qidx, bd, qm, &q);
    orc_fwd_txfm(s, t, resid.data(), &q, levels.data(), targets.data());
    if (c.rflags & 2) {  // trellis_optimize_ctx in the candidate's rate context
        orc_rate_ctx tc;
        std::memset(&tc, 0, sizeof(tc));
        tc.qctx = c.qctx;
        tc.cls = (tx_dim_log2(w) - 2 + tx_dim_log2(h) - 2 + 1) >> 1;
        tc.plane = c.rplane;
        tc.w = std::min(w, 32);
        tc.h = std::min(h, 32);
        tc.eob_bin = c.eob_bin;
        tc.dcs_ctx = c.dcs_ctx;
        orc_trellis_ctx(levels.data(), targets.data(), (float)q.dc, (float)q.ac, orc_scan(s),
                        orc_trellis_lambda0(c.qidx, bd, (c.rflags >> 2) & 3), &tc, q.qm_level, 0.0f);
    }
    orc_inv_txfm(s, t, levels.data(), &q, inv.data());
    if (levels_out) std::memcpy(levels_out, levels.data(), levels.size() * sizeof(int32_t));
    const int vw = std::min(std::max(sp.vis_w - (int)c.x, 0), w);
    const int vh = std::min(std::max(sp.vis_h - (int)c.y, 0), h);
    r.sse = orc_sse_recon_vis(pred.data(), inv.data(), src, sp.stride, w, vw, vh, bd);
    if (recon)
        for (int i = 0; i < w * h; ++i) recon[i] = std::min(std::max(pred[i] + inv[i], 0), (1 << bd) - 1);
    if (c.wsse_scale > 0.0f) {
        const int maxv = (1 << bd) - 1;
        std::vector<int32_t> err((size_t)std::max(vw * vh, 1));
        for (int y = 0; y < vh; ++y)
            for (int x = 0; x < vw; ++x) {
                const int i = y * w + x;
                const int rec = std::min(std::max(pred[i] + inv[i], 0), maxv);
                err[y * vw + x] = (int32_t)src[(size_t)y * sp.stride + x] - rec;
            }
        r.dist = orc_weighted_dist(err.data(), vw, vh, c.wsse_scale, 1);
    } else {
        r.dist = r.sse;
    }
    r.satd = orc_luma_satd(src, sp.stride, w, h, vw, vh, bd, pred.data(), inv.data());
    orc_rate_ctx rc;
    std::memset(&rc, 0, sizeof(rc));
    rc.qctx = c.qctx;
    rc.cls = (tx_dim_log2(w) - 2 + tx_dim_log2(h) - 2 + 1) >> 1;
    rc.plane = c.rplane;
    rc.w = std::min(w, 32);
    rc.h = std::min(h, 32);
    rc.eob_bin = c.eob_bin;
    rc.skip_ctx = c.skip_ctx;
    rc.dcs_ctx = c.dcs_ctx;
    rc.txtp_set = c.txtp_set;
    rc.txtp_mode = cfl ? 0 : c.txtp_mode;  // CfL: txtp_mode names the luma plane
    rc.txtp = c.txtp;
    rc.use_skip_bits = c.rflags & 1;
    rc.skip_one = c.skip_one;
    rc.skip_zero = c.skip_zero;
    const float bits = orc_block_rate(levels.data(), orc_scan(s), &rc, __builtin_inff());
    r.bits = bits + c.extra_bits;
    r.cost = (float)r.dist + c.lambda * r.bits;
    uint32_t nnz = 0;
    for (int32_t v : levels) nnz += v != 0;
    r.nnz = nnz;
    r.status = 0;
}

}  // namespace

extern "C" void orc_eval_candidate(const orc_cand* c, const orc_plane* planes, int nplanes,
                                   orc_candres* r) {
    eval_one(*c, planes, nplanes, *r);
}

extern "C" void orc_eval_candidate_recon(const orc_cand* c, const orc_plane* planes, int nplanes,
                                         orc_candres* r, int32_t* recon) {
    eval_one(*c, planes, nplanes, *r, recon);
}

extern "C" void orc_eval_candidate_recon_levels(const orc_cand* c, const orc_plane* planes, int nplanes,
                                                orc_candres* r, int32_t* recon, int32_t* levels) {
    eval_one(*c, planes, nplanes, *r, recon, levels);
}

extern "C" void orc_eval_batch(const orc_cand* c, int n, const orc_plane* planes, int nplanes,
                               orc_candres* r, int nthreads) {
    if (nthreads <= 1) {
        for (int i = 0; i < n; ++i) eval_one(c[i], planes, nplanes, r[i]);
        return;
    }
    std::vector<std::thread> th;
    for (int k = 0; k < nthreads; ++k)
        th.emplace_back([=] {
            for (int i = k; i < n; i += nthreads) eval_one(c[i], planes, nplanes, r[i]);
        });
    for (auto& x : th) x.join();
}
