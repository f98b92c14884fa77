// The encoder's entry (mt_encode_rgb8): the reference's colour conversion of an 8-bit RGB
// image on the GPU, then the whole-frame RDO at the configuration's quantizer, then the
// in-loop deblocking filter on the reconstruction.
//
// Colour conversion (src/encoder.rs:114-127, :398-410, :560-602): full-range BT.601 in Q13
// integers, Y = clamp((Y_R r + Y_G g + Y_B b + 2^12) >> 13); at 4:4:4 Cb / Cr the same with
// their coefficients and a 2^(bd-1) << 13 offset; at 4:2:0 the four Q13 values of a 2 x 2
// quad are summed, + 2^14, >> 15.  One thread per 2 x 2 quad (4:2:0) or per sample (4:4:4):
// rows are read and written contiguously across the lanes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/mtgpu.h"
#include "mt_common.h"

namespace {

constexpr int kQ = 13;
constexpr int kYR = 2449, kYG = 4809, kYB = 934;       // round(0.299 / 0.587 / 0.114 * 8192)
constexpr int kCbR = -1382, kCbG = -2714, kCbB = 4096;  // round(-0.168736 / -0.331264 / 0.5 * 8192)
constexpr int kCrR = 4096, kCrG = -3430, kCrB = -666;   // round(0.5 / -0.418688 / -0.081312 * 8192)

__device__ __forceinline__ int clamp8(int v, int mx) { return v < 0 ? 0 : v > mx ? mx : v; }

__global__ void __launch_bounds__(256) csc420_kernel(const uint8_t* __restrict__ g, const uint8_t* __restrict__ b,
                                                     const uint8_t* __restrict__ r, int w, int h,
                                                     uint16_t* __restrict__ y, uint16_t* __restrict__ u,
                                                     uint16_t* __restrict__ v) {
    const int cw = w >> 1, ch = h >> 1;
    const int cx = blockIdx.x * blockDim.x + threadIdx.x, cy = blockIdx.y;
    if (cx >= cw || cy >= ch) return;
    const int mx = 255, off_q = 128 << kQ;
    int scb = 0, scr = 0;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const size_t i = (size_t)(2 * cy + dy) * w + 2 * cx + dx;
            const int ri = r[i], gi = g[i], bi = b[i];
            y[i] = (uint16_t)clamp8((kYR * ri + kYG * gi + kYB * bi + (1 << (kQ - 1))) >> kQ, mx);
            scb += kCbR * ri + kCbG * gi + kCbB * bi + off_q;
            scr += kCrR * ri + kCrG * gi + kCrB * bi + off_q;
        }
    const size_t o = (size_t)cy * cw + cx;
    u[o] = (uint16_t)clamp8((scb + (1 << (kQ + 1))) >> (kQ + 2), mx);
    v[o] = (uint16_t)clamp8((scr + (1 << (kQ + 1))) >> (kQ + 2), mx);
}

__global__ void __launch_bounds__(256) csc444_kernel(const uint8_t* __restrict__ g, const uint8_t* __restrict__ b,
                                                     const uint8_t* __restrict__ r, int w, int h,
                                                     uint16_t* __restrict__ y, uint16_t* __restrict__ u,
                                                     uint16_t* __restrict__ v) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y;
    if (x >= w || row >= h) return;
    const int mx = 255, off_q = 128 << kQ, half = 1 << (kQ - 1);
    const size_t i = (size_t)row * w + x;
    const int ri = r[i], gi = g[i], bi = b[i];
    y[i] = (uint16_t)clamp8((kYR * ri + kYG * gi + kYB * bi + half) >> kQ, mx);
    u[i] = (uint16_t)clamp8((kCbR * ri + kCbG * gi + kCbB * bi + off_q + half) >> kQ, mx);
    v[i] = (uint16_t)clamp8((kCrR * ri + kCrG * gi + kCrB * bi + off_q + half) >> kQ, mx);
}

// Per (device, stream): the deblock maps of the RDO's decisions and the deblock status word.
struct EncWork {
    uint8_t* maps = nullptr;
    size_t maps_bytes = 0;
    uint32_t* status = nullptr;
    int32_t* coef = nullptr;   // CDEF's skip flags: the RDO's level planes (luma, U, V)
    size_t coef_n = 0;
    uint8_t* skip8 = nullptr;
    size_t skip_n = 0;
};

template <typename T> int grow_buf(T*& p, size_t& have, size_t n) {
    if (have >= n) return MT_OK;
    if (p) MT_HIP_TRY(hipFree(p));
    p = nullptr;
    have = 0;
    MT_HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    have = n;
    return MT_OK;
}
std::mutex g_enc_mu;
std::map<mtlib::StreamKey, EncWork> g_enc;

int enc_work(int dev, hipStream_t s, size_t maps_bytes, EncWork** out) {
    std::lock_guard<std::mutex> lk(g_enc_mu);
    EncWork& w = g_enc[mtlib::stream_key(dev, s)];
    if (!w.status) MT_HIP_TRY(hipMalloc(&w.status, sizeof(uint32_t)));
    if (w.maps_bytes < maps_bytes) {
        if (w.maps) MT_HIP_TRY(hipFree(w.maps));
        w.maps = nullptr;
        w.maps_bytes = 0;
        MT_HIP_TRY(hipMalloc(&w.maps, maps_bytes));
        w.maps_bytes = maps_bytes;
    }
    *out = &w;
    return MT_OK;
}

}  // namespace

extern "C" void mt_encode_config_default(mt_encode_config* cfg) {
    if (!cfg) return;
    *cfg = mt_encode_config{};
    cfg->quality = 80;
    cfg->chroma = MT_ENC_YUV420;
    cfg->speed = MT_SPEED_SLOW;
    cfg->quantization_matrices = 1;
    cfg->qmatrix_level = -1;
    cfg->adaptive_quant = 1;
    cfg->variance_boost = 1;
    cfg->dark_aq = 1;
    cfg->cdef = 0;
    cfg->wiener = 0;
    cfg->screen_content = 1;
    cfg->intrabc = 1;
    cfg->updating_cdf = 1;
}

// quality_to_q (src/avif.rs:637-666): piecewise linear between the anchors, integer division
// truncating toward zero.
extern "C" int mt_quality_to_qindex(int quality) {
    static const int kAnchors[14][2] = {{1, 215}, {30, 140}, {40, 116}, {50, 91}, {60, 70}, {70, 53}, {80, 35},
                                        {85, 28}, {90, 17},  {93, 12},  {96, 7},  {98, 5},  {99, 3},  {100, 1}};
    const int q = std::min(std::max(quality, 1), 100);
    for (int k = 1; k < 14; ++k) {
        if (q <= kAnchors[k][0]) {
            const int l0 = kAnchors[k - 1][0], q0 = kAnchors[k - 1][1], l1 = kAnchors[k][0], q1 = kAnchors[k][1];
            const int t = q - l0;
            return std::min(std::max(q0 + (q1 - q0) * t / std::max(l1 - l0, 1), 1), 255);
        }
    }
    return 1;
}

extern "C" int mt_encode_rgb8(const uint8_t* g_d, const uint8_t* b_d, const uint8_t* r_d, int width, int height,
                              const mt_encode_config* cfg, mt_encode_out* out, void* stream) {
    if (!g_d || !b_d || !r_d || !cfg || !out || !out->y || !out->recon_y || !out->nodes)
        return mtlib::set_error(MT_ERR_ARG, "mt_encode_rgb8: bad pointer");
    if (width < 8 || height < 8 || width > 16383 || height > 16383 || (width & 7) || (height & 7))
        return mtlib::set_error(MT_ERR_ARG, "mt_encode_rgb8: width and height must be multiples of 8 in 8..16383");
    if (cfg->quality < 1 || cfg->quality > 100 || cfg->speed > MT_SPEED_FAST || cfg->qmatrix_level > 15 ||
        cfg->qmatrix_level < -1)
        return mtlib::set_error(MT_ERR_ARG, "mt_encode_rgb8: bad configuration");
    if (cfg->chroma != MT_ENC_YUV420 && cfg->chroma != MT_ENC_YUV444)
        return mtlib::set_error(MT_ERR_UNSUPPORTED, "mt_encode_rgb8: only 4:2:0 and 4:4:4 are supported");
    if (!out->u || !out->v || !out->recon_u || !out->recon_v)
        return mtlib::set_error(MT_ERR_ARG, "mt_encode_rgb8: chroma planes required");
    hipStream_t s = (hipStream_t)stream;
    const bool c420 = cfg->chroma == MT_ENC_YUV420;
    const int sub = c420 ? 2 : 0;
    const int cw = c420 ? width / 2 : width, chh = c420 ? height / 2 : height;
    if (c420) {
        hipLaunchKernelGGL(csc420_kernel, dim3((cw + 255) / 256, chh), dim3(256), 0, s, g_d, b_d, r_d, width, height,
                           out->y, out->u, out->v);
    } else {
        hipLaunchKernelGGL(csc444_kernel, dim3((width + 255) / 256, height), dim3(256), 0, s, g_d, b_d, r_d, width,
                           height, out->y, out->u, out->v);
    }
    MT_HIP_TRY(hipGetLastError());

    const int q = mt_quality_to_qindex(cfg->quality);
    int dc = 0, ac = 0;
    int rc = mt_chroma_q_deltas(q, sub, &dc, &ac);
    if (rc != MT_OK) return rc;
    int qm = 15, uvqm = 15;
    if (cfg->quantization_matrices) {
        if (cfg->qmatrix_level >= 0) {
            qm = uvqm = cfg->qmatrix_level;
        } else {
            qm = mt_qm_level_law(q, sub);
            uvqm = mt_qm_chroma_level_law(q, sub);
        }
    }
    mt_rdo_frame f{};
    f.src = mt_plane{out->y, width, width, height, width, height, 0};
    f.recon = out->recon_y;
    f.recon_stride = width;
    f.bd = 8;
    f.qidx = (uint8_t)q;
    f.qm_level = (uint8_t)qm;
    f.chroma = c420 ? MT_RDO_CHROMA_420 : MT_RDO_CHROMA_444;
    f.lambda = mt_rdo_lambda(q, 8, sub);
    f.wsse_scale = 1.0f;
    f.prune = nullptr;
    f.nodes = out->nodes;
    f.src_u = mt_plane{out->u, cw, cw, chh, cw, chh, 0};
    f.src_v = mt_plane{out->v, cw, cw, chh, cw, chh, 0};
    f.recon_u = out->recon_u;
    f.recon_v = out->recon_v;
    f.recon_uv_stride = cw;
    f.uv_dc_delta = (int8_t)dc;
    f.uv_ac_delta = (int8_t)ac;
    f.uv_qm_level = (uint8_t)uvqm;
    int flags = MT_RDO_ANGLE_DELTAS | MT_RDO_FILTER_INTRA | MT_RDO_MIN4 | MT_RDO_CFL;
    if (cfg->speed <= MT_SPEED_SLOW) flags |= MT_RDO_TRELLIS;
    // CDEF (on with cdef or at Slow / ExtraSlow, src/avif.rs:449-451) needs the deblocked
    // planes, its own outputs and the skip flags of the coded levels
    const bool cdef_asked = cfg->cdef || cfg->speed <= MT_SPEED_SLOW;
    const bool cdef_run = cdef_asked && out->cdef_y && out->cdef_u && out->cdef_v && out->dbk_y && out->dbk_u &&
                          out->dbk_v;
    EncWork* cw_wk = nullptr;
    const size_t nl = (size_t)width * height, nc = (size_t)cw * chh;
    if (cdef_run) {
        int dev = 0;
        MT_HIP_TRY(hipGetDevice(&dev));
        if ((rc = enc_work(dev, s, 0, &cw_wk)) != MT_OK) return rc;
        if ((rc = grow_buf(cw_wk->coef, cw_wk->coef_n, nl + 2 * nc)) != MT_OK) return rc;
        f.coef = cw_wk->coef;
        f.coef_u = cw_wk->coef + nl;
        f.coef_v = cw_wk->coef + nl + nc;
        f.coef_stride = width;
        f.coef_uv_stride = cw;
    }
    rc = mt_intra_rdo(&f, flags, &out->stats, stream);
    if (rc != MT_OK) return rc;
    out->qidx = (uint8_t)q;
    out->qm_level = (uint8_t)qm;
    out->uv_qm_level = (uint8_t)uvqm;
    out->rdo_flags = flags;
    uint32_t na = 0;
    if (cfg->adaptive_quant) na |= MT_ENC_NA_AQ;
    if (cfg->variance_boost) na |= MT_ENC_NA_VARIANCE_BOOST;
    if (cfg->dark_aq) na |= MT_ENC_NA_DARK_AQ;
    if (cdef_asked && !cdef_run) na |= MT_ENC_NA_CDEF;
    if (cfg->wiener) na |= MT_ENC_NA_WIENER;
    if (cfg->screen_content) na |= MT_ENC_NA_PALETTE;
    if (cfg->intrabc) na |= MT_ENC_NA_INTRABC;
    if (cfg->updating_cdf) na |= MT_ENC_NA_UPDATING_CDF;
    if (cfg->speed == MT_SPEED_EXTRA_SLOW) na |= MT_ENC_NA_EXTRA_SLOW;
    out->lf_level_y = out->lf_level_u = out->lf_level_v = out->lf_sharpness = 0;
    out->dbk_status = 0;
    if (out->dbk_y && out->dbk_u && out->dbk_v) {
        // frame_deblock on the reconstruction at loop_filter_levels (the frame has no IntraBC,
        // so the reference filters it: src/coder/frame_encode.rs:303)
        int ly = 0, luv = 0, sharp = 0;
        rc = mt_deblock_levels(q, sub, &ly, &luv, &sharp);
        if (rc != MT_OK) return rc;
        int dev = 0;
        MT_HIP_TRY(hipGetDevice(&dev));
        const size_t n4 = (size_t)(width / 4) * (height / 4);
        EncWork* wk = nullptr;
        rc = enc_work(dev, s, 6 * n4, &wk);
        if (rc != MT_OK) return rc;
        MT_HIP_TRY(hipMemsetAsync(wk->status, 0, sizeof(uint32_t), s));
        rc = mt_rdo_deblock_maps(out->nodes, width, height, wk->maps, stream);
        if (rc != MT_OK) return rc;
        mt_deblock_geom g{};
        g.w8 = width;
        g.h8 = height;
        g.disp_w = width;
        g.disp_h = height;
        g.blk4 = wk->maps;
        g.blk4h = wk->maps + n4;
        g.blk4v = wk->maps + 2 * n4;
        g.blk4t = wk->maps + 3 * n4;
        g.pblk4 = wk->maps + 4 * n4;
        g.pblk4h = wk->maps + 5 * n4;
        g.sub_x = g.sub_y = c420 ? 1 : 0;
        g.sharp = (uint8_t)sharp;
        g.bd = 8;
        rc = mt_deblock_frame(&g, out->recon_y, out->recon_u, out->recon_v, width, cw, out->dbk_y, out->dbk_u,
                              out->dbk_v, ly, luv, luv, wk->status, stream);
        if (rc != MT_OK) return rc;
        uint32_t st = 0;
        MT_HIP_TRY(hipMemcpyAsync(&st, wk->status, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        MT_HIP_TRY(hipStreamSynchronize(s));
        out->lf_level_y = (uint8_t)ly;
        out->lf_level_u = out->lf_level_v = (uint8_t)luv;
        out->lf_sharpness = (uint8_t)sharp;
        out->dbk_status = st;
    } else {
        na |= MT_ENC_NA_DEBLOCK;
    }
    out->cdef = mt_cdef_params{};
    if (cdef_run) {
        // the skip flags of the coded frame (mt_av1_skip8 over the decisions and levels on the
        // host), then frame_cdef on the deblocked planes
        const int nsb = ((width + 63) / 64) * ((height + 63) / 64);
        std::vector<mt_rdo_node> nodes((size_t)nsb * MT_RDO_NODES);
        std::vector<int32_t> lv(nl + 2 * nc);
        MT_HIP_TRY(hipMemcpyAsync(nodes.data(), out->nodes, nodes.size() * sizeof(mt_rdo_node), hipMemcpyDeviceToHost, s));
        MT_HIP_TRY(hipMemcpyAsync(lv.data(), cw_wk->coef, lv.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        MT_HIP_TRY(hipStreamSynchronize(s));
        mt_av1_frame af{};
        af.width = width;
        af.height = height;
        af.bd = 8;
        af.chroma = (uint8_t)f.chroma;
        af.qidx = (uint8_t)q;
        af.qm_level = (uint8_t)qm;
        af.uv_qm_level = (uint8_t)uvqm;
        af.uv_dc_delta = (int8_t)dc;
        af.uv_ac_delta = (int8_t)ac;
        af.filter_intra = 1;
        af.cicp[0] = af.cicp[1] = af.cicp[2] = 2;
        af.nodes = nodes.data();
        af.coef_y = lv.data();
        af.coef_u = lv.data() + nl;
        af.coef_v = lv.data() + nl + nc;
        af.coef_stride = width;
        af.coef_uv_stride = cw;
        const size_t nb = (size_t)(width / 8) * (height / 8);
        std::vector<uint8_t> skip(nb);
        if ((rc = mt_av1_skip8(&af, skip.data())) != MT_OK) return rc;
        if ((rc = grow_buf(cw_wk->skip8, cw_wk->skip_n, nb)) != MT_OK) return rc;
        MT_HIP_TRY(hipMemcpyAsync(cw_wk->skip8, skip.data(), nb, hipMemcpyHostToDevice, s));
        mt_cdef_frame_args ca{};
        ca.width = width;
        ca.height = height;
        ca.bd = 8;
        ca.chroma = (uint8_t)f.chroma;
        ca.slow = cfg->speed <= MT_SPEED_SLOW ? 1 : 0;
        ca.qidx = q;
        ca.wsse_scale = f.wsse_scale;
        ca.lambda = f.lambda;
        ca.rec_y = out->dbk_y;
        ca.rec_u = out->dbk_u;
        ca.rec_v = out->dbk_v;
        ca.y_stride = width;
        ca.uv_stride = cw;
        ca.src_y = out->y;
        ca.src_u = out->u;
        ca.src_v = out->v;
        ca.src_y_stride = width;
        ca.src_uv_stride = cw;
        ca.skip8 = cw_wk->skip8;
        ca.dst_y = out->cdef_y;
        ca.dst_u = out->cdef_u;
        ca.dst_v = out->cdef_v;
        if ((rc = mt_cdef_frame(&ca, &out->cdef, out->cdef_grid, nullptr, stream)) != MT_OK) return rc;
    }
    out->not_applied = na;
    return MT_OK;
}
