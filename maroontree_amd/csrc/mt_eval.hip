// Batched intra candidate evaluation: one call runs the reference's candidate loop body
// (src/coder/partition_search.rs:471-520 and its siblings) for every candidate of a batch:
//   predict (predict_nd_ad / dc_pred / filter_predict) + residual_pred
//   -> forward transform + quant (+QM) -> dequant + inverse          (the transform chain)
//   -> sse_recon / weighted distortion / luma_satd                  (distortion)
//   -> real_block_bits_bounded + the caller's mode bits             (rate)
//   -> rd_cost = distortion + lambda * bits                         (src/partition_rd.rs:140)
// Each stage is the library's own batched kernel; the stages pass work through device pools
// owned by the library (predictions, residuals, levels, reconstructed residuals), so a
// candidate never leaves HBM between stages.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/mtgpu.h"
#include "mt_common.h"
#include "mt_trellis.h"

namespace {

using mtlib::PlaneSet;

__global__ void __launch_bounds__(64) zero_u32_kernel(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += 64) p[i] = 0;
}

__device__ __forceinline__ int tx_w(int s) {
    constexpr int w[19] = {4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64};
    return s < 19 ? w[s] : 0;
}
__device__ __forceinline__ int tx_h(int s) {
    constexpr int h[19] = {4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16};
    return s < 19 ? h[s] : 0;
}

// Splits each candidate into the stage descriptors; a candidate outside its planes or its
// pool slot gets status MT_ERR_ARG and descriptors the stages reject.  CfL candidates
// predict DC_PRED first and are listed for the CfL stage (cfl_list, *cfl_count).
__global__ void __launch_bounds__(256) eval_setup_kernel(const mt_cand* __restrict__ cands, int n,
                                                         PlaneSet planes, uint64_t pool_elems,
                                                         mt_predblk* __restrict__ pb,
                                                         mt_txblk* __restrict__ tb,
                                                         mt_distblk* __restrict__ db,
                                                         mt_rateblk* __restrict__ rb,
                                                         int8_t* __restrict__ qd,
                                                         mt_trellisblk* __restrict__ trd,
                                                         int32_t* __restrict__ status,
                                                         uint32_t* __restrict__ cfl_list,
                                                         uint32_t* __restrict__ cfl_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const mt_cand c = cands[i];
    const int w = tx_w(c.tx_size), h = tx_h(c.tx_size);
    const int si = (c.planes & 3) | ((c.planes_hi & 15) << 2);
    const int ri = ((c.planes >> 2) & 3) | ((c.planes_hi >> 4) << 2);
    const int bdi = (c.bd_qm >> 4) & 3;
    const int bd = bdi == 0 ? 8 : bdi == 1 ? 10 : 12;
    bool ok = w > 0 && bdi < 3 && si < planes.n && ri < planes.n &&
              (uint64_t)c.offset + (uint64_t)(w * h) <= pool_elems;
    if (ok) {
        const mt_plane& s = planes.p[si];
        const mt_plane& r = planes.p[ri];
        ok = (int)c.x + w <= s.width && (int)c.y + h <= s.height && (int)c.x + w <= r.width &&
             (int)c.y + h <= r.height;
    }
    const bool cfl = c.mode == MT_PRED_CFL;
    if (ok && cfl) {  // square 4..32, a luma plane holding the covered luma, no tx-type symbol
        const int sx = (c.pflags & MT_CAND_CFL_SS_HOR) ? 1 : 0, sy = (c.pflags & MT_CAND_CFL_SS_VER) ? 1 : 0;
        ok = w == h && w <= 32 && c.txtp_set == 0 && c.txtp_mode < planes.n && c.angle_delta >= -16 &&
             c.angle_delta <= 16;
        if (ok) {
            const mt_plane& l = planes.p[c.txtp_mode];
            ok = (((int)c.x + w) << sx) <= l.width && (((int)c.y + h) << sy) <= l.height;
        }
        if (ok) cfl_list[atomicAdd(cfl_count, 1u)] = (uint32_t)i;
    }
    status[i] = ok ? MT_OK : MT_ERR_ARG;
    mt_predblk p{};
    p.out_offset = c.offset;
    p.x = c.x;
    p.y = c.y;
    p.bw = ok ? (uint8_t)w : 0;  // bw 0: the predictor rejects it
    p.bh = (uint8_t)h;
    p.mode = cfl ? (uint8_t)MT_PRED_DC : c.mode;
    p.angle_delta = cfl ? (int8_t)0 : c.angle_delta;
    p.flags = cfl ? (uint8_t)0 : c.pflags;
    p.bd = (uint8_t)bd;
    p.plane = (uint8_t)ri;
    p.pad = (uint8_t)si;  // source plane of the residual the predictor writes alongside
    pb[i] = p;
    mt_txblk t;
    t.offset = c.offset;
    t.tx_size = ok ? c.tx_size : 0xff;  // 0xff: the chain reports it unsupported
    t.tx_type = c.tx_type;
    t.qidx = c.qidx;
    t.bd_qm = c.bd_qm;
    tb[i] = t;
    qd[2 * (size_t)i] = c.dc_delta;
    qd[2 * (size_t)i + 1] = c.ac_delta;
    mt_distblk d{};
    d.pred_offset = c.offset;
    d.resid_offset = c.offset;
    d.x = c.x;
    d.y = c.y;
    d.bw = ok ? (uint8_t)w : 0;
    d.bh = (uint8_t)h;
    d.bd = (uint8_t)bd;
    d.plane = (uint8_t)si;
    d.wsse_scale = c.wsse_scale;
    db[i] = d;
    mt_rateblk r{};
    r.offset = c.offset;
    r.tx_size = ok ? c.tx_size : 0xff;
    r.plane = c.rplane;
    r.eob_bin = c.eob_bin;
    r.skip_ctx = c.skip_ctx;
    r.dcs_ctx = c.dcs_ctx;
    r.txtp_set = c.txtp_set;
    r.txtp_mode = cfl ? (uint8_t)0 : c.txtp_mode;  // CfL: txtp_mode names the luma plane
    r.txtp = c.txtp;
    r.qctx = c.qctx;
    r.flags = c.rflags & 1;
    r.skip_one = c.skip_one;
    r.skip_zero = c.skip_zero;
    r.bound = __int_as_float(0x7f800000);  // +inf: price fully
    rb[i] = r;
    // the trellis in the rate context above (lambda0 0: the block is left as quantized)
    mt_trellisblk tr{};
    tr.offset = c.offset;
    tr.tx_size = ok ? c.tx_size : 0xff;
    tr.qidx = c.qidx;
    tr.bd_qm = c.bd_qm;
    tr.flags = MT_TRELLIS_CTX;
    tr.dc_delta = c.dc_delta;
    tr.ac_delta = c.ac_delta;
    tr.qctx = c.qctx;
    tr.eob_bin = c.eob_bin;
    tr.dcs_ctx = c.dcs_ctx;
    tr.plane = c.rplane;
    tr.lambda0 = (c.rflags & MT_CAND_TRELLIS) ? trellis_lambda0(c.qidx, bd, (c.rflags >> 2) & 3) : 0.0f;
    tr.band_tilt = 0.0f;
    trd[i] = tr;
}

__global__ void __launch_bounds__(256) eval_combine_kernel(const mt_cand* __restrict__ cands, int n,
                                                           const int32_t* __restrict__ status,
                                                           const int32_t* __restrict__ pstatus,
                                                           const mt_txres* __restrict__ txr,
                                                           const mt_distres* __restrict__ dr,
                                                           const float* __restrict__ bits,
                                                           const int32_t* __restrict__ alpha,
                                                           mt_candres* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mt_candres r{};
    int st = status[i];
    if (st == MT_OK && pstatus[i] != MT_OK) st = MT_ERR_ARG;
    if (st == MT_OK && txr[i].sse < 0) st = MT_ERR_UNSUPPORTED;
    if (st == MT_OK && (dr[i].sse < 0 || bits[i] != bits[i])) st = MT_ERR_ARG;
    r.status = st;
    if (st == MT_OK) {
        const mt_cand c = cands[i];
        r.sse = dr[i].sse;
        r.dist = dr[i].dist;
        r.satd = dr[i].satd;
        r.nnz = txr[i].nnz;
        r.bits = __fadd_rn(bits[i], c.extra_bits);
        // rd_cost (src/partition_rd.rs:140): distortion + lambda * rate, in f32
        r.cost = __fadd_rn((float)r.dist, __fmul_rn(c.lambda, r.bits));
        r.alpha = c.mode == MT_PRED_CFL ? alpha[i] : 0;
    } else {
        r.sse = -1;
        r.dist = -1;
        r.cost = __int_as_float(0x7f800000);
    }
    out[i] = r;
}

// Library-owned stage buffers of one device, grown on demand.
struct EvalWork {
    size_t cap_n = 0, cap_pool = 0;
    mt_predblk* pb = nullptr;
    mt_txblk* tb = nullptr;
    mt_distblk* db = nullptr;
    mt_rateblk* rb = nullptr;
    int8_t* qd = nullptr;
    mt_trellisblk* trd = nullptr;
    int32_t* status = nullptr;
    int32_t* pstatus = nullptr;
    uint32_t* cfl_list = nullptr;
    uint32_t* cfl_count = nullptr;
    int32_t* alpha = nullptr;
    mt_txres* txr = nullptr;
    mt_distres* dr = nullptr;
    float* bits = nullptr;
    int32_t* pred = nullptr;
    int16_t* resid = nullptr;
    int32_t* levels = nullptr;
    int32_t* recon = nullptr;
    float* targets = nullptr;  // the trellis chain only (allocated on its first use)
    size_t cap_targets = 0;
};
std::mutex g_eval_mu;
std::map<mtlib::StreamKey, EvalWork> g_work;

template <typename T> int grow(T*& p, size_t n) {
    if (p) MT_HIP_TRY(hipFree(p));
    p = nullptr;
    MT_HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    return MT_OK;
}

int ensure(EvalWork& w, size_t n, size_t pool) {
    int rc = MT_OK;
    if (w.cap_n < n) {
        const size_t c = n < 4096 ? 4096 : n;
        if ((rc = grow(w.pb, c)) || (rc = grow(w.tb, c)) || (rc = grow(w.db, c)) ||
            (rc = grow(w.rb, c)) || (rc = grow(w.qd, 2 * c)) || (rc = grow(w.trd, c)) || (rc = grow(w.status, c)) ||
            (rc = grow(w.pstatus, c)) ||
            (rc = grow(w.txr, c)) || (rc = grow(w.dr, c)) || (rc = grow(w.bits, c)) ||
            (rc = grow(w.cfl_list, c)) || (rc = grow(w.alpha, c)))
            return rc;
        if (!w.cfl_count && (rc = grow(w.cfl_count, 1))) return rc;
        w.cap_n = c;
    }
    if (w.cap_pool < pool) {
        const size_t c = pool < (1u << 20) ? (1u << 20) : pool;
        if ((rc = grow(w.pred, c)) || (rc = grow(w.resid, c)) || (rc = grow(w.levels, c)) ||
            (rc = grow(w.recon, c)))
            return rc;
        w.cap_pool = c;
    }
    return MT_OK;
}

}  // namespace

namespace mtlib {
int eval_reserve(size_t n, uint64_t pool_elems, hipStream_t s, bool trellis) {
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_eval_mu);
        EvalWork& w = g_work[mtlib::stream_key(dev, s)];
        if ((rc = ensure(w, n, (size_t)pool_elems))) return rc;
        if (trellis && w.cap_targets < w.cap_pool) {
            if ((rc = grow(w.targets, w.cap_pool))) return rc;
            w.cap_targets = w.cap_pool;
        }
    }
    if ((rc = pred_reserve(n, s)) || (rc = dist_reserve(n, s)) || (rc = rate_reserve(n, s)) ||
        (rc = chain_reserve(n, s)))
        return rc;
    return trellis ? trellis_reserve(n, s) : MT_OK;
}

int eval_launch(const mt_cand* cands_d, int n, const PlaneSet& ps, uint64_t pool_elems,
                mt_candres* res_d, hipStream_t s, bool trellis) {
    if (n == 0) return MT_OK;
    int dev = 0, cus = 0;
    int rc = device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    std::lock_guard<std::mutex> lk(g_eval_mu);
    EvalWork& w = g_work[mtlib::stream_key(dev, s)];
    rc = ensure(w, (size_t)n, (size_t)pool_elems);
    if (rc != MT_OK) return rc;
    if (trellis && w.cap_targets < w.cap_pool) {
        if ((rc = grow(w.targets, w.cap_pool))) return rc;
        w.cap_targets = w.cap_pool;
    }
    ProfScope prof(kProfEval, s);
    const int g256 = (n + 255) / 256;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, w.cfl_count, 1);
    hipLaunchKernelGGL(eval_setup_kernel, dim3(g256), dim3(256), 0, s, cands_d, n, ps, pool_elems,
                       w.pb, w.tb, w.db, w.rb, w.qd, w.trd, w.status, w.cfl_list, w.cfl_count);
    MT_HIP_TRY(hipGetLastError());
    // prediction and residual_pred in one pass, then chroma-from-luma on the listed candidates
    if ((rc = intra_pred_launch(w.pb, n, ps, w.pred, w.pstatus, w.resid, s))) return rc;
    if ((rc = cfl_eval_launch(cands_d, w.cfl_list, w.cfl_count, n, ps, w.pred, w.resid, w.alpha, s))) return rc;
    if (trellis) {  // forward + quant, the trellis, dequant + inverse
        if ((rc = tx_chain_launch(w.tb, n, w.resid, w.txr, w.levels, w.targets, w.recon, w.qd, s, w.trd)))
            return rc;
    } else if ((rc = tx_chain_launch(w.tb, n, w.resid, w.txr, w.levels, nullptr, w.recon, w.qd, s))) {
        return rc;
    }
    if ((rc = distortion_launch(w.db, n, ps, w.pred, w.recon, w.dr, s))) return rc;
    if ((rc = rate_launch_with_nnz(w.rb, n, w.levels, w.bits, w.txr, s))) return rc;
    hipLaunchKernelGGL(eval_combine_kernel, dim3(g256), dim3(256), 0, s, cands_d, n, w.status,
                       w.pstatus, w.txr, w.dr, w.bits, w.alpha, res_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
}  // namespace mtlib

extern "C" int mt_eval_candidates(const mt_cand* cands_d, int n, const mt_plane* planes,
                                  int nplanes, uint64_t pool_elems, mt_candres* res_d,
                                  void* stream) {
    if (n < 0 || (n > 0 && (!cands_d || !res_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_eval_candidates: bad pointer or count");
    if (pool_elems > (uint64_t)UINT32_MAX + 1)
        return mtlib::set_error(MT_ERR_ARG, "mt_eval_candidates: pool offsets are 32-bit");
    PlaneSet ps{};
    int rc = mtlib::planes_from_host(planes, nplanes, "mt_eval_candidates", &ps);
    if (rc != MT_OK) return rc;
    return mtlib::eval_launch(cands_d, n, ps, pool_elems, res_d, (hipStream_t)stream, false);
}

extern "C" int mt_eval_candidates_trellis(const mt_cand* cands_d, int n, const mt_plane* planes,
                                          int nplanes, uint64_t pool_elems, mt_candres* res_d,
                                          void* stream) {
    if (n < 0 || (n > 0 && (!cands_d || !res_d)))
        return mtlib::set_error(MT_ERR_ARG, "mt_eval_candidates_trellis: bad pointer or count");
    if (pool_elems > (uint64_t)UINT32_MAX + 1)
        return mtlib::set_error(MT_ERR_ARG, "mt_eval_candidates_trellis: pool offsets are 32-bit");
    PlaneSet ps{};
    int rc = mtlib::planes_from_host(planes, nplanes, "mt_eval_candidates_trellis", &ps);
    if (rc != MT_OK) return rc;
    return mtlib::eval_launch(cands_d, n, ps, pool_elems, res_d, (hipStream_t)stream, true);
}
