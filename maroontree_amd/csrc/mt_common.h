// Shared host-side helpers of the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>

#include "../../include/mtgpu.h"

namespace mtlib {
// Library-owned workspaces are kept per (device, stream): calls on different streams never
// share one, and the calls on one stream are ordered by it.
using StreamKey = std::pair<int, uintptr_t>;
// WSSE_DCT4 and the normalised WSSE_WEIGHTS (src/rd_sse.rs:1402-1420) as f32 host tables
void wsse_tables(float wd[4][4], float ww[4][4]);
inline StreamKey stream_key(int dev, hipStream_t s) { return {dev, reinterpret_cast<uintptr_t>(s)}; }
// Records the calling thread's last error message and returns `code`.
int set_error(int code, const std::string& what);
// Per-device initialisation hooks of the other translation units, called by mt_init.
int rate_init_device(int device);
int pred_init_device(int device);
// The current device if mt_init has set it up (else MT_ERR_NOT_INIT), and its CU count.
int device_ready(int* dev, int* cus);
// The planes of one launch, passed to the kernel by value.
struct PlaneSet {
    mt_plane p[MT_MAX_PLANES];
    int n;
};
// Copies a host plane array into a PlaneSet, validating it.
int planes_from_host(const mt_plane* planes, int nplanes, const char* who, PlaneSet* ps);
int dist_init_device(int device);
// mt_tx_chain with an optional per-block (dc, ac) qindex delta pair for MT_TXBLK_CHROMA blocks.
int tx_chain_launch(const mt_txblk* blocks_d, int n, const int16_t* pool_d, mt_txres* res_d,
                    int32_t* levels_d, float* targets_d, int32_t* recon_d, const int8_t* qdelta_d,
                    hipStream_t stream, const mt_trellisblk* trellis_d = nullptr);
// the trellis over levels in place (mt_rate.hip; mt_trellis without argument checks)
int trellis_launch(const mt_trellisblk* blocks_d, int n, int32_t* levels_d, const float* targets_d,
                   int32_t* status_d, hipStream_t s);
// mt_intra_pred with an optional residual output (src - pred, int16, at each out_offset; the
// source plane index in each descriptor's pad byte).
int intra_pred_launch(const mt_predblk* blocks_d, int n, const PlaneSet& ps, int32_t* pred_d,
                      int32_t* status_d, int16_t* resid_d, hipStream_t s);
// The CfL candidates list_d[0 .. *count_d) (count <= n_max) of a candidate evaluation: the
// CfL prediction from the DC prediction already in pred_d, src - prediction into resid_d,
// the alpha into alpha_d[candidate].
int cfl_eval_launch(const mt_cand* cands_d, const uint32_t* list_d, const uint32_t* count_d, int n_max,
                    const PlaneSet& ps, int32_t* pred_d, int16_t* resid_d, int32_t* alpha_d, hipStream_t s);
// mt_block_rate whose eob search uses the chain's nonzero counts.
int rate_launch_with_nnz(const mt_rateblk* blocks_d, int n, const int32_t* levels_d, float* bits_d,
                         const mt_txres* nnz_d, hipStream_t s);

// mt_distortion on a prepared plane set.
int distortion_launch(const mt_distblk* blocks_d, int n, const PlaneSet& ps, const int32_t* pred_d,
                      const int32_t* resid_d, mt_distres* out_d, hipStream_t s);
// Grow the (device, stream) workspaces of the candidate evaluation and of every stage it
// launches (prediction, distortion and rate class lists, the chain's order, the trellis's
// classes) to n candidates and pool_elems pool elements, so that launches up to that size
// allocate and free nothing (what a stream capture needs: a buffer freed while being
// captured would be left in the graph).
int eval_reserve(size_t n, uint64_t pool_elems, hipStream_t s, bool trellis);
int pred_reserve(size_t n, hipStream_t s);
int dist_reserve(size_t n, hipStream_t s);
int rate_reserve(size_t n, hipStream_t s);
int trellis_reserve(size_t n, hipStream_t s);
int chain_reserve(size_t n, hipStream_t s);
// mt_eval_candidates on a prepared plane set.
int eval_launch(const mt_cand* cands_d, int n, const PlaneSet& ps, uint64_t pool_elems,
                mt_candres* res_d, hipStream_t s, bool trellis = false);
// One frame of the RDO's reconstruction: its source, reconstruction plane (rstride samples
// per row, frame size of src) and quantizer (base_q_idx, bd_qm as in mt_txblk).
struct ReconFrame {
    mt_plane src;
    uint16_t* rec;
    int32_t rstride;
    int32_t qidx, bd_qm, pad;
    // chroma: U and V sources and reconstructions, the chroma quantizer, the subsampling
    mt_plane src_uv[2];
    uint16_t* rec_uv[2];
    int32_t uv_stride, uv_bd_qm, uv_dc, uv_ac, ss, pad2;
    // the trellis between quantization and dequantization (the RDO's MT_RDO_TRELLIS): on / off,
    // and the chroma subsampling of its lambda law (0 4:4:4, 2 4:2:0)
    int32_t trellis, tsub;
    // optional level planes (frequency (fx, fy) of the block at (x, y) at sample (x + fx, y + fy))
    int32_t* coef;
    int32_t* coef_uv[2];
    int32_t coef_stride, coef_uv_stride;
};
// The RDO's superblock reconstruction: one wave per superblock b walks the leaves
// leaves_d[first_d[b] .. + count_d[b]) in coding order, predicting from and writing to the
// reconstruction plane of frame frames_d[leaf.plane] (device array).
int sb_recon_launch(const mt_predblk* leaves_d, const uint32_t* first_d, const uint32_t* count_d,
                    int nsb, const ReconFrame* frames_d, hipStream_t s);
// qctx of a frame qindex: the coefficient CDF snapshot the RDO prices with (Cdfs::new(qctx))
__host__ __device__ inline int rdo_qctx(int qidx) { return qidx <= 20 ? 0 : qidx <= 60 ? 1 : qidx <= 120 ? 2 : 3; }
// The RDO's chroma reconstruction: one wave per superblock b walks the chroma transform
// blocks desc_d[first_d[b] .. + count_d[b]) (descriptor fields as sb_recon_uv_kernel reads
// them) of frame frames_d[desc.plane].
int sb_recon_uv_launch(const mt_predblk* desc_d, const uint32_t* first_d, const uint32_t* count_d, int nsb,
                       const ReconFrame* frames_d, hipStream_t s);
// Cost in bits of symbol s of a raw AV1 CDF of nsym symbols (nsym - 1 values), through the
// reference's cost_q_table (host).
float cdf_cost_raw(const uint16_t* raw_cdf, int nsym, int s);
int rdo_init_device(int device);

// Kernel timing slots of mt_profile_end.
enum {
    kProfBucket = 0,   // 1..13: the transform chain per size
    kProfRate = 14,
    kProfPred = 15,
    kProfDist = 16,
    kProfEval = 17,
    kProfRdo = 18,     // the RDO's candidate generation, argmin and decision kernels
    kProfRecon = 19,   // the RDO's superblock reconstruction
    kProfTrellis = 20,
    kProfSlots = 21
};
// While profiling is on (mt_profile_begin), brackets a launch with HIP events on `s`.
struct ProfScope {
    int slot;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(int slot_, hipStream_t s_);
    ~ProfScope();
};
}  // namespace mtlib

#define MT_HIP_TRY(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return mtlib::set_error(MT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
