// Whole-frame luma intra RDO on the GPU: the superblock anti-diagonal wavefront of
// mt_intra_rdo (the procedure is documented at its declaration in include/mtgpu.h).
//
// Per wavefront step (every SB (r, c) with c + 2r = step):
//   rdo_plan_kernel    one wave per SB: node flags (level by level), per-node candidate
//                      counts and their prefix in node order;
//   rdo_gen_kernel     one wave per (SB, node): a ballot over the 66 mode ids writes the
//                      allowed candidates (compacted, in id order) with their
//                      decision-time contexts and mode syntax bits;
//   eval_launch        the batched candidate evaluation (prediction ... rd_cost);
//   rdo_argmin_kernel  one thread per candidate: atomicMin of (cost bits, mode id) per node
//                      (costs are >= 0, so their bit patterns order like the values);
//   rdo_decide_kernel  one wave per SB: bottom-up NONE / SPLIT level by level, leaves in
//                      coding order, the final mode / partition maps and leaf descriptors;
//   sb_recon_launch    one wave per SB reconstructs its leaves (mt_sbrecon.h).
// With chroma, the chroma of step k runs one step behind, inside step k + 1 (its
// candidates need step k's luma reconstruction and step k - 1's chroma, nothing of step
// k + 1), so one batched evaluation per step serves both:
//   rdo_uv_gen_kernel     one wave per SB of step k: the chroma blocks of the decided leaves
//                         in coding order and every (uv mode, plane, transform block)
//                         candidate, in a fixed slot range per SB after step k + 1's luma
//                         candidates (unused slots are marked invalid);
//   rdo_uv_decide_kernel  one wave per SB: per chroma block the cheapest uv mode (U + V +
//                         lambda * uv mode bits), the uv map and the reconstruction list;
//   sb_recon_uv_launch    one wave per (SB, plane) reconstructs the chroma blocks.
// Afterwards every leaf and chroma block is re-priced with the final contexts in one batched
// evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/mtgpu.h"
#include "av1_cdf_tables.h"
#include "mt_common.h"

namespace {

using mtlib::PlaneSet;

constexpr int kNodes = MT_RDO_NODES;
constexpr int kModeIds = 66;
constexpr int kMaxLeaves = 256;
constexpr int kMaxUvBlocks = 256;  // chroma blocks of one SB (4:4:4 with 4x4 leaves)
constexpr int kMaxUvDesc = 512;    // chroma reconstruction blocks of one SB (U and V)

// Symbol bits of the frame-initial decision CDFs (Cdfs::new, src/coder/cdfs.rs:400).
struct SymBits {
    float kf[5][5][13];   // kf_y mode [above ctx][left ctx][mode]
    float angle[8][7];    // angle delta [mode - V_PRED][delta + 3]
    float fuse[5][2];     // use_filter_intra at 4x4 .. 64x64 (av1 block sizes 0, 3, 6, 9, 12)
    float fmode[5];       // filter intra mode
    float part[4][4][2];  // [bl - 1][ctx][NONE, SPLIT]: 64, 32, 16 (10 symbols), 8 (4)
    float uv[2][13][14];  // uv_mode [CfL allowed][luma mode][uv mode] (13 or 14 symbols)
    float cfl_sign[8];    // CfL joint sign - 1
    float cfl_alpha[6][16];
};
__constant__ SymBits c_sym;

__host__ __device__ inline int level_base(int L) { return L == 0 ? 0 : L == 1 ? 1 : L == 2 ? 5 : L == 3 ? 21 : 85; }
__host__ __device__ inline int node_level(int id) {
    return id >= 85 ? 4 : id >= 21 ? 3 : id >= 5 ? 2 : id >= 1 ? 1 : 0;
}
__host__ __device__ inline int node_id(int L, int ix, int iy) { return level_base(L) + iy * (1 << L) + ix; }

// Coding order of the s-aligned cell (ix, iy) inside a superblock: bit-interleaved, y high.
__host__ __device__ inline int zidx(int ix, int iy) {
    int z = 0;
    for (int b = 0; b < 4; ++b) z |= (((ix >> b) & 1) << (2 * b)) | (((iy >> b) & 1) << (2 * b + 1));
    return z;
}
// Top-right / bottom-left neighbour of the same size coded before the s x s block at
// (xin, yin) of its SB (the frame edge clips the sample counts inside the predictor).
__host__ __device__ inline bool have_tr(int xin, int yin, int s) {
    if (yin == 0) return true;
    if (xin + s >= 64) return false;
    return zidx(xin / s + 1, yin / s - 1) < zidx(xin / s, yin / s);
}
__host__ __device__ inline bool have_bl(int xin, int yin, int s) {
    if (yin + s >= 64) return false;
    if (xin == 0) return true;
    return zidx(xin / s - 1, yin / s + 1) < zidx(xin / s, yin / s);
}

__host__ __device__ inline int id_mode(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 16 + id - 61; }
__host__ __device__ inline int id_delta(int id) {
    if (id < 13 || id > 60) return 0;
    const int j = (id - 13) % 6;
    return j < 3 ? j - 3 : j - 2;
}
__host__ __device__ inline int id_base(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 0; }
__host__ __device__ inline int id_bit(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 13; }
__host__ __device__ inline bool id_valid(int id, int s, int flags) {
    if (id <= 12) return true;
    if (id <= 60) return (flags & MT_RDO_ANGLE_DELTAS) && s >= 8;
    return (flags & MT_RDO_FILTER_INTRA) && s <= 32;
}
__host__ __device__ inline bool id_allowed(int id, uint16_t mask) {
    return id == 0 || ((mask >> id_bit(id)) & 1);  // DC_PRED is always evaluated
}

enum { kSearched = 1, kInside = 2, kNoneOk = 4, kMaySplit = 8 };

struct FrameArgs {
    mt_plane src;
    const uint16_t* prune;
    mt_rdo_node* nodes;
    uint8_t* mode4;  // (W / 4) x (H / 4) luma modes (filter intra as DC_PRED)
    uint8_t* part8;  // (W / 8) x (H / 8) partition context bytes
    int W, H, C;
    int bd, qidx, qm, flags, minsz;
    float lambda, wsse;
    int frame;     // index in the batch (its ReconFrame)
    int pbase;     // its planes in the batch's plane set: pbase source, pbase + 1 reconstruction,
                   // with chroma pbase + 2 / + 3 U source / reconstruction, pbase + 4 / + 5 V
    uint32_t sb0;  // index of the frame's first SB in the batch's SB arrays
    // chroma
    int chroma;    // MT_RDO_LUMA_ONLY, MT_RDO_CHROMA_420, MT_RDO_CHROMA_444
    int ss;        // 1 at 4:2:0 (both directions), 0 at 4:4:4
    int Wc, Hc;    // chroma plane size
    int uv_dc, uv_ac, uv_qm;
    uint8_t* uvmode4;  // (Wc / 4) x (Hc / 4) chosen uv modes (CfL as 13)
};

// One chroma block of an SB: the node it is coded at, its place and size in the chroma
// plane, its transform blocks (4 at a 4:4:4 64x64, else 1), the luma mode of its uv_mode
// context, the luma block its availability follows (SB-relative, luma units) and where its
// candidates, pool slots and reconstruction descriptors start inside the SB's ranges.
struct UvBlk {
    uint16_t node, cx, cy;
    uint8_t cs, nt, nm, ymode;
    uint8_t flags;  // 1 CfL allowed (and searched when kUvCfl), 2 angle delta coded (>= 8x8)
    uint8_t ls;
    uint8_t lx, ly;
    uint16_t desc;
    uint32_t cand, pool;
};
enum { kUvCflAllowed = 1, kUvAngle = 2 };

// One SB of a step: its frame, its raster index in the frame and where its candidates /
// pool slots start.
struct SbTask {
    uint32_t sb;
    uint32_t cand_base;
    uint32_t pool_base;
    uint32_t frame;
};

// Flags of every node of SB (r, c) into LDS, level by level (a node's search depends on its
// parent); called by all 64 lanes of the wave.
__device__ void node_flags(const FrameArgs& f, int r, int c, uint8_t* fl) {
    const uint16_t* mask = f.prune ? f.prune + (size_t)(r * f.C + c) * kNodes : nullptr;
    for (int L = 0; L < 5; ++L) {
        const int n = 1 << L, s = 64 >> L;
        for (int k = threadIdx.x; k < n * n; k += 64) {
            const int id = level_base(L) + k, ix = k % n, iy = k / n;
            const int x = c * 64 + ix * s, y = r * 64 + iy * s;
            const uint16_t m = mask ? mask[id] : (uint16_t)MT_RDO_ALL_MODES;
            const bool sr = id == 0 || ((fl[node_id(L - 1, ix >> 1, iy >> 1)] & kMaySplit) && x < f.W && y < f.H);
            const bool in = x + s <= f.W && y + s <= f.H;
            const bool none_ok = sr && in && (s == f.minsz || !(m & MT_RDO_FORCE_SPLIT) || (m & MT_RDO_NO_SPLIT));
            const bool may_split = sr && s > f.minsz && (!in || !(m & MT_RDO_NO_SPLIT));
            fl[id] = (uint8_t)((sr ? kSearched : 0) | (in ? kInside : 0) | (none_ok ? kNoneOk : 0) |
                               (may_split ? kMaySplit : 0));
        }
        __syncthreads();
    }
}

// Number of mode ids node `id` (flags fl) evaluates under mask m: the allowed intra modes
// (DC always), 6 angle deltas per allowed directional mode at >= 8x8, 5 filter-intra modes.
__device__ inline int node_count(int id, uint8_t fl, uint16_t m, int flags) {
    if (!(fl & kNoneOk)) return 0;
    const int s = 64 >> node_level(id);
    const uint32_t mm = (uint32_t)m | 1u;
    int k = __popc(mm & 0x1FFFu);
    if ((flags & MT_RDO_ANGLE_DELTAS) && s >= 8) k += 6 * __popc(mm & 0x1FEu);
    if ((flags & MT_RDO_FILTER_INTRA) && s <= 32 && (mm & (1u << 13))) k += 5;
    return k;
}

__device__ inline int map_mode(const FrameArgs& f, int x, int y) {
    return f.mode4[(size_t)(y >> 2) * (f.W >> 2) + (x >> 2)];
}
__device__ inline int map_part(const FrameArgs& f, int x, int y) {
    return f.part8[(size_t)(y >> 3) * (f.W >> 3) + (x >> 3)];
}

// Partition symbol bits of a node at (x, y), level L (bl = L + 1); decision-time contexts
// read only the neighbours outside the SB at (x0, y0).
__device__ float part_bits(const FrameArgs& f, int x, int y, int L, int sym, int x0, int y0, bool fin) {
    const int bl = L + 1, sh = 4 - bl;
    const int a = (y > 0 && (fin || y == y0)) ? map_part(f, x, y - 1) : 0;
    const int l = (x > 0 && (fin || x == x0)) ? map_part(f, x - 1, y) : 0;
    const int ctx = ((a >> sh) & 1) + (((l >> sh) & 1) << 1);
    return c_sym.part[bl - 1][ctx][sym];
}

__constant__ int kIntraModeCtx[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};
__constant__ int kFilterDir[5] = {0, 1, 2, 6, 0};  // Filter_Intra_Mode_To_Intra_Dir

// The candidates' trellis bits: MT_CAND_TRELLIS with the chroma subsampling of the trellis
// lambda law (4:4:4 frames 0; 4:2:0 and luma-only frames 2, as their decision lambda).
__device__ inline uint8_t trellis_rflags(const FrameArgs& f) {
    if (!(f.flags & MT_RDO_TRELLIS)) return 0;
    return (uint8_t)(MT_CAND_TRELLIS | ((f.chroma == MT_RDO_CHROMA_444 ? 0 : 2) << 2));
}

// The candidate of mode id `id` on the s x s block at (x, y) of the SB at (x0, y0).
__device__ mt_cand make_cand(const FrameArgs& f, int x, int y, int s, int id, int x0, int y0, bool fin,
                             uint32_t offset) {
    mt_cand c{};
    c.offset = offset;
    c.x = (uint16_t)x;
    c.y = (uint16_t)y;
    c.tx_size = (uint8_t)(s == 4 ? MT_TX_4X4 : s == 8 ? MT_TX_8X8 : s == 16 ? MT_TX_16X16
                          : s == 32 ? MT_TX_32X32 : MT_TX_64X64);
    c.tx_type = MT_TXTP_DCT_DCT;
    c.mode = (uint8_t)id_mode(id);
    c.angle_delta = (int8_t)id_delta(id);
    const bool above_ok = y > 0 && (fin || y == y0);
    const bool left_ok = x > 0 && (fin || x == x0);
    const int am = above_ok ? map_mode(f, x, y - 1) : 0, lm = left_ok ? map_mode(f, x - 1, y) : 0;
    const bool smooth = (am >= MT_PRED_SMOOTH && am <= MT_PRED_SMOOTH_H) ||
                        (lm >= MT_PRED_SMOOTH && lm <= MT_PRED_SMOOTH_H);
    c.pflags = (uint8_t)((have_tr(x - x0, y - y0, s) ? MT_PRED_HAVE_TR : 0) |
                         (have_bl(x - x0, y - y0, s) ? MT_PRED_HAVE_BL : 0) | (smooth ? MT_PRED_SMOOTH_NB : 0));
    const int sp = f.pbase, rp = f.pbase + 1;
    c.planes = (uint8_t)((sp & 3) | ((rp & 3) << 2));
    c.planes_hi = (uint8_t)((sp >> 2) | ((rp >> 2) << 4));
    c.qidx = (uint8_t)f.qidx;
    const int bdi = f.bd == 8 ? 0 : f.bd == 10 ? 1 : 2;
    c.bd_qm = (uint8_t)(f.qm | (bdi << 4));
    c.rplane = 0;
    c.eob_bin = (uint8_t)(s == 4 ? 0 : s == 8 ? 5 : s == 16 ? 11 : 15);
    c.txtp_set = (uint8_t)(s == 4 ? 1 : s == 8 ? 2 : s == 16 ? 3 : 0);
    c.txtp_mode = (uint8_t)(id >= 61 ? kFilterDir[id - 61] : id_base(id));
    c.txtp = (uint8_t)(c.txtp_set ? 1 : 0);  // DCT_DCT's index in the intra tx sets
    c.qctx = (uint8_t)mtlib::rdo_qctx(f.qidx);
    c.rflags = trellis_rflags(f);
    c.wsse_scale = f.wsse;
    c.lambda = f.lambda;
    // luma_intra_syntax_bits (src/coder/superblock.rs:192)
    const int base = id_base(id);
    const int lv = s == 64 ? 4 : s == 32 ? 3 : s == 16 ? 2 : s == 8 ? 1 : 0;
    float bits = c_sym.kf[kIntraModeCtx[am]][kIntraModeCtx[lm]][base];
    if (base >= MT_PRED_V && base <= MT_PRED_D67 && s >= 8) {
        bits = __fadd_rn(bits, c_sym.angle[base - 1][id_delta(id) + 3]);
    } else if ((f.flags & MT_RDO_FILTER_INTRA) && base == MT_PRED_DC && s <= 32) {
        const int use = id >= 61;
        bits = __fadd_rn(bits, c_sym.fuse[lv][use]);
        if (use) bits = __fadd_rn(bits, c_sym.fmode[id - 61]);
    }
    c.extra_bits = bits;
    return c;
}

// Candidate count and pool elements of every SB of the batch (blockIdx = position in the
// batch's SB list).
__global__ __launch_bounds__(64) void rdo_count_kernel(const FrameArgs* __restrict__ frames,
                                                       const SbTask* __restrict__ all,
                                                       uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ pools) {
    __shared__ uint8_t fl[kNodes];
    const int g = blockIdx.x, lane = threadIdx.x;
    const SbTask t = all[g];
    const FrameArgs& f = frames[t.frame];
    const int sb = (int)t.sb;
    const int r = sb / f.C, c = sb % f.C;
    node_flags(f, r, c, fl);
    const uint16_t* mask = f.prune ? f.prune + (size_t)sb * kNodes : nullptr;
    uint32_t cnt = 0, pool = 0;
    for (int id = lane; id < kNodes; id += 64) {
        const int s = 64 >> node_level(id);
        const uint32_t k = (uint32_t)node_count(id, fl[id], mask ? mask[id] : (uint16_t)MT_RDO_ALL_MODES, f.flags);
        cnt += k;
        pool += k * (uint32_t)(s * s);
    }
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o);
        pool += __shfl_xor(pool, o);
    }
    if (lane == 0) {
        counts[g] = cnt;
        pools[g] = pool;
    }
}

// Per-node plan of a step's SB: where its candidates and pool slots start, how many, its
// flags and mask.
struct NodePlan {
    uint32_t cand_base, pool_base;
    uint16_t mask;
    uint8_t count, flags;
};

// One wave per SB of the step: node flags, per-node counts and their prefix in node order.
__global__ __launch_bounds__(64) void rdo_plan_kernel(const FrameArgs* __restrict__ frames, const SbTask* __restrict__ tasks,
                                                      NodePlan* __restrict__ plan) {
    __shared__ uint8_t fl[kNodes];
    __shared__ uint32_t cnt[kNodes], pl[kNodes];
    const int local = blockIdx.x, lane = threadIdx.x;
    const SbTask t = tasks[local];
    const FrameArgs& f = frames[t.frame];
    const int r = t.sb / f.C, c = t.sb % f.C;
    node_flags(f, r, c, fl);
    const uint16_t* mask = f.prune ? f.prune + (size_t)t.sb * kNodes : nullptr;
    for (int id = lane; id < kNodes; id += 64) {
        const int s = 64 >> node_level(id);
        const uint16_t m = mask ? mask[id] : (uint16_t)MT_RDO_ALL_MODES;
        cnt[id] = (uint32_t)node_count(id, fl[id], m, f.flags);
        pl[id] = cnt[id] * (uint32_t)(s * s);
    }
    __syncthreads();
    {  // exclusive prefix in node order: lane l sums nodes 6l .. 6l + 5, then a wave scan
        uint32_t a = 0, b = 0;
        for (int j = 0; j < 6; ++j) {
            const int id = 6 * lane + j;
            if (id < kNodes) {
                a += cnt[id];
                b += pl[id];
            }
        }
        uint32_t ia = a, ib = b;  // inclusive scan over lanes
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t ua = __shfl_up(ia, o), ub = __shfl_up(ib, o);
            if (lane >= o) {
                ia += ua;
                ib += ub;
            }
        }
        a = t.cand_base + ia - a;
        b = t.pool_base + ib - b;
        for (int j = 0; j < 6; ++j) {
            const int id = 6 * lane + j;
            if (id < kNodes) {
                const uint32_t ca = cnt[id], pa = pl[id];
                cnt[id] = a;
                pl[id] = b;
                a += ca;
                b += pa;
            }
        }
    }
    __syncthreads();
    for (int id = lane; id < kNodes; id += 64) {
        NodePlan p;
        p.cand_base = cnt[id];
        p.pool_base = pl[id];
        p.mask = mask ? mask[id] : (uint16_t)MT_RDO_ALL_MODES;
        p.flags = fl[id];
        p.count = (uint8_t)node_count(id, fl[id], p.mask, f.flags);
        plan[(size_t)local * kNodes + id] = p;
    }
}

// Writes the candidates of the step: one wave per (SB, node); lane j takes mode ids j and
// j + 64 and the allowed ones are compacted in id order.  refs[k] = (local SB << 16) |
// (node << 7) | id.
__global__ __launch_bounds__(64) void rdo_gen_kernel(const FrameArgs* __restrict__ frames, const SbTask* __restrict__ tasks,
                                                     const NodePlan* __restrict__ plan,
                                                     mt_cand* __restrict__ cands,
                                                     uint32_t* __restrict__ refs) {
    const int local = blockIdx.x / kNodes, id = blockIdx.x % kNodes, lane = threadIdx.x;
    const NodePlan p = plan[blockIdx.x];
    if (p.count == 0) return;
    const SbTask t = tasks[local];
    const FrameArgs& f = frames[t.frame];
    const int r = t.sb / f.C, c = t.sb % f.C, x0 = c * 64, y0 = r * 64;
    const int L = node_level(id), kk = id - level_base(L), n = 1 << L, s = 64 >> L;
    const int x = x0 + (kk % n) * s, y = y0 + (kk / n) * s;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t k0 = p.cand_base, p0 = p.pool_base;
    for (int j = lane; j < 128; j += 64) {
        const bool ok = j < kModeIds && id_valid(j, s, f.flags) && id_allowed(j, p.mask);
        const uint64_t b = __ballot(ok);
        if (ok) {
            const uint32_t rank = (uint32_t)__popcll(b & below);
            cands[k0 + rank] = make_cand(f, x, y, s, j, x0, y0, false, p0 + rank * (uint32_t)(s * s));
            refs[k0 + rank] = ((uint32_t)local << 16) | ((uint32_t)id << 7) | (uint32_t)j;
        }
        const uint32_t tot = (uint32_t)__popcll(b);
        k0 += tot;
        p0 += tot * (uint32_t)(s * s);
    }
}

__global__ void __launch_bounds__(256) rdo_fill_u64_kernel(unsigned long long* p, size_t n, unsigned long long v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void __launch_bounds__(256) rdo_fill_u8_kernel(uint8_t* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
__global__ void __launch_bounds__(256) rdo_copy_plane_kernel(mt_plane src, uint16_t* __restrict__ dst, int dstride) {
    const int W = src.width;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)W * src.height;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t y = i / W, x = i % W;
        dst[y * dstride + x] = src.data[y * src.stride + x];
    }
}

__global__ void __launch_bounds__(256) rdo_argmin_kernel(const uint32_t* __restrict__ refs,
                                                         const mt_candres* __restrict__ res, int n,
                                                         unsigned long long* __restrict__ keys) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t ref = refs[k];
    const uint32_t local = ref >> 16, node = (ref >> 7) & 511, id = ref & 127;
    const float cost = res[k].cost;  // >= 0, or +inf for an invalid candidate
    const unsigned long long key = ((unsigned long long)__float_as_uint(cost) << 32) | id;
    atomicMin(&keys[(size_t)local * kNodes + node], key);
}

// Decision of one SB per wave: bottom-up NONE / SPLIT level by level (lanes over the
// level's nodes), the leaves in coding order (lane 0 walks the tree in LDS), the final maps
// (lanes over the SB's 4x4 and 8x8 units) and the leaf descriptors (lanes over leaves).
__global__ __launch_bounds__(64) void rdo_decide_kernel(const FrameArgs* __restrict__ frames, const SbTask* __restrict__ tasks,
                                                        const unsigned long long* __restrict__ keys,
                                                        const NodePlan* __restrict__ plan,
                                                        mt_predblk* __restrict__ leaves,
                                                        uint32_t* __restrict__ first,
                                                        uint32_t* __restrict__ count,
                                                        uint32_t* __restrict__ sb_leaves,
                                                        uint32_t* __restrict__ sb_leaf_pool) {
    __shared__ float cost[kNodes];
    __shared__ uint8_t split[kNodes];
    __shared__ uint8_t mode[kNodes];
    __shared__ uint16_t lv[kMaxLeaves];
    __shared__ int s_nl;
    const int local = blockIdx.x, lane = threadIdx.x;
    const SbTask t = tasks[local];
    const FrameArgs& f = frames[t.frame];
    const int r = t.sb / f.C, c = t.sb % f.C, x0 = c * 64, y0 = r * 64;
    mt_rdo_node* out = f.nodes + (size_t)t.sb * kNodes;
    const float inf = __int_as_float(0x7f800000);
    for (int L = 4; L >= 0; --L) {
        const int n = 1 << L, s = 64 >> L;
        for (int kk = lane; kk < n * n; kk += 64) {
            const int id = level_base(L) + kk, ix = kk % n, iy = kk / n, x = x0 + ix * s, y = y0 + iy * s;
            const uint8_t fl = plan[(size_t)local * kNodes + id].flags;
            const unsigned long long key = keys[(size_t)local * kNodes + id];
            const float best = key == ~0ull ? inf : __uint_as_float((uint32_t)(key >> 32));
            const int bid = key == ~0ull ? 0 : (int)(key & 127);
            mode[id] = (uint8_t)bid;
            mt_rdo_node nd{};
            nd.mode = (uint8_t)bid;
            nd.uv_mode = 255;
            if (!(fl & kSearched)) {
                cost[id] = 0.0f;
                split[id] = 2;
                nd.cost = inf;
                nd.none_cost = inf;
                nd.split = 2;
                out[id] = nd;
                continue;
            }
            float none = inf, spl = inf;
            if ((fl & kNoneOk) && best < inf) {
                none = best;
                if (s >= 8) none = __fadd_rn(none, __fmul_rn(f.lambda, part_bits(f, x, y, L, 0, x0, y0, false)));
            }
            if (fl & kMaySplit) {
                float sum = 0.0f;
                for (int q = 0; q < 4; ++q) sum = __fadd_rn(sum, cost[node_id(L + 1, 2 * ix + (q & 1), 2 * iy + (q >> 1))]);
                spl = (fl & kInside) ? __fadd_rn(sum, __fmul_rn(f.lambda, part_bits(f, x, y, L, 1, x0, y0, false)))
                                     : sum;
            }
            // a 4x4 node, or one that may not split, is a leaf whatever its costs compare
            // like (a NaN cost must not send a walk below level 4)
            const bool leaf = L == 4 || !(fl & kMaySplit) || none <= spl;
            cost[id] = leaf ? none : spl;
            split[id] = leaf ? 0 : 1;
            nd.cost = cost[id];
            nd.none_cost = none;
            nd.split = split[id];
            out[id] = nd;
        }
        __syncthreads();
    }
    if (lane == 0) {  // leaves in coding order (children pushed in reverse: they pop 0, 1, 2, 3)
        int nl = 0, sp = 0;
        uint16_t stack[32];
        stack[sp++] = 0;
        while (sp) {
            const int id = stack[--sp];
            if (split[id] == 2) continue;
            const int L = node_level(id), kk = id - level_base(L), n = 1 << L;
            if (split[id] == 1 && L < 4 && sp + 4 <= 32) {
                for (int q = 3; q >= 0; --q)
                    stack[sp++] = (uint16_t)node_id(L + 1, 2 * (kk % n) + (q & 1), 2 * (kk / n) + (q >> 1));
                continue;
            }
            if (split[id] == 1) continue;  // unreachable: level 4 never splits
            lv[nl++] = (uint16_t)id;
        }
        s_nl = nl;
    }
    // the final maps: each 4x4 (and 8x8) unit inside the frame finds its leaf from the root;
    // partition context bytes of part_ctx_bytes (src/coder/superblock.rs:1062): NONE at
    // 64 / 32 / 16 / 8, and an 8x8 split into 4x4
    const uint8_t pbyte[5] = {0x10, 0x18, 0x1c, 0x1e, 0x1f};
    for (int u = lane; u < 256 + 64; u += 64) {
        const bool u4 = u < 256;
        const int g = u4 ? 16 : 8, k = u4 ? u : u - 256, cell = u4 ? 4 : 8;
        const int ux = (k % g) * cell, uy = (k / g) * cell;
        if (x0 + ux >= f.W || y0 + uy >= f.H) continue;
        int id = 0, L = 0;
        while (split[id] == 1 && L < 4) {
            ++L;
            const int s = 64 >> L;
            id = node_id(L, ux / s, uy / s);
        }
        if (split[id] != 0) continue;
        if (u4) f.mode4[(size_t)((y0 + uy) >> 2) * (f.W >> 2) + ((x0 + ux) >> 2)] = (uint8_t)id_base(mode[id]);
        else f.part8[(size_t)((y0 + uy) >> 3) * (f.W >> 3) + ((x0 + ux) >> 3)] = pbyte[L];
    }
    __syncthreads();  // the maps and the leaf list are complete
    const int nl = s_nl;
    uint32_t pool = 0;
    for (int i = lane; i < nl; i += 64) {
        const int id = lv[i];
        const int L = node_level(id), kk = id - level_base(L), n = 1 << L, s = 64 >> L;
        const int x = x0 + (kk % n) * s, y = y0 + (kk / n) * s;
        const mt_cand cd = make_cand(f, x, y, s, mode[id], x0, y0, true, 0);
        mt_predblk p{};
        p.x = cd.x;
        p.y = cd.y;
        p.bw = (uint8_t)s;
        p.bh = (uint8_t)s;
        p.mode = cd.mode;
        p.angle_delta = cd.angle_delta;
        p.flags = cd.pflags;
        p.bd = (uint8_t)f.bd;
        p.plane = (uint8_t)f.frame;
        leaves[(size_t)local * kMaxLeaves + i] = p;
        pool += (uint32_t)(s * s);
    }
    for (int o = 32; o > 0; o >>= 1) pool += __shfl_xor(pool, o);
    if (lane == 0) {
        first[local] = (uint32_t)local * kMaxLeaves;
        count[local] = (uint32_t)nl;
        sb_leaves[f.sb0 + t.sb] = (uint32_t)nl;
        sb_leaf_pool[f.sb0 + t.sb] = pool;
    }
}

// ---- chroma ----------------------------------------------------------------------------

// The transform type the decoder derives from a uv mode (chroma_tx_for_mode,
// src/coder/chroma_tx.rs:46); DCT_DCT at 32x32, where the intra set has no other.
__host__ __device__ inline int uv_tx_type(int m, int cs) {
    if (cs >= 32) return MT_TXTP_DCT_DCT;
    switch (m) {
        case MT_PRED_PAETH: case MT_PRED_SMOOTH: case MT_PRED_D135: return MT_TXTP_ADST_ADST;
        case MT_PRED_SMOOTH_V: case MT_PRED_V: case MT_PRED_D113: case MT_PRED_D67: return MT_TXTP_ADST_DCT;
        case MT_PRED_SMOOTH_H: case MT_PRED_H: case MT_PRED_D157: case MT_PRED_D203: return MT_TXTP_DCT_ADST;
        default: return MT_TXTP_DCT_DCT;
    }
}

// uv_mode_bits (src/coder/lossy_state.rs:1302) of uv mode m (13 CfL with alphas au, av)
// under luma mode ym: the uv_mode symbol from the CfL-allowed or the no-CfL CDF, the delta-0
// angle symbol of a directional mode where angle deltas are coded, the CfL joint sign and
// the alpha magnitudes (CfL with both alphas zero is priced as DC_PRED, :1333).
__device__ float uv_mode_bits(int ym, int m, int au, int av, int flags) {
    const int allowed = (flags & kUvCflAllowed) ? 1 : 0;
    if (m == MT_PRED_CFL) {
        const int su = au == 0 ? 0 : au < 0 ? 1 : 2, sv = av == 0 ? 0 : av < 0 ? 1 : 2;
        if (su == 0 && sv == 0) return c_sym.uv[1][ym][MT_PRED_DC];  // both zero: DC_PRED
        float b = __fadd_rn(c_sym.uv[1][ym][MT_PRED_CFL], c_sym.cfl_sign[su * 3 + sv - 1]);
        if (su) b = __fadd_rn(b, c_sym.cfl_alpha[(su == 2) * 3 + sv][abs(au) - 1]);
        if (sv) b = __fadd_rn(b, c_sym.cfl_alpha[(sv == 2) * 3 + su][abs(av) - 1]);
        return b;
    }
    float b = c_sym.uv[allowed][ym][m];
    if ((flags & kUvAngle) && m >= MT_PRED_V && m <= MT_PRED_D67) b = __fadd_rn(b, c_sym.angle[m - 1][3]);
    return b;
}

__device__ inline int uv_map(const FrameArgs& f, int cx, int cy) {
    return f.uvmode4[(size_t)(cy >> 2) * (f.Wc >> 2) + (cx >> 2)];
}

// Visits the chroma blocks of the SB at (x0, y0) in coding order from its decided nodes.
template <class V>
__device__ void uv_walk(const FrameArgs& f, const mt_rdo_node* nd, int x0, int y0, bool cfl_search, V visit) {
    int sp = 0;
    uint16_t stack[32];
    stack[sp++] = 0;
    while (sp) {
        const int id = stack[--sp];
        const int sp_ = nd[id].split;
        if (sp_ == 2) continue;
        const int L = node_level(id), kk = id - level_base(L), n = 1 << L, s = 64 >> L;
        const int ix = kk % n, iy = kk / n;
        const bool group = f.ss && s == 8 && sp_ == 1;  // 4:2:0: 8x8 split into 4x4 leaves
        if (sp_ == 1 && !group) {
            if (L < 4 && sp + 4 <= 32)
                for (int q = 3; q >= 0; --q) stack[sp++] = (uint16_t)node_id(L + 1, 2 * ix + (q & 1), 2 * iy + (q >> 1));
            continue;
        }
        UvBlk b{};
        b.node = (uint16_t)id;
        b.cx = (uint16_t)((x0 + ix * s) >> f.ss);
        b.cy = (uint16_t)((y0 + iy * s) >> f.ss);
        b.cs = (uint8_t)(f.ss ? (s == 8 ? 4 : s >> 1) : s);
        b.nt = (uint8_t)(b.cs == 64 ? 4 : 1);
        b.lx = (uint8_t)(ix * s);
        b.ly = (uint8_t)(iy * s);
        b.ls = (uint8_t)s;
        b.ymode = (uint8_t)id_base(nd[group ? node_id(4, 2 * ix + 1, 2 * iy + 1) : id].mode);
        const int luma_bs = group ? 4 : s;  // the block the chroma is coded with
        b.flags = (uint8_t)((luma_bs <= 32 ? kUvCflAllowed : 0) | (luma_bs >= 8 ? kUvAngle : 0));
        // uv modes 0..12, CfL (13), then CfL's walk candidates (14 + j: alpha step -(j + 1)
        // toward zero for j < S, j - S + 1 away from zero after that)
        b.nm = (uint8_t)(13 + ((cfl_search && (b.flags & kUvCflAllowed)) ? 1 + 2 * MT_RDO_CFL_STEPS_OF(f.flags) : 0));
        visit(b);
    }
}

// The alpha step of CfL walk candidate m >= 14 under S walk steps.
__device__ inline int cfl_walk_step(int m, int S) {
    const int j = m - 14;
    return j < S ? -(j + 1) : j - S + 1;
}

// Candidate of uv mode m (13 CfL, 14.. CfL walk steps), plane p (0 U, 1 V), transform block t
// of chroma block b; with the final contexts, alpha gives CfL its alpha.
__device__ mt_cand make_uv_cand(const FrameArgs& f, const UvBlk& b, int m, int p, int t, int x0, int y0, bool fin,
                                int alpha, uint32_t offset) {
    mt_cand c{};
    const int ts = b.nt == 4 ? 32 : b.cs;
    const int tx = b.cx + (t & 1) * ts, ty = b.cy + (t >> 1) * ts;
    c.offset = offset;
    c.x = (uint16_t)tx;
    c.y = (uint16_t)ty;
    c.tx_size = (uint8_t)(ts == 4 ? MT_TX_4X4 : ts == 8 ? MT_TX_8X8 : ts == 16 ? MT_TX_16X16 : MT_TX_32X32);
    c.tx_type = (uint8_t)(m >= MT_PRED_CFL ? MT_TXTP_DCT_DCT : uv_tx_type(m, ts));
    c.mode = (uint8_t)(m > MT_PRED_CFL ? MT_PRED_CFL : m);
    if (m >= MT_PRED_CFL) {
        c.pflags = (uint8_t)(f.ss ? (MT_CAND_CFL_SS_HOR | MT_CAND_CFL_SS_VER) : 0) |
                   (uint8_t)(fin ? MT_CAND_CFL_ALPHA_GIVEN : 0) | (uint8_t)(m > MT_PRED_CFL ? MT_CAND_CFL_ALPHA_STEP : 0);
        c.angle_delta = (int8_t)(fin ? alpha : m > MT_PRED_CFL ? cfl_walk_step(m, MT_RDO_CFL_STEPS_OF(f.flags)) : 0);
        c.txtp_mode = (uint8_t)(f.pbase + 1);
    } else {
        // availability: the luma block's (4:4:4 64x64: each 32x32 transform block's)
        const int ax = b.nt == 4 ? (t & 1) * 32 : b.lx, ay = b.nt == 4 ? (t >> 1) * 32 : b.ly;
        const int as = b.nt == 4 ? 32 : b.ls;
        const int cy0 = y0 >> f.ss, cx0 = x0 >> f.ss;
        const bool above_ok = ty > 0 && (fin || ty == cy0);
        const bool left_ok = tx > 0 && (fin || tx == cx0);
        const int am = above_ok ? uv_map(f, tx, ty - 1) : 0, lm = left_ok ? uv_map(f, tx - 1, ty) : 0;
        const bool smooth = (am >= MT_PRED_SMOOTH && am <= MT_PRED_SMOOTH_H) ||
                            (lm >= MT_PRED_SMOOTH && lm <= MT_PRED_SMOOTH_H);
        c.pflags = (uint8_t)((have_tr(ax, ay, as) ? MT_PRED_HAVE_TR : 0) | (have_bl(ax, ay, as) ? MT_PRED_HAVE_BL : 0) |
                             (smooth ? MT_PRED_SMOOTH_NB : 0));
    }
    const int sp = f.pbase + 2 + 2 * p, rp = sp + 1;
    c.planes = (uint8_t)((sp & 3) | ((rp & 3) << 2));
    c.planes_hi = (uint8_t)((sp >> 2) | ((rp >> 2) << 4));
    c.qidx = (uint8_t)f.qidx;
    const int bdi = f.bd == 8 ? 0 : f.bd == 10 ? 1 : 2;
    c.bd_qm = (uint8_t)(f.uv_qm | (bdi << 4) | MT_TXBLK_CHROMA);
    c.dc_delta = (int8_t)f.uv_dc;
    c.ac_delta = (int8_t)f.uv_ac;
    c.rplane = 1;
    c.eob_bin = (uint8_t)(ts == 4 ? 1 : ts == 8 ? 6 : ts == 16 ? 12 : 16);  // the chroma eob_bin CDFs
    c.skip_ctx = 7;
    c.qctx = (uint8_t)mtlib::rdo_qctx(f.qidx);
    c.rflags = trellis_rflags(f);
    c.wsse_scale = 0.0f;
    c.lambda = f.lambda;
    return c;
}

// One wave per SB of the step: its chroma blocks and their candidates in slots
// [local * maxc, + maxc) of the candidates / [local * maxp, + maxp) of the pool; the slots it
// does not use get an invalid transform size (the evaluation skips them).
__global__ __launch_bounds__(64) void rdo_uv_gen_kernel(const FrameArgs* __restrict__ frames,
                                                        const SbTask* __restrict__ tasks, uint32_t maxc,
                                                        uint32_t maxp, uint32_t pool_base,
                                                        mt_cand* __restrict__ cands,
                                                        UvBlk* __restrict__ blks, uint32_t* __restrict__ nblks) {
    __shared__ UvBlk s_b[kMaxUvBlocks];
    __shared__ int s_nb;
    __shared__ uint32_t s_total;
    const int local = blockIdx.x, lane = threadIdx.x;
    const SbTask t = tasks[local];
    const FrameArgs& f = frames[t.frame];
    const int r = t.sb / f.C, c = t.sb % f.C, x0 = c * 64, y0 = r * 64;
    if (lane == 0) {
        int nb = 0;
        uint32_t co = 0, po = 0;
        if (f.chroma) {
            uv_walk(f, f.nodes + (size_t)t.sb * kNodes, x0, y0, (f.flags & MT_RDO_CFL) != 0,
                    [&](const UvBlk& b) { s_b[nb++] = b; });
            uint32_t de = 0;
            for (int i = 0; i < nb; ++i) {
                UvBlk& b = s_b[i];
                const int ts = b.nt == 4 ? 32 : b.cs;
                b.cand = co;
                b.pool = po;
                b.desc = (uint16_t)de;
                co += (uint32_t)b.nm * 2u * b.nt;
                po += (uint32_t)b.nm * 2u * b.nt * (uint32_t)(ts * ts);
                de += 2u * b.nt;
            }
        }
        s_nb = nb;
        s_total = co;
    }
    __syncthreads();
    const int nb = s_nb;
    mt_cand* out = cands + (size_t)local * maxc;
    const uint32_t pool0 = pool_base + (uint32_t)local * maxp;
    for (int i = 0; i < nb; ++i) {
        const UvBlk b = s_b[i];
        const int ts = b.nt == 4 ? 32 : b.cs, per = 2 * b.nt;
        for (int k = lane; k < b.nm * per; k += 64) {
            const int m = k / per, p = (k / b.nt) & 1, tt = k % b.nt;
            out[b.cand + k] = make_uv_cand(f, b, m, p, tt, x0, y0, false, 0, pool0 + b.pool + (uint32_t)k * (ts * ts));
        }
        if (lane == 0) blks[(size_t)local * kMaxUvBlocks + i] = b;
    }
    for (uint32_t k = s_total + lane; k < maxc; k += 64) out[k].tx_size = 0xff;
    if (lane == 0) nblks[local] = (uint32_t)nb;
}

// CfL of chroma block b from its candidates q (q[0] U, q[1] V at cfl_best_alpha's alphas,
// then the 2S walk candidates j of plane p at q[2 + 2j + p]): with S > 0, cfl_alpha_rd_refine's
// walk (src/cost.rs:332) per plane, U first against V's energy optimum, then V against U's
// result, an alpha scored as dist + lambda * (coefficient bits + uv_mode_bits of the pair)
// (cfl_alpha_walk_cost, :324).  Returns the summed rd_cost of the chosen candidates.
__device__ float cfl_decide(const FrameArgs& f, const UvBlk& b, const mt_candres* q, int& au, int& av) {
    const int S = MT_RDO_CFL_STEPS_OF(f.flags);
    int a[2] = {q[0].alpha, q[1].alpha};
    float body[2] = {q[0].cost, q[1].cost};
    for (int pl = 0; pl < 2 && S > 0; ++pl) {
        const mt_candres* st = q + 2 + pl;  // walk candidate j at st[2 j]
        auto walk_cost = [&](const mt_candres& r) {
            const int pu = pl ? a[0] : r.alpha, pv = pl ? r.alpha : a[1];
            return __fadd_rn((float)r.dist,
                             __fmul_rn(f.lambda, __fadd_rn(r.bits, uv_mode_bits(b.ymode, MT_PRED_CFL, pu, pv, b.flags))));
        };
        const int a0 = a[pl];
        if (a0 == 0) continue;
        const int dir = a0 > 0 ? 1 : -1;
        int best = -1;  // the walk candidate chosen; -1: the energy optimum
        float best_c = walk_cost(q[pl]);
        for (int s = 1; s <= S; ++s) {  // toward zero
            if ((a0 - dir * s) * dir < 0) break;
            const float c = walk_cost(st[2 * (s - 1)]);
            if (!(c < best_c)) break;
            best = s - 1;
            best_c = c;
        }
        if (best < 0)
            for (int s = 1; s <= S; ++s) {  // away from zero
                if (abs(a0 + dir * s) > 16) break;
                const float c = walk_cost(st[2 * (S + s - 1)]);
                if (!(c < best_c)) break;
                best = S + s - 1;
                best_c = c;
            }
        if (best >= 0) {
            a[pl] = st[2 * best].alpha;
            body[pl] = st[2 * best].cost;
        }
    }
    au = a[0];
    av = a[1];
    return __fadd_rn(__fadd_rn(0.0f, body[0]), body[1]);
}

// One wave per SB: per chroma block (lanes over blocks) the cheapest uv mode, the uv map and
// the node record; then (final contexts) the reconstruction descriptors in coding order:
// per block U then V, each over its transform blocks.  Descriptor fields: x, y, bw = bh,
// mode (13 CfL), angle_delta (CfL alpha), flags, bd, plane = the frame's index;
// out_offset = plane (0 U, 1 V) | transform type << 1.
__global__ __launch_bounds__(64) void rdo_uv_decide_kernel(const FrameArgs* __restrict__ frames,
                                                           const SbTask* __restrict__ tasks, uint32_t maxc,
                                                           const mt_candres* __restrict__ res,
                                                           const UvBlk* __restrict__ blks,
                                                           const uint32_t* __restrict__ nblks,
                                                           mt_predblk* __restrict__ desc,
                                                           uint32_t* __restrict__ first, uint32_t* __restrict__ count,
                                                           uint32_t* __restrict__ sb_uv_cands,
                                                           uint32_t* __restrict__ sb_uv_pool,
                                                           uint32_t* __restrict__ sb_uv_dec) {
    __shared__ uint8_t s_m[kMaxUvBlocks];
    __shared__ int8_t s_au[kMaxUvBlocks], s_av[kMaxUvBlocks];
    const int local = blockIdx.x, lane = threadIdx.x;
    const SbTask t = tasks[local];
    const FrameArgs& f = frames[t.frame];
    const int r = t.sb / f.C, c = t.sb % f.C, x0 = c * 64, y0 = r * 64;
    const int nb = (int)nblks[local];
    const UvBlk* bl = blks + (size_t)local * kMaxUvBlocks;
    const mt_candres* rs = res + (size_t)local * maxc;
    mt_rdo_node* nd = f.nodes + (size_t)t.sb * kNodes;
    const float inf = __int_as_float(0x7f800000);
    uint32_t dec = 0;
    for (int i = lane; i < nb; i += 64) {
        const UvBlk b = bl[i];
        const int per = 2 * b.nt;
        float best = inf;
        int bm = 0, ba = 0, bv = 0;
        for (int m = 0; m < min((int)b.nm, MT_PRED_CFL + 1); ++m) {
            const mt_candres* q = rs + b.cand + m * per;
            float sum = 0.0f;
            int au = 0, av = 0;
            if (m == MT_PRED_CFL) {  // one transform block per plane
                sum = cfl_decide(f, b, q, au, av);
                if (au == 0 && av == 0) continue;  // CfL with both alphas 0 is DC_PRED
            } else {
                for (int k = 0; k < per; ++k) sum = __fadd_rn(sum, q[k].cost);
            }
            if (!(sum < inf)) continue;
            const float cost = __fadd_rn(sum, __fmul_rn(f.lambda, uv_mode_bits(b.ymode, m, au, av, b.flags)));
            if (cost < best) {
                best = cost;
                bm = m;
                ba = au;
                bv = av;
            }
        }
        dec += (uint32_t)(b.nm * per);
        s_m[i] = (uint8_t)bm;
        s_au[i] = (int8_t)ba;
        s_av[i] = (int8_t)bv;
        nd[b.node].uv_mode = (uint8_t)bm;
        nd[b.node].cfl_alpha_u = (int8_t)ba;
        nd[b.node].cfl_alpha_v = (int8_t)bv;
        for (int yy = 0; yy < b.cs; yy += 4)
            for (int xx = 0; xx < b.cs; xx += 4)
                f.uvmode4[(size_t)((b.cy + yy) >> 2) * (f.Wc >> 2) + ((b.cx + xx) >> 2)] = (uint8_t)bm;
    }
    __syncthreads();  // the uv map of the whole SB is final
    uint32_t cands = 0, pool = 0;
    for (int i = lane; i < nb; i += 64) {
        const UvBlk b = bl[i];
        const int ts = b.nt == 4 ? 32 : b.cs;
        for (int p = 0; p < 2; ++p)
            for (int tt = 0; tt < b.nt; ++tt) {
                const mt_cand cd = make_uv_cand(f, b, s_m[i], p, tt, x0, y0, true, p ? s_av[i] : s_au[i], 0);
                mt_predblk d{};
                d.out_offset = (uint32_t)p | ((uint32_t)cd.tx_type << 1);
                d.x = cd.x;
                d.y = cd.y;
                d.bw = (uint8_t)ts;
                d.bh = (uint8_t)ts;
                d.mode = cd.mode;
                d.angle_delta = cd.angle_delta;
                d.flags = cd.mode == MT_PRED_CFL ? (uint8_t)0 : cd.pflags;
                d.bd = (uint8_t)f.bd;
                d.plane = (uint8_t)f.frame;
                desc[(size_t)local * kMaxUvDesc + b.desc + p * b.nt + tt] = d;
            }
        cands += 2u * b.nt;
        pool += 2u * b.nt * (uint32_t)(ts * ts);
    }
    for (int o = 32; o > 0; o >>= 1) {
        cands += __shfl_xor(cands, o);
        pool += __shfl_xor(pool, o);
        dec += __shfl_xor(dec, o);
    }
    if (lane == 0) {
        uint32_t nd_ = 0;
        if (nb) nd_ = (uint32_t)bl[nb - 1].desc + 2u * bl[nb - 1].nt;
        first[local] = (uint32_t)local * kMaxUvDesc;
        count[local] = nd_;
        sb_uv_cands[f.sb0 + t.sb] = cands;
        sb_uv_pool[f.sb0 + t.sb] = pool;
        sb_uv_dec[f.sb0 + t.sb] = dec;
    }
}

// The SB's leaves with the final contexts, at leaf_base[g] / pool_base[g] (one thread per SB
// of the batch's SB list).
__global__ void __launch_bounds__(64) rdo_final_gen_kernel(const FrameArgs* __restrict__ frames,
                                                          const SbTask* __restrict__ all, int nall,
                                                          const uint32_t* __restrict__ leaf_base,
                                                          const uint32_t* __restrict__ pool_base,
                                                          mt_cand* __restrict__ cands) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nall) return;
    const SbTask t = all[g];
    const FrameArgs& f = frames[t.frame];
    const int sb = (int)t.sb;
    const int r = sb / f.C, c = sb % f.C, x0 = c * 64, y0 = r * 64;
    const mt_rdo_node* nd = f.nodes + (size_t)sb * kNodes;
    uint32_t k = leaf_base[g], p = pool_base[g];
    uint16_t stack[32];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const int id = stack[--sp];
        const int sp_ = nd[id].split;
        if (sp_ == 2) continue;
        const int L = node_level(id), kk = id - level_base(L), n = 1 << L, s = 64 >> L;
        if (sp_ == 1) {  // decided nodes split only above level 4
            if (L < 4 && sp + 4 <= 32)
                for (int q = 3; q >= 0; --q) stack[sp++] = (uint16_t)node_id(L + 1, 2 * (kk % n) + (q & 1), 2 * (kk / n) + (q >> 1));
            continue;
        }
        const int x = x0 + (kk % n) * s, y = y0 + (kk / n) * s;
        cands[k++] = make_cand(f, x, y, s, nd[id].mode, x0, y0, true, p);
        p += (uint32_t)(s * s);
    }
    if (!f.chroma) return;
    // the chroma blocks: U then V over the transform blocks, the uv mode bits on the first
    uv_walk(f, nd, x0, y0, (f.flags & MT_RDO_CFL) != 0, [&](const UvBlk& b) {
        const int ts = b.nt == 4 ? 32 : b.cs;
        const mt_rdo_node& o = nd[b.node];
        for (int pl = 0; pl < 2; ++pl)
            for (int tt = 0; tt < b.nt; ++tt) {
                mt_cand cd = make_uv_cand(f, b, o.uv_mode, pl, tt, x0, y0, true, pl ? o.cfl_alpha_v : o.cfl_alpha_u, p);
                if (pl == 0 && tt == 0) cd.extra_bits = uv_mode_bits(b.ymode, o.uv_mode, o.cfl_alpha_u, o.cfl_alpha_v, b.flags);
                cands[k++] = cd;
                p += (uint32_t)(ts * ts);
            }
    });
}

// Per SB, in coding order: partition symbol bits of every coded node inside the frame and
// the re-priced bits and SSE of every leaf (f64 sum of the f32 values; i64 SSE).
__global__ void __launch_bounds__(64) rdo_final_sum_kernel(const FrameArgs* __restrict__ frames,
                                                          const SbTask* __restrict__ all, int nall,
                                                          const uint32_t* __restrict__ leaf_base,
                                                          const mt_candres* __restrict__ res,
                                                          double* __restrict__ sb_bits,
                                                          long long* __restrict__ sb_sse,
                                                          double* __restrict__ sb_bits_uv,
                                                          long long* __restrict__ sb_sse_uv) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nall) return;
    const SbTask t = all[g];
    const FrameArgs& f = frames[t.frame];
    const int sb = (int)t.sb;
    const int r = sb / f.C, c = sb % f.C, x0 = c * 64, y0 = r * 64;
    const mt_rdo_node* nd = f.nodes + (size_t)sb * kNodes;
    uint32_t k = leaf_base[g];
    double b = 0.0;
    long long e = 0;
    uint16_t stack[32];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const int id = stack[--sp];
        const int sp_ = nd[id].split;
        if (sp_ == 2) continue;
        const int L = node_level(id), kk = id - level_base(L), n = 1 << L, s = 64 >> L;
        const int x = x0 + (kk % n) * s, y = y0 + (kk / n) * s;
        if (s >= 8 && x + s <= f.W && y + s <= f.H)
            b = __dadd_rn(b, (double)part_bits(f, x, y, L, sp_, x0, y0, true));
        if (sp_ == 1) {  // decided nodes split only above level 4
            if (L < 4 && sp + 4 <= 32)
                for (int q = 3; q >= 0; --q) stack[sp++] = (uint16_t)node_id(L + 1, 2 * (kk % n) + (q & 1), 2 * (kk / n) + (q >> 1));
            continue;
        }
        b = __dadd_rn(b, (double)res[k].bits);
        e += res[k].sse;
        ++k;
    }
    double buv = 0.0;
    long long eu = 0, ev = 0;
    if (f.chroma) {
        uv_walk(f, nd, x0, y0, (f.flags & MT_RDO_CFL) != 0, [&](const UvBlk& blk) {
            for (int pl = 0; pl < 2; ++pl)
                for (int tt = 0; tt < blk.nt; ++tt) {
                    b = __dadd_rn(b, (double)res[k].bits);
                    buv = __dadd_rn(buv, (double)res[k].bits);
                    (pl ? ev : eu) += res[k].sse;
                    ++k;
                }
        });
    }
    sb_bits[g] = b;
    sb_sse[g] = e;
    sb_bits_uv[g] = buv;
    sb_sse_uv[2 * g] = eu;
    sb_sse_uv[2 * g + 1] = ev;
}

// Library-owned buffers of one (device, stream), grown on demand; mu serialises the calls
// that share them.
struct RdoWork {
    std::mutex mu;
    hipStream_t capture = nullptr;  // MT_RDO_GRAPH: the stream the steps are captured from
    size_t cap_sb = 0, cap_step_sb = 0, cap_cand = 0, cap_final = 0, cap_map = 0, cap_frames = 0;
    uint32_t *sb_uv_cands = nullptr, *sb_uv_pool = nullptr, *sb_uv_dec = nullptr;
    double* sb_bits_uv = nullptr;
    long long* sb_sse_uv = nullptr;
    UvBlk* uvblk = nullptr;
    uint32_t* nuvblk = nullptr;
    mt_predblk* uvdesc = nullptr;
    uint32_t *uvfirst = nullptr, *uvcount = nullptr;
    uint32_t *counts = nullptr, *pools = nullptr, *sb_leaves = nullptr, *sb_leaf_pool = nullptr;
    uint32_t *leaf_base = nullptr, *pool_base = nullptr;
    double* sb_bits = nullptr;
    long long* sb_sse = nullptr;
    SbTask* all = nullptr;    // the batch's SBs, frame by frame, raster order
    SbTask* tasks = nullptr;  // the same SBs in step order
    unsigned long long* keys = nullptr;
    NodePlan* plan = nullptr;
    mt_predblk* leaves = nullptr;
    uint32_t *first = nullptr, *count = nullptr;
    mt_cand* cands = nullptr;
    uint32_t* refs = nullptr;
    mt_candres* cres = nullptr;
    mt_cand* fcands = nullptr;
    mt_candres* fres = nullptr;
    uint8_t* maps = nullptr;  // every frame's mode and partition maps
    FrameArgs* fargs = nullptr;
    mtlib::ReconFrame* rframes = nullptr;
};
std::mutex g_rdo_mu;  // guards the map
std::map<mtlib::StreamKey, RdoWork> g_rdo;

template <typename T> int regrow(T*& p, size_t n) {
    if (p) MT_HIP_TRY(hipFree(p));
    p = nullptr;
    MT_HIP_TRY(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    return MT_OK;
}

// Frames of one batch: at most 16, and the eval's plane set holds their planes (2 per frame,
// 6 with chroma).
constexpr int kBatchFrames = 16;
int frame_planes(const mt_rdo_frame& fr) { return fr.chroma ? 6 : 2; }

int check_frame(const mt_rdo_frame& fr) {
    const int W = fr.src.width, H = fr.src.height;
    if (!fr.src.data || !fr.recon || !fr.nodes || W <= 0 || H <= 0 || (W & 7) || (H & 7) || W > 65535 ||
        H > 65535 || fr.src.stride < W || fr.recon_stride < W)
        return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: bad planes (sizes must be multiples of 8)");
    if ((fr.bd != 8 && fr.bd != 10 && fr.bd != 12) || fr.qm_level > 15 || !(fr.lambda > 0.0f) ||
        !std::isfinite(fr.lambda) || !(fr.wsse_scale >= 0.0f) || !std::isfinite(fr.wsse_scale))
        return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: bad parameters");
    if (fr.chroma > MT_RDO_CHROMA_444) return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: bad chroma format");
    if (fr.chroma) {
        const int ss = fr.chroma == MT_RDO_CHROMA_420 ? 1 : 0, Wc = W >> ss, Hc = H >> ss;
        for (const mt_plane* c : {&fr.src_u, &fr.src_v})
            if (!c->data || c->width != Wc || c->height != Hc || c->stride < Wc)
                return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: chroma sources must be the chroma size");
        if (!fr.recon_u || !fr.recon_v || fr.recon_uv_stride < Wc || fr.uv_qm_level > 15 || fr.uv_dc_delta < -64 ||
            fr.uv_dc_delta > 63 || fr.uv_ac_delta < -64 || fr.uv_ac_delta > 63)
            return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: bad chroma parameters");
        if ((fr.coef_u || fr.coef_v) && (!fr.coef_u || !fr.coef_v || fr.coef_uv_stride < Wc))
            return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: chroma level planes need both planes and a stride");
    }
    if (fr.coef && fr.coef_stride < W)
        return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: level plane stride below the width");
    return MT_OK;
}

// The wavefront of up to kBatchFrames frames at once: step k covers step k of every frame.
int run_batch(RdoWork& w, const mt_rdo_frame* frames, int nf, int flags, mt_rdo_stats* stats, int cus,
              hipStream_t s) {
    int rc = MT_OK;
    std::vector<FrameArgs> fa(nf);
    std::vector<mtlib::ReconFrame> rf(nf);
    std::vector<int> C(nf), Rr(nf), nsteps(nf);
    size_t map_bytes = 0;
    uint32_t nall = 0;
    int max_steps = 0, nplanes = 0;
    bool any_chroma = false;
    uint32_t maxc = 0, maxp = 0;  // chroma candidate / pool slots per SB of a step
    for (int i = 0; i < nf; ++i) {
        const mt_rdo_frame& fr = frames[i];
        const int W = fr.src.width, H = fr.src.height;
        C[i] = (W + 63) / 64;
        Rr[i] = (H + 63) / 64;
        nsteps[i] = (C[i] - 1) + 2 * (Rr[i] - 1) + 1;
        max_steps = std::max(max_steps, nsteps[i]);
        FrameArgs& f = fa[i];
        f = FrameArgs{};
        f.src = fr.src;
        f.src.vis_w = W;
        f.src.vis_h = H;
        f.prune = fr.prune;
        f.nodes = (mt_rdo_node*)fr.nodes;
        f.W = W;
        f.H = H;
        f.C = C[i];
        f.bd = fr.bd;
        f.qidx = fr.qidx;
        f.qm = fr.qm_level;
        f.flags = flags;
        f.minsz = (flags & MT_RDO_MIN4) ? 4 : 8;
        f.lambda = fr.lambda;
        f.wsse = fr.wsse_scale;
        f.frame = i;
        f.pbase = nplanes;
        nplanes += frame_planes(fr);
        f.sb0 = nall;
        nall += (uint32_t)(C[i] * Rr[i]);
        map_bytes += (size_t)(W / 4) * (H / 4) + (size_t)(W / 8) * (H / 8);
        const int bdi = fr.bd == 8 ? 0 : fr.bd == 10 ? 1 : 2;
        rf[i] = mtlib::ReconFrame{};
        rf[i].src = f.src;
        rf[i].rec = fr.recon;
        rf[i].rstride = fr.recon_stride;
        rf[i].qidx = fr.qidx;
        rf[i].bd_qm = fr.qm_level | (bdi << 4);
        rf[i].trellis = (flags & MT_RDO_TRELLIS) ? 1 : 0;
        rf[i].coef = fr.coef;
        rf[i].coef_stride = fr.coef_stride;
        rf[i].coef_uv[0] = fr.coef_u;
        rf[i].coef_uv[1] = fr.coef_v;
        rf[i].coef_uv_stride = fr.coef_uv_stride;
        rf[i].tsub = fr.chroma == MT_RDO_CHROMA_444 ? 0 : 2;
        f.chroma = fr.chroma;
        if (fr.chroma) {
            any_chroma = true;
            f.ss = fr.chroma == MT_RDO_CHROMA_420 ? 1 : 0;
            f.Wc = W >> f.ss;
            f.Hc = H >> f.ss;
            f.uv_dc = fr.uv_dc_delta;
            f.uv_ac = fr.uv_ac_delta;
            f.uv_qm = fr.uv_qm_level;
            map_bytes += (size_t)(f.Wc / 4) * (f.Hc / 4);
            rf[i].src_uv[0] = fr.src_u;
            rf[i].src_uv[1] = fr.src_v;
            for (int k = 0; k < 2; ++k) {
                rf[i].src_uv[k].vis_w = f.Wc;
                rf[i].src_uv[k].vis_h = f.Hc;
            }
            rf[i].rec_uv[0] = fr.recon_u;
            rf[i].rec_uv[1] = fr.recon_v;
            rf[i].uv_stride = fr.recon_uv_stride;
            rf[i].uv_bd_qm = fr.uv_qm_level | (bdi << 4) | MT_TXBLK_CHROMA;
            rf[i].uv_dc = fr.uv_dc_delta;
            rf[i].uv_ac = fr.uv_ac_delta;
            rf[i].ss = f.ss;
            // 4:2:0: <= 64 chroma blocks of <= 14 + 2S modes x 2 planes; 4:4:4: <= 256 (4x4
            // leaves)
            const uint32_t per = 2u * (14u + 2u * (uint32_t)MT_RDO_CFL_STEPS_OF(flags));
            const uint32_t c1 = (f.ss || f.minsz == 8) ? 64u * per : 256u * per;
            const uint32_t p1 = f.ss ? per * 1024u : per * 4096u;
            maxc = std::max(maxc, c1);
            maxp = std::max(maxp, p1);
        }
    }
    if (w.cap_sb < nall) {
        if ((rc = regrow(w.counts, nall)) || (rc = regrow(w.pools, nall)) || (rc = regrow(w.sb_leaves, nall)) ||
            (rc = regrow(w.sb_leaf_pool, nall)) || (rc = regrow(w.leaf_base, nall)) ||
            (rc = regrow(w.pool_base, nall)) || (rc = regrow(w.sb_bits, nall)) || (rc = regrow(w.sb_sse, nall)) ||
            (rc = regrow(w.all, nall)) || (rc = regrow(w.tasks, nall)) || (rc = regrow(w.sb_uv_cands, nall)) ||
            (rc = regrow(w.sb_uv_pool, nall)) || (rc = regrow(w.sb_uv_dec, nall)) ||
            (rc = regrow(w.sb_bits_uv, nall)) || (rc = regrow(w.sb_sse_uv, 2 * (size_t)nall)))
            return rc;
        w.cap_sb = nall;
    }
    if (w.cap_map < map_bytes) {
        if ((rc = regrow(w.maps, map_bytes))) return rc;
        w.cap_map = map_bytes;
    }
    if (w.cap_frames < (size_t)nf) {
        if ((rc = regrow(w.fargs, kBatchFrames)) || (rc = regrow(w.rframes, kBatchFrames))) return rc;
        w.cap_frames = kBatchFrames;
    }
    {
        size_t off = 0;
        for (int i = 0; i < nf; ++i) {
            fa[i].mode4 = w.maps + off;
            off += (size_t)(fa[i].W / 4) * (fa[i].H / 4);
            fa[i].part8 = w.maps + off;
            off += (size_t)(fa[i].W / 8) * (fa[i].H / 8);
            if (fa[i].chroma) {
                fa[i].uvmode4 = w.maps + off;
                off += (size_t)(fa[i].Wc / 4) * (fa[i].Hc / 4);
            }
        }
    }
    std::vector<SbTask> all(nall);
    for (int i = 0; i < nf; ++i)
        for (int sb = 0; sb < C[i] * Rr[i]; ++sb) all[fa[i].sb0 + sb] = SbTask{(uint32_t)sb, 0, 0, (uint32_t)i};
    MT_HIP_TRY(hipMemcpyAsync(w.fargs, fa.data(), nf * sizeof(FrameArgs), hipMemcpyHostToDevice, s));
    MT_HIP_TRY(hipMemcpyAsync(w.rframes, rf.data(), nf * sizeof(mtlib::ReconFrame), hipMemcpyHostToDevice, s));
    MT_HIP_TRY(hipMemcpyAsync(w.all, all.data(), nall * sizeof(SbTask), hipMemcpyHostToDevice, s));
    PlaneSet ps{};
    ps.n = nplanes;
    for (int i = 0; i < nf; ++i) {
        const int b = fa[i].pbase;
        ps.p[b] = fa[i].src;
        ps.p[b + 1] = mt_plane{frames[i].recon, frames[i].recon_stride, fa[i].W, fa[i].H, fa[i].W, fa[i].H, 0};
        if (fa[i].chroma) {
            const int Wc = fa[i].Wc, Hc = fa[i].Hc, us = frames[i].recon_uv_stride;
            ps.p[b + 2] = rf[i].src_uv[0];
            ps.p[b + 3] = mt_plane{frames[i].recon_u, us, Wc, Hc, Wc, Hc, 0};
            ps.p[b + 4] = rf[i].src_uv[1];
            ps.p[b + 5] = mt_plane{frames[i].recon_v, us, Wc, Hc, Wc, Hc, 0};
        }
    }

    // each reconstruction starts as its source (the decision's stand-in inside each SB)
    {
        mtlib::ProfScope prof(mtlib::kProfRdo, s);
        for (int i = 0; i < nf; ++i) {
            const size_t npx = (size_t)fa[i].W * fa[i].H;
            hipLaunchKernelGGL(rdo_copy_plane_kernel, dim3(std::max(1, std::min(cus * 4, (int)((npx + 255) / 256)))),
                               dim3(256), 0, s, fa[i].src, frames[i].recon, frames[i].recon_stride);
            if (fa[i].chroma) {
                const size_t cpx = (size_t)fa[i].Wc * fa[i].Hc;
                const dim3 g(std::max(1, std::min(cus * 4, (int)((cpx + 255) / 256))));
                hipLaunchKernelGGL(rdo_copy_plane_kernel, g, dim3(256), 0, s, rf[i].src_uv[0], frames[i].recon_u,
                                   frames[i].recon_uv_stride);
                hipLaunchKernelGGL(rdo_copy_plane_kernel, g, dim3(256), 0, s, rf[i].src_uv[1], frames[i].recon_v,
                                   frames[i].recon_uv_stride);
            }
        }
        // no chroma stage runs for luma-only SBs: their chroma counts stay 0
        hipLaunchKernelGGL(rdo_fill_u8_kernel, dim3(std::max(1, std::min(cus * 4, (int)((nall * 4 + 255) / 256)))),
                           dim3(256), 0, s, reinterpret_cast<uint8_t*>(w.sb_uv_cands), (size_t)nall * 4);
        hipLaunchKernelGGL(rdo_fill_u8_kernel, dim3(std::max(1, std::min(cus * 4, (int)((nall * 4 + 255) / 256)))),
                           dim3(256), 0, s, reinterpret_cast<uint8_t*>(w.sb_uv_pool), (size_t)nall * 4);
        hipLaunchKernelGGL(rdo_fill_u8_kernel, dim3(std::max(1, std::min(cus * 4, (int)((nall * 4 + 255) / 256)))),
                           dim3(256), 0, s, reinterpret_cast<uint8_t*>(w.sb_uv_dec), (size_t)nall * 4);
        hipLaunchKernelGGL(rdo_fill_u8_kernel, dim3(std::max(1, std::min(cus * 4, (int)((map_bytes + 255) / 256)))),
                           dim3(256), 0, s, w.maps, map_bytes);
        hipLaunchKernelGGL(rdo_count_kernel, dim3(nall), dim3(64), 0, s, w.fargs, w.all, w.counts, w.pools);
        MT_HIP_TRY(hipGetLastError());
    }
    std::vector<uint32_t> counts(nall), pools(nall);
    MT_HIP_TRY(hipMemcpyAsync(counts.data(), w.counts, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(pools.data(), w.pools, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipStreamSynchronize(s));

    // steps: the SBs of step k of every frame, their candidate / pool bases
    std::vector<SbTask> tasks;
    tasks.reserve(nall);
    std::vector<int> step_first(max_steps + 1);
    std::vector<uint32_t> step_n(max_steps), step_pool(max_steps);
    std::vector<int64_t> frame_cands(nf, 0);
    size_t max_cand = 0, max_sb = 0;
    for (int k = 0; k < max_steps; ++k) {
        step_first[k] = (int)tasks.size();
        uint64_t nc = 0, np = 0;
        for (int i = 0; i < nf; ++i) {
            for (int r = 0; r < Rr[i]; ++r) {
                const int c = k - 2 * r;
                if (c < 0 || c >= C[i]) continue;
                const uint32_t sb = (uint32_t)(r * C[i] + c), g = fa[i].sb0 + sb;
                tasks.push_back(SbTask{sb, (uint32_t)nc, (uint32_t)np, (uint32_t)i});
                nc += counts[g];
                np += pools[g];
                frame_cands[i] += counts[g];
            }
        }
        if (np > UINT32_MAX) return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: batch step too large");
        step_n[k] = (uint32_t)nc;
        step_pool[k] = (uint32_t)np;
        max_cand = std::max(max_cand, (size_t)nc);
        max_sb = std::max(max_sb, tasks.size() - step_first[k]);
    }
    step_first[max_steps] = (int)tasks.size();
    if (max_sb >= 65536) return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: batch step too large");
    // each step's evaluation: its luma candidates, then the chroma slots of the step before
    auto step_nk = [&](int k) { return (k < 0 || k >= max_steps) ? 0 : step_first[k + 1] - step_first[k]; };
    std::vector<uint64_t> eval_n(max_steps + 1), eval_pool(max_steps + 1);
    for (int k = 0; k <= max_steps; ++k) {
        const uint64_t pk = any_chroma ? (uint64_t)step_nk(k - 1) : 0;
        eval_n[k] = (k < max_steps ? step_n[k] : 0) + pk * maxc;
        eval_pool[k] = (k < max_steps ? step_pool[k] : 0) + pk * maxp;
        if (eval_pool[k] > UINT32_MAX || eval_n[k] > (uint64_t)INT32_MAX)
            return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: batch step too large");
        max_cand = std::max(max_cand, (size_t)eval_n[k]);
    }
    if (w.cap_step_sb < max_sb) {
        if ((rc = regrow(w.keys, max_sb * kNodes)) || (rc = regrow(w.plan, max_sb * kNodes)) ||
            (rc = regrow(w.leaves, max_sb * kMaxLeaves)) || (rc = regrow(w.first, max_sb)) ||
            (rc = regrow(w.count, max_sb)) || (rc = regrow(w.uvblk, max_sb * kMaxUvBlocks)) ||
            (rc = regrow(w.nuvblk, max_sb)) || (rc = regrow(w.uvdesc, max_sb * kMaxUvDesc)) ||
            (rc = regrow(w.uvfirst, max_sb)) || (rc = regrow(w.uvcount, max_sb)))
            return rc;
        w.cap_step_sb = max_sb;
    }
    if (w.cap_cand < max_cand) {
        if ((rc = regrow(w.cands, max_cand)) || (rc = regrow(w.refs, max_cand)) || (rc = regrow(w.cres, max_cand)))
            return rc;
        w.cap_cand = max_cand;
    }
    MT_HIP_TRY(hipMemcpyAsync(w.tasks, tasks.data(), tasks.size() * sizeof(SbTask), hipMemcpyHostToDevice, s));

    // The steps' launches: every size is known now, so with MT_RDO_GRAPH they are captured
    // into one graph (relaxed mode: a workspace that grows allocates outside the graph) and
    // launched at once.  Fills are kernels, never memset nodes.
    // The capture runs on a stream of the workspace (the caller's may be the null stream,
    // which cannot be captured); it only records, and the graph runs on the caller's stream.
    const bool as_graph = (flags & MT_RDO_GRAPH) != 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const hipStream_t caller = s;
    if (as_graph) {
        if (!w.capture) MT_HIP_TRY(hipStreamCreateWithFlags(&w.capture, hipStreamNonBlocking));
        // every workspace the steps use, at its largest step, before the capture: a buffer
        // that grew during it would free memory the graph's earlier nodes still point at
        uint64_t nmax = 0, pmax = 0;
        for (int k = 0; k <= max_steps; ++k) {
            nmax = std::max(nmax, eval_n[k]);
            pmax = std::max(pmax, eval_pool[k]);
        }
        if ((rc = mtlib::eval_reserve((size_t)nmax, pmax, w.capture, (flags & MT_RDO_TRELLIS) != 0))) return rc;
        MT_HIP_TRY(hipStreamBeginCapture(w.capture, hipStreamCaptureModeRelaxed));
    }
    rc = [&]() -> int {
    const hipStream_t s = as_graph ? w.capture : caller;
    int rc = MT_OK;
    for (int k = 0; k <= max_steps; ++k) {
        const int nk = step_nk(k), pk = any_chroma ? step_nk(k - 1) : 0;
        if (nk == 0 && pk == 0) continue;
        const SbTask* tk = w.tasks + step_first[std::min(k, max_steps)];
        const SbTask* tp = pk ? w.tasks + step_first[k - 1] : nullptr;
        const int n = k < max_steps ? (int)step_n[k] : 0;
        {
            mtlib::ProfScope prof(mtlib::kProfRdo, s);
            if (nk) {
                hipLaunchKernelGGL(rdo_fill_u64_kernel, dim3(std::max(1, (nk * kNodes + 255) / 256)), dim3(256), 0, s,
                                   w.keys, (size_t)nk * kNodes, ~0ull);
                hipLaunchKernelGGL(rdo_plan_kernel, dim3(nk), dim3(64), 0, s, w.fargs, tk, w.plan);
                hipLaunchKernelGGL(rdo_gen_kernel, dim3(nk * kNodes), dim3(64), 0, s, w.fargs, tk, w.plan, w.cands,
                                   w.refs);
            }
            if (pk)  // the chroma candidates of step k - 1 after this step's luma candidates
                hipLaunchKernelGGL(rdo_uv_gen_kernel, dim3(pk), dim3(64), 0, s, w.fargs, tp, maxc, maxp,
                                   (uint32_t)(k < max_steps ? step_pool[k] : 0), w.cands + n, w.uvblk, w.nuvblk);
            MT_HIP_TRY(hipGetLastError());
        }
        if ((rc = mtlib::eval_launch(w.cands, (int)eval_n[k], ps, eval_pool[k], w.cres, s,
                                     (flags & MT_RDO_TRELLIS) != 0)))
            return rc;
        if (nk) {
            {
                mtlib::ProfScope prof(mtlib::kProfRdo, s);
                if (n > 0)
                    hipLaunchKernelGGL(rdo_argmin_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w.refs, w.cres, n,
                                       w.keys);
                hipLaunchKernelGGL(rdo_decide_kernel, dim3(nk), dim3(64), 0, s, w.fargs, tk, w.keys, w.plan, w.leaves,
                                   w.first, w.count, w.sb_leaves, w.sb_leaf_pool);
                MT_HIP_TRY(hipGetLastError());
            }
            mtlib::ProfScope prof(mtlib::kProfRecon, s);
            if ((rc = mtlib::sb_recon_launch(w.leaves, w.first, w.count, nk, w.rframes, s))) return rc;
        }
        if (pk) {
            {
                mtlib::ProfScope prof(mtlib::kProfRdo, s);
                hipLaunchKernelGGL(rdo_uv_decide_kernel, dim3(pk), dim3(64), 0, s, w.fargs, tp, maxc, w.cres + n,
                                   w.uvblk, w.nuvblk, w.uvdesc, w.uvfirst, w.uvcount, w.sb_uv_cands, w.sb_uv_pool,
                                   w.sb_uv_dec);
                MT_HIP_TRY(hipGetLastError());
            }
            mtlib::ProfScope prof(mtlib::kProfRecon, s);
            if ((rc = mtlib::sb_recon_uv_launch(w.uvdesc, w.uvfirst, w.uvcount, pk, w.rframes, s))) return rc;
        }
    }
    return rc;
    }();
    if (as_graph) {
        const hipError_t e = hipStreamEndCapture(w.capture, &graph);
        if (rc == MT_OK && e != hipSuccess)
            rc = mtlib::set_error(MT_ERR_DEVICE, std::string("mt_intra_rdo: graph capture: ") + hipGetErrorString(e));
        if (rc == MT_OK && hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess)
            rc = mtlib::set_error(MT_ERR_DEVICE, "mt_intra_rdo: graph instantiation failed");
        if (rc == MT_OK && hipGraphLaunch(gexec, s) != hipSuccess)
            rc = mtlib::set_error(MT_ERR_DEVICE, "mt_intra_rdo: graph launch failed");
        if (rc == MT_OK && hipStreamSynchronize(s) != hipSuccess)
            rc = mtlib::set_error(MT_ERR_DEVICE, "mt_intra_rdo: graph execution failed");
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    if (rc) return rc;

    // final accounting with the final contexts
    std::vector<uint32_t> nleaf(nall), lpool(nall), lbase(nall), pbase(nall), uvc(nall), uvp(nall), uvd(nall);
    MT_HIP_TRY(hipMemcpyAsync(nleaf.data(), w.sb_leaves, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(lpool.data(), w.sb_leaf_pool, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(uvc.data(), w.sb_uv_cands, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(uvp.data(), w.sb_uv_pool, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(uvd.data(), w.sb_uv_dec, nall * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipStreamSynchronize(s));
    uint64_t tl = 0, tp = 0;
    for (uint32_t i = 0; i < nall; ++i) {
        lbase[i] = (uint32_t)tl;
        pbase[i] = (uint32_t)tp;
        tl += nleaf[i] + uvc[i];
        tp += lpool[i] + uvp[i];
    }
    if (tp > UINT32_MAX) return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: batch too large");
    if (w.cap_final < tl) {
        if ((rc = regrow(w.fcands, tl)) || (rc = regrow(w.fres, tl))) return rc;
        w.cap_final = tl;
    }
    MT_HIP_TRY(hipMemcpyAsync(w.leaf_base, lbase.data(), nall * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    MT_HIP_TRY(hipMemcpyAsync(w.pool_base, pbase.data(), nall * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    {
        mtlib::ProfScope prof(mtlib::kProfRdo, s);
        hipLaunchKernelGGL(rdo_final_gen_kernel, dim3((nall + 63) / 64), dim3(64), 0, s, w.fargs, w.all, (int)nall,
                           w.leaf_base, w.pool_base, w.fcands);
        MT_HIP_TRY(hipGetLastError());
    }
    if ((rc = mtlib::eval_launch(w.fcands, (int)tl, ps, tp, w.fres, s, (flags & MT_RDO_TRELLIS) != 0))) return rc;
    {
        mtlib::ProfScope prof(mtlib::kProfRdo, s);
        hipLaunchKernelGGL(rdo_final_sum_kernel, dim3((nall + 63) / 64), dim3(64), 0, s, w.fargs, w.all, (int)nall,
                           w.leaf_base, w.fres, w.sb_bits, w.sb_sse, w.sb_bits_uv, w.sb_sse_uv);
        MT_HIP_TRY(hipGetLastError());
    }
    std::vector<double> bits(nall), bits_uv(nall);
    std::vector<long long> sse(nall), sse_uv(2 * (size_t)nall);
    MT_HIP_TRY(hipMemcpyAsync(bits.data(), w.sb_bits, nall * sizeof(double), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(sse.data(), w.sb_sse, nall * sizeof(long long), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(bits_uv.data(), w.sb_bits_uv, nall * sizeof(double), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipMemcpyAsync(sse_uv.data(), w.sb_sse_uv, 2 * nall * sizeof(long long), hipMemcpyDeviceToHost, s));
    MT_HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < nf; ++i) {
        mt_rdo_stats& st = stats[i];
        std::memset(&st, 0, sizeof(st));
        const uint32_t n = (uint32_t)(C[i] * Rr[i]);
        int64_t uv_dec = 0;
        for (uint32_t sb = 0; sb < n; ++sb) {
            const uint32_t g = fa[i].sb0 + sb;
            st.bits += bits[g];
            st.sse += sse[g];
            st.leaves += (int32_t)nleaf[g];
            st.bits_uv += bits_uv[g];
            st.sse_u += sse_uv[2 * (size_t)g];
            st.sse_v += sse_uv[2 * (size_t)g + 1];
            uv_dec += uvd[g];
        }
        st.candidates = frame_cands[i] + uv_dec;
        st.sbs = (int32_t)n;
        st.steps = nsteps[i];
    }
    return MT_OK;
}

}  // namespace

namespace mtlib {
int rdo_init_device(int /*device*/) {
    SymBits b;
    for (int a = 0; a < 5; ++a)
        for (int l = 0; l < 5; ++l)
            for (int m = 0; m < 13; ++m) b.kf[a][l][m] = cdf_cost_raw(MT_KF_Y_MODE_CDF[a][l], 13, m);
    for (int m = 0; m < 8; ++m)
        for (int j = 0; j < 7; ++j) b.angle[m][j] = cdf_cost_raw(MT_ANGLE_DELTA_CDF[m], 7, j);
    const int bsz[5] = {0, 3, 6, 9, 12};  // av1_block_size_index (src/coder.rs:279) of 4x4 .. 64x64
    for (int L = 0; L < 5; ++L)
        for (int u = 0; u < 2; ++u) b.fuse[L][u] = cdf_cost_raw(&MT_FILTER_INTRA_CDF[bsz[L]], 2, u);
    for (int k = 0; k < 5; ++k) b.fmode[k] = cdf_cost_raw(MT_FILTER_INTRA_MODE_CDF, 5, k);
    for (int ctx = 0; ctx < 4; ++ctx) {
        for (int bl = 1; bl <= 3; ++bl) {
            b.part[bl - 1][ctx][0] = cdf_cost_raw(MT_PART_SPLIT_CDF[bl - 1][ctx], 10, 0);
            b.part[bl - 1][ctx][1] = cdf_cost_raw(MT_PART_SPLIT_CDF[bl - 1][ctx], 10, 3);
        }
        b.part[3][ctx][0] = cdf_cost_raw(MT_PART_BL8_CDF[ctx], 4, 0);
        b.part[3][ctx][1] = cdf_cost_raw(MT_PART_BL8_CDF[ctx], 4, 3);
    }
    for (int y = 0; y < 13; ++y) {
        for (int m = 0; m < 14; ++m) {
            b.uv[0][y][m] = m < 13 ? cdf_cost_raw(MT_UV_MODE_NOCFL_CDF[y], 13, m) : 0.0f;
            b.uv[1][y][m] = cdf_cost_raw(MT_UV_MODE_CFL_CDF[y], 14, m);
        }
    }
    for (int k = 0; k < 8; ++k) b.cfl_sign[k] = cdf_cost_raw(MT_CFL_SIGN_CDF, 8, k);
    for (int c = 0; c < 6; ++c)
        for (int k = 0; k < 16; ++k) b.cfl_alpha[c][k] = cdf_cost_raw(MT_CFL_ALPHA_CDF[c], 16, k);
    MT_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_sym), &b, sizeof(b)));
    return MT_OK;
}
}  // namespace mtlib

extern "C" float mt_mode_lambda(int qidx, int bd, int sub) {
    if (qidx < 0 || qidx > 255 || (bd != 8 && bd != 10 && bd != 12)) return -1.0f;
    // mode_lambda_q (src/cost.rs:285) on the luma DC step of Quant::new (src/quant.rs:350)
    mt_quant q;
    if (mt_quant_new(qidx, bd, 15, &q) != MT_OK) return -1.0f;
    const float dc_q = (float)q.dc;
    const float dc_q8 = dc_q / (float)(1u << ((bd > 8 ? bd : 8) - 8));
    const float mult = 3.3f + 0.0015f * dc_q8;  // def_kf_rd_multiplier (src/cost.rs:116)
    const float lq = 0.009005174719460433f * 1.0f * dc_q * dc_q * mult;
    // mode_lambda_weight (src/cost.rs:132) with the shipped tuning (src/tuning.rs:761-763,
    // :913-915): per-format scale 0.8 / 0.65 / 0.6, top-band tilt floor 0, full 72, weight 3
    const float scale = sub == 2 ? 0.8f : sub == 1 ? 0.65f : 0.6f;
    int wi = ((255 - qidx) * 3) / 4;
    wi = wi < 0 ? 0 : wi > 72 ? 72 : wi;
    const float w = (float)(wi + 128) * (1.0f / 128.0f);  // aom_ssimulacra2_rdmult_weight (:125)
    float tw;
    if (sub == 1) {
        tw = w * scale;
    } else {
        float r = ((float)qidx - 0.0f) / (72.0f - 0.0f);
        r = r < 0.0f ? 0.0f : r > 1.0f ? 1.0f : r;
        tw = (3.0f + (w - 3.0f) * r) * scale;
    }
    return lq * tw;
}

extern "C" float mt_rdo_lambda(int qidx, int bd, int sub) {
    const float m = mt_mode_lambda(qidx, bd, sub);
    return m < 0.0f ? m : m / 2.775f;  // K_PART_LAMBDA (src/cost.rs:278)
}

extern "C" int mt_intra_rdo_batch(const mt_rdo_frame* frames, int nframes, int flags, mt_rdo_stats* stats,
                                  void* stream) {
    if (!frames || !stats || nframes < 0) return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: null argument");
    if (flags & ~(MT_RDO_ANGLE_DELTAS | MT_RDO_FILTER_INTRA | MT_RDO_MIN4 | MT_RDO_CFL | MT_RDO_TRELLIS | MT_RDO_GRAPH |
                  MT_RDO_CFL_STEPS(15)))
        return mtlib::set_error(MT_ERR_ARG, "mt_intra_rdo: bad flags");
    int rc = MT_OK;
    for (int i = 0; i < nframes; ++i)
        if ((rc = check_frame(frames[i]))) return rc;
    if (nframes == 0) return MT_OK;
    int dev = 0, cus = 0;
    rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    RdoWork* wp;
    {
        std::lock_guard<std::mutex> lk(g_rdo_mu);
        wp = &g_rdo[mtlib::stream_key(dev, (hipStream_t)stream)];
    }
    RdoWork& w = *wp;
    std::lock_guard<std::mutex> lk(w.mu);
    for (int i0 = 0; i0 < nframes;) {
        int nf = 0, planes = 0;
        while (i0 + nf < nframes && nf < kBatchFrames && planes + frame_planes(frames[i0 + nf]) <= MT_MAX_PLANES)
            planes += frame_planes(frames[i0 + nf++]);
        if ((rc = run_batch(w, frames + i0, nf, flags, stats + i0, cus, (hipStream_t)stream))) return rc;
        i0 += nf;
    }
    return MT_OK;
}

extern "C" int mt_intra_rdo(const mt_rdo_frame* frame, int flags, mt_rdo_stats* stats, void* stream) {
    return mt_intra_rdo_batch(frame, 1, flags, stats, stream);
}
