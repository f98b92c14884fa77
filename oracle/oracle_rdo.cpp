// CPU oracle: the whole-frame intra RDO of the library (mt_intra_rdo), restated on the CPU
// with the oracle's candidate evaluation.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  The decision procedure is the library's own
// (the reference's ~35k-line decision heuristics are not part of the path); every quantity
// it combines is a reference kernel already restated here: the candidate loop body
// (src/coder/partition_search.rs:471-520, orc_eval_candidate), the mode syntax bits
// (luma_intra_syntax_bits, src/coder/superblock.rs:192; mode_bits,
// src/coder/lossy_state.rs:1290), the partition symbol bits (partition_bits_bl,
// src/coder/lossy_state.rs:1718 with get_partition_ctx, src/coeffs.rs:2917, and the context
// bytes of part_ctx_bytes, src/coder/superblock.rs:1062), rd_cost
// (src/partition_rd.rs:140) and the decision lambda (src/coder/lossy_state.rs:1769-1790).
//
// Procedure, per 64x64 superblock (SB) in wavefront order (SB (r, c) in step c + 2r):
//  1. every quadtree node (64 .. 8, or .. 4) inside the frame evaluates every allowed mode
//     (13 intra modes, the 6 angle deltas of the 8 directional modes on blocks of >= 8 in
//     both directions, the 5 filter-intra modes up to 32 in both) as one DCT_DCT transform
//     block, predicting from the reconstruction plane; samples of the SB itself still hold
//     the source there (the decision sees the source inside its own SB and the final
//     reconstruction outside); with the rectangular flag, the two halves of PARTITION_HORZ
//     and of PARTITION_VERT of every 32x32 / 16x16 (and, down to 4x4, 8x8) node likewise;
//  2. bottom-up, each node keeps PARTITION_NONE (its best mode), HORZ / VERT (the sums of
//     their halves' best) or SPLIT (the sum of its children), each priced with its
//     partition symbol; ties keep the first of NONE, HORZ, VERT, SPLIT;
//  3. the chosen leaves are reconstructed in coding order from the reconstruction.
// Decision-time contexts read the final mode / partition maps outside the SB and count
// neighbours inside the SB as DC_PRED / unsplit.  The final accounting re-evaluates every
// leaf with the final contexts.
// Chroma (4:2:0 / 4:4:4), after the SB's luma reconstruction: per chroma block (4:2:0: the
// half-size block of each leaf of >= 8 in both directions, one 4x4 block per 8x8 node
// split into 4x4 / 8x4 / 4x8 leaves; 4:4:4: one per leaf, a 64x64 as four 32x32 transform
// blocks), the 13 uv modes and CfL (where the luma block is <= 32x32) on U and V with the
// decoder's transform type for the uv mode (chroma_tx_for_mode, src/coder/chroma_tx.rs:46),
// the chroma quantizer, plain SSE and the plane-1 rate (txb skip context 7); cost = U + V
// rd costs + lambda * uv_mode_bits (src/coder/lossy_state.rs:1302), ties to the lower
// mode; then the chroma reconstruction in coding order and the final re-pricing.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "../maroontree_amd/csrc/av1_cdf_tables.h"

namespace {

constexpr int kNodes = 341;
constexpr int kLevelBase[5] = {0, 1, 5, 21, 85};
enum { kAngle = 1, kFilter = 2, kMin4 = 4, kCfl = 8, kRect = 16, kTrellis = 32 };
enum { kForceSplit = 0x4000, kNoSplit = 0x8000 };
// node decisions (the split field of a node record)
enum { kNone = 0, kSplit = 1, kUncoded = 2, kHorz = 3, kVert = 4 };

struct Geo {
    int L, ix, iy, s;
};
Geo node_geo(int id) {
    const int L = id >= 85 ? 4 : id >= 21 ? 3 : id >= 5 ? 2 : id >= 1 ? 1 : 0;
    const int k = id - kLevelBase[L], n = 1 << L;
    return {L, k % n, k / n, 64 >> L};
}
int node_id(int L, int ix, int iy) { return kLevelBase[L] + iy * (1 << L) + ix; }

// Coding order of the s-aligned cell (ix, iy) inside a superblock: bit-interleaved, y high.
int zidx(int ix, int iy) {
    int z = 0;
    for (int b = 0; b < 4; ++b) z |= (((ix >> b) & 1) << (2 * b)) | (((iy >> b) & 1) << (2 * b + 1));
    return z;
}
// Whether the s x s block at (xin, yin) inside its SB has its top-right / bottom-left
// neighbour of the same size coded before it (the frame edge clips the count separately).
bool have_tr(int xin, int yin, int s) {
    if (yin == 0) return true;
    if (xin + s >= 64) return false;
    return zidx(xin / s + 1, yin / s - 1) < zidx(xin / s, yin / s);
}
bool have_bl(int xin, int yin, int s) {
    if (yin + s >= 64) return false;
    if (xin == 0) return true;
    return zidx(xin / s - 1, yin / s + 1) < zidx(xin / s, yin / s);
}

// A block to predict: position, size, and whether its top-right / bottom-left neighbours
// are coded before it.
struct Blk {
    int x, y, w, h;
    bool tr, bl;
};
// Half k of a node at (x, y) of size s with availability (ntr, nbl): 0 / 1 the top / bottom
// of PARTITION_HORZ, 2 / 3 the left / right of PARTITION_VERT.  The top half's bottom-left
// and the left half's top-right lie in blocks coded before the node; the bottom half's
// top-right and the right half's bottom-left in blocks coded after it.
Blk rect_blk(int x, int y, int s, int k, bool ntr, bool nbl) {
    const int h = s / 2;
    switch (k) {
        case 0: return {x, y, s, h, ntr, true};
        case 1: return {x, y + h, s, h, false, nbl};
        case 2: return {x, y, h, s, true, nbl};
        default: return {x + h, y, h, s, ntr, false};
    }
}

// Mode ids 0..65: 0..12 the intra modes; 13 + 6 (m - 1) + j the directional mode m with
// delta {-3, -2, -1, 1, 2, 3}[j]; 61..65 filter intra 0..4.
constexpr int kModeIds = 66;
int id_mode(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 16 + id - 61; }
int id_delta(int id) {
    static const int d[6] = {-3, -2, -1, 1, 2, 3};
    return id >= 13 && id <= 60 ? d[(id - 13) % 6] : 0;
}
int id_base(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 0; }
int id_bit(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 13; }
// angle deltas are coded on blocks of >= 8x8 (av1_use_angle_delta: both sides >= 8 here),
// filter intra on blocks up to 32x32
bool id_valid(int id, int w, int h, int flags) {
    if (id <= 12) return true;
    if (id <= 60) return (flags & kAngle) && w >= 8 && h >= 8;
    return (flags & kFilter) && w <= 32 && h <= 32;
}

// AV1 BlockSize index of a w x h block (4x4 0, 4x8 1, 8x4 2, 8x8 3, ..., 64x64 12).
int bsize_index(int w, int h) {
    const int lw = __builtin_ctz((unsigned)w) - 2, lh = __builtin_ctz((unsigned)h) - 2;
    return 3 * std::min(lw, lh) + (lw == lh ? 0 : lw < lh ? 1 : 2);
}
// AV1 TxSize of a w x h transform (the 13 reference sizes).
int tx_of(int w, int h) {
    if (w == h) return w == 4 ? 0 : w == 8 ? 1 : w == 16 ? 2 : w == 32 ? 3 : 4;
    if (w == 4 && h == 8) return 5;
    if (w == 8 && h == 4) return 6;
    if (w == 8 && h == 16) return 7;
    if (w == 16 && h == 8) return 8;
    if (w == 16 && h == 32) return 9;
    return 10;  // 32x16
}
// eob_bin CDF index of a block of n coefficients (numbering of the rate descriptor)
int eob_luma(int n) { return n == 16 ? 0 : n == 32 ? 3 : n == 64 ? 5 : n == 128 ? 8 : n == 256 ? 11 : n == 512 ? 13 : 15; }
int eob_chroma(int n) { return n == 16 ? 1 : n == 32 ? 4 : n == 64 ? 6 : n == 128 ? 9 : n == 256 ? 12 : n == 512 ? 14 : 16; }
// the intra tx-type symbol set of a w x h luma transform: 1 / 2 the 7-type set with the
// 4x4 / 8x8 tables (by the smaller side), 3 the 5-type set at 16x16, 0 none (32 and up)
int txtp_set_of(int w, int h) {
    if (w >= 32 || h >= 32) return 0;
    if (w == 16 && h == 16) return 3;
    return std::min(w, h) == 4 ? 1 : 2;
}

// Symbol costs of the frame-initial decision CDFs (Cdfs::new, src/coder/cdfs.rs:400).
struct SymBits {
    float kf[5][5][13];
    float angle[8][7];
    float fuse[13][2];    // use_filter_intra by block size index
    float fmode[5];
    float part[4][4][4];  // [bl - 1][ctx][NONE, HORZ, VERT, SPLIT]
    float uv[2][13][14];  // uv_mode [CfL allowed][luma mode][uv mode]
    float cfl_sign[8];
    float cfl_alpha[6][16];
};
const SymBits& sym_bits() {
    static SymBits t = [] {
        SymBits b;
        for (int a = 0; a < 5; ++a)
            for (int l = 0; l < 5; ++l)
                for (int m = 0; m < 13; ++m) b.kf[a][l][m] = orc_cdf_cost_raw(MT_KF_Y_MODE_CDF[a][l], 13, m);
        for (int m = 0; m < 8; ++m)
            for (int j = 0; j < 7; ++j) b.angle[m][j] = orc_cdf_cost_raw(MT_ANGLE_DELTA_CDF[m], 7, j);
        for (int k = 0; k < 13; ++k)
            for (int u = 0; u < 2; ++u) b.fuse[k][u] = orc_cdf_cost_raw(&MT_FILTER_INTRA_CDF[k], 2, u);
        for (int k = 0; k < 5; ++k) b.fmode[k] = orc_cdf_cost_raw(MT_FILTER_INTRA_MODE_CDF, 5, k);
        for (int ctx = 0; ctx < 4; ++ctx) {
            for (int bl = 1; bl <= 3; ++bl)
                for (int sym = 0; sym < 4; ++sym)
                    b.part[bl - 1][ctx][sym] = orc_cdf_cost_raw(MT_PART_SPLIT_CDF[bl - 1][ctx], 10, sym);
            for (int sym = 0; sym < 4; ++sym) b.part[3][ctx][sym] = orc_cdf_cost_raw(MT_PART_BL8_CDF[ctx], 4, sym);
        }
        for (int y = 0; y < 13; ++y)
            for (int m = 0; m < 14; ++m) {
                b.uv[0][y][m] = m < 13 ? orc_cdf_cost_raw(MT_UV_MODE_NOCFL_CDF[y], 13, m) : 0.0f;
                b.uv[1][y][m] = orc_cdf_cost_raw(MT_UV_MODE_CFL_CDF[y], 14, m);
            }
        for (int k = 0; k < 8; ++k) b.cfl_sign[k] = orc_cdf_cost_raw(MT_CFL_SIGN_CDF, 8, k);
        for (int c = 0; c < 6; ++c)
            for (int k = 0; k < 16; ++k) b.cfl_alpha[c][k] = orc_cdf_cost_raw(MT_CFL_ALPHA_CDF[c], 16, k);
        return b;
    }();
    return t;
}

const int kIntraModeCtx[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};  // src/intrapred.rs:1088
const int kFilterDir[5] = {0, 1, 2, 6, 0};  // Filter_Intra_Mode_To_Intra_Dir
// partition context byte of a block dimension (part_ctx_bytes, src/coder/superblock.rs:1062):
// the above context takes the width's, the left context the height's
uint8_t part_byte(int d) { return d == 64 ? 0x10 : d == 32 ? 0x18 : d == 16 ? 0x1c : d == 8 ? 0x1e : 0x1f; }
// the partition symbol of a node decision
int part_sym(int split) { return split == kHorz ? 1 : split == kVert ? 2 : split == kSplit ? 3 : 0; }

int bdi_of(int bd) { return bd == 8 ? 0 : bd == 10 ? 1 : 2; }
int qctx_of(int q) { return q <= 20 ? 0 : q <= 60 ? 1 : q <= 120 ? 2 : 3; }

// luma_intra_syntax_bits (src/coder/superblock.rs:192) of mode id `id` on a w x h block.
float mode_extra_bits(int actx, int lctx, int id, int w, int h, int flags) {
    const SymBits& T = sym_bits();
    const int base = id_base(id);
    float bits = T.kf[actx][lctx][base];
    if (base >= 1 && base <= 8 && w >= 8 && h >= 8) {
        bits = bits + T.angle[base - 1][id_delta(id) + 3];
    } else if ((flags & kFilter) && base == 0 && w <= 32 && h <= 32) {
        const int use = id >= 61;
        bits = bits + T.fuse[bsize_index(w, h)][use];
        if (use) bits = bits + T.fmode[id - 61];
    }
    return bits;
}

// chroma_tx_for_mode (src/coder/chroma_tx.rs:46): ADST_ADST 3, ADST_DCT 1, DCT_ADST 2 (the
// AV1 TxType numbering), DCT_DCT otherwise and at 32 and up, where the intra set has no other.
int uv_txtp(int m, int w, int h) {
    if (w >= 32 || h >= 32) return 0;
    if (m == 12 || m == 9 || m == 4) return 3;
    if (m == 10 || m == 1 || m == 5 || m == 8) return 1;
    if (m == 11 || m == 2 || m == 6 || m == 7) return 2;
    return 0;
}

// uv_mode_bits (src/coder/lossy_state.rs:1302): the uv_mode symbol (CfL-allowed or no-CfL
// CDF as AV1 codes it), the delta-0 angle symbol where angle deltas are coded, the CfL joint
// sign and alpha magnitudes.
float uv_bits(int ym, int m, int au, int av, bool cfl_allowed, bool angle_ok) {
    const SymBits& T = sym_bits();
    if (m == 13) {
        const int su = au == 0 ? 0 : au < 0 ? 1 : 2, sv = av == 0 ? 0 : av < 0 ? 1 : 2;
        if (su == 0 && sv == 0) return T.uv[1][ym][0];  // both zero: DC_PRED (:1333)
        float b = T.uv[1][ym][13] + T.cfl_sign[su * 3 + sv - 1];
        if (su) b = b + T.cfl_alpha[(su == 2) * 3 + sv][std::abs(au) - 1];
        if (sv) b = b + T.cfl_alpha[(sv == 2) * 3 + su][std::abs(av) - 1];
        return b;
    }
    float b = T.uv[cfl_allowed ? 1 : 0][ym][m];
    if (angle_ok && m >= 1 && m <= 8) b = b + T.angle[m - 1][3];
    return b;
}

// A chroma block: the node it is coded at (and which half of a rectangular node), chroma
// origin and size, transform blocks, luma mode context, whether CfL is allowed / angle
// deltas are coded, and its availability.
struct UvB {
    int node, half, cx, cy, cw, ch, nt, ym;
    bool cfl_ok, angle_ok, tr, bl;
};

// A luma leaf in coding order.
struct Leaf {
    Blk b;
    int id;  // mode id
};

struct Frame {
    const uint16_t* src;
    int sstride, W, H, bd, qidx, qm, flags;
    float lambda, wsse;
    std::vector<uint16_t> rec;     // W x H
    std::vector<uint8_t> mode4;    // (W/4) x (H/4) luma modes
    std::vector<uint8_t> parta8;   // (W/8) x (H/8) above partition context bytes
    std::vector<uint8_t> partl8;   // (W/8) x (H/8) left partition context bytes
    orc_plane planes[6];  // source, reconstruction; with chroma U source / reconstruction, V
    int nplanes = 2;
    int minsz;
    int chroma = 0, ss = 0, Wc = 0, Hc = 0, uv_dc = 0, uv_ac = 0, uv_qm = 15;
    std::vector<uint16_t> rec_uv[2];  // Wc x Hc
    std::vector<uint8_t> uvmode4;     // (Wc/4) x (Hc/4) uv modes
    // optional: the reconstruction's levels, frequency (fx, fy) of the transform block at
    // (x, y) at sample (x + fx, y + fy) (W x H, chroma Wc x Hc)
    int32_t* coef = nullptr;
    int32_t* coef_uv[2] = {nullptr, nullptr};
    void put_levels(int32_t* plane, int pw, int x, int y, int n, const int32_t* lv) const {
        if (!plane) return;
        const int nc = n > 32 ? 32 : n;  // 64x64 codes its low 32x32 (x-major, 32 high)
        for (int fx = 0; fx < nc; ++fx)
            for (int fy = 0; fy < nc; ++fy) plane[(size_t)(y + fy) * pw + x + fx] = lv[fx * nc + fy];
    }

    int mode_at(int x, int y) const { return mode4[(size_t)(y >> 2) * (W >> 2) + (x >> 2)]; }
    int pa_at(int x, int y) const { return parta8[(size_t)(y >> 3) * (W >> 3) + (x >> 3)]; }
    int pl_at(int x, int y) const { return partl8[(size_t)(y >> 3) * (W >> 3) + (x >> 3)]; }
    int uv_at(int cx, int cy) const { return uvmode4[(size_t)(cy >> 2) * (Wc >> 2) + (cx >> 2)]; }

    bool rect_on() const { return (flags & kRect) != 0; }
    // the trellis bits of every candidate (rflags bit 1, the lambda law's subsampling in bits
    // 2-3: 0 at 4:4:4, 2 at 4:2:0 and for luma-only frames)
    uint8_t trellis_rflags() const { return (flags & kTrellis) ? (uint8_t)(2 | ((chroma == 2 ? 0 : 2) << 2)) : 0; }

    // the luma candidate of mode id `id` on block b of the SB at (x0, y0)
    orc_cand cand(const Blk& b, int id, int x0, int y0, bool final_ctx) const {
        orc_cand c;
        std::memset(&c, 0, sizeof(c));
        c.x = (uint16_t)b.x;
        c.y = (uint16_t)b.y;
        c.tx_size = (uint8_t)tx_of(b.w, b.h);
        c.tx_type = 0;
        c.mode = (uint8_t)id_mode(id);
        c.angle_delta = (int8_t)id_delta(id);
        // neighbours: final maps, except inside the SB at decision time (DC, not smooth)
        const bool above_ok = b.y > 0 && (final_ctx || b.y == y0);
        const bool left_ok = b.x > 0 && (final_ctx || b.x == x0);
        const int am = above_ok ? mode_at(b.x, b.y - 1) : 0, lm = left_ok ? mode_at(b.x - 1, b.y) : 0;
        const bool smooth = (am >= 9 && am <= 11) || (lm >= 9 && lm <= 11);
        c.pflags = (uint8_t)((b.tr ? 1 : 0) | (b.bl ? 2 : 0) | (smooth ? 4 : 0));
        c.planes = 0 | (1 << 2);
        c.qidx = (uint8_t)qidx;
        c.bd_qm = (uint8_t)(qm | (bdi_of(bd) << 4));
        c.rplane = 0;
        c.eob_bin = (uint8_t)eob_luma(std::min(b.w, 32) * std::min(b.h, 32));
        c.txtp_set = (uint8_t)txtp_set_of(b.w, b.h);
        c.txtp_mode = (uint8_t)(id >= 61 ? kFilterDir[id - 61] : id_base(id));
        c.txtp = (uint8_t)(c.txtp_set ? 1 : 0);  // DCT_DCT in the intra sets
        c.qctx = (uint8_t)qctx_of(qidx);
        c.rflags = trellis_rflags();
        c.wsse_scale = wsse;
        c.lambda = lambda;
        c.extra_bits = mode_extra_bits(kIntraModeCtx[am], kIntraModeCtx[lm], id, b.w, b.h, flags);
        return c;
    }

    // partition symbol `sym` (0 NONE, 1 HORZ, 2 VERT, 3 SPLIT) of the level-L node at (x, y)
    float part_bits(int x, int y, int L, int sym, int x0, int y0, bool final_ctx) const {
        const int bl = L + 1, sh = 4 - bl;
        const int a = (y > 0 && (final_ctx || y == y0)) ? pa_at(x, y - 1) : 0;
        const int l = (x > 0 && (final_ctx || x == x0)) ? pl_at(x - 1, y) : 0;
        const int ctx = ((a >> sh) & 1) + (((l >> sh) & 1) << 1);
        return sym_bits().part[bl - 1][ctx][sym];
    }

    // the leaves of the SB at (x0, y0) in coding order, from its node records
    std::vector<Leaf> leaves(const orc_rdo_node* nd, int x0, int y0) const {
        std::vector<Leaf> out;
        int stack[64], sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const int id = stack[--sp];
            const int d = nd[id].split;
            if (d == kUncoded) continue;
            const Geo g = node_geo(id);
            if (d == kSplit) {
                for (int k = 3; k >= 0; --k) stack[sp++] = node_id(g.L + 1, 2 * g.ix + (k & 1), 2 * g.iy + (k >> 1));
                continue;
            }
            const int x = x0 + g.ix * g.s, y = y0 + g.iy * g.s;
            const bool ntr = have_tr(x - x0, y - y0, g.s), nbl = have_bl(x - x0, y - y0, g.s);
            if (d == kNone) {
                out.push_back({{x, y, g.s, g.s, ntr, nbl}, nd[id].mode});
            } else {
                const int k0 = d == kHorz ? 0 : 2;
                out.push_back({rect_blk(x, y, g.s, k0, ntr, nbl), nd[id].rmode0});
                out.push_back({rect_blk(x, y, g.s, k0 + 1, ntr, nbl), nd[id].rmode1});
            }
        }
        return out;
    }

    // the chroma blocks of the SB at (x0, y0) in coding order
    std::vector<UvB> uv_blocks(const orc_rdo_node* nd, int x0, int y0) const {
        std::vector<UvB> out;
        int stack[64], sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const int id = stack[--sp];
            const int d = nd[id].split;
            if (d == kUncoded) continue;
            const Geo g = node_geo(id);
            const bool group = ss && g.s == 8 && d != kNone;  // 4:2:0, an 8x8 coded as smaller blocks
            if (d == kSplit && !group) {
                for (int k = 3; k >= 0; --k) stack[sp++] = node_id(g.L + 1, 2 * g.ix + (k & 1), 2 * g.iy + (k >> 1));
                continue;
            }
            const int x = x0 + g.ix * g.s, y = y0 + g.iy * g.s;
            const bool ntr = have_tr(x - x0, y - y0, g.s), nbl = have_bl(x - x0, y - y0, g.s);
            if (d == kNone || group) {
                UvB b;
                b.node = id;
                b.half = 0;
                b.cx = x >> ss;
                b.cy = y >> ss;
                b.cw = b.ch = group ? 4 : g.s >> ss;
                b.nt = b.cw == 64 ? 4 : 1;
                // the chroma of a split 8x8 is coded with (and under the mode of) its last block
                b.ym = !group ? id_base(nd[id].mode)
                       : d == kSplit ? id_base(nd[node_id(4, 2 * g.ix + 1, 2 * g.iy + 1)].mode)
                                     : id_base(nd[id].rmode1);
                b.cfl_ok = group || g.s <= 32;
                b.angle_ok = !group && g.s >= 8;
                b.tr = ntr;
                b.bl = nbl;
                out.push_back(b);
                continue;
            }
            const int k0 = d == kHorz ? 0 : 2;
            for (int h = 0; h < 2; ++h) {
                const Blk r = rect_blk(x, y, g.s, k0 + h, ntr, nbl);
                UvB b;
                b.node = id;
                b.half = h;
                b.cx = r.x >> ss;
                b.cy = r.y >> ss;
                b.cw = r.w >> ss;
                b.ch = r.h >> ss;
                b.nt = 1;
                b.ym = id_base(h ? nd[id].rmode1 : nd[id].rmode0);
                b.cfl_ok = true;
                b.angle_ok = r.w >= 8 && r.h >= 8;
                b.tr = r.tr;
                b.bl = r.bl;
                out.push_back(b);
            }
        }
        return out;
    }

    // cfl_alpha_rd_steps (src/tuning.rs:343): flag bits 8-11
    int cfl_steps() const { return (flags >> 8) & 15; }
    // uv modes 0..12, CfL (13), then CfL's walk candidates (14 + j: alpha step -(j + 1) toward
    // zero for j < S, j - S + 1 away from zero after that)
    int uv_modes(const UvB& b) const { return 13 + ((flags & kCfl) && b.cfl_ok ? 1 + 2 * cfl_steps() : 0); }
    int cfl_step(int m) const {
        const int j = m - 14, S = cfl_steps();
        return j < S ? -(j + 1) : j - S + 1;
    }

    // candidate of uv mode m, plane p (0 U, 1 V), transform block t; fin: final contexts and
    // the given CfL alpha
    orc_cand uv_cand(const UvB& b, int m, int p, int t, int x0, int y0, bool fin, int alpha) const {
        orc_cand c;
        std::memset(&c, 0, sizeof(c));
        const int tw = b.nt == 4 ? 32 : b.cw, th = b.nt == 4 ? 32 : b.ch;
        const int tx = b.cx + (t & 1) * tw, ty = b.cy + (t >> 1) * th;
        c.x = (uint16_t)tx;
        c.y = (uint16_t)ty;
        c.tx_size = (uint8_t)tx_of(tw, th);
        c.tx_type = (uint8_t)(m >= 13 ? 0 : uv_txtp(m, tw, th));
        c.mode = (uint8_t)(m > 13 ? 13 : m);
        if (m >= 13) {
            c.pflags = (uint8_t)((ss ? 8 | 16 : 0) | (fin ? 32 : 0) | (m > 13 ? 64 : 0));
            c.angle_delta = (int8_t)(fin ? alpha : m > 13 ? cfl_step(m) : 0);
            c.txtp_mode = 1;  // the luma reconstruction plane
        } else {
            const bool tr = b.nt == 4 ? have_tr((t & 1) * 32, (t >> 1) * 32, 32) : b.tr;
            const bool bl = b.nt == 4 ? have_bl((t & 1) * 32, (t >> 1) * 32, 32) : b.bl;
            const bool above_ok = ty > 0 && (fin || ty == (y0 >> ss));
            const bool left_ok = tx > 0 && (fin || tx == (x0 >> ss));
            const int am = above_ok ? uv_at(tx, ty - 1) : 0, lm = left_ok ? uv_at(tx - 1, ty) : 0;
            const bool smooth = (am >= 9 && am <= 11) || (lm >= 9 && lm <= 11);
            c.pflags = (uint8_t)((tr ? 1 : 0) | (bl ? 2 : 0) | (smooth ? 4 : 0));
        }
        const int sp = 2 + 2 * p, rp = sp + 1;
        c.planes = (uint8_t)((sp & 3) | ((rp & 3) << 2));
        c.planes_hi = (uint8_t)((sp >> 2) | ((rp >> 2) << 4));
        c.qidx = (uint8_t)qidx;
        c.bd_qm = (uint8_t)(uv_qm | (bdi_of(bd) << 4) | 0x40);
        c.dc_delta = (int8_t)uv_dc;
        c.ac_delta = (int8_t)uv_ac;
        c.rplane = 1;
        c.eob_bin = (uint8_t)eob_chroma(tw * th);  // the chroma eob_bin CDFs
        c.skip_ctx = 7;
        c.qctx = (uint8_t)qctx_of(qidx);
        c.rflags = trellis_rflags();
        c.wsse_scale = 0.0f;
        c.lambda = lambda;
        return c;
    }

    // CfL of chroma block b (one transform block per plane): each plane's cfl_best_alpha
    // candidate, then with S = cfl_steps() > 0 the walk of cfl_alpha_rd_refine
    // (src/coder/block32.rs:2350-2391, src/cost.rs:332) over the 2S step candidates, U first
    // against V's energy optimum, then V against U's result, each alpha scored by
    // cfl_alpha_walk_cost's default (src/cost.rs:324): dist + lambda * (coefficient bits +
    // uv_mode_bits of the alpha pair).  Returns the summed body cost (rd_cost of each plane's
    // chosen candidate) and the alphas.
    float cfl_decide(const UvB& b, int x0, int y0, int& au, int& av, int64_t& ncand) const {
        orc_candres base[2];
        for (int pl = 0; pl < 2; ++pl) {
            const orc_cand cd = uv_cand(b, 13, pl, 0, x0, y0, false, 0);
            orc_eval_candidate(&cd, planes, nplanes, &base[pl]);
            ++ncand;
        }
        int a[2] = {base[0].alpha, base[1].alpha};
        float body[2] = {base[0].cost, base[1].cost};
        const int S = cfl_steps();
        if (S > 0) {
            std::vector<orc_candres> st(4 * (size_t)S);
            for (int pl = 0; pl < 2; ++pl)
                for (int j = 0; j < 2 * S; ++j) {
                    const orc_cand cd = uv_cand(b, 14 + j, pl, 0, x0, y0, false, 0);
                    orc_eval_candidate(&cd, planes, nplanes, &st[(size_t)pl * 2 * S + j]);
                    ++ncand;
                }
            for (int pl = 0; pl < 2; ++pl) {
                auto walk_cost = [&](const orc_candres& r, int al) {
                    const int pu = pl ? a[0] : al, pv = pl ? al : a[1];
                    return (float)r.dist + lambda * (r.bits + uv_bits(b.ym, 13, pu, pv, true, b.angle_ok));
                };
                float cost[33], bc[33];
                for (int k = 0; k < 33; ++k) cost[k] = bc[k] = __builtin_inff();
                for (int j = 0; j < 2 * S; ++j) {
                    const orc_candres& r = st[(size_t)pl * 2 * S + j];
                    cost[r.alpha + 16] = walk_cost(r, r.alpha);
                    bc[r.alpha + 16] = r.cost;
                }
                const int a1 = orc_cfl_alpha_rd_refine(a[pl], walk_cost(base[pl], a[pl]), S, cost);
                if (a1 != a[pl]) body[pl] = bc[a1 + 16];
                a[pl] = a1;
            }
        }
        au = a[0];
        av = a[1];
        return (0.0f + body[0]) + body[1];
    }

    void put_leaf(const Leaf& lf) {
        const Blk& b = lf.b;
        const int base = id_base(lf.id);
        for (int yy = b.y; yy < b.y + b.h; yy += 4)
            for (int xx = b.x; xx < b.x + b.w; xx += 4) mode4[(size_t)(yy >> 2) * (W >> 2) + (xx >> 2)] = (uint8_t)base;
        for (int yy = b.y; yy < b.y + b.h; yy += 4)
            for (int xx = b.x; xx < b.x + b.w; xx += 4) {
                const size_t u = (size_t)(yy >> 3) * (W >> 3) + (xx >> 3);
                parta8[u] = part_byte(b.w);
                partl8[u] = part_byte(b.h);
            }
    }
};

struct SbOut {
    int64_t candidates = 0;
    int32_t leaves = 0;
};

// best candidate of block b over the mode ids allowed by mask m
void best_of(const Frame& f, const Blk& b, uint16_t m, int x0, int y0, float& best, int& best_id, SbOut& out) {
    best = __builtin_inff();
    best_id = 0;
    for (int mid = 0; mid < kModeIds; ++mid) {
        if (!id_valid(mid, b.w, b.h, f.flags)) continue;
        if (mid != 0 && !((m >> id_bit(mid)) & 1)) continue;
        const orc_cand cd = f.cand(b, mid, x0, y0, false);
        orc_candres res;
        orc_eval_candidate(&cd, f.planes, f.nplanes, &res);
        ++out.candidates;
        if (res.cost < best) {
            best = res.cost;
            best_id = mid;
        }
    }
}

// Decide and reconstruct one superblock (its neighbours of earlier steps are final).
void do_sb(Frame& f, int r, int c, const uint16_t* mask, orc_rdo_node* nodes, SbOut& out) {
    const int x0 = c * 64, y0 = r * 64;
    bool searched[kNodes], inside[kNodes], none_ok[kNodes], may_split[kNodes], rect_ok[kNodes];
    float best[kNodes], rbest[kNodes][4];
    int best_id[kNodes], rbest_id[kNodes][4];
    for (int id = 0; id < kNodes; ++id) {
        const Geo g = node_geo(id);
        const int x = x0 + g.ix * g.s, y = y0 + g.iy * g.s;
        const uint16_t m = mask ? mask[id] : 0x3FFF;
        bool sr;
        if (id == 0) {
            sr = true;
        } else {
            const int pid = node_id(g.L - 1, g.ix >> 1, g.iy >> 1);
            sr = may_split[pid] && x < f.W && y < f.H;
        }
        searched[id] = sr;
        inside[id] = x + g.s <= f.W && y + g.s <= f.H;
        none_ok[id] = sr && inside[id] && (g.s == f.minsz || !(m & kForceSplit) || (m & kNoSplit));
        may_split[id] = sr && g.s > f.minsz && (!inside[id] || !(m & kNoSplit));
        // the halves of HORZ / VERT: 32x32 .. 8x8 nodes whose halves are not below the minimum
        rect_ok[id] = f.rect_on() && sr && inside[id] && g.L >= 1 && g.L <= 3 && g.s / 2 >= f.minsz &&
                      !(m & kForceSplit);
        best[id] = __builtin_inff();
        best_id[id] = 0;
        for (int k = 0; k < 4; ++k) {
            rbest[id][k] = __builtin_inff();
            rbest_id[id][k] = 0;
        }
        const bool ntr = have_tr(x - x0, y - y0, g.s), nbl = have_bl(x - x0, y - y0, g.s);
        if (none_ok[id]) best_of(f, {x, y, g.s, g.s, ntr, nbl}, m, x0, y0, best[id], best_id[id], out);
        if (rect_ok[id])
            for (int k = 0; k < 4; ++k)
                best_of(f, rect_blk(x, y, g.s, k, ntr, nbl), m, x0, y0, rbest[id][k], rbest_id[id][k], out);
    }
    // bottom-up decision
    float cost[kNodes];
    for (int id = kNodes - 1; id >= 0; --id) {
        const Geo g = node_geo(id);
        const int x = x0 + g.ix * g.s, y = y0 + g.iy * g.s;
        orc_rdo_node& nd = nodes[id];
        std::memset(&nd, 0, sizeof(nd));
        nd.mode = (uint8_t)best_id[id];
        nd.uv_mode = 255;
        nd.uv_mode1 = 255;
        if (!searched[id]) {
            cost[id] = 0.0f;
            nd.cost = __builtin_inff();
            nd.none_cost = __builtin_inff();
            nd.split = kUncoded;
            continue;
        }
        const float inf = __builtin_inff();
        float none = inf, horz = inf, vert = inf, spl = inf;
        if (none_ok[id] && best[id] < inf) {
            none = best[id];
            if (g.s >= 8) none = none + f.lambda * f.part_bits(x, y, g.L, 0, x0, y0, false);
        }
        if (rect_ok[id]) {
            if (rbest[id][0] < inf && rbest[id][1] < inf)
                horz = (rbest[id][0] + rbest[id][1]) + f.lambda * f.part_bits(x, y, g.L, 1, x0, y0, false);
            if (rbest[id][2] < inf && rbest[id][3] < inf)
                vert = (rbest[id][2] + rbest[id][3]) + f.lambda * f.part_bits(x, y, g.L, 2, x0, y0, false);
        }
        if (may_split[id]) {
            float sum = 0.0f;
            for (int k = 0; k < 4; ++k) sum = sum + cost[node_id(g.L + 1, 2 * g.ix + (k & 1), 2 * g.iy + (k >> 1))];
            spl = inside[id] ? sum + f.lambda * f.part_bits(x, y, g.L, 3, x0, y0, false) : sum;
        }
        float bc = none;
        int d = kNone;
        if (horz < bc) {
            bc = horz;
            d = kHorz;
        }
        if (vert < bc) {
            bc = vert;
            d = kVert;
        }
        if (spl < bc) {
            bc = spl;
            d = kSplit;
        }
        cost[id] = bc;
        nd.cost = bc;
        nd.none_cost = none;
        nd.split = (uint8_t)d;
        if (d == kHorz || d == kVert) {
            const int k0 = d == kHorz ? 0 : 2;
            nd.rmode0 = (uint8_t)rbest_id[id][k0];
            nd.rmode1 = (uint8_t)rbest_id[id][k0 + 1];
        }
    }
    // leaves in coding order: maps, reconstruction
    const std::vector<Leaf> lv = f.leaves(nodes, x0, y0);
    out.leaves = (int32_t)lv.size();
    for (const Leaf& lf : lv) {
        f.put_leaf(lf);
        const orc_cand cd = f.cand(lf.b, lf.id, x0, y0, true);
        orc_candres res;
        std::vector<int32_t> rc((size_t)lf.b.w * lf.b.h), lvl(1024);
        orc_eval_candidate_recon_levels(&cd, f.planes, f.nplanes, &res, rc.data(), lvl.data());
        if (lf.b.w == lf.b.h) f.put_levels(f.coef, f.W, lf.b.x, lf.b.y, lf.b.w, lvl.data());
        for (int yy = 0; yy < lf.b.h; ++yy)
            for (int xx = 0; xx < lf.b.w; ++xx)
                f.rec[(size_t)(lf.b.y + yy) * f.W + lf.b.x + xx] = (uint16_t)rc[yy * lf.b.w + xx];
    }
    if (!f.chroma) return;
    // chroma decisions (decision-time contexts), the uv map, then the reconstruction
    const std::vector<UvB> blocks = f.uv_blocks(nodes, x0, y0);
    std::vector<int> um(blocks.size()), ua(blocks.size()), uav(blocks.size());
    for (size_t i = 0; i < blocks.size(); ++i) {
        const UvB& b = blocks[i];
        float bestc = __builtin_inff();
        int bm = 0, ba = 0, bv = 0;
        for (int m = 0; m < std::min(f.uv_modes(b), 14); ++m) {
            float sum = 0.0f;
            int au = 0, av = 0;
            if (m == 13) {
                sum = f.cfl_decide(b, x0, y0, au, av, out.candidates);
            } else {
                for (int pl = 0; pl < 2; ++pl)
                    for (int t = 0; t < b.nt; ++t) {
                        const orc_cand cd = f.uv_cand(b, m, pl, t, x0, y0, false, 0);
                        orc_candres res;
                        orc_eval_candidate(&cd, f.planes, f.nplanes, &res);
                        ++out.candidates;
                        sum = sum + res.cost;
                    }
            }
            if (m == 13 && au == 0 && av == 0) continue;
            if (!(sum < __builtin_inff())) continue;
            const float cost = sum + f.lambda * uv_bits(b.ym, m, au, av, b.cfl_ok, b.angle_ok);
            if (cost < bestc) {
                bestc = cost;
                bm = m;
                ba = au;
                bv = av;
            }
        }
        um[i] = bm;
        ua[i] = ba;
        uav[i] = bv;
        orc_rdo_node& nd = nodes[b.node];
        if (b.half) {
            nd.uv_mode1 = (uint8_t)bm;
            nd.cfl_alpha_u1 = (int8_t)ba;
            nd.cfl_alpha_v1 = (int8_t)bv;
        } else {
            nd.uv_mode = (uint8_t)bm;
            nd.cfl_alpha_u = (int8_t)ba;
            nd.cfl_alpha_v = (int8_t)bv;
        }
        for (int yy = 0; yy < b.ch; yy += 4)
            for (int xx = 0; xx < b.cw; xx += 4)
                f.uvmode4[(size_t)((b.cy + yy) >> 2) * (f.Wc >> 2) + ((b.cx + xx) >> 2)] = (uint8_t)bm;
    }
    for (size_t i = 0; i < blocks.size(); ++i) {
        const UvB& b = blocks[i];
        const int tw = b.nt == 4 ? 32 : b.cw, th = b.nt == 4 ? 32 : b.ch;
        for (int pl = 0; pl < 2; ++pl)
            for (int t = 0; t < b.nt; ++t) {
                const orc_cand cd = f.uv_cand(b, um[i], pl, t, x0, y0, true, pl ? uav[i] : ua[i]);
                orc_candres res;
                std::vector<int32_t> rc((size_t)tw * th), lvl(1024);
                orc_eval_candidate_recon_levels(&cd, f.planes, f.nplanes, &res, rc.data(), lvl.data());
                if (tw == th) f.put_levels(f.coef_uv[pl], f.Wc, cd.x, cd.y, tw, lvl.data());
                std::vector<uint16_t>& R = f.rec_uv[pl];
                for (int yy = 0; yy < th; ++yy)
                    for (int xx = 0; xx < tw; ++xx)
                        R[(size_t)(cd.y + yy) * f.Wc + cd.x + xx] = (uint16_t)rc[yy * tw + xx];
            }
    }
}

}  // namespace

extern "C" int orc_rdo_have_tr(int xin, int yin, int s) { return have_tr(xin, yin, s) ? 1 : 0; }
extern "C" int orc_rdo_have_bl(int xin, int yin, int s) { return have_bl(xin, yin, s) ? 1 : 0; }
extern "C" int orc_rdo_rect_avail(int xin, int yin, int s, int k) {
    const Blk b = rect_blk(xin, yin, s, k, have_tr(xin, yin, s), have_bl(xin, yin, s));
    return (b.tr ? 1 : 0) | (b.bl ? 2 : 0);
}

extern "C" float orc_mode_lambda(int qidx, int bd, int sub) {
    // mode_lambda_q (src/cost.rs:285) on the luma DC step of Quant::new (src/quant.rs:350)
    orc_quant q;
    orc_quant_new(qidx, bd, 15, &q);
    const float dc_q = (float)q.dc;
    const float dc_q8 = dc_q / (float)(1u << ((bd > 8 ? bd : 8) - 8));
    const float mult = 3.3f + 0.0015f * dc_q8;  // def_kf_rd_multiplier (src/cost.rs:116)
    const float lq = 0.009005174719460433f * 1.0f * dc_q * dc_q * mult;
    // mode_lambda_weight (src/cost.rs:132) with the shipped tuning (src/tuning.rs:761-763,
    // :913-915): scale 0.8 / 0.65 / 0.6, top-band tilt floor 0, full 72, weight 3
    const float scale = sub == 2 ? 0.8f : sub == 1 ? 0.65f : 0.6f;
    int wi = ((255 - qidx) * 3) / 4;
    wi = wi < 0 ? 0 : wi > 72 ? 72 : wi;
    const float w = (float)(wi + 128) * (1.0f / 128.0f);  // aom_ssimulacra2_rdmult_weight
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

extern "C" float orc_rdo_lambda(int qidx, int bd, int sub) {
    return orc_mode_lambda(qidx, bd, sub) / 2.775f;  // lambda_d = mlam / K_PART_LAMBDA
}

extern "C" int orc_intra_rdo(const uint16_t* src, int src_stride, const orc_rdo_params* p,
                             const uint16_t* prune, uint16_t* recon, orc_rdo_node* nodes,
                             uint8_t* modes, orc_rdo_stats* stats, int nthreads) {
    if (!p || p->chroma != 0) return -1;
    return orc_intra_rdo_yuv(src, src_stride, nullptr, nullptr, 0, p, prune, recon, nullptr, nullptr, nodes, modes,
                             stats, nthreads);
}

extern "C" int orc_intra_rdo_yuv(const uint16_t* src, int src_stride, const uint16_t* src_u, const uint16_t* src_v,
                                 int uv_stride, const orc_rdo_params* p, const uint16_t* prune, uint16_t* recon,
                                 uint16_t* recon_u, uint16_t* recon_v, orc_rdo_node* nodes, uint8_t* modes,
                                 orc_rdo_stats* stats, int nthreads) {
    return orc_intra_rdo_yuv_coef(src, src_stride, src_u, src_v, uv_stride, p, prune, recon, recon_u, recon_v, nodes,
                                  modes, stats, nthreads, nullptr, nullptr, nullptr);
}

extern "C" int orc_intra_rdo_yuv_coef(const uint16_t* src, int src_stride, const uint16_t* src_u,
                                      const uint16_t* src_v, int uv_stride, const orc_rdo_params* p,
                                      const uint16_t* prune, uint16_t* recon, uint16_t* recon_u, uint16_t* recon_v,
                                      orc_rdo_node* nodes, uint8_t* modes, orc_rdo_stats* stats, int nthreads,
                                      int32_t* coef, int32_t* coef_u, int32_t* coef_v) {
    if (!p || !src || !recon || !nodes || !stats) return -1;
    if (p->chroma < 0 || p->chroma > 2) return -1;
    if (p->chroma && (!src_u || !src_v || !recon_u || !recon_v || p->uv_qm_level < 0 || p->uv_qm_level > 15))
        return -1;
    if (p->width <= 0 || p->height <= 0 || (p->width & 7) || (p->height & 7) || p->width > 65535 ||
        p->height > 65535 || src_stride < p->width)
        return -1;
    if (p->bd != 8 && p->bd != 10 && p->bd != 12) return -1;
    if (p->qidx < 0 || p->qidx > 255 || p->qm_level < 0 || p->qm_level > 15) return -1;
    Frame f;
    f.src = src;
    f.sstride = src_stride;
    f.coef = coef;
    f.coef_uv[0] = coef_u;
    f.coef_uv[1] = coef_v;
    f.W = p->width;
    f.H = p->height;
    f.bd = p->bd;
    f.qidx = p->qidx;
    f.qm = p->qm_level;
    f.flags = p->flags;
    f.lambda = p->lambda;
    f.wsse = p->wsse_scale;
    f.minsz = (p->flags & kMin4) ? 4 : 8;
    f.rec.resize((size_t)f.W * f.H);
    for (int y = 0; y < f.H; ++y) std::memcpy(&f.rec[(size_t)y * f.W], src + (size_t)y * src_stride, f.W * 2);
    f.mode4.assign((size_t)(f.W / 4) * (f.H / 4), 0);
    f.parta8.assign((size_t)(f.W / 8) * (f.H / 8), 0);
    f.partl8.assign((size_t)(f.W / 8) * (f.H / 8), 0);
    f.planes[0] = {src, src_stride, f.W, f.H, f.W, f.H, 0};
    f.planes[1] = {f.rec.data(), f.W, f.W, f.H, f.W, f.H, 0};
    f.chroma = p->chroma;
    if (f.chroma) {
        f.ss = f.chroma == 1 ? 1 : 0;
        f.Wc = f.W >> f.ss;
        f.Hc = f.H >> f.ss;
        if (uv_stride < f.Wc) return -1;
        f.uv_dc = p->uv_dc_delta;
        f.uv_ac = p->uv_ac_delta;
        f.uv_qm = p->uv_qm_level;
        const uint16_t* cs[2] = {src_u, src_v};
        for (int k = 0; k < 2; ++k) {
            f.rec_uv[k].resize((size_t)f.Wc * f.Hc);
            for (int y = 0; y < f.Hc; ++y)
                std::memcpy(&f.rec_uv[k][(size_t)y * f.Wc], cs[k] + (size_t)y * uv_stride, f.Wc * 2);
            f.planes[2 + 2 * k] = {cs[k], uv_stride, f.Wc, f.Hc, f.Wc, f.Hc, 0};
            f.planes[3 + 2 * k] = {f.rec_uv[k].data(), f.Wc, f.Wc, f.Hc, f.Wc, f.Hc, 0};
        }
        f.uvmode4.assign((size_t)(f.Wc / 4) * (f.Hc / 4), 0);
        f.nplanes = 6;
    }
    const int C = (f.W + 63) / 64, R = (f.H + 63) / 64, nsb = C * R;
    std::vector<SbOut> outs(nsb);
    const int nsteps = (C - 1) + 2 * (R - 1) + 1;
    for (int k = 0; k < nsteps; ++k) {
        std::vector<int> sbs;
        for (int r = 0; r < R; ++r) {
            const int c = k - 2 * r;
            if (c >= 0 && c < C) sbs.push_back(r * C + c);
        }
        auto run = [&](int i) {
            const int sb = sbs[i];
            do_sb(f, sb / C, sb % C, prune ? prune + (size_t)sb * kNodes : nullptr,
                  nodes + (size_t)sb * kNodes, outs[sb]);
        };
        if (nthreads <= 1 || sbs.size() == 1) {
            for (size_t i = 0; i < sbs.size(); ++i) run((int)i);
        } else {
            std::vector<std::thread> th;
            const int nt = std::min<int>(nthreads, (int)sbs.size());
            for (int t = 0; t < nt; ++t)
                th.emplace_back([&, t] {
                    for (int i = t; i < (int)sbs.size(); i += nt) run(i);
                });
            for (auto& x : th) x.join();
        }
    }
    // final accounting with the final contexts
    double bits = 0.0;
    int64_t sse = 0, cands = 0;
    int32_t leaves = 0;
    std::vector<double> sb_bits(nsb, 0.0), sb_bits_uv(nsb, 0.0);
    std::vector<int64_t> sb_sse(nsb, 0), sb_sse_u(nsb, 0), sb_sse_v(nsb, 0);
    auto account = [&](int sb) {
        const int r = sb / C, c = sb % C, x0 = c * 64, y0 = r * 64;
        const orc_rdo_node* nd = nodes + (size_t)sb * kNodes;
        double b = 0.0;
        int64_t e = 0;
        int stack[64], sp = 0;
        stack[sp++] = 0;
        // per coded node in coding order: its partition symbol, then its leaf or halves
        while (sp) {
            const int id = stack[--sp];
            const int d = nd[id].split;
            if (d == kUncoded) continue;
            const Geo g = node_geo(id);
            const int x = x0 + g.ix * g.s, y = y0 + g.iy * g.s;
            const bool in = x + g.s <= f.W && y + g.s <= f.H;
            if (g.s >= 8 && in) b += (double)f.part_bits(x, y, g.L, part_sym(d), x0, y0, true);
            if (d == kSplit) {
                for (int k = 3; k >= 0; --k) stack[sp++] = node_id(g.L + 1, 2 * g.ix + (k & 1), 2 * g.iy + (k >> 1));
                continue;
            }
            const bool ntr = have_tr(x - x0, y - y0, g.s), nbl = have_bl(x - x0, y - y0, g.s);
            Leaf lf[2];
            int nl = 1;
            if (d == kNone) {
                lf[0] = {{x, y, g.s, g.s, ntr, nbl}, nd[id].mode};
            } else {
                const int k0 = d == kHorz ? 0 : 2;
                lf[0] = {rect_blk(x, y, g.s, k0, ntr, nbl), nd[id].rmode0};
                lf[1] = {rect_blk(x, y, g.s, k0 + 1, ntr, nbl), nd[id].rmode1};
                nl = 2;
            }
            for (int i = 0; i < nl; ++i) {
                const orc_cand cd = f.cand(lf[i].b, lf[i].id, x0, y0, true);
                orc_candres res;
                orc_eval_candidate(&cd, f.planes, f.nplanes, &res);
                b += (double)res.bits;
                e += res.sse;
            }
        }
        double buv = 0.0;
        int64_t euv[2] = {0, 0};
        if (f.chroma) {
            for (const UvB& ub : f.uv_blocks(nd, x0, y0)) {
                const orc_rdo_node& o = nd[ub.node];
                const int um = ub.half ? o.uv_mode1 : o.uv_mode;
                const int au = ub.half ? o.cfl_alpha_u1 : o.cfl_alpha_u, av = ub.half ? o.cfl_alpha_v1 : o.cfl_alpha_v;
                for (int pl = 0; pl < 2; ++pl)
                    for (int t = 0; t < ub.nt; ++t) {
                        orc_cand cd = f.uv_cand(ub, um, pl, t, x0, y0, true, pl ? av : au);
                        if (pl == 0 && t == 0) cd.extra_bits = uv_bits(ub.ym, um, au, av, ub.cfl_ok, ub.angle_ok);
                        orc_candres res;
                        orc_eval_candidate(&cd, f.planes, f.nplanes, &res);
                        b += (double)res.bits;
                        buv += (double)res.bits;
                        euv[pl] += res.sse;
                    }
            }
        }
        sb_bits[sb] = b;
        sb_sse[sb] = e;
        sb_bits_uv[sb] = buv;
        sb_sse_u[sb] = euv[0];
        sb_sse_v[sb] = euv[1];
    };
    {
        std::vector<std::thread> th;
        const int nt = nthreads <= 1 ? 1 : std::min(nthreads, nsb);
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (int sb = t; sb < nsb; sb += nt) account(sb);
            });
        for (auto& x : th) x.join();
    }
    double bits_uv = 0.0;
    int64_t sse_u = 0, sse_v = 0;
    for (int sb = 0; sb < nsb; ++sb) {
        bits += sb_bits[sb];
        sse += sb_sse[sb];
        bits_uv += sb_bits_uv[sb];
        sse_u += sb_sse_u[sb];
        sse_v += sb_sse_v[sb];
        cands += outs[sb].candidates;
        leaves += outs[sb].leaves;
    }
    std::memcpy(recon, f.rec.data(), f.rec.size() * 2);
    if (f.chroma) {
        std::memcpy(recon_u, f.rec_uv[0].data(), f.rec_uv[0].size() * 2);
        std::memcpy(recon_v, f.rec_uv[1].data(), f.rec_uv[1].size() * 2);
    }
    stats->bits_uv = bits_uv;
    stats->sse_u = sse_u;
    stats->sse_v = sse_v;
    if (modes) std::memcpy(modes, f.mode4.data(), f.mode4.size());
    stats->bits = bits;
    stats->sse = sse;
    stats->candidates = cands;
    stats->leaves = leaves;
    stats->sbs = nsb;
    return 0;
}
