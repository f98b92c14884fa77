// The AV1 bitstream of the RDO's decisions (host): entropy tokens, the range coder, the
// OBUs and the AVIF container.
//
// The RDO (mt_intra_rdo) leaves the decisions in its node records and, on request, the
// levels of every transform block of the reconstruction.  Everything the coder needs is
// therefore final before it starts, so the token stream of each superblock depends only on
// those arrays: the contexts of a block are read from the final decisions and levels of its
// neighbours.  The superblocks' tokens are generated in parallel (std::thread) and the range
// coder replays them in raster order.  The frame is coded with static frame-initial CDFs
// (disable_cdf_update = 1), the CDFs the RDO priced its decisions against, and one tile.
//
// The syntax, the contexts and the symbol coding follow the AV1 specification: sections 5.9
// (frame header), 5.11 (tile: partition, intra frame mode info, residual, coefficients),
// 8.2 (symbol coding) and 8.3 (CDF selection); the reference writes the same syntax in
// src/obu.rs (headers), src/coder/superblock.rs / src/coder/tile_encode.rs (blocks) and
// src/msac_enc.rs / src/odec.rs (symbols and its EntropyTokens stream).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mtgpu.h"
#include "av1_cdf_tables.h"

namespace mtlib {
int set_error(int code, const std::string& what);
}

namespace {

// ---------------------------------------------------------------- CDF store and tokens
// Every CDF the coder uses, copied once into one array; a token names a CDF by its offset.
struct CdfStore {
    std::vector<uint16_t> v;
    uint32_t add(const uint16_t* p, size_t n) {
        const uint32_t off = (uint32_t)v.size();
        v.insert(v.end(), p, p + n);
        return off;
    }
};

// Base offsets of each table in the store.
struct Tabs {
    CdfStore st;
    uint32_t skip_tx[5], base_tok[5][2], br_tok[4][2], eob_base[5][2], eob_hi[5][2], dc_sign;
    uint32_t eob16[2], eob64[2], eob256[2], eob1024[2];
    uint32_t kf_y, uv_cfl, uv_nocfl, txtp1_4, txtp1_8, txtp2_16, angle, fi_use, fi_mode, part_split, part8,
        cfl_sign, cfl_alpha, skip;
    Tabs() {
        const uint16_t* skips[5] = {&MT_SKIP_TX4[0][0], &MT_SKIP_TX8[0][0], &MT_SKIP_TX16[0][0], &MT_SKIP_TX32[0][0],
                                    &MT_SKIP_TX64[0][0]};
        for (int t = 0; t < 5; ++t) skip_tx[t] = st.add(skips[t], 4 * 13);
        base_tok[0][0] = st.add(&MT_BASE_TOK_TX4_LUMA_Q[0][0][0], 4 * 41 * 3);
        base_tok[0][1] = st.add(&MT_BASE_TOK_TX4_CHROMA_Q[0][0][0], 4 * 41 * 3);
        base_tok[1][0] = st.add(&MT_BASE_TOK_TX8_LUMA_Q[0][0][0], 4 * 41 * 3);
        base_tok[1][1] = st.add(&MT_BASE_TOK_TX8_CHROMA_Q[0][0][0], 4 * 41 * 3);
        base_tok[2][0] = st.add(&MT_BASE_TOK_TX16_LUMA_Q[0][0][0], 4 * 42 * 3);
        base_tok[2][1] = st.add(&MT_BASE_TOK_TX16_CHROMA_Q[0][0][0], 4 * 42 * 3);
        base_tok[3][0] = st.add(&MT_BASE_TOK_TX32_LUMA_Q[0][0][0], 4 * 42 * 3);
        base_tok[3][1] = st.add(&MT_BASE_TOK_TX32_CHROMA_Q[0][0][0], 4 * 42 * 3);
        base_tok[4][0] = st.add(&MT_BASE_TOK_TX64_LUMA_Q[0][0][0], 4 * 42 * 3);
        base_tok[4][1] = base_tok[4][0];
        br_tok[0][0] = st.add(&MT_BR_TOK_TX4_LUMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[0][1] = st.add(&MT_BR_TOK_TX4_CHROMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[1][0] = st.add(&MT_BR_TOK_TX8_LUMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[1][1] = st.add(&MT_BR_TOK_TX8_CHROMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[2][0] = st.add(&MT_BR_TOK_TX16_LUMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[2][1] = st.add(&MT_BR_TOK_TX16_CHROMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[3][0] = st.add(&MT_BR_TOK_TX32_LUMA_Q[0][0][0], 4 * 21 * 3);
        br_tok[3][1] = st.add(&MT_BR_TOK_TX32_CHROMA_Q[0][0][0], 4 * 21 * 3);
        eob_base[0][0] = st.add(&MT_EOB_BASE_TX4_LUMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[0][1] = st.add(&MT_EOB_BASE_TX4_CHROMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[1][0] = st.add(&MT_EOB_BASE_TX8_LUMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[1][1] = st.add(&MT_EOB_BASE_TX8_CHROMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[2][0] = st.add(&MT_EOB_BASE_TX16_LUMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[2][1] = st.add(&MT_EOB_BASE_TX16_CHROMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[3][0] = st.add(&MT_EOB_BASE_TX32_LUMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[3][1] = st.add(&MT_EOB_BASE_TX32_CHROMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[4][0] = st.add(&MT_EOB_BASE_TX64_LUMA_Q[0][0][0], 4 * 4 * 2);
        eob_base[4][1] = eob_base[4][0];
        eob_hi[0][0] = st.add(&MT_EOB_HI_TX4_LUMA[0][0], 4 * 11);
        eob_hi[0][1] = st.add(&MT_EOB_HI_TX4_CHROMA[0][0], 4 * 11);
        eob_hi[1][0] = st.add(&MT_EOB_HI_TX8_LUMA[0][0], 4 * 11);
        eob_hi[1][1] = st.add(&MT_EOB_HI_TX8_CHROMA[0][0], 4 * 11);
        eob_hi[2][0] = st.add(&MT_EOB_HI_TX16_LUMA[0][0], 4 * 11);
        eob_hi[2][1] = st.add(&MT_EOB_HI_TX16_CHROMA[0][0], 4 * 11);
        eob_hi[3][0] = st.add(&MT_EOB_HI_TX32_LUMA[0][0], 4 * 11);
        eob_hi[3][1] = st.add(&MT_EOB_HI_TX32_CHROMA[0][0], 4 * 11);
        eob_hi[4][0] = st.add(&MT_EOB_HI_TX64_LUMA[0][0], 4 * 11);
        eob_hi[4][1] = eob_hi[4][0];
        dc_sign = st.add(&MT_DC_SIGN_Q[0][0][0], 4 * 2 * 3);
        eob16[0] = st.add(&MT_EOB_BIN_16_LUMA[0][0], 4 * 4);
        eob16[1] = st.add(&MT_EOB_BIN_16_CHROMA[0][0], 4 * 4);
        eob64[0] = st.add(&MT_EOB_BIN_64_LUMA[0][0], 4 * 6);
        eob64[1] = st.add(&MT_EOB_BIN_64_CHROMA[0][0], 4 * 6);
        eob256[0] = st.add(&MT_EOB_BIN_256_LUMA[0][0], 4 * 8);
        eob256[1] = st.add(&MT_EOB_BIN_256_CHROMA[0][0], 4 * 8);
        eob1024[0] = st.add(&MT_EOB_BIN_1024_LUMA[0][0], 4 * 10);
        eob1024[1] = st.add(&MT_EOB_BIN_1024_CHROMA[0][0], 4 * 10);
        kf_y = st.add(&MT_KF_Y_MODE_CDF[0][0][0], 5 * 5 * 12);
        uv_cfl = st.add(&MT_UV_MODE_CFL_CDF[0][0], 13 * 13);
        uv_nocfl = st.add(&MT_UV_MODE_NOCFL_CDF[0][0], 13 * 12);
        txtp1_4 = st.add(&MT_TXTP_INTRA1_TX4[0][0], 13 * 6);
        txtp1_8 = st.add(&MT_TXTP_INTRA1_TX8[0][0], 13 * 6);
        txtp2_16 = st.add(&MT_TXTP_INTRA2_TX16[0][0], 13 * 4);
        angle = st.add(&MT_ANGLE_DELTA_CDF[0][0], 8 * 6);
        fi_use = st.add(&MT_FILTER_INTRA_CDF[0], 22);
        fi_mode = st.add(&MT_FILTER_INTRA_MODE_CDF[0], 4);
        part_split = st.add(&MT_PART_SPLIT_CDF[0][0][0], 3 * 4 * 9);
        part8 = st.add(&MT_PART_BL8_CDF[0][0], 4 * 3);
        cfl_sign = st.add(&MT_CFL_SIGN_CDF[0], 7);
        cfl_alpha = st.add(&MT_CFL_ALPHA_CDF[0][0], 6 * 15);
        skip = st.add(&MT_SKIP_CDF[0], 3);
    }
};
const Tabs& tabs() {
    static Tabs t;
    return t;
}

// Token: bit 31 clear: a symbol of a stored CDF: offset << 8 | (nsym - 1) << 4 | symbol;
// bit 31 set: a bool with an explicit probability: 1 << 31 | cdf0 << 1 | value, cdf0 =
// 32768 * P(value = 0) (literal bits: cdf0 = 16384).
inline uint32_t tok_sym(uint32_t off, int nsym, int s) {
    return off << 8 | (uint32_t)(nsym - 1) << 4 | (uint32_t)s;
}
inline uint32_t tok_bool(int cdf0, int v) { return 1u << 31 | (uint32_t)cdf0 << 1 | (uint32_t)(v & 1); }

// ---------------------------------------------------------------- the range coder
// AV1 symbol coder, the inverse of the specification's decoding process (8.2.6): for symbol
// s of an N-symbol CDF, the decoder's threshold of symbol k is
//   cur(k) = ((rng >> 8) * (f(k) >> 6) >> 1) + 4 * (N - k - 1),  f(k) = 32768 - cdf[k],
// and symbol s owns [cur(s), cur(s - 1)) of the decoder's (complemented) value, so the coder
// adds rng - cur(s - 1) to the low end and keeps cur(s - 1) - cur(s) as the range.  The low
// end is kept as the bytes already written (which a carry may still increment) plus a
// window of the newest bits.
struct RangeEnc {
    std::vector<uint8_t> out;
    uint64_t win = 0;  // the newest nb bits of low
    int nb = 15;       // bits in win; the low 16 are the active window
    uint32_t rng = 0x8000;

    void carry() {
        size_t i = out.size();
        while (i > 0 && out[i - 1] == 0xFF) out[--i] = 0;
        if (i > 0) ++out[i - 1];
    }
    void add_low(uint32_t x) {
        win += x;
        if (win >> nb) {  // a carry into the bytes already written
            win &= (1ull << nb) - 1;
            carry();
        }
    }
    void renorm() {
        const int d = 15 - (31 - __builtin_clz(rng));
        rng <<= d;
        win <<= d;
        nb += d;
        while (nb >= 24) {
            out.push_back((uint8_t)(win >> (nb - 8)));
            nb -= 8;
            win &= (1ull << nb) - 1;
        }
    }
    // symbol s of an N-symbol CDF given by its N - 1 cumulative values
    void symbol(const uint16_t* cdf, int N, int s) {
        const uint32_t r8 = rng >> 8;
        auto cur = [&](int k) -> uint32_t {
            if (k < 0) return rng;
            const uint32_t f = 32768u - (k < N - 1 ? cdf[k] : 32768u);
            return ((r8 * (f >> 6)) >> 1) + 4u * (uint32_t)(N - k - 1);
        };
        const uint32_t hi = cur(s - 1), lo = cur(s);
        add_low(rng - hi);
        rng = hi - lo;
        renorm();
    }
    void boolean(int cdf0, int v) {
        const uint16_t c = (uint16_t)cdf0;
        symbol(&c, 2, v);
    }
    // The end of the tile: a value inside the final interval whose bits below position 14
    // are zero and whose bit 14 is one (the trailing one bit the specification's exit
    // process expects, followed by zero padding), written out to the byte holding bit 14.
    void finish() {
        const uint64_t m = 0x3FFF;
        uint64_t e = ((win + m) & ~m) | (m + 1);
        if (e >> nb) {
            e &= (1ull << nb) - 1;
            carry();
        }
        win = e;
        int keep = nb;  // bits of e from the top down to bit 14 (nb >= 15 always)
        while (keep > 14) {
            out.push_back((uint8_t)(win >> (keep - 8)));
            keep -= 8;
            win &= (1ull << keep) - 1;
        }
    }
};

// ---------------------------------------------------------------- bit writer (headers)
struct BitWriter {
    std::vector<uint8_t> b;
    int nbits = 0;
    void put(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; --i) {
            if ((nbits & 7) == 0) b.push_back(0);
            if ((v >> i) & 1) b.back() |= (uint8_t)(0x80 >> (nbits & 7));
            ++nbits;
        }
    }
    void flag(bool f) { put(f ? 1 : 0, 1); }
    void su(int v, int n) { put((uint32_t)v & ((1u << n) - 1), n); }
    void trailing() {  // trailing_bits(): a one and zeros to the byte boundary
        put(1, 1);
        while (nbits & 7) put(0, 1);
    }
    void byte_align() {
        while (nbits & 7) put(0, 1);
    }
};

void leb128(std::vector<uint8_t>& o, size_t v) {
    do {
        uint8_t c = v & 0x7F;
        v >>= 7;
        if (v) c |= 0x80;
        o.push_back(c);
    } while (v);
}
void obu(std::vector<uint8_t>& o, int type, const std::vector<uint8_t>& payload) {
    o.push_back((uint8_t)(type << 3 | 2));  // obu_has_size_field
    leb128(o, payload.size());
    o.insert(o.end(), payload.begin(), payload.end());
}

// ---------------------------------------------------------------- the frame model
constexpr int kNodes = MT_RDO_NODES;
inline int level_base(int L) { return L == 0 ? 0 : L == 1 ? 1 : L == 2 ? 5 : L == 3 ? 21 : 85; }
inline int node_id(int L, int ix, int iy) { return level_base(L) + iy * (1 << L) + ix; }

// the RDO's mode ids (mt_rdo_node.mode): 0..12 the intra modes, 13..60 the directional
// modes 1..8 with angle deltas -3..-1, +1..+3, 61..65 filter intra modes 0..4
inline int id_ymode(int id) { return id <= 12 ? id : id <= 60 ? 1 + (id - 13) / 6 : 0; }
inline int id_delta(int id) {
    if (id < 13 || id > 60) return 0;
    const int j = (id - 13) % 6;
    return j < 3 ? j - 3 : j - 2;
}
inline int id_filter(int id) { return id >= 61 ? id - 61 : -1; }

// AV1 BLOCK_SIZE of a square of side s (4..64)
inline int bsize_sq(int s) { return s == 4 ? 0 : s == 8 ? 3 : s == 16 ? 6 : s == 32 ? 9 : 12; }
constexpr int kIntraModeCtx[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};
constexpr int kFilterDir[5] = {0, 1, 2, 6, 0};  // Filter_Intra_Mode_To_Intra_Dir
// Mode_To_Txfm of the uv modes (13 = UV_CFL_PRED): 0 DCT_DCT, 1 ADST_DCT, 2 DCT_ADST, 3 ADST_ADST
constexpr int kModeToTxfm[14] = {0, 1, 2, 0, 3, 1, 2, 2, 1, 3, 1, 2, 3, 0};
// Coeff_Base_Ctx_Offset of the square sizes and Coeff_Base_Pos... (2-D class only)
constexpr int kBaseCtxOff[5][5] = {{0, 1, 6, 6, 21}, {1, 6, 6, 21, 21}, {6, 6, 21, 21, 21},
                                   {6, 21, 21, 21, 21}, {21, 21, 21, 21, 21}};

struct Frame {
    const mt_av1_frame* f;
    int W, H, mi_cols, mi_rows, sb_cols, sb_rows, ss, planes, qctx;
    // per 4x4 of luma (mi units): block size (side), y mode, skip
    std::vector<uint8_t> mi_side, mi_ymode, mi_skip;
    // per plane, per 4x4 of that plane: the covering transform block's level and DC context
    std::vector<uint8_t> cul[3], dcc[3];
    int pw4[3], ph4[3];
    // the scans of the square sizes in the specification's row-major layout
    std::vector<uint16_t> scan[4];

    const mt_rdo_node& node(int sb, int id) const { return f->nodes[(size_t)sb * kNodes + id]; }
    int32_t level(int plane, int x, int y) const {  // level of the coefficient plane at (x, y)
        if (plane == 0) return f->coef_y[(size_t)y * f->coef_stride + x];
        return (plane == 1 ? f->coef_u : f->coef_v)[(size_t)y * f->coef_uv_stride + x];
    }
};

// One transform block: its levels in the specification's layout (row-major, tw x tw), eob,
// and the contexts it leaves behind.
struct TxBlk {
    int plane, x, y, n;  // position and side in plane samples (n <= 64; 64 codes 32x32)
    int tw;              // coded width (min(n, 32))
    int eob;             // 0: all zero
    int cul, dccat;
    std::vector<int32_t> q;       // tw * tw, row-major (row = vertical frequency)
};

void tx_levels(const Frame& F, TxBlk& t) {
    const int tw = t.tw = std::min(t.n, 32);
    t.q.assign((size_t)tw * tw, 0);
    for (int fy = 0; fy < tw; ++fy)
        for (int fx = 0; fx < tw; ++fx) t.q[(size_t)fy * tw + fx] = F.level(t.plane, t.x + fx, t.y + fy);
    const int lg = __builtin_ctz((unsigned)tw) - 2;
    const std::vector<uint16_t>& sc = F.scan[lg];
    int eob = 0;
    for (int c = 0; c < tw * tw; ++c)
        if (t.q[sc[c]] != 0) eob = c + 1;
    t.eob = eob;
    int cul = 0;
    for (int c = 0; c < eob; ++c) cul += std::abs(t.q[sc[c]]) & 0xFFFFF;
    t.cul = std::min(63, cul);
    t.dccat = t.q[0] == 0 ? 0 : t.q[0] < 0 ? 1 : 2;
}

// The transform blocks of the coded block at luma (x, y), side s, that carry plane p.
void block_txs(const Frame& F, int x, int y, int s, int p, std::vector<TxBlk>& out) {
    if (p == 0) {
        out.push_back(TxBlk{0, x, y, s, 0, 0, 0, 0, {}});
        return;
    }
    if (F.ss) {  // 4:2:0: the chroma of an 8x8 area (4x4 luma blocks) or of the block
        const int cs = s == 4 ? 4 : s / 2;
        const int cx = s == 4 ? (x >> 3) * 4 : x / 2, cy = s == 4 ? (y >> 3) * 4 : y / 2;
        out.push_back(TxBlk{p, cx, cy, cs, 0, 0, 0, 0, {}});
        return;
    }
    const int ts = std::min(s, 32);  // 4:4:4: chroma transforms are at most 32x32
    for (int ty = 0; ty < s; ty += ts)
        for (int tx = 0; tx < s; tx += ts) out.push_back(TxBlk{p, x + tx, y + ty, ts, 0, 0, 0, 0, {}});
}

// Does the block at luma mi (r, c) of side s code chroma (HasChroma)?
bool has_chroma(const Frame& F, int r, int c, int s) {
    if (F.planes == 1) return false;
    if (F.ss && s == 4 && (!(r & 1) || !(c & 1))) return false;
    return true;
}

// ---------------------------------------------------------------- per-superblock tokens
struct SbCoder {
    const Frame& F;
    const Tabs& T;
    std::vector<uint32_t>& tok;
    int sb, r0, c0;  // superblock index and origin (mi)
    bool cdef_done = false;  // the SB's cdef_idx is coded

    void sym(uint32_t off, int nsym, int s) { tok.push_back(tok_sym(off, nsym, s)); }
    void lit(int v) { tok.push_back(tok_bool(16384, v)); }

    int above_mi_side(int r, int c) const { return F.mi_side[(size_t)(r - 1) * F.mi_cols + c]; }
    int left_mi_side(int r, int c) const { return F.mi_side[(size_t)r * F.mi_cols + c - 1]; }

    // partition symbols (5.11.4) and the recursion over the quadtree
    void partition(int r, int c, int L) {
        if (r >= F.mi_rows || c >= F.mi_cols) return;
        const int s = 64 >> L, n4 = s / 4, half = n4 / 2;
        const int ix = (c - c0) / n4, iy = (r - r0) / n4;
        const mt_rdo_node& nd = F.node(sb, node_id(L, ix, iy));
        if (s == 4) {
            block(r, c, 4, sb, node_id(L, ix, iy));
            return;
        }
        const bool has_rows = r + half < F.mi_rows, has_cols = c + half < F.mi_cols;
        const bool split = nd.split == 1;
        // partition context: neighbours narrower / shorter than this node
        const int bsl = L == 0 ? 4 : L == 1 ? 3 : L == 2 ? 2 : 1;  // Mi_Width_Log2
        const int above = r > 0 && __builtin_ctz((unsigned)above_mi_side(r, c) / 4) < bsl;
        const int left = c > 0 && __builtin_ctz((unsigned)left_mi_side(r, c) / 4) < bsl;
        const int ctx = left * 2 + above;
        const uint16_t* pc = s == 8 ? &T.st.v[T.part8 + ctx * 3] : &T.st.v[T.part_split + (L * 4 + ctx) * 9];
        const int nsym = s == 8 ? 4 : 10;
        if (has_rows && has_cols) {
            sym(s == 8 ? T.part8 + ctx * 3 : T.part_split + (L * 4 + ctx) * 9, nsym, split ? 3 : 0);
        } else if (s >= 16 && (has_cols || has_rows)) {
            // split_or_horz (bottom edge) / split_or_vert (right edge): a bool whose
            // probability of 1 gathers the partitions that split the node that way
            auto P = [&](int k) -> int {  // probability of partition k (x 32768)
                const int hi = k < nsym - 1 ? pc[k] : 32768, lo = k > 0 ? pc[k - 1] : 0;
                return hi - lo;
            };
            int psum;
            if (has_cols)  // split_or_horz: VERT, SPLIT, HORZ_A, VERT_A, VERT_B, VERT_4
                psum = P(2) + P(3) + P(4) + P(6) + P(7) + P(9);
            else           // split_or_vert: HORZ, SPLIT, HORZ_A, HORZ_B, VERT_A, HORZ_4
                psum = P(1) + P(3) + P(4) + P(5) + P(6) + P(8);
            tok.push_back(tok_bool(32768 - psum, 1));
        }
        if (split || !(has_rows && has_cols)) {
            for (int q = 0; q < 4; ++q) partition(r + (q >> 1) * half, c + (q & 1) * half, L + 1);
        } else {
            block(r, c, s, sb, node_id(L, ix, iy));
        }
    }

    // one coded block (5.11.5 - 5.11.7 intra frame mode info, 5.11.34 residual)
    void block(int r, int c, int s, int sbi, int id) {
        const mt_rdo_node& nd = F.node(sbi, id);
        const int x = c * 4, y = r * 4;
        const bool hc = has_chroma(F, r, c, s);
        // chroma decisions sit at the node, or at 4:2:0 at the 8x8 node of split 8x8s
        const mt_rdo_node& un = (F.ss && s == 4) ? F.node(sbi, 21 + ((id - 85) / 16 / 2) * 8 + ((id - 85) % 16) / 2) : nd;
        std::vector<TxBlk> tx;
        for (int p = 0; p < (hc ? 3 : 1); ++p) block_txs(F, x, y, s, p, tx);
        bool any = false;
        for (TxBlk& t : tx) {
            tx_levels(F, t);
            any = any || t.eob > 0;
        }
        const int skip = any ? 0 : 1;
        // skip
        const int sctx = (r > 0 ? F.mi_skip[(size_t)(r - 1) * F.mi_cols + c] : 0) +
                         (c > 0 ? F.mi_skip[(size_t)r * F.mi_cols + c - 1] : 0);
        sym(T.skip + sctx, 2, skip);
        // read_cdef: the SB's strength entry at its first non-skip block
        if (F.f->cdef_on && !skip && !cdef_done) {
            const int e = F.f->cdef_bits ? F.f->cdef_grid[sb] : 0;
            for (int b = F.f->cdef_bits - 1; b >= 0; --b) lit((e >> b) & 1);
            cdef_done = true;
        }
        // y mode, angle, chroma, filter intra
        const int ym = id_ymode(nd.mode), fi = id_filter(nd.mode);
        const int am = r > 0 ? F.mi_ymode[(size_t)(r - 1) * F.mi_cols + c] : 0;
        const int lm = c > 0 ? F.mi_ymode[(size_t)r * F.mi_cols + c - 1] : 0;
        sym(T.kf_y + (kIntraModeCtx[am] * 5 + kIntraModeCtx[lm]) * 12, 13, ym);
        if (s >= 8 && ym >= 1 && ym <= 8) sym(T.angle + (ym - 1) * 6, 7, id_delta(nd.mode) + 3);
        int uvm = 0;
        if (hc) {
            uvm = un.uv_mode > 13 ? 0 : un.uv_mode;
            const bool cfl_ok = s <= 32;
            if (cfl_ok) sym(T.uv_cfl + ym * 13, 14, uvm);
            else sym(T.uv_nocfl + ym * 12, 13, uvm);
            if (uvm == 13) {
                const int au = un.cfl_alpha_u, av = un.cfl_alpha_v;
                const int su = au == 0 ? 0 : au < 0 ? 1 : 2, sv = av == 0 ? 0 : av < 0 ? 1 : 2;
                sym(T.cfl_sign, 8, su * 3 + sv - 1);
                if (su) sym(T.cfl_alpha + ((su - 1) * 3 + sv) * 15, 16, std::abs(au) - 1);
                if (sv) sym(T.cfl_alpha + ((sv - 1) * 3 + su) * 15, 16, std::abs(av) - 1);
            }
            if (s >= 8 && uvm >= 1 && uvm <= 8) sym(T.angle + (uvm - 1) * 6, 7, 3);
        }
        if (F.f->filter_intra && ym == 0 && s <= 32) {
            sym(T.fi_use + bsize_sq(s), 2, fi >= 0 ? 1 : 0);
            if (fi >= 0) sym(T.fi_mode, 5, fi);
        }
        // residual
        if (skip) return;
        const int intra_dir = fi >= 0 ? kFilterDir[fi] : ym;
        for (TxBlk& t : tx) coeffs(t, intra_dir, t.plane ? kModeToTxfm[uvm] : 0);
    }

    // the coefficients of one transform block (5.11.39)
    void coeffs(const TxBlk& t, int intra_dir, int uv_type) {
        const int p = t.plane, ptype = p > 0;
        const int n = t.n, lgn = __builtin_ctz((unsigned)n), w4 = n / 4;
        const int x4 = t.x / 4, y4 = t.y / 4;
        const int tsc = lgn - 2;  // txSzCtx of a square size: 0 (4x4) .. 4 (64x64)
        const int maxx4 = F.pw4[p], maxy4 = F.ph4[p];
        // all_zero: luma blocks are as large as their transform (context 0); chroma reads the
        // level and DC contexts above and to the left
        int zctx;
        if (p == 0) {
            zctx = 0;
        } else {
            int a = 0, l = 0;
            for (int i = 0; i < w4; ++i)
                if (x4 + i < maxx4 && y4 > 0) a |= F.cul[p][(size_t)(y4 - 1) * F.pw4[p] + x4 + i] | F.dcc[p][(size_t)(y4 - 1) * F.pw4[p] + x4 + i];
            for (int i = 0; i < w4; ++i)
                if (y4 + i < maxy4 && x4 > 0)
                    l |= F.cul[p][(size_t)(y4 + i) * F.pw4[p] + x4 - 1] | F.dcc[p][(size_t)(y4 + i) * F.pw4[p] + x4 - 1];
            zctx = 7 + (a != 0) + (l != 0);
            // the chroma block is larger than its transform (4:4:4 64x64: four 32x32)
            if (!F.ss && n == 32 && chroma_block_side(t) == 64) zctx += 3;
        }
        sym(T.skip_tx[tsc] + F.qctx * 13 + zctx, 2, t.eob == 0 ? 1 : 0);
        if (t.eob == 0) return;
        // the luma transform type (DCT_DCT: index 1 of the intra sets) below 32x32
        if (p == 0 && n <= 16) {
            if (n == 4) sym(T.txtp1_4 + intra_dir * 6, 7, 1);
            else if (n == 8) sym(T.txtp1_8 + intra_dir * 6, 7, 1);
            else sym(T.txtp2_16 + intra_dir * 4, 5, 1);
        }
        (void)uv_type;
        const int tw = t.tw, bwl = __builtin_ctz((unsigned)tw);
        const std::vector<uint16_t>& sc = F.scan[bwl - 2];
        const int eob = t.eob;
        // eob_pt: eob in [start(pt), start(pt + 1)) with start = 1, 2, 3, 5, 9, 17, ...
        int eob_pt = 1;
        {
            int pt = 1;
            while (true) {
                const int lo = pt < 2 ? pt : (1 << (pt - 2)) + 1;
                const int hi = pt + 1 < 2 ? pt + 1 : (1 << (pt - 1)) + 1;
                if (eob >= lo && eob < hi) break;
                ++pt;
            }
            eob_pt = pt;
        }
        const int area = tw * tw;
        if (area == 16) sym(T.eob16[ptype] + F.qctx * 4, 5, eob_pt - 1);
        else if (area == 64) sym(T.eob64[ptype] + F.qctx * 6, 7, eob_pt - 1);
        else if (area == 256) sym(T.eob256[ptype] + F.qctx * 8, 9, eob_pt - 1);
        else sym(T.eob1024[ptype] + F.qctx * 10, 11, eob_pt - 1);
        if (eob_pt >= 3) {
            const int start = (1 << (eob_pt - 2)) + 1, off = eob - start;
            const int nbits = eob_pt - 2;  // off < 2^nbits
            const int hi = (off >> (nbits - 1)) & 1;
            // eob_extra's context is eobPt - 3; the tables keep it at index eobPt - 1
            sym(T.eob_hi[tsc][ptype] + F.qctx * 11 + (eob_pt - 1), 2, hi);
            for (int i = nbits - 2; i >= 0; --i) lit((off >> i) & 1);
        }
        // base levels and ranges, last to first
        const int brs = std::min(tsc, 3);
        for (int c = eob - 1; c >= 0; --c) {
            const int pos = sc[c], row = pos >> bwl, col = pos & (tw - 1);
            const int lv = std::abs(t.q[pos]);
            const int base = std::min(lv, 3);
            if (c == eob - 1) {
                const int ectx = c == 0 ? 0 : c <= area / 8 ? 1 : c <= area / 4 ? 2 : 3;
                sym(T.eob_base[tsc][ptype] + (F.qctx * 4 + ectx) * 2, 3, base - 1);
            } else {
                int mag = 0;
                static const int off5[5][2] = {{0, 1}, {1, 0}, {1, 1}, {0, 2}, {2, 0}};
                for (const auto& o : off5) {
                    const int rr = row + o[0], cc = col + o[1];
                    if (rr < tw && cc < tw) mag += std::min(std::abs(t.q[(size_t)rr * tw + cc]), 3);
                }
                const int bctx = (row == 0 && col == 0) ? 0
                                                        : std::min((mag + 1) >> 1, 4) + kBaseCtxOff[std::min(row, 4)][std::min(col, 4)];
                const int nctx = tsc <= 1 ? 41 : 42;
                sym(T.base_tok[tsc][ptype] + (F.qctx * nctx + bctx) * 3, 4, base);
            }
            if (lv > 2) {
                int mag = 0;
                static const int off3[3][2] = {{0, 1}, {1, 0}, {1, 1}};
                for (const auto& o : off3) {
                    const int rr = row + o[0], cc = col + o[1];
                    if (rr < tw && cc < tw) mag += std::min(std::abs(t.q[(size_t)rr * tw + cc]), 15);
                }
                mag = std::min((mag + 1) >> 1, 6);
                const int rctx = pos == 0 ? mag : (row < 2 && col < 2) ? mag + 7 : mag + 14;
                int rem = std::min(lv, 15) - 3;
                for (int k = 0; k < 4; ++k) {
                    const int v = std::min(rem, 3);
                    sym(T.br_tok[brs][ptype] + (F.qctx * 21 + rctx) * 3, 4, v);
                    rem -= v;
                    if (v < 3) break;
                }
            }
        }
        // signs and Golomb remainders, first to last
        for (int c = 0; c < eob; ++c) {
            const int pos = sc[c];
            const int32_t v = t.q[pos];
            if (v == 0) continue;
            if (c == 0) {
                int dc = 0;
                for (int i = 0; i < w4; ++i) {
                    if (x4 + i < F.pw4[p] && y4 > 0) {
                        const int k = F.dcc[p][(size_t)(y4 - 1) * F.pw4[p] + x4 + i];
                        dc += k == 1 ? -1 : k == 2 ? 1 : 0;
                    }
                    if (y4 + i < F.ph4[p] && x4 > 0) {
                        const int k = F.dcc[p][(size_t)(y4 + i) * F.pw4[p] + x4 - 1];
                        dc += k == 1 ? -1 : k == 2 ? 1 : 0;
                    }
                }
                const int dctx = dc < 0 ? 1 : dc > 0 ? 2 : 0;
                sym(T.dc_sign + (F.qctx * 2 + ptype) * 3 + dctx, 2, v < 0 ? 1 : 0);
            } else {
                lit(v < 0 ? 1 : 0);
            }
            const int a = std::abs(v);
            if (a > 14) {
                const uint32_t g = (uint32_t)(a - 14);  // >= 1
                const int len = 32 - __builtin_clz(g);
                for (int i = 0; i < len - 1; ++i) lit(0);
                lit(1);
                for (int i = len - 2; i >= 0; --i) lit((g >> i) & 1);
            }
        }
    }
    int chroma_block_side(const TxBlk& t) const {
        // the luma block holding chroma position (t.x, t.y) at 4:4:4
        return F.mi_side[(size_t)(t.y / 4) * F.mi_cols + t.x / 4];
    }
};

// ---------------------------------------------------------------- frame setup
int build_frame(const mt_av1_frame* f, Frame& F) {
    if (!f || !f->nodes || !f->coef_y || f->width <= 0 || f->height <= 0 || (f->width & 7) || (f->height & 7))
        return mtlib::set_error(MT_ERR_ARG, "mt_av1: bad frame (sizes must be multiples of 8)");
    if (f->bd != 8 && f->bd != 10 && f->bd != 12) return mtlib::set_error(MT_ERR_ARG, "mt_av1: bad bit depth");
    if (f->chroma > MT_RDO_CHROMA_444) return mtlib::set_error(MT_ERR_ARG, "mt_av1: bad chroma format");
    if (f->chroma && (!f->coef_u || !f->coef_v))
        return mtlib::set_error(MT_ERR_ARG, "mt_av1: chroma needs the chroma level planes");
    if (f->qidx == 0) return mtlib::set_error(MT_ERR_ARG, "mt_av1: base_q_idx 0 (lossless) is not coded here");
    F.f = f;
    F.W = f->width;
    F.H = f->height;
    F.mi_cols = F.W / 4;
    F.mi_rows = F.H / 4;
    F.sb_cols = (F.W + 63) / 64;
    F.sb_rows = (F.H + 63) / 64;
    if (F.sb_cols > 64 || F.sb_cols * F.sb_rows > 2304)
        return mtlib::set_error(MT_ERR_UNSUPPORTED,
                                "mt_av1: frames wider than 4096 or above 4096x2304 need several tiles");
    F.ss = f->chroma == MT_RDO_CHROMA_420 ? 1 : 0;
    F.planes = f->chroma ? 3 : 1;
    F.qctx = f->qidx <= 20 ? 0 : f->qidx <= 60 ? 1 : f->qidx <= 120 ? 2 : 3;
    // scans: the reference's (x-major) scans transposed to row-major positions
    const uint32_t* sc[4] = {MT_SCAN_4X4, MT_SCAN_8X8, MT_SCAN_16X16, MT_SCAN_32X32};
    for (int k = 0; k < 4; ++k) {
        const int n = 4 << k;
        F.scan[k].resize((size_t)n * n);
        for (int c = 0; c < n * n; ++c) {
            const int rc = (int)sc[k][c], fx = rc / n, fy = rc % n;
            F.scan[k][c] = (uint16_t)(fy * n + fx);
        }
    }
    return MT_OK;
}

// the decisions' maps: block side, y mode and skip per 4x4, and every transform block's
// contexts
void build_maps(Frame& F) {
    const size_t nmi = (size_t)F.mi_cols * F.mi_rows;
    F.mi_side.assign(nmi, 0);
    F.mi_ymode.assign(nmi, 0);
    F.mi_skip.assign(nmi, 0);
    for (int p = 0; p < F.planes; ++p) {
        const int s = p ? F.ss : 0;
        F.pw4[p] = F.mi_cols >> s;
        F.ph4[p] = F.mi_rows >> s;
        F.cul[p].assign((size_t)F.pw4[p] * F.ph4[p], 0);
        F.dcc[p].assign((size_t)F.pw4[p] * F.ph4[p], 0);
    }
    struct Walk {
        Frame& F;
        int sb;
        void leaf(int r, int c, int s, int id) {
            const mt_rdo_node& nd = F.node(sb, id);
            const int n4 = s / 4;
            std::vector<TxBlk> tx;
            const bool hc = has_chroma(F, r, c, s);
            for (int p = 0; p < (hc ? 3 : 1); ++p) block_txs(F, c * 4, r * 4, s, p, tx);
            bool any = false;
            for (TxBlk& t : tx) {
                tx_levels(F, t);
                any = any || t.eob > 0;
            }
            for (int yy = 0; yy < n4 && r + yy < F.mi_rows; ++yy)
                for (int xx = 0; xx < n4 && c + xx < F.mi_cols; ++xx) {
                    const size_t i = (size_t)(r + yy) * F.mi_cols + c + xx;
                    F.mi_side[i] = (uint8_t)s;
                    F.mi_ymode[i] = (uint8_t)id_ymode(nd.mode);
                    F.mi_skip[i] = any ? 0 : 1;
                }
            for (const TxBlk& t : tx) {
                const int p = t.plane, n4t = t.n / 4;
                for (int yy = 0; yy < n4t && t.y / 4 + yy < F.ph4[p]; ++yy)
                    for (int xx = 0; xx < n4t && t.x / 4 + xx < F.pw4[p]; ++xx) {
                        const size_t i = (size_t)(t.y / 4 + yy) * F.pw4[p] + t.x / 4 + xx;
                        F.cul[p][i] = any ? (uint8_t)t.cul : 0;
                        F.dcc[p][i] = any ? (uint8_t)t.dccat : 0;
                    }
            }
        }
        void node(int r, int c, int L, int r0, int c0) {
            if (r >= F.mi_rows || c >= F.mi_cols) return;
            const int s = 64 >> L, n4 = s / 4, half = n4 / 2;
            const int id = node_id(L, (c - c0) / n4, (r - r0) / n4);
            if (s > 4 && (F.node(sb, id).split == 1 || !(r + half < F.mi_rows && c + half < F.mi_cols))) {
                for (int q = 0; q < 4; ++q) node(r + (q >> 1) * half, c + (q & 1) * half, L + 1, r0, c0);
                return;
            }
            leaf(r, c, s, id);
        }
    };
    for (int sr = 0; sr < F.sb_rows; ++sr)
        for (int sc = 0; sc < F.sb_cols; ++sc) {
            Walk w{F, sr * F.sb_cols + sc};
            w.node(sr * 16, sc * 16, 0, sr * 16, sc * 16);
        }
}

// the luma 8x8 skip map: an 8x8 is skip when every 4x4 (mi) in it is
int skip8_map(const mt_av1_frame* f, uint8_t* out) {
    Frame F;
    int rc = build_frame(f, F);
    if (rc) return rc;
    build_maps(F);
    const int nbx = f->width / 8, nby = f->height / 8;
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            int all = 1;
            for (int r = 2 * by; r < 2 * by + 2; ++r)
                for (int c = 2 * bx; c < 2 * bx + 2; ++c) all &= F.mi_skip[(size_t)r * F.mi_cols + c];
            out[(size_t)by * nbx + bx] = (uint8_t)all;
        }
    return MT_OK;
}

int gen_tokens(const mt_av1_frame* f, Frame& F, std::vector<uint32_t>& all, int nthreads) {
    int rc = build_frame(f, F);
    if (rc) return rc;
    build_maps(F);
    const int nsb = F.sb_cols * F.sb_rows;
    std::vector<std::vector<uint32_t>> per(nsb);
    const Tabs& T = tabs();
    auto work = [&](int k, int step) {
        for (int sb = k; sb < nsb; sb += step) {
            const int sr = sb / F.sb_cols, sc = sb % F.sb_cols;
            SbCoder co{F, T, per[sb], sb, sr * 16, sc * 16};
            co.partition(sr * 16, sc * 16, 0);
        }
    };
    nthreads = std::max(1, std::min(nthreads, nsb));
    if (nthreads == 1) {
        work(0, 1);
    } else {
        std::vector<std::thread> th;
        for (int k = 0; k < nthreads; ++k) th.emplace_back(work, k, nthreads);
        for (auto& t : th) t.join();
    }
    size_t tot = 0;
    for (auto& v : per) tot += v.size();
    all.clear();
    all.reserve(tot);
    for (auto& v : per) all.insert(all.end(), v.begin(), v.end());
    return MT_OK;
}

std::vector<uint8_t> range_code(const std::vector<uint32_t>& toks) {
    const Tabs& T = tabs();
    RangeEnc e;
    e.out.reserve(toks.size() / 2 + 64);
    for (uint32_t t : toks) {
        if (t >> 31) {
            e.boolean((int)((t >> 1) & 0x7FFF), (int)(t & 1));
        } else {
            const uint32_t off = t >> 8;
            e.symbol(&T.st.v[off], (int)((t >> 4) & 15) + 1, (int)(t & 15));
        }
    }
    e.finish();
    return std::move(e.out);
}

// ---------------------------------------------------------------- OBUs
int seq_profile(const mt_av1_frame* f) {
    if (f->bd == 12) return 2;
    return f->chroma == MT_RDO_CHROMA_444 ? 1 : 0;
}

std::vector<uint8_t> sequence_header(const mt_av1_frame* f) {
    BitWriter w;
    const int profile = seq_profile(f);
    w.put((uint32_t)profile, 3);
    w.flag(true);   // still_picture
    w.flag(true);   // reduced_still_picture_header
    w.put(31, 5);   // seq_level_idx[0]: 31, no level constraints
    auto bits_for = [](uint32_t v) { int n = 1; while ((v >> n) != 0) ++n; return n; };
    const int wb = bits_for((uint32_t)f->width - 1), hb = bits_for((uint32_t)f->height - 1);
    w.put((uint32_t)wb - 1, 4);
    w.put((uint32_t)hb - 1, 4);
    w.put((uint32_t)f->width - 1, wb);
    w.put((uint32_t)f->height - 1, hb);
    w.flag(false);                 // use_128x128_superblock
    w.flag(f->filter_intra != 0);  // enable_filter_intra
    w.flag(true);                  // enable_intra_edge_filter
    w.flag(false);                 // enable_superres
    w.flag(f->cdef_on != 0);       // enable_cdef
    w.flag(false);                 // enable_restoration
    // color_config
    const bool hbd = f->bd > 8;
    w.flag(hbd);
    if (profile == 2 && hbd) w.flag(f->bd == 12);
    const bool mono = f->chroma == MT_RDO_LUMA_ONLY;
    if (profile != 1) w.flag(mono);
    const bool desc = f->cicp[0] != 2 || f->cicp[1] != 2 || f->cicp[2] != 2;
    w.flag(desc);
    if (desc) {
        w.put(f->cicp[0], 8);
        w.put(f->cicp[1], 8);
        w.put(f->cicp[2], 8);
    }
    if (mono) {
        w.flag(f->full_range != 0);
    } else if (f->cicp[0] == 1 && f->cicp[1] == 13 && f->cicp[2] == 0) {
        // sRGB primaries and transfer with the identity matrix: full range, 4:4:4, not coded
    } else {
        w.flag(f->full_range != 0);
        if (profile == 2) {  // 12-bit: the subsampling is coded
            w.flag(f->chroma == MT_RDO_CHROMA_420);
            if (f->chroma == MT_RDO_CHROMA_420) w.flag(true);
        }
        if (f->chroma == MT_RDO_CHROMA_420) w.put(0, 2);  // chroma_sample_position: unknown
    }
    if (!mono) w.flag(false);  // separate_uv_delta_q
    w.flag(false);             // film_grain_params_present
    w.trailing();
    return w.b;
}

void frame_header(const mt_av1_frame* f, int sb_cols, int sb_rows, BitWriter& w) {
    // reduced still picture header: KEY_FRAME, shown, no frame size override
    w.flag(true);   // disable_cdf_update
    w.flag(false);  // allow_screen_content_tools
    w.flag(false);  // render_and_frame_size_different
    // tile_info: uniform spacing, the minimum tile counts (one tile)
    w.flag(true);  // uniform_tile_spacing_flag
    auto tile_log2 = [](int blk, int target) { int k = 0; while ((blk << k) < target) ++k; return k; };
    if (tile_log2(1, std::min(sb_cols, 64)) > 0) w.flag(false);  // increment_tile_cols_log2
    if (tile_log2(1, std::min(sb_rows, 64)) > 0) w.flag(false);  // increment_tile_rows_log2
    // quantization_params
    w.put(f->qidx, 8);
    w.flag(true);  // DeltaQYDc coded: the reference's luma DC offset -8
    w.su(-8, 7);
    const bool mono = f->chroma == MT_RDO_LUMA_ONLY;
    if (!mono) {
        w.flag(f->uv_dc_delta != 0);
        if (f->uv_dc_delta) w.su(f->uv_dc_delta, 7);
        w.flag(f->uv_ac_delta != 0);
        if (f->uv_ac_delta) w.su(f->uv_ac_delta, 7);
    }
    const bool qm = f->qm_level < 15 || (!mono && f->uv_qm_level < 15);
    w.flag(qm);
    if (qm) {
        w.put(f->qm_level, 4);
        w.put(mono ? 15 : f->uv_qm_level, 4);  // qm_u (qm_v = qm_u)
    }
    w.flag(false);  // segmentation_enabled
    w.flag(false);  // delta_q_present
    // loop_filter_params
    w.put(f->lf_level[0], 6);
    w.put(f->lf_level[1], 6);
    if (!mono && (f->lf_level[0] || f->lf_level[1])) {
        w.put(f->lf_level[2], 6);
        w.put(f->lf_level[3], 6);
    }
    w.put(f->lf_sharpness, 3);
    w.flag(false);  // loop_filter_delta_enabled
    if (f->cdef_on) {  // cdef_params
        w.put((uint32_t)f->cdef_damping - 3, 2);
        w.put(f->cdef_bits, 2);
        auto sec_code = [](int s) { return s == 4 ? 3u : (uint32_t)s; };
        for (int i = 0; i < (1 << f->cdef_bits); ++i) {
            w.put(f->cdef_y_pri[i], 4);
            w.put(sec_code(f->cdef_y_sec[i]), 2);
            if (!mono) {
                w.put(f->cdef_uv_pri[i], 4);
                w.put(sec_code(f->cdef_uv_sec[i]), 2);
            }
        }
    }
    w.flag(false);  // tx_mode_select: TX_MODE_LARGEST
    w.flag(false);  // reduced_tx_set
}

int check_frame_args(const mt_av1_frame* f) {
    if (!f) return mtlib::set_error(MT_ERR_ARG, "mt_av1: no frame");
    if (f->qm_level > 15 || f->uv_qm_level > 15 || f->lf_sharpness > 7 || f->uv_dc_delta < -64 ||
        f->uv_dc_delta > 63 || f->uv_ac_delta < -64 || f->uv_ac_delta > 63)
        return mtlib::set_error(MT_ERR_ARG, "mt_av1: bad quantizer or filter parameters");
    for (int i = 0; i < 4; ++i)
        if (f->lf_level[i] > 63) return mtlib::set_error(MT_ERR_ARG, "mt_av1: loop filter level above 63");
    if (f->cdef_on) {
        if (f->cdef_damping < 3 || f->cdef_damping > 6 || f->cdef_bits > 3 || (f->cdef_bits && !f->cdef_grid))
            return mtlib::set_error(MT_ERR_ARG, "mt_av1: bad CDEF parameters");
        for (int i = 0; i < (1 << f->cdef_bits); ++i)
            if (f->cdef_y_pri[i] > 15 || f->cdef_uv_pri[i] > 15 || f->cdef_y_sec[i] == 3 || f->cdef_y_sec[i] > 4 ||
                f->cdef_uv_sec[i] == 3 || f->cdef_uv_sec[i] > 4)
                return mtlib::set_error(MT_ERR_ARG, "mt_av1: bad CDEF strength");
    }
    if (f->cicp[2] == 0 && f->chroma != MT_RDO_CHROMA_444)
        return mtlib::set_error(MT_ERR_ARG, "mt_av1: the identity matrix needs 4:4:4");
    return MT_OK;
}

int encode_obus(const mt_av1_frame* f, std::vector<uint8_t>& seq, std::vector<uint8_t>& obus, size_t* ntok) {
    int rc = check_frame_args(f);
    if (rc) return rc;
    Frame F;
    std::vector<uint32_t> toks;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    if ((rc = gen_tokens(f, F, toks, (int)std::min(hw, 16u)))) return rc;
    if (ntok) *ntok = toks.size();
    const std::vector<uint8_t> tile = range_code(toks);
    seq = sequence_header(f);
    BitWriter fh;
    frame_header(f, F.sb_cols, F.sb_rows, fh);
    fh.byte_align();  // byte_alignment() after the frame header; one tile: no tile group header
    std::vector<uint8_t> frame = fh.b;
    frame.insert(frame.end(), tile.begin(), tile.end());
    obus.clear();
    obu(obus, 1, seq);     // OBU_SEQUENCE_HEADER
    obu(obus, 6, frame);   // OBU_FRAME
    return MT_OK;
}

// ---------------------------------------------------------------- AVIF (ISOBMFF)
struct Box {
    std::vector<uint8_t> b;
    void u8(uint32_t v) { b.push_back((uint8_t)v); }
    void u16(uint32_t v) { u8(v >> 8); u8(v); }
    void u32(uint32_t v) { u16(v >> 16); u16(v & 0xFFFF); }
    void fourcc(const char* s) { for (int i = 0; i < 4; ++i) u8((uint8_t)s[i]); }
    void bytes(const std::vector<uint8_t>& v) { b.insert(b.end(), v.begin(), v.end()); }
};
std::vector<uint8_t> box(const char* type, const std::vector<uint8_t>& body) {
    Box o;
    o.u32((uint32_t)(8 + body.size()));
    o.fourcc(type);
    o.bytes(body);
    return o.b;
}
std::vector<uint8_t> fullbox(const char* type, int version, uint32_t flags, const std::vector<uint8_t>& body) {
    Box o;
    o.u8((uint32_t)version);
    o.u8(flags >> 16);
    o.u16(flags & 0xFFFF);
    o.bytes(body);
    return box(type, o.b);
}

std::vector<uint8_t> avif_file(const mt_av1_frame* f, const std::vector<uint8_t>& seq, const std::vector<uint8_t>& obus) {
    Box ftyp;
    ftyp.fourcc("avif");
    ftyp.u32(0);
    ftyp.fourcc("avif");
    ftyp.fourcc("mif1");
    ftyp.fourcc("miaf");
    const std::vector<uint8_t> ftyp_b = box("ftyp", ftyp.b);

    Box hdlr;
    hdlr.u32(0);
    hdlr.fourcc("pict");
    hdlr.u32(0);
    hdlr.u32(0);
    hdlr.u32(0);
    hdlr.u8(0);
    Box pitm;
    pitm.u16(1);
    Box infe;
    infe.u16(1);
    infe.u16(0);
    infe.fourcc("av01");
    infe.u8(0);
    Box iinf;
    iinf.u16(1);
    iinf.bytes(fullbox("infe", 2, 0, infe.b));
    // properties: ispe, av1C, pixi, colr
    Box ispe;
    ispe.u32((uint32_t)f->width);
    ispe.u32((uint32_t)f->height);
    Box av1c;
    const int profile = seq_profile(f);
    av1c.u8(0x81);
    av1c.u8((uint32_t)(profile << 5 | 31));
    const bool mono = f->chroma == MT_RDO_LUMA_ONLY;
    const int ssx = f->chroma == MT_RDO_CHROMA_444 ? 0 : 1, ssy = ssx;
    av1c.u8((uint32_t)(0 << 7 | (f->bd > 8) << 6 | (f->bd == 12) << 5 | mono << 4 | ssx << 3 | ssy << 2 | 0));
    av1c.u8(0);
    std::vector<uint8_t> seq_obu;
    obu(seq_obu, 1, seq);
    av1c.bytes(seq_obu);
    Box pixi;
    const int nch = mono ? 1 : 3;
    pixi.u8((uint32_t)nch);
    for (int i = 0; i < nch; ++i) pixi.u8(f->bd);
    Box colr;
    colr.fourcc("nclx");
    colr.u16(f->cicp[0]);
    colr.u16(f->cicp[1]);
    colr.u16(f->cicp[2]);
    colr.u8(f->full_range ? 0x80 : 0);
    Box ipco;
    ipco.bytes(fullbox("ispe", 0, 0, ispe.b));
    ipco.bytes(box("av1C", av1c.b));
    ipco.bytes(fullbox("pixi", 0, 0, pixi.b));
    ipco.bytes(box("colr", colr.b));
    Box ipma;
    ipma.u32(1);
    ipma.u16(1);
    ipma.u8(4);
    ipma.u8(1);         // ispe
    ipma.u8(0x80 | 2);  // av1C, essential
    ipma.u8(3);         // pixi
    ipma.u8(4);         // colr
    Box iprp;
    iprp.bytes(box("ipco", ipco.b));
    iprp.bytes(fullbox("ipma", 0, 0, ipma.b));

    // iloc with a placeholder offset, fixed once the meta box size is known
    auto meta_with = [&](uint32_t off) {
        Box iloc;
        iloc.u8(0x44);  // offset_size 4, length_size 4
        iloc.u8(0x00);  // base_offset_size 0
        iloc.u16(1);
        iloc.u16(1);    // item 1
        iloc.u16(0);    // data_reference_index
        iloc.u16(1);    // one extent
        iloc.u32(off);
        iloc.u32((uint32_t)obus.size());
        Box meta;
        meta.bytes(fullbox("hdlr", 0, 0, hdlr.b));
        meta.bytes(fullbox("pitm", 0, 0, pitm.b));
        meta.bytes(fullbox("iloc", 0, 0, iloc.b));
        meta.bytes(fullbox("iinf", 0, 0, iinf.b));
        meta.bytes(box("iprp", iprp.b));
        return fullbox("meta", 0, 0, meta.b);
    };
    const size_t meta_size = meta_with(0).size();
    const uint32_t data_off = (uint32_t)(ftyp_b.size() + meta_size + 8);
    std::vector<uint8_t> file = ftyp_b;
    const std::vector<uint8_t> meta = meta_with(data_off);
    file.insert(file.end(), meta.begin(), meta.end());
    const std::vector<uint8_t> mdat = box("mdat", obus);
    file.insert(file.end(), mdat.begin(), mdat.end());
    return file;
}

int copy_out(const std::vector<uint8_t>& v, uint8_t* out, size_t cap, size_t* len) {
    if (len) *len = v.size();
    if (!out) return MT_OK;
    if (cap < v.size()) return mtlib::set_error(MT_ERR_ARG, "mt_av1: output buffer too small (*len holds the size)");
    std::memcpy(out, v.data(), v.size());
    return MT_OK;
}

}  // namespace

extern "C" int mt_av1_tokens(const mt_av1_frame* f, uint32_t* tokens, size_t cap, size_t* n) {
    int rc = check_frame_args(f);
    if (rc) return rc;
    Frame F;
    std::vector<uint32_t> toks;
    if ((rc = gen_tokens(f, F, toks, 16))) return rc;
    if (n) *n = toks.size();
    if (!tokens) return MT_OK;
    if (cap < toks.size()) return mtlib::set_error(MT_ERR_ARG, "mt_av1_tokens: token buffer too small");
    std::memcpy(tokens, toks.data(), toks.size() * sizeof(uint32_t));
    return MT_OK;
}

extern "C" int mt_av1_skip8(const mt_av1_frame* f, uint8_t* skip8) {
    int rc = check_frame_args(f);
    if (rc) return rc;
    if (!skip8) return mtlib::set_error(MT_ERR_ARG, "mt_av1_skip8: no output");
    return skip8_map(f, skip8);
}

extern "C" int mt_av1_range_code(const uint32_t* tokens, size_t n, uint8_t* out, size_t cap, size_t* len) {
    if (n && !tokens) return mtlib::set_error(MT_ERR_ARG, "mt_av1_range_code: no tokens");
    const Tabs& T = tabs();
    for (size_t i = 0; i < n; ++i)
        if (!(tokens[i] >> 31)) {
            const uint32_t off = tokens[i] >> 8;
            const int nsym = (int)((tokens[i] >> 4) & 15) + 1, s = (int)(tokens[i] & 15);
            if (nsym < 2 || s >= nsym || off + (uint32_t)nsym - 1 > T.st.v.size())
                return mtlib::set_error(MT_ERR_ARG, "mt_av1_range_code: token names no stored CDF");
        } else if (((tokens[i] >> 1) & 0x7FFF) == 0) {
            return mtlib::set_error(MT_ERR_ARG, "mt_av1_range_code: bool token with probability 0");
        }
    return copy_out(range_code(std::vector<uint32_t>(tokens, tokens + n)), out, cap, len);
}

extern "C" int mt_av1_cdf(uint32_t offset, int nsym, uint16_t* cdf) {
    const Tabs& T = tabs();
    if (nsym < 2 || nsym > 16 || !cdf || offset + (uint32_t)nsym - 1 > T.st.v.size())
        return mtlib::set_error(MT_ERR_ARG, "mt_av1_cdf: no such CDF");
    std::memcpy(cdf, &T.st.v[offset], (size_t)(nsym - 1) * sizeof(uint16_t));
    return MT_OK;
}

extern "C" int mt_av1_encode(const mt_av1_frame* f, uint8_t* out, size_t cap, size_t* len) {
    std::vector<uint8_t> seq, obus;
    int rc = encode_obus(f, seq, obus, nullptr);
    if (rc) return rc;
    return copy_out(obus, out, cap, len);
}

extern "C" int mt_avif_encode(const mt_av1_frame* f, uint8_t* out, size_t cap, size_t* len) {
    std::vector<uint8_t> seq, obus;
    int rc = encode_obus(f, seq, obus, nullptr);
    if (rc) return rc;
    return copy_out(avif_file(f, seq, obus), out, cap, len);
}
