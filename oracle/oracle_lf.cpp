// Oracle (test infrastructure only; the library never links it): the reference's in-loop
// deblocking filter on the CPU, in the reference's own sequential order, for checking the
// library's parallel deblocking.
//   orc_lf_limits        limits (src/loopfilter.rs:33-44): E / I / H of a level and sharpness
//   orc_loop_filter      loop_filter_scalar (src/loopfilter.rs:106-285): one edge, four lines
//   orc_filter_plane     filter_plane_impl (src/loopfilter.rs:483-655) with the scalar dispatch
//                        (LoopFilterDispatch::scalar, :367-373): per superblock row, its
//                        vertical edges (rows top to bottom, edges left to right), then its
//                        horizontal edges; edges clipped to the visible 4 x 4 grid
//   orc_chroma_deblock_geom  chroma_deblock_geom (src/coder/frame_deblock.rs:91-150) with
//                        cdblk_variant() == 1: sizes from the prediction maps, no edge flags
//   orc_frame_deblock    frame_deblock (src/coder/frame_deblock.rs:206-222)
//   orc_lf_levels        loop_filter_levels / loop_filter_sharpness (src/obu.rs:306-334) at
//                        the shipped tuning (src/tuning.rs:764-768)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "oracle.h"

namespace {

inline int32_t clip3(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : v > hi ? hi : v; }

// Width class of a transform dimension in 4-sample units (luma_cls / chroma_cls,
// src/loopfilter.rs:423-440) and the filter width of an edge between two of them.
int edge_width(int a4, int b4, bool luma) {
    if (luma) {
        const int ca = a4 <= 1 ? 0 : a4 == 2 ? 1 : 2, cb = b4 <= 1 ? 0 : b4 == 2 ? 1 : 2;
        return 4 << (ca < cb ? ca : cb);
    }
    const int ca = a4 >= 2 ? 1 : 0, cb = b4 >= 2 ? 1 : 0;
    return 4 + 2 * (ca < cb ? ca : cb);
}

// One line across the edge.  s[k + 7] is the sample at offset k (-7..6) from the edge along
// the line; only the offsets the width reads are valid on entry.  Writes the filtered samples
// back into s and returns a bit mask (bit k + 7) of the offsets written.
uint32_t filter_line(int32_t* s, int e, int ilim, int hth, int wd, int bd) {
    const int sh = bd - 8;
    const int32_t flat = 1 << sh, mx = (1 << bd) - 1;
    const int32_t lo = -128 * (1 << sh), hi = 128 * (1 << sh) - 1;
    e <<= sh;
    ilim <<= sh;
    hth <<= sh;
    auto P = [&](int k) -> int32_t& { return s[7 - 1 - k]; };  // p0 = offset -1
    auto Q = [&](int k) -> int32_t& { return s[7 + k]; };      // q0 = offset 0
    const int32_t p0 = P(0), p1 = P(1), q0 = Q(0), q1 = Q(1);
    bool fm = std::abs(p1 - p0) <= ilim && std::abs(q1 - q0) <= ilim &&
              2 * std::abs(p0 - q0) + (std::abs(p1 - q1) >> 1) <= e;
    int32_t p2 = 0, q2 = 0, p3 = 0, q3 = 0;
    if (wd > 4) {
        p2 = P(2);
        q2 = Q(2);
        fm = fm && std::abs(p2 - p1) <= ilim && std::abs(q2 - q1) <= ilim;
        if (wd > 6) {
            p3 = P(3);
            q3 = Q(3);
            fm = fm && std::abs(p3 - p2) <= ilim && std::abs(q3 - q2) <= ilim;
        }
    }
    if (!fm) return 0;
    int32_t p4 = 0, p5 = 0, p6 = 0, q4 = 0, q5 = 0, q6 = 0;
    bool outer = false;
    if (wd >= 16) {
        p4 = P(4), p5 = P(5), p6 = P(6), q4 = Q(4), q5 = Q(5), q6 = Q(6);
        outer = std::abs(p6 - p0) <= flat && std::abs(p5 - p0) <= flat && std::abs(p4 - p0) <= flat &&
                std::abs(q4 - q0) <= flat && std::abs(q5 - q0) <= flat && std::abs(q6 - q0) <= flat;
    }
    bool inner = false;
    if (wd >= 6)
        inner = std::abs(p2 - p0) <= flat && std::abs(p1 - p0) <= flat && std::abs(q1 - q0) <= flat &&
                std::abs(q2 - q0) <= flat;
    if (wd >= 8) inner = inner && std::abs(p3 - p0) <= flat && std::abs(q3 - q0) <= flat;

    uint32_t wrote = 0;
    auto setp = [&](int k, int32_t v) { P(k) = v; wrote |= 1u << (6 - k); };
    auto setq = [&](int k, int32_t v) { Q(k) = v; wrote |= 1u << (7 + k); };
    if (wd >= 16 && outer && inner) {
        // 15-tap smoothing of p5..q5 (sums of 16)
        setp(5, (7 * p6 + 2 * p5 + 2 * p4 + p3 + p2 + p1 + p0 + q0 + 8) >> 4);
        setp(4, (5 * p6 + 2 * p5 + 2 * p4 + 2 * p3 + p2 + p1 + p0 + q0 + q1 + 8) >> 4);
        setp(3, (4 * p6 + p5 + 2 * p4 + 2 * p3 + 2 * p2 + p1 + p0 + q0 + q1 + q2 + 8) >> 4);
        setp(2, (3 * p6 + p5 + p4 + 2 * p3 + 2 * p2 + 2 * p1 + p0 + q0 + q1 + q2 + q3 + 8) >> 4);
        setp(1, (2 * p6 + p5 + p4 + p3 + 2 * p2 + 2 * p1 + 2 * p0 + q0 + q1 + q2 + q3 + q4 + 8) >> 4);
        setp(0, (p6 + p5 + p4 + p3 + p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + q2 + q3 + q4 + q5 + 8) >> 4);
        setq(0, (p5 + p4 + p3 + p2 + p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + q3 + q4 + q5 + q6 + 8) >> 4);
        setq(1, (p4 + p3 + p2 + p1 + p0 + 2 * q0 + 2 * q1 + 2 * q2 + q3 + q4 + q5 + 2 * q6 + 8) >> 4);
        setq(2, (p3 + p2 + p1 + p0 + q0 + 2 * q1 + 2 * q2 + 2 * q3 + q4 + q5 + 3 * q6 + 8) >> 4);
        setq(3, (p2 + p1 + p0 + q0 + q1 + 2 * q2 + 2 * q3 + 2 * q4 + q5 + 4 * q6 + 8) >> 4);
        setq(4, (p1 + p0 + q0 + q1 + q2 + 2 * q3 + 2 * q4 + 2 * q5 + 5 * q6 + 8) >> 4);
        setq(5, (p0 + q0 + q1 + q2 + q3 + 2 * q4 + 2 * q5 + 7 * q6 + 8) >> 4);
    } else if (wd >= 8 && inner) {
        setp(2, (3 * p3 + 2 * p2 + p1 + p0 + q0 + 4) >> 3);
        setp(1, (2 * p3 + p2 + 2 * p1 + p0 + q0 + q1 + 4) >> 3);
        setp(0, (p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2 + 4) >> 3);
        setq(0, (p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4) >> 3);
        setq(1, (p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4) >> 3);
        setq(2, (p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4) >> 3);
    } else if (wd == 6 && inner) {
        setp(1, (3 * p2 + 2 * p1 + 2 * p0 + q0 + 4) >> 3);
        setp(0, (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
        setq(0, (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
        setq(1, (p0 + 2 * q0 + 2 * q1 + 3 * q2 + 4) >> 3);
    } else {
        // narrow filter: the step across the edge, clamped to the signed sample range
        const bool hev = std::abs(p1 - p0) > hth || std::abs(q1 - q0) > hth;
        int32_t f = hev ? clip3(p1 - q1, lo, hi) : 0;
        f = clip3(3 * (q0 - p0) + f, lo, hi);
        const int32_t f1 = (f + 4 < hi ? f + 4 : hi) >> 3;
        const int32_t f2 = (f + 3 < hi ? f + 3 : hi) >> 3;
        setp(0, clip3(p0 + f2, 0, mx));
        setq(0, clip3(q0 - f1, 0, mx));
        if (!hev) {
            const int32_t f3 = (f1 + 1) >> 1;
            setp(1, clip3(p1 + f3, 0, mx));
            setq(1, clip3(q1 - f3, 0, mx));
        }
    }
    return wrote;
}

int reach(int wd) { return wd == 16 ? 7 : wd == 8 ? 4 : wd == 6 ? 3 : 2; }

}  // namespace

extern "C" void orc_lf_limits(int level, int sharp, int* e, int* i, int* h) {
    int il = level;
    if (sharp > 0) {
        il >>= (sharp + 3) >> 2;
        if (il > 9 - sharp) il = 9 - sharp;
    }
    if (il < 1) il = 1;
    *e = 2 * (level + 2) + il;
    *i = il;
    *h = level >> 4;
}

extern "C" void orc_loop_filter(uint16_t* px, int64_t base, int e, int ilim, int hth, int64_t sa, int64_t sb,
                                int wd, int bd) {
    const int r = reach(wd);
    for (int line = 0; line < 4; ++line) {
        const int64_t o = base + line * sa;
        int32_t s[14] = {0};
        for (int k = -r; k < r; ++k) s[k + 7] = px[o + k * sb];
        const uint32_t wrote = filter_line(s, e, ilim, hth, wd, bd);
        for (int k = -7; k < 7; ++k)
            if (wrote >> (k + 7) & 1) px[o + k * sb] = (uint16_t)s[k + 7];
    }
}

extern "C" void orc_filter_plane(uint16_t* px, int w, int h, int vis_w, int vis_h, const uint8_t* bw4,
                                 const uint8_t* bh4, const uint8_t* vedge4, const uint8_t* hedge4, int nc4,
                                 int level, int sharp, int is_luma, int sb_rows4, int bd) {
    if (level <= 0) return;
    int e, il, hth;
    orc_lf_limits(level, sharp, &e, &il, &hth);
    const int w4 = (vis_w + 3) / 4, h4 = (vis_h + 3) / 4;
    const bool luma = is_luma != 0;
    for (int top = 0; top < h4; top += sb_rows4) {
        const int bot = top + sb_rows4 < h4 ? top + sb_rows4 : h4;
        // vertical edges of the superblock row
        for (int r4 = top; r4 < bot; ++r4) {
            if (r4 * 4 >= h) continue;
            for (int c4 = 1; c4 < w4; ++c4) {
                if (c4 * 4 >= w) continue;
                const size_t idx = (size_t)r4 * nc4 + c4;
                const bool present = vedge4 ? vedge4[idx] != 0 : c4 % bw4[idx] == 0;
                if (!present) continue;
                orc_loop_filter(px, (int64_t)r4 * 4 * w + c4 * 4, e, il, hth, w, 1, edge_width(bw4[idx], bw4[idx - 1], luma),
                                bd);
            }
        }
        // horizontal edges of the superblock row (the frame's top edge is not filtered)
        for (int r4 = top > 0 ? top : 1; r4 < bot; ++r4) {
            for (int c4 = 0; c4 < w4; ++c4) {
                const size_t idx = (size_t)r4 * nc4 + c4;
                const bool present = hedge4 ? hedge4[idx] != 0 : r4 % bh4[idx] == 0;
                if (!present || r4 * 4 >= h || c4 * 4 >= w) continue;
                orc_loop_filter(px, (int64_t)r4 * 4 * w + c4 * 4, e, il, hth, 1, w,
                                edge_width(bh4[idx], bh4[idx - nc4], luma), bd);
            }
        }
    }
}

extern "C" void orc_chroma_deblock_geom(const uint8_t* pblk4, const uint8_t* pblk4h, int nc4, int cw8, int ch8,
                                        int sub_x, int sub_y, uint8_t* cbw4, uint8_t* cbh4) {
    const int cnc4 = cw8 / 4, cnr4 = ch8 / 4;
    for (int cr = 0; cr < cnr4; ++cr)
        for (int cc = 0; cc < cnc4; ++cc) {
            const size_t li = ((size_t)cr << sub_y) * nc4 + ((size_t)cc << sub_x);
            int bw = pblk4[li] >> sub_x, bh = pblk4h[li] >> sub_y;
            // at least one 4-unit; at most a 32-sample transform (the largest chroma one)
            bw = bw < 1 ? 1 : bw > 8 ? 8 : bw;
            bh = bh < 1 ? 1 : bh > 8 ? 8 : bh;
            cbw4[(size_t)cr * cnc4 + cc] = (uint8_t)bw;
            cbh4[(size_t)cr * cnc4 + cc] = (uint8_t)bh;
        }
}

extern "C" void orc_frame_deblock(uint16_t* y, uint16_t* u, uint16_t* v, int w8, int h8, int disp_w, int disp_h,
                                  const uint8_t* blk4, const uint8_t* blk4h, const uint8_t* blk4v,
                                  const uint8_t* blk4t, const uint8_t* pblk4, const uint8_t* pblk4h, int sub_x,
                                  int sub_y, int mono, int sharp, int bd, int lvl_y, int lvl_u, int lvl_v) {
    const int nc4 = w8 / 4;
    orc_filter_plane(y, w8, h8, disp_w, disp_h, blk4, blk4h, blk4v, blk4t, nc4, lvl_y, sharp, 1, 16, bd);
    if (mono || lvl_y <= 0 || (lvl_u <= 0 && lvl_v <= 0)) return;
    const int cw8 = w8 >> sub_x, ch8 = h8 >> sub_y;
    std::vector<uint8_t> cbw((size_t)(cw8 / 4) * (ch8 / 4)), cbh(cbw.size());
    orc_chroma_deblock_geom(pblk4, pblk4h, nc4, cw8, ch8, sub_x, sub_y, cbw.data(), cbh.data());
    const int vw = (disp_w + (1 << sub_x) - 1) >> sub_x, vh = (disp_h + (1 << sub_y) - 1) >> sub_y;
    orc_filter_plane(u, cw8, ch8, vw, vh, cbw.data(), cbh.data(), nullptr, nullptr, cw8 / 4, lvl_u, sharp, 0,
                     16 >> sub_y, bd);
    orc_filter_plane(v, cw8, ch8, vw, vh, cbw.data(), cbh.data(), nullptr, nullptr, cw8 / 4, lvl_v, sharp, 0,
                     16 >> sub_y, bd);
}

extern "C" void orc_lf_levels(int qidx, int sub, int* lvl_y, int* lvl_uv, int* sharp) {
    const float scale = 1.0f, uv_scale = sub == 2 ? 0.5f : 1.0f;
    int y = (int)std::round((float)(qidx / 8) * scale);
    y = y < 0 ? 0 : y > 40 ? 40 : y;
    int uv = 0;
    if (y > 0) {
        uv = (int)std::round((float)(qidx / 10) * uv_scale);
        uv = uv < 1 ? 1 : uv > 32 ? 32 : uv;
    }
    *lvl_y = y;
    *lvl_uv = uv;
    *sharp = 7;
}
