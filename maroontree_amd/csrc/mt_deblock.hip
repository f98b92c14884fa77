// The in-loop deblocking filter (frame_deblock, src/coder/frame_deblock.rs:206;
// filter_plane_impl and loop_filter_scalar, src/loopfilter.rs:483, :106) on the GPU.
//
// Order.  The reference filters per 64-row superblock row: its vertical edges, then its
// horizontal edges.  On a partition's transform grid no two edges of one direction touch a
// common sample (an edge of width 4 / 6 / 8 / 16 reads 2 / 3 / 4 / 7 samples on each side,
// never more than half of the transform next to it), and no horizontal edge reads a row of
// the next superblock row.  So every vertical edge of the frame can run at once, then every
// horizontal edge: AV1's own order, with the same result.  Each tile checks those conditions
// for the edges it owns and reports maps that break them in a status word.
//
// Kernel.  One 512-thread workgroup per 128 x 64 output tile.  The tile and a halo of 16
// columns and 12 rows are loaded into LDS (88 x 160 samples); the vertical edges whose writes
// reach the tile are filtered over all 88 rows, then the horizontal edges of the tile's own
// columns; the tile's interior is stored to the output plane.  The filter is out of place
// (halos read the input while other tiles write the output).  Edges are listed per filter
// width, so the lanes of a wave run one width.  Tiles of one XCD (workgroup b runs on XCD
// b % 8) form a contiguous run of the raster, so the halos a tile shares with the tiles above
// and below it are in that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <map>
#include <mutex>

#include "../../include/mtgpu.h"
#include "mt_common.h"

namespace {

constexpr int TW = 128, TH = 64;  // output tile
constexpr int HX = 16, HY = 12;  // halo columns / rows
constexpr int LW = TW + 2 * HX;  // LDS columns (160)
constexpr int LH = TH + 2 * HY;  // LDS rows (88)
constexpr int NVR = LH / 4;      // 4-unit rows with vertical edges (22)
constexpr int NVC = TW / 4 + 3;  // vertical-edge columns: the tile's 4-unit columns -1 .. +33 (35)
constexpr int NHR = TH / 4 + 3;  // horizontal-edge rows: the tile's 4-unit rows -1 .. +17 (19)
constexpr int NHC = TW / 4;      // horizontal-edge columns: the tile's own (32)
constexpr int kThreads = 512;
constexpr int kEnt = 6;          // edge-list entries: r << kEnt | c
static_assert(HY == 12 && HX >= 12 && HX % 4 == 0, "halo: 7-sample reach of edges 4 past the tile");
static_assert(NVC <= (1 << kEnt) && NHC <= (1 << kEnt) && NVR * (1 << kEnt) < 65536, "edge-list packing");

struct LfPlane {
    const uint16_t* src;
    uint16_t* dst;
    const uint8_t *bw4, *bh4, *ve, *he;
    int32_t stride, w, h, w4, h4, nc4;
    int32_t derive, sx, sy;   // derive: sizes from a luma prediction map (chroma_deblock_geom)
    int32_t e, il, hth;       // limits at the bit depth
    int32_t bd, luma, sb_rows4, filter;
    int32_t tiles_x, tile0;
};
struct LfArgs {
    LfPlane p[3];
    int32_t np, ntiles, per_xcd, pad;
    uint32_t* status;
};

__device__ __forceinline__ int reach(int wd) { return wd == 16 ? 7 : wd == 8 ? 4 : wd == 6 ? 3 : 2; }

__device__ __forceinline__ int edge_width(int a4, int b4, bool luma) {
    const int m = a4 < b4 ? a4 : b4;
    if (luma) return m <= 1 ? 4 : m == 2 ? 8 : 16;
    return m >= 2 ? 6 : 4;
}

// transform width / height (4-units) at (r4, c4) of the plane's grid
__device__ __forceinline__ int size_w(const LfPlane& P, int r4, int c4) {
    if (!P.derive) return P.bw4[(size_t)r4 * P.nc4 + c4];
    const int v = P.bw4[((size_t)r4 << P.sy) * P.nc4 + ((size_t)c4 << P.sx)] >> P.sx;
    return v < 1 ? 1 : v > 8 ? 8 : v;
}
__device__ __forceinline__ int size_h(const LfPlane& P, int r4, int c4) {
    if (!P.derive) return P.bh4[(size_t)r4 * P.nc4 + c4];
    const int v = P.bh4[((size_t)r4 << P.sy) * P.nc4 + ((size_t)c4 << P.sx)] >> P.sy;
    return v < 1 ? 1 : v > 8 ? 8 : v;
}

// The tile's window of the maps in LDS: 4-unit rows ty0 / 4 - 4 .. + TH / 4 + 3 and columns
// tx0 / 4 - 4 .. + TW / 4 + 3 (every edge a tile lists or checks, with its left / top
// neighbour), entries outside the plane 0.
constexpr int MWR = TH / 4 + 8, MWC = TW / 4 + 8;
struct Win {
    const uint8_t *w, *h, *ve, *he;  // sizes and edge flags, MWR x MWC
    int r0, c0;                      // the window's first 4-unit row / column
};

// filter width of the vertical edge at (r4, c4), 0 where there is none
__device__ __forceinline__ int vwd(const LfPlane& P, const Win& W, int r4, int c4, uint32_t& st) {
    if (r4 < 0 || r4 >= P.h4 || c4 < 1 || c4 >= P.w4) return 0;
    const int i = (r4 - W.r0) * MWC + (c4 - W.c0);
    const int cur = W.w[i];
    bool present;
    if (P.ve) {
        present = W.ve[i] != 0;
    } else {
        if (cur == 0) {
            st |= MT_DEBLOCK_BAD_MAP;
            return 0;
        }
        present = c4 % cur == 0;
    }
    return present ? edge_width(cur, W.w[i - 1], P.luma) : 0;
}
__device__ __forceinline__ int hwd(const LfPlane& P, const Win& W, int r4, int c4, uint32_t& st) {
    if (r4 < 1 || r4 >= P.h4 || c4 < 0 || c4 >= P.w4) return 0;
    const int i = (r4 - W.r0) * MWC + (c4 - W.c0);
    const int cur = W.h[i];
    bool present;
    if (P.he) {
        present = W.he[i] != 0;
    } else {
        if (cur == 0) {
            st |= MT_DEBLOCK_BAD_MAP;
            return 0;
        }
        present = r4 % cur == 0;
    }
    return present ? edge_width(cur, W.h[i - MWC], P.luma) : 0;
}

// The conditions under which the parallel order equals the reference's, for one edge against
// the edges before it (the pairs with the edges after it are checked by those).
__device__ uint32_t check_v(const LfPlane& P, const Win& W, int r4, int c4, int wd) {
    const int R = reach(wd);
    uint32_t st = 0;
    if (c4 * 4 - R < 0 || c4 * 4 + R > P.w) st |= MT_DEBLOCK_ORDER;
    for (int k = 1; k <= 3; ++k) {
        const int w2 = vwd(P, W, r4, c4 - k, st);
        if (w2 && 4 * k < R + reach(w2)) st |= MT_DEBLOCK_ORDER;
    }
    return st;
}
__device__ uint32_t check_h(const LfPlane& P, const Win& W, int r4, int c4, int wd) {
    const int R = reach(wd);
    uint32_t st = 0;
    if (r4 * 4 - R < 0 || r4 * 4 + R > P.h) st |= MT_DEBLOCK_ORDER;
    const int next = (r4 / P.sb_rows4 + 1) * P.sb_rows4;  // the next superblock row
    if (next < P.h4 && r4 * 4 + R > next * 4) st |= MT_DEBLOCK_ORDER;
    for (int k = 1; k <= 3; ++k) {
        const int w2 = hwd(P, W, r4 - k, c4, st);
        if (w2 && 4 * k < R + reach(w2)) st |= MT_DEBLOCK_ORDER;
    }
    return st;
}

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ int clip3(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

// One line across an edge of width WD.  q points at q0 (the first sample past the edge) in
// LDS, st is the step across the edge; e / il / hth are the limits at the bit depth.
template <int WD>
__device__ __forceinline__ void lf_line(uint16_t* q, int st, int e, int il, int hth, int bd) {
    const int p0 = q[-st], p1 = q[-2 * st], q0 = q[0], q1 = q[st];
    bool fm = iabs(p1 - p0) <= il && iabs(q1 - q0) <= il && 2 * iabs(p0 - q0) + (iabs(p1 - q1) >> 1) <= e;
    int p2 = 0, q2 = 0, p3 = 0, q3 = 0;
    if (WD > 4) {
        p2 = q[-3 * st];
        q2 = q[2 * st];
        fm = fm && iabs(p2 - p1) <= il && iabs(q2 - q1) <= il;
    }
    if (WD > 6) {
        p3 = q[-4 * st];
        q3 = q[3 * st];
        fm = fm && iabs(p3 - p2) <= il && iabs(q3 - q2) <= il;
    }
    if (!fm) return;
    const int flat = 1 << (bd - 8);
    bool inner = false;
    if (WD >= 6) inner = iabs(p2 - p0) <= flat && iabs(p1 - p0) <= flat && iabs(q1 - q0) <= flat && iabs(q2 - q0) <= flat;
    if (WD >= 8) inner = inner && iabs(p3 - p0) <= flat && iabs(q3 - q0) <= flat;
    if (WD == 16 && inner) {
        const int p4 = q[-5 * st], p5 = q[-6 * st], p6 = q[-7 * st];
        const int q4 = q[4 * st], q5 = q[5 * st], q6 = q[6 * st];
        if (iabs(p6 - p0) <= flat && iabs(p5 - p0) <= flat && iabs(p4 - p0) <= flat && iabs(q4 - q0) <= flat &&
            iabs(q5 - q0) <= flat && iabs(q6 - q0) <= flat) {
            // the 15-tap smoothing of p5..q5 as a running sum of 16 taps
            int s = 7 * p6 + 2 * p5 + 2 * p4 + p3 + p2 + p1 + p0 + q0;
            q[-6 * st] = (uint16_t)((s + 8) >> 4);
            s += -2 * p6 + p3 + q1;
            q[-5 * st] = (uint16_t)((s + 8) >> 4);
            s += -p6 - p5 + p2 + q2;
            q[-4 * st] = (uint16_t)((s + 8) >> 4);
            s += -p6 - p4 + p1 + q3;
            q[-3 * st] = (uint16_t)((s + 8) >> 4);
            s += -p6 - p3 + p0 + q4;
            q[-2 * st] = (uint16_t)((s + 8) >> 4);
            s += -p6 - p2 + q0 + q5;
            q[-st] = (uint16_t)((s + 8) >> 4);
            s += -p6 - p1 + q1 + q6;
            q[0] = (uint16_t)((s + 8) >> 4);
            s += -p5 - p0 + q2 + q6;
            q[st] = (uint16_t)((s + 8) >> 4);
            s += -p4 - q0 + q3 + q6;
            q[2 * st] = (uint16_t)((s + 8) >> 4);
            s += -p3 - q1 + q4 + q6;
            q[3 * st] = (uint16_t)((s + 8) >> 4);
            s += -p2 - q2 + q5 + q6;
            q[4 * st] = (uint16_t)((s + 8) >> 4);
            s += -p1 - q3 + 2 * q6;
            q[5 * st] = (uint16_t)((s + 8) >> 4);
            return;
        }
    }
    if (WD >= 8 && inner) {
        q[-3 * st] = (uint16_t)((3 * p3 + 2 * p2 + p1 + p0 + q0 + 4) >> 3);
        q[-2 * st] = (uint16_t)((2 * p3 + p2 + 2 * p1 + p0 + q0 + q1 + 4) >> 3);
        q[-st] = (uint16_t)((p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2 + 4) >> 3);
        q[0] = (uint16_t)((p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4) >> 3);
        q[st] = (uint16_t)((p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4) >> 3);
        q[2 * st] = (uint16_t)((p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4) >> 3);
        return;
    }
    if (WD == 6 && inner) {
        q[-2 * st] = (uint16_t)((3 * p2 + 2 * p1 + 2 * p0 + q0 + 4) >> 3);
        q[-st] = (uint16_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
        q[0] = (uint16_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
        q[st] = (uint16_t)((p0 + 2 * q0 + 2 * q1 + 3 * q2 + 4) >> 3);
        return;
    }
    // narrow filter
    const int sh = bd - 8, lo = -128 << sh, hi = (128 << sh) - 1, mx = (1 << bd) - 1;
    const bool hev = iabs(p1 - p0) > hth || iabs(q1 - q0) > hth;
    int f = hev ? clip3(p1 - q1, lo, hi) : 0;
    f = clip3(3 * (q0 - p0) + f, lo, hi);
    const int f1 = min(f + 4, hi) >> 3, f2 = min(f + 3, hi) >> 3;
    q[-st] = (uint16_t)clip3(p0 + f2, 0, mx);
    q[0] = (uint16_t)clip3(q0 - f1, 0, mx);
    if (!hev) {
        const int f3 = (f1 + 1) >> 1;
        q[-2 * st] = (uint16_t)clip3(p1 + f3, 0, mx);
        q[st] = (uint16_t)clip3(q1 - f3, 0, mx);
    }
}

// the vertical edges of one width: entries (r, c) -> 4 lines each
template <int WD>
__device__ __forceinline__ void run_v(uint16_t* px, const uint16_t* list, int n, const LfPlane& P) {
    for (int k = threadIdx.x; k < 4 * n; k += kThreads) {
        const int ent = list[k >> 2], r = ent >> kEnt, c = ent & ((1 << kEnt) - 1);
        lf_line<WD>(px + (4 * r + (k & 3)) * LW + HX - 4 + 4 * c, 1, P.e, P.il, P.hth, P.bd);
    }
}
template <int WD>
__device__ __forceinline__ void run_h(uint16_t* px, const uint16_t* list, int n, const LfPlane& P) {
    for (int k = threadIdx.x; k < 4 * n; k += kThreads) {
        const int ent = list[k >> 2], r = ent >> kEnt, c = ent & ((1 << kEnt) - 1);
        lf_line<WD>(px + (4 * r + HY - 4) * LW + HX + 4 * c + (k & 3), LW, P.e, P.il, P.hth, P.bd);
    }
}

__global__ void __launch_bounds__(kThreads) deblock_kernel(LfArgs a) {
    const int b = blockIdx.x;
    const int t = (b & 7) * a.per_xcd + (b >> 3);
    if (t >= a.ntiles) return;
    const int pl = (a.np > 2 && t >= a.p[2].tile0) ? 2 : (a.np > 1 && t >= a.p[1].tile0) ? 1 : 0;
    const LfPlane& P = a.p[pl];
    const int lt = t - P.tile0, tyi = lt / P.tiles_x, txi = lt - tyi * P.tiles_x;
    const int tx0 = txi * TW, ty0 = tyi * TH, gx0 = tx0 - HX, gy0 = ty0 - HY;

    __shared__ __align__(16) uint16_t s_px[LH * LW];
    __shared__ uint16_t s_vl[4][NVR * NVC];
    __shared__ uint16_t s_hl[4][NHR * NHC];
    __shared__ uint8_t s_map[4][MWR * MWC];
    __shared__ int s_cnt[8];
    const int tid = threadIdx.x;

    // every global load of the tile (samples and maps) is issued before the first LDS store
    constexpr int NPX = LH * (LW / 4), PIT = (NPX + kThreads - 1) / kThreads;
    uint2 pv[PIT];
#pragma unroll
    for (int i = 0; i < PIT; ++i) {
        const int k = tid + i * kThreads;
        pv[i] = make_uint2(0u, 0u);
        if (k < NPX) {
            const int r = k / (LW / 4), c = k - r * (LW / 4);
            const int gy = gy0 + r, gx = gx0 + 4 * c;
            if (gy >= 0 && gy < P.h && gx >= 0 && gx < P.w)
                pv[i] = *reinterpret_cast<const uint2*>(P.src + (size_t)gy * P.stride + gx);
        }
    }
    constexpr int NM = MWR * MWC, MIT = (NM + kThreads - 1) / kThreads;
    const int r0 = ty0 / 4 - 4, c0 = tx0 / 4 - 4;
    uint32_t mv[MIT];
#pragma unroll
    for (int i = 0; i < MIT; ++i) {
        const int k = tid + i * kThreads;
        mv[i] = 0;
        if (P.filter && k < NM) {
            const int r4 = r0 + k / MWC, c4 = c0 + k % MWC;
            if (r4 >= 0 && r4 < P.h / 4 && c4 >= 0 && c4 < P.w / 4) {
                const size_t j = (size_t)r4 * P.nc4 + c4;
                mv[i] = (uint32_t)size_w(P, r4, c4) | (uint32_t)size_h(P, r4, c4) << 8 |
                        (P.ve ? (uint32_t)P.ve[j] << 16 : 0u) | (P.he ? (uint32_t)P.he[j] << 24 : 0u);
            }
        }
    }
    if (tid < 8) s_cnt[tid] = 0;
#pragma unroll
    for (int i = 0; i < PIT; ++i) {
        const int k = tid + i * kThreads;
        if (k < NPX) {
            const int r = k / (LW / 4), c = k - r * (LW / 4);
            *reinterpret_cast<uint2*>(s_px + r * LW + 4 * c) = pv[i];
        }
    }
#pragma unroll
    for (int i = 0; i < MIT; ++i) {
        const int k = tid + i * kThreads;
        if (k < NM) {
            s_map[0][k] = (uint8_t)mv[i];
            s_map[1][k] = (uint8_t)(mv[i] >> 8);
            s_map[2][k] = (uint8_t)(mv[i] >> 16);
            s_map[3][k] = (uint8_t)(mv[i] >> 24);
        }
    }
    __syncthreads();
    uint32_t st = 0;
    if (P.filter) {
        const Win W{s_map[0], s_map[1], s_map[2], s_map[3], r0, c0};
        const int r40 = ty0 / 4 - 3, c40 = tx0 / 4 - 1;
        for (int k = tid; k < NVR * NVC; k += kThreads) {
            const int r = k / NVC, c = k - r * NVC, r4 = r40 + r, c4 = c40 + c;
            const int wd = vwd(P, W, r4, c4, st);
            if (wd) {
                const int cls = wd == 4 ? 0 : wd == 6 ? 1 : wd == 8 ? 2 : 3;
                s_vl[cls][atomicAdd(&s_cnt[cls], 1)] = (uint16_t)(r << kEnt | c);
                if (r4 >= ty0 / 4 && r4 < ty0 / 4 + TH / 4 && c4 >= tx0 / 4 && c4 < tx0 / 4 + TW / 4)
                    st |= check_v(P, W, r4, c4, wd);
            }
        }
        const int h40 = ty0 / 4 - 1;
        for (int k = tid; k < NHR * NHC; k += kThreads) {
            const int r = k / NHC, c = k - r * NHC, r4 = h40 + r, c4 = tx0 / 4 + c;
            const int wd = hwd(P, W, r4, c4, st);
            if (wd) {
                const int cls = wd == 4 ? 0 : wd == 6 ? 1 : wd == 8 ? 2 : 3;
                s_hl[cls][atomicAdd(&s_cnt[4 + cls], 1)] = (uint16_t)(r << kEnt | c);
                if (r4 >= ty0 / 4 && r4 < ty0 / 4 + TH / 4) st |= check_h(P, W, r4, c4, wd);
            }
        }
        __syncthreads();
        run_v<16>(s_px, s_vl[3], s_cnt[3], P);
        run_v<8>(s_px, s_vl[2], s_cnt[2], P);
        run_v<6>(s_px, s_vl[1], s_cnt[1], P);
        run_v<4>(s_px, s_vl[0], s_cnt[0], P);
        __syncthreads();
        run_h<16>(s_px, s_hl[3], s_cnt[7], P);
        run_h<8>(s_px, s_hl[2], s_cnt[6], P);
        run_h<6>(s_px, s_hl[1], s_cnt[5], P);
        run_h<4>(s_px, s_hl[0], s_cnt[4], P);
        __syncthreads();
    }
    for (int k = tid; k < TH * (TW / 4); k += kThreads) {
        const int r = k / (TW / 4), c = k - r * (TW / 4);
        const int gy = ty0 + r, gx = tx0 + 4 * c;
        if (gy < P.h && gx < P.w)
            *reinterpret_cast<uint2*>(P.dst + (size_t)gy * P.stride + gx) =
                *reinterpret_cast<const uint2*>(s_px + (r + HY) * LW + HX + 4 * c);
    }
    if (st && a.status) atomicOr(a.status, st);
}

// The reference's deblock maps (record_blk, src/coder/superblock.rs:548-562) of the RDO's
// decisions: one thread per 4 x 4 unit walks its superblock's quadtree to its leaf.
__global__ void __launch_bounds__(256) rdo_maps_kernel(const mt_rdo_node* __restrict__ nodes, int w, int h,
                                                       uint8_t* __restrict__ maps) {
    const int nc4 = w / 4, nr4 = h / 4, n = nc4 * nr4;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int r4 = idx / nc4, c4 = idx - r4 * nc4;
    const int sbc = (w + 63) / 64;
    const mt_rdo_node* sb = nodes + (size_t)((r4 >> 4) * sbc + (c4 >> 4)) * MT_RDO_NODES;
    const int lx = c4 & 15, ly = r4 & 15;
    int d = 0, bx = 0, by = 0, base = 0;
    for (int L = 0; L <= 4; ++L) {
        const int s4 = 16 >> L, ix = lx / s4, iy = ly / s4;
        const int split = sb[base + iy * (1 << L) + ix].split;
        if (split == 0) {
            d = s4;
            bx = ix * s4;
            by = iy * s4;
            break;
        }
        if (split != 1) break;
        base += 1 << (2 * L);
    }
    const size_t np = (size_t)n;
    maps[idx] = (uint8_t)d;
    maps[np + idx] = (uint8_t)d;
    maps[2 * np + idx] = (uint8_t)(d && lx == bx);
    maps[3 * np + idx] = (uint8_t)(d && ly == by);
    maps[4 * np + idx] = (uint8_t)d;
    maps[5 * np + idx] = (uint8_t)d;
}

int limits_at(int level, int sharp, int bd, LfPlane& P) {
    int il = level;
    if (sharp > 0) {
        il >>= (sharp + 3) >> 2;
        il = std::min(il, 9 - sharp);
    }
    il = std::max(il, 1);
    const int sh = bd - 8;
    P.e = (2 * (level + 2) + il) << sh;
    P.il = il << sh;
    P.hth = (level >> 4) << sh;
    return MT_OK;
}

bool aligned8(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 7) == 0; }

int check_plane(const uint16_t* src, const uint16_t* dst, int stride, int w, int h, const char* who) {
    if (!src || !dst) return mtlib::set_error(MT_ERR_ARG, std::string(who) + ": null plane");
    if (w < 4 || h < 4 || (w & 3) || (h & 3) || stride < w || (stride & 3) || !aligned8(src) || !aligned8(dst))
        return mtlib::set_error(MT_ERR_ARG, std::string(who) +
                                                ": sizes and stride must be multiples of 4, planes 8-byte aligned");
    const uintptr_t s0 = reinterpret_cast<uintptr_t>(src), d0 = reinterpret_cast<uintptr_t>(dst);
    const uintptr_t bytes = (uintptr_t)stride * h * 2;
    if (s0 < d0 + bytes && d0 < s0 + bytes)
        return mtlib::set_error(MT_ERR_ARG, std::string(who) + ": dst overlaps src (the filter is out of place)");
    return MT_OK;
}

int launch(LfArgs& a, hipStream_t s) {
    int tiles = 0;
    for (int i = 0; i < a.np; ++i) {
        LfPlane& P = a.p[i];
        P.tiles_x = (P.w + TW - 1) / TW;
        P.tile0 = tiles;
        tiles += P.tiles_x * ((P.h + TH - 1) / TH);
    }
    a.ntiles = tiles;
    a.per_xcd = (tiles + 7) / 8;
    hipLaunchKernelGGL(deblock_kernel, dim3(8 * a.per_xcd), dim3(kThreads), 0, s, a);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}

}  // namespace

extern "C" int mt_deblock_levels(int qidx, int sub, int* lvl_y, int* lvl_uv, int* sharp) {
    if (qidx < 0 || qidx > 255 || sub < 0 || sub > 2 || !lvl_y || !lvl_uv || !sharp)
        return mtlib::set_error(MT_ERR_ARG, "mt_deblock_levels: bad argument");
    // loop_filter_levels (src/obu.rs:306-330) at the shipped scales: luma 1.0, chroma 0.5 at
    // 4:2:0 (sub 2) and 1.0 otherwise; f32 products rounded half away from zero
    const float uv_scale = sub == 2 ? 0.5f : 1.0f;
    const int y = std::min(std::max((int)std::round((float)(qidx / 8) * 1.0f), 0), 40);
    int uv = 0;
    if (y > 0) uv = std::min(std::max((int)std::round((float)(qidx / 10) * uv_scale), 1), 32);
    *lvl_y = y;
    *lvl_uv = uv;
    *sharp = 7;  // loop_filter_sharpness (:332): the shipped lf_sharpness
    return MT_OK;
}

extern "C" int mt_deblock_plane(const mt_lf_plane* p, int level, int sharp, int is_luma, int sb_rows4, int bd,
                                uint32_t* status_d, void* stream) {
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    if (!p) return mtlib::set_error(MT_ERR_ARG, "mt_deblock_plane: null plane");
    rc = check_plane(p->src, p->dst, p->stride, p->w, p->h, "mt_deblock_plane");
    if (rc != MT_OK) return rc;
    if (level > 63 || sharp < 0 || sharp > 7 || sb_rows4 < 1 || (bd != 8 && bd != 10 && bd != 12) || !p->bw4 ||
        !p->bh4 || p->nc4 < p->w / 4 || p->vis_w < 0 || p->vis_h < 0)
        return mtlib::set_error(MT_ERR_ARG, "mt_deblock_plane: bad argument");
    LfArgs a{};
    LfPlane& P = a.p[0];
    P.src = p->src;
    P.dst = p->dst;
    P.bw4 = p->bw4;
    P.bh4 = p->bh4;
    P.ve = p->vedge4;
    P.he = p->hedge4;
    P.stride = p->stride;
    P.w = p->w;
    P.h = p->h;
    P.w4 = std::min((p->vis_w + 3) / 4, p->w / 4);
    P.h4 = std::min((p->vis_h + 3) / 4, p->h / 4);
    P.nc4 = p->nc4;
    P.bd = bd;
    P.luma = is_luma ? 1 : 0;
    P.sb_rows4 = sb_rows4;
    P.filter = level > 0;
    limits_at(std::max(level, 0), sharp, bd, P);
    a.np = 1;
    a.status = status_d;
    return launch(a, (hipStream_t)stream);
}

extern "C" int mt_deblock_frame(const mt_deblock_geom* g, const uint16_t* src_y, const uint16_t* src_u,
                                const uint16_t* src_v, int32_t y_stride, int32_t uv_stride, uint16_t* dst_y,
                                uint16_t* dst_u, uint16_t* dst_v, int lvl_y, int lvl_u, int lvl_v,
                                uint32_t* status_d, void* stream) {
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    if (!g || !g->blk4 || !g->blk4h || !g->blk4v || !g->blk4t || !g->pblk4 || !g->pblk4h)
        return mtlib::set_error(MT_ERR_ARG, "mt_deblock_frame: null geometry");
    if (g->w8 < 8 || g->h8 < 8 || (g->w8 & 7) || (g->h8 & 7) || g->disp_w < 1 || g->disp_h < 1 ||
        g->disp_w > g->w8 || g->disp_h > g->h8 || g->sub_x > 1 || g->sub_y > 1 || g->sharp > 7 ||
        (g->bd != 8 && g->bd != 10 && g->bd != 12) || lvl_y > 63 || lvl_u > 63 || lvl_v > 63)
        return mtlib::set_error(MT_ERR_ARG, "mt_deblock_frame: bad geometry or level");
    rc = check_plane(src_y, dst_y, y_stride, g->w8, g->h8, "mt_deblock_frame (luma)");
    if (rc != MT_OK) return rc;
    const int cw8 = g->w8 >> g->sub_x, ch8 = g->h8 >> g->sub_y;
    if (!g->mono) {
        rc = check_plane(src_u, dst_u, uv_stride, cw8, ch8, "mt_deblock_frame (U)");
        if (rc == MT_OK) rc = check_plane(src_v, dst_v, uv_stride, cw8, ch8, "mt_deblock_frame (V)");
        if (rc != MT_OK) return rc;
    }
    LfArgs a{};
    LfPlane& Y = a.p[0];
    Y.src = src_y;
    Y.dst = dst_y;
    Y.bw4 = g->blk4;
    Y.bh4 = g->blk4h;
    Y.ve = g->blk4v;
    Y.he = g->blk4t;
    Y.stride = y_stride;
    Y.w = g->w8;
    Y.h = g->h8;
    Y.w4 = (g->disp_w + 3) / 4;
    Y.h4 = (g->disp_h + 3) / 4;
    Y.nc4 = g->w8 / 4;
    Y.bd = g->bd;
    Y.luma = 1;
    Y.sb_rows4 = 16;
    Y.filter = lvl_y > 0;
    limits_at(std::max(lvl_y, 0), g->sharp, g->bd, Y);
    a.np = 1;
    if (!g->mono) {
        // frame_deblock: chroma is filtered only while luma is, and not when both levels are 0
        const bool chroma = lvl_y > 0 && (lvl_u > 0 || lvl_v > 0);
        const int vw = (g->disp_w + (1 << g->sub_x) - 1) >> g->sub_x;
        const int vh = (g->disp_h + (1 << g->sub_y) - 1) >> g->sub_y;
        for (int c = 0; c < 2; ++c) {
            LfPlane& C = a.p[1 + c];
            C.src = c ? src_v : src_u;
            C.dst = c ? dst_v : dst_u;
            C.bw4 = g->pblk4;
            C.bh4 = g->pblk4h;
            C.derive = 1;
            C.sx = g->sub_x;
            C.sy = g->sub_y;
            C.stride = uv_stride;
            C.w = cw8;
            C.h = ch8;
            C.w4 = (vw + 3) / 4;
            C.h4 = (vh + 3) / 4;
            C.nc4 = g->w8 / 4;
            C.bd = g->bd;
            C.luma = 0;
            C.sb_rows4 = 16 >> g->sub_y;
            const int lv = c ? lvl_v : lvl_u;
            C.filter = chroma && lv > 0;
            limits_at(std::max(lv, 0), g->sharp, g->bd, C);
        }
        a.np = 3;
    }
    a.status = status_d;
    return launch(a, (hipStream_t)stream);
}

extern "C" int mt_rdo_deblock_maps(const void* nodes_d, int width, int height, uint8_t* maps_d, void* stream) {
    int dev = 0, cus = 0;
    int rc = mtlib::device_ready(&dev, &cus);
    if (rc != MT_OK) return rc;
    if (!nodes_d || !maps_d || width < 8 || height < 8 || (width & 7) || (height & 7))
        return mtlib::set_error(MT_ERR_ARG, "mt_rdo_deblock_maps: bad argument");
    const int n = (width / 4) * (height / 4);
    hipLaunchKernelGGL(rdo_maps_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const mt_rdo_node*>(nodes_d), width, height, maps_d);
    MT_HIP_TRY(hipGetLastError());
    return MT_OK;
}
