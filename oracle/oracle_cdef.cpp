// CDEF restated from the reference (TEST INFRASTRUCTURE ONLY: the checker of the HIP CDEF
// kernels and of the frame decision; the library never calls it).
//
// src/cdef.rs: the direction search (cdef_direction_partials / cdef_direction_cost,
// :103-179), adjust_pri (:181), the block filter (cdef_filter_block_generic, :514-606, with
// out-of-plane taps excluded as CDEF_VERY_LARGE, :608) and the twelve signalled strength
// pairs (CAND_PRI x CAND_SEC, :622-649).  src/coder/frame_cdef.rs: frame_cdef (:49-558): the
// per-64x64-unit distortion tables of every strength pair (:622-720 luma, :727-800
// chroma, metric wSSE of the shipped tuning with plain SSE where the dispatch measures
// plain SSE), the directional-variance gate, the Slow two-stage candidate list or the
// Fast full list, the greedy strength-set construction, the mildest-entry assignment with
// the 40 per-mille margin, the cdef_bits choice on rate, and the apply pass
// (apply_cdef_plane / apply_cdef_chroma, :995-1125).  The tuning values are the shipped
// ones (src/tuning.rs:772-776: margin 40, gate 15000, chroma weight 1, wSSE, lambda_d).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int kVeryLarge = 0x4000;
const int kDirs[8][2][2] = {{{-1, 1}, {-2, 2}}, {{0, 1}, {-1, 2}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}},
                            {{1, 1}, {2, 2}},   {{1, 0}, {2, 1}},  {{1, 0}, {2, 0}}, {{1, 0}, {2, -1}}};
const int kPriTaps[2][2] = {{4, 2}, {3, 3}};
const int kSecTaps[2] = {2, 1};
const int kCandPri[4] = {0, 1, 2, 4};
const int kCandSec[3] = {0, 1, 2};

int log2_floor(int v) { return v <= 0 ? 0 : 31 - __builtin_clz((unsigned)v); }

int constrain(int diff, int strength, int damping) {
    if (strength == 0) return 0;
    const int shift = std::max(0, damping - log2_floor(strength));
    const int a = std::abs(diff);
    const int c = std::min(std::max(strength - (a >> shift), 0), a);
    return diff < 0 ? -c : c;
}

int sample(const uint16_t* p, int stride, int w, int h, int x, int y) {
    if (x < 0 || y < 0 || x >= w || y >= h) return kVeryLarge;
    return p[(size_t)y * stride + x];
}

// the strength pair's filter of one sample (cdef_filter_block_generic's body)
int filter_px(const uint16_t* p, int stride, int w, int h, int cx, int cy, int pri, int sec, int dir, int damping,
              int bd) {
    const int centre = p[(size_t)cy * stride + cx];
    const int cs = bd - 8;
    const int* pt = kPriTaps[(pri >> cs) & 1];
    const bool ep = pri != 0, es = sec != 0, clip = ep && es;
    const int damp = std::max(damping, 1);
    int sum = 0, mn = centre, mx = centre;
    for (int k = 0; k < 2; ++k) {
        if (ep)
            for (int sg = 1; sg >= -1; sg -= 2) {
                const int v = sample(p, stride, w, h, cx + sg * kDirs[dir][k][1], cy + sg * kDirs[dir][k][0]);
                if (v == kVeryLarge) continue;
                sum += pt[k] * constrain(v - centre, pri, damp);
                if (clip) {
                    mx = std::max(mx, v);
                    mn = std::min(mn, v);
                }
            }
        if (es)
            for (int off : {2, 6}) {
                const int sd = (dir + off) & 7;
                for (int sg = 1; sg >= -1; sg -= 2) {
                    const int v = sample(p, stride, w, h, cx + sg * kDirs[sd][k][1], cy + sg * kDirs[sd][k][0]);
                    if (v == kVeryLarge) continue;
                    sum += kSecTaps[k] * constrain(v - centre, sec, damp);
                    if (clip) {
                        mx = std::max(mx, v);
                        mn = std::min(mn, v);
                    }
                }
            }
    }
    int out = centre + ((8 + sum - (sum < 0 ? 1 : 0)) >> 4);
    if (clip) out = std::min(std::max(out, mn), mx);
    return std::min(std::max(out, 0), (1 << bd) - 1);
}

}  // namespace

extern "C" {

int orc_cdef_dir(const uint16_t* p, int stride, int w, int h, int x, int y, int bd, int32_t* var) {
    int64_t part[8][15];
    std::memset(part, 0, sizeof(part));
    const int cs = bd - 8;
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
            const int yy = std::min(y + i, h - 1), xx = std::min(x + j, w - 1);
            const int v = (p[(size_t)yy * stride + xx] >> cs) - 128;
            part[0][i + j] += v;
            part[1][i + j / 2] += v;
            part[2][i] += v;
            part[3][3 + i - j / 2] += v;
            part[4][7 + i - j] += v;
            part[5][3 - i / 2 + j] += v;
            part[6][j] += v;
            part[7][i / 2 + j] += v;
        }
    static const int64_t kDiv[8] = {0, 840, 420, 280, 210, 168, 140, 120};
    int64_t cost[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 8; ++i) {
        cost[2] += part[2][i] * part[2][i];
        cost[6] += part[6][i] * part[6][i];
    }
    cost[2] *= 105;
    cost[6] *= 105;
    for (int i = 0; i < 7; ++i) {
        cost[0] += (part[0][i] * part[0][i] + part[0][14 - i] * part[0][14 - i]) * kDiv[i + 1];
        cost[4] += (part[4][i] * part[4][i] + part[4][14 - i] * part[4][14 - i]) * kDiv[i + 1];
    }
    cost[0] += part[0][7] * part[0][7] * 105;
    cost[4] += part[4][7] * part[4][7] * 105;
    for (int s = 1; s < 8; s += 2) {
        for (int i = 0; i < 5; ++i) cost[s] += part[s][3 + i] * part[s][3 + i] * 105;
        for (int i = 0; i < 3; ++i)
            cost[s] += (part[s][i] * part[s][i] + part[s][10 - i] * part[s][10 - i]) * kDiv[2 * i + 2];
    }
    int best = 0;
    for (int d = 1; d < 8; ++d)
        if (cost[d] > cost[best]) best = d;
    *var = (int32_t)((cost[best] - cost[(best + 4) & 7]) >> 10);
    return best;
}

int32_t orc_cdef_adjust_pri(int32_t pri, int32_t var) {
    if (var == 0) return 0;
    const int vs = var >> 6;
    const int l = vs ? std::min(log2_floor(vs), 12) : 0;
    return (pri * (4 + l) + 8) >> 4;
}

// One bw x bh block (taps outside the w x h plane excluded) with effective strengths
// (pri, sec, already bit-depth scaled and adjusted) into dst at the block's own place.
void orc_cdef_filter(const uint16_t* p, int stride, int w, int h, int x, int y, int bw, int bh, int pri, int sec,
                     int dir, int damping, int bd, uint16_t* dst, int dstride) {
    for (int i = 0; i < bh; ++i)
        for (int j = 0; j < bw; ++j) {
            if (x + j >= w || y + i >= h) continue;
            dst[(size_t)(y + i) * dstride + x + j] =
                (uint16_t)filter_px(p, stride, w, h, x + j, y + i, pri, sec, dir, damping, bd);
        }
}

// The twelve signalled strength pairs of one block, candidate si * 4 + pi = (CAND_PRI[pi],
// CAND_SEC[si]); var < 0: chroma (no variance adjustment).  out[c * bw * bh + i * bw + j].
void orc_cdef_candidates(const uint16_t* p, int stride, int w, int h, int x, int y, int bw, int bh, int dir,
                         int32_t var, int damping, int bd, int32_t* out) {
    const int cs = bd - 8;
    for (int si = 0; si < 3; ++si)
        for (int pi = 0; pi < 4; ++pi) {
            const int sp = kCandPri[pi] << cs;
            const int pri = var >= 0 ? orc_cdef_adjust_pri(sp, var) : sp;
            const int sec = kCandSec[si] << cs;
            const int d = kCandPri[pi] == 0 ? 0 : dir;
            int32_t* o = out + (size_t)(si * 4 + pi) * bw * bh;
            for (int i = 0; i < bh; ++i)
                for (int j = 0; j < bw; ++j)
                    o[i * bw + j] = filter_px(p, stride, w, h, x + j, y + i, pri, sec, d, damping, bd);
        }
}

}  // extern "C"

namespace {

int cand_index(int pri, int sec) {
    int pi = 0, si = 0;
    while (kCandPri[pi] != pri) ++pi;
    while (kCandSec[si] != sec) ++si;
    return si * 4 + pi;
}

// RdDispatch::wsse_err: the scaled weighted SSE (AVX2 order), or the plain SSE
int64_t wsse_err(const int32_t* err, int w, int h, float wsse_scale) {
    if (wsse_scale > 0.0f) return (int64_t)std::round(orc_weighted_sse_avx2(err, w, h) * wsse_scale);
    int64_t s = 0;
    for (int i = 0; i < w * h; ++i) s += (int64_t)err[i] * err[i];
    return s;
}

}  // namespace

extern "C" int orc_cdef_frame(uint16_t* rec_y, uint16_t* rec_u, uint16_t* rec_v, const uint16_t* src_y,
                              const uint16_t* src_u, const uint16_t* src_v, int w, int h, int sub_x, int sub_y,
                              int mono, const uint8_t* skip8, int qidx, int bd, int slow, float wsse_scale,
                              float lambda, orc_cdef_params* prm, uint8_t* grid_out, int32_t* dirs_out,
                              int32_t* vars_out, int64_t* tables_out) {
    // frame dimensions are multiples of 8: the coded (mi) area is the frame
    if (w <= 0 || h <= 0 || (w & 7) || (h & 7) || bd < 8 || bd > 12) return -1;
    const int sig_damp = 3 + std::min(qidx / 64, 3);  // cdef_damping (src/coder/frame_cdef.rs:34)
    const int damping = sig_damp + (bd - 8);
    const int nbx = w / 8, nby = h / 8, uc = (w + 63) / 64, ur = (h + 63) / 64, nu = uc * ur;
    const int cw = mono ? 0 : w >> sub_x, ch = mono ? 0 : h >> sub_y;
    std::vector<int> dirs((size_t)nbx * nby), vars((size_t)nbx * nby);
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            int32_t v;
            dirs[(size_t)by * nbx + bx] = orc_cdef_dir(rec_y, w, w, h, bx * 8, by * 8, bd, &v);
            vars[(size_t)by * nbx + bx] = v;
        }
    if (dirs_out)
        for (size_t i = 0; i < dirs.size(); ++i) {
            dirs_out[i] = dirs[i];
            vars_out[i] = vars[i];
        }
    std::vector<char> sig(nu, 0);
    int n_sig = 0;
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx)
            if (!skip8[(size_t)by * nbx + bx]) sig[(by / 8) * uc + bx / 8] = 1;
    for (int u = 0; u < nu; ++u) n_sig += sig[u];
    std::memset(prm, 0, sizeof(*prm));
    prm->damping = sig_damp;
    prm->n = 1;
    if (grid_out && nu > 0) std::memset(grid_out, 0, (size_t)nu);
    if (n_sig == 0) return 0;
    const int64_t gate = 15000, margin = 40;
    // directional-variance gate
    std::vector<int64_t> dvs(nu, 0), dvc(nu, 0);
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx)
            if (!skip8[(size_t)by * nbx + bx]) {
                const int u = (by / 8) * uc + bx / 8;
                dvs[u] += vars[(size_t)by * nbx + bx];
                dvc[u] += 1;
            }
    int n_trusted = 0;
    for (int u = 0; u < nu; ++u) n_trusted += (gate == 0 || dvs[u] / std::max<int64_t>(dvc[u], 1) >= gate) ? 1 : 0;
    // per-unit distortion of every strength pair: luma, then U + V
    std::vector<int64_t> ltab((size_t)nu * 12, 0), ctab((size_t)nu * 12, 0);
    std::vector<int32_t> cand(12 * 64);
    int32_t err[64];
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            const size_t bi = (size_t)by * nbx + bx;
            if (skip8[bi]) continue;
            const int x = bx * 8, y = by * 8, u = (by / 8) * uc + bx / 8;
            orc_cdef_candidates(rec_y, w, w, h, x, y, 8, 8, dirs[bi], vars[bi], damping, bd, cand.data());
            for (int c = 0; c < 12; ++c) {
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j)
                        err[i * 8 + j] = cand[c * 64 + i * 8 + j] - (int32_t)src_y[(size_t)(y + i) * w + x + j];
                ltab[(size_t)u * 12 + c] += wsse_err(err, 8, 8, wsse_scale);
            }
            if (mono) continue;
            const int cbw = 8 >> sub_x, cbh = 8 >> sub_y, cx = x >> sub_x, cy = y >> sub_y;
            for (int pl = 0; pl < 2; ++pl) {
                const uint16_t* rp = pl ? rec_v : rec_u;
                const uint16_t* sp = pl ? src_v : src_u;
                orc_cdef_candidates(rp, cw, cw, ch, cx, cy, cbw, cbh, dirs[bi], -1, damping - 1, bd, cand.data());
                for (int c = 0; c < 12; ++c) {
                    for (int i = 0; i < cbh; ++i)
                        for (int j = 0; j < cbw; ++j)
                            err[i * cbw + j] = cand[c * cbw * cbh + i * cbw + j] - (int32_t)sp[(size_t)(cy + i) * cw + cx + j];
                    ctab[(size_t)u * 12 + c] += wsse_err(err, cbw, cbh, wsse_scale);
                }
            }
        }
    if (tables_out) {
        std::memcpy(tables_out, ltab.data(), ltab.size() * sizeof(int64_t));
        std::memcpy(tables_out + ltab.size(), ctab.data(), ctab.size() * sizeof(int64_t));
    }
    auto clears = [&](int64_t off, int64_t on) -> int64_t {
        const int64_t thr = off - off * margin / 1000;
        return on < thr ? off - on : 0;
    };
    // candidate strength pairs (index 0 = unfiltered)
    std::vector<std::pair<int, int>> cands{{0, 0}};
    if (slow) {
        const int pl[3] = {1, 2, 4};
        std::vector<std::pair<int64_t, int>> ranked;
        for (int i = 0; i < 3; ++i) {
            const int c = cand_index(pl[i], 0);
            int64_t g = 0;
            for (int u = 0; u < nu; ++u) g += clears(ltab[(size_t)u * 12], ltab[(size_t)u * 12 + c]);
            ranked.push_back({g, i});
            cands.push_back({pl[i], 0});
        }
        std::stable_sort(ranked.begin(), ranked.end(), [](const std::pair<int64_t, int>& a, const std::pair<int64_t, int>& b) {
            return a.first != b.first ? a.first > b.first : a.second < b.second;
        });
        for (int sec = 1; sec <= 2; ++sec) {
            cands.push_back({0, sec});
            for (int k = 0; k < 3; ++k)
                if (ranked[k].first > 0) cands.push_back({pl[ranked[k].second], sec});
        }
    } else {
        for (int pri : kCandPri)
            for (int sec : kCandSec)
                if (pri || sec) cands.push_back({pri, sec});
    }
    const int ny = (int)cands.size(), nc = mono ? 1 : ny;
    auto dy = [&](int i, int u) { return ltab[(size_t)u * 12 + cand_index(cands[i].first, cands[i].second)]; };
    auto dc = [&](int i, int u) -> int64_t {
        return mono ? 0 : ctab[(size_t)u * 12 + cand_index(cands[i].first, cands[i].second)];
    };
    auto dpair = [&](int p, int u) { return dy(p / nc, u) + dc(p % nc, u); };
    std::vector<int64_t> doff(nu), thr(nu);
    int64_t total_off = 0;
    for (int u = 0; u < nu; ++u) {
        doff[u] = dy(0, u) + dc(0, u);
        thr[u] = doff[u] - doff[u] * margin / 1000;
        total_off += doff[u];
    }
    const int npairs = ny * nc;
    // global entry: the raw frame-total minimizer
    int64_t prefix[4] = {INT64_MAX, INT64_MAX, INT64_MAX, INT64_MAX};
    int global_entry = 0;
    for (int p = 0; p < npairs; ++p) {
        int64_t t = 0;
        for (int u = 0; u < nu; ++u) t += sig[u] ? dpair(p, u) : doff[u];
        if (t < prefix[0]) {
            prefix[0] = t;
            global_entry = p;
        }
    }
    // greedy set construction seeded with the all-zero entry
    std::vector<int64_t> cur = doff;
    std::vector<int> entries{0};
    while (entries.size() < 8) {
        int best_p = 0;
        int64_t best_t = INT64_MAX;
        for (int p = 0; p < npairs; ++p) {
            int64_t t = 0;
            for (int u = 0; u < nu; ++u) {
                if (!sig[u]) {
                    t += doff[u];
                    continue;
                }
                const int64_t b = std::min(cur[u], dpair(p, u));
                t += b >= thr[u] ? doff[u] : b;
            }
            if (t < best_t) {
                best_t = t;
                best_p = p;
            }
        }
        entries.push_back(best_p);
        for (int u = 0; u < nu; ++u) cur[u] = std::min(cur[u], dpair(best_p, u));
        const size_t n = entries.size();
        if ((n & (n - 1)) == 0) prefix[__builtin_ctz((unsigned)n)] = best_t;
    }
    auto mag = [&](int p) {
        return cands[p / nc].first + cands[p / nc].second + (mono ? 0 : cands[p % nc].first + cands[p % nc].second);
    };
    auto assign = [&](int nb, std::vector<uint8_t>& grid) -> int64_t {
        int z = -1;
        for (int e = 0; e < nb; ++e)
            if (entries[e] == 0) {
                z = e;
                break;
            }
        int64_t total = 0;
        grid.assign(nu, 0);
        for (int u = 0; u < nu; ++u) {
            if (!sig[u]) {
                total += doff[u];
                continue;
            }
            int pm = -1, pe = 0;
            int64_t pd = 0, rb = INT64_MAX;
            int re = 0;
            for (int e = 0; e < nb; ++e) {
                const int64_t d = dpair(entries[e], u);
                if (d < rb) {
                    rb = d;
                    re = e;
                }
                if (d < thr[u]) {
                    const int m = mag(entries[e]);
                    if (pm < 0 || m < pm || (m == pm && d < pd)) {
                        pm = m;
                        pd = d;
                        pe = e;
                    }
                }
            }
            if (pm >= 0) {
                total += pd;
                grid[u] = (uint8_t)pe;
            } else if (z >= 0) {
                total += doff[u];
                grid[u] = (uint8_t)z;
            } else {
                total += rb;
                grid[u] = (uint8_t)re;
            }
        }
        return total;
    };
    const float per_entry = mono ? 6.0f : 12.0f;
    float best_cost = (float)total_off;
    int best_bits = -1;
    std::vector<uint8_t> g;
    for (int b = 0; b <= 3; ++b) {
        const int nb = 1 << b;
        const int64_t total = b == 0 ? prefix[0] : assign(nb, g);
        if (total == INT64_MAX) continue;
        if (b == 0 && (global_entry == 0 || n_trusted * 4 < n_sig * 3 || total >= total_off - total_off * margin / 1000))
            continue;
        const float cost = (float)total + lambda * (per_entry * (float)(nb - 1) + (float)b * (float)n_sig);
        if (cost < best_cost) {
            best_cost = cost;
            best_bits = b;
        }
    }
    if (best_bits < 0) return 0;
    std::vector<int> set;
    std::vector<uint8_t> grid(nu, 0);
    if (best_bits == 0) {
        set.push_back(global_entry);
    } else {
        set.assign(entries.begin(), entries.begin() + (1 << best_bits));
        assign(1 << best_bits, grid);
    }
    prm->bits = best_bits;
    prm->n = (int)set.size();
    for (size_t e = 0; e < set.size(); ++e) {
        const int p = set[e];
        prm->y_pri[e] = (uint8_t)cands[p / nc].first;
        prm->y_sec[e] = (uint8_t)cands[p / nc].second;
        prm->uv_pri[e] = (uint8_t)(mono ? 0 : cands[p % nc].first);
        prm->uv_sec[e] = (uint8_t)(mono ? 0 : cands[p % nc].second);
    }
    if (grid_out) std::memcpy(grid_out, grid.data(), (size_t)nu);
    // apply, every tap from the pre-CDEF planes
    const std::vector<uint16_t> snap_y(rec_y, rec_y + (size_t)w * h);
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            const size_t bi = (size_t)by * nbx + bx;
            if (skip8[bi]) continue;
            const int e = grid[(by / 8) * uc + bx / 8], pri = prm->y_pri[e], sec = prm->y_sec[e];
            if (!pri && !sec) continue;
            orc_cdef_filter(snap_y.data(), w, w, h, bx * 8, by * 8, 8, 8,
                            orc_cdef_adjust_pri(pri << (bd - 8), vars[bi]), sec << (bd - 8), pri ? dirs[bi] : 0,
                            damping, bd, rec_y, w);
        }
    if (!mono)
        for (int pl = 0; pl < 2; ++pl) {
            uint16_t* rp = pl ? rec_v : rec_u;
            const std::vector<uint16_t> snap(rp, rp + (size_t)cw * ch);
            const int cbw = 8 >> sub_x, cbh = 8 >> sub_y;
            for (int by = 0; by < nby; ++by)
                for (int bx = 0; bx < nbx; ++bx) {
                    const size_t bi = (size_t)by * nbx + bx;
                    if (skip8[bi]) continue;
                    const int e = grid[(by / 8) * uc + bx / 8], pri = prm->uv_pri[e], sec = prm->uv_sec[e];
                    if (!pri && !sec) continue;
                    orc_cdef_filter(snap.data(), cw, cw, ch, (bx * 8) >> sub_x, (by * 8) >> sub_y, cbw, cbh,
                                    pri << (bd - 8), sec << (bd - 8), pri ? dirs[bi] : 0, damping - 1, bd, rp, cw);
                }
        }
    return 1;
}
