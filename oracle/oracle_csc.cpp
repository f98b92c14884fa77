// Oracle (test infrastructure only; the library never links it): the reference's RGB ->
// YCbCr conversion and quality map on the CPU, for checking mt_encode_rgb8.
//   orc_csc_rgb8      encode_still_lossy_420_with_cdf (src/encoder.rs:530-605) and
//                     encode_still_lossy_with_cdf (:398-410): Q13 BT.601 full range with
//                     the constants of :114-127; 4:2:0 chroma from the 2 x 2 sums of the Q13
//                     values with the last row / column repeated past an odd size
//                     (:590-601)
//   orc_quality_to_q  quality_to_q (src/avif.rs:637-666)
#include <algorithm>
#include <cstdint>
#include <vector>

#include "oracle.h"

namespace {
constexpr int32_t Q = 13, HALF = 1 << (Q - 1);
constexpr int32_t Y_R = 2449, Y_G = 4809, Y_B = 934;
constexpr int32_t CB_R = -1382, CB_G = -2714, CB_B = 4096;
constexpr int32_t CR_R = 4096, CR_G = -3430, CR_B = -666;
}  // namespace

extern "C" void orc_csc_rgb8(const uint8_t* r, const uint8_t* g, const uint8_t* b, int w, int h, int c420,
                             uint16_t* y, uint16_t* cb, uint16_t* cr) {
    const int32_t mx = 255, off_q = 128 << Q;
    std::vector<int32_t> fcb(c420 ? (size_t)w * h : 0), fcr(c420 ? (size_t)w * h : 0);
    for (size_t j = 0; j < (size_t)w * h; ++j) {
        const int32_t ri = r[j], gi = g[j], bi = b[j];
        y[j] = (uint16_t)std::clamp((Y_R * ri + Y_G * gi + Y_B * bi + HALF) >> Q, 0, mx);
        const int32_t cbq = CB_R * ri + CB_G * gi + CB_B * bi + off_q;
        const int32_t crq = CR_R * ri + CR_G * gi + CR_B * bi + off_q;
        if (c420) {
            fcb[j] = cbq;
            fcr[j] = crq;
        } else {
            cb[j] = (uint16_t)std::clamp((cbq + HALF) >> Q, 0, mx);
            cr[j] = (uint16_t)std::clamp((crq + HALF) >> Q, 0, mx);
        }
    }
    if (!c420) return;
    const int cw = (w + 1) / 2, ch = (h + 1) / 2;
    const int32_t half_avg = 1 << (Q + 1);
    for (int row = 0; row < ch; ++row)
        for (int c = 0; c < cw; ++c) {
            const int x0 = 2 * c, x1 = std::min(2 * c + 1, w - 1);
            const int y0 = 2 * row, y1 = std::min(2 * row + 1, h - 1);
            auto avg = [&](const std::vector<int32_t>& f) {
                return f[(size_t)y0 * w + x0] + f[(size_t)y0 * w + x1] + f[(size_t)y1 * w + x0] + f[(size_t)y1 * w + x1];
            };
            cb[(size_t)row * cw + c] = (uint16_t)std::clamp((avg(fcb) + half_avg) >> (Q + 2), 0, mx);
            cr[(size_t)row * cw + c] = (uint16_t)std::clamp((avg(fcr) + half_avg) >> (Q + 2), 0, mx);
        }
}

extern "C" int orc_quality_to_q(int quality) {
    static const int anchors[14][2] = {{1, 215}, {30, 140}, {40, 116}, {50, 91}, {60, 70}, {70, 53}, {80, 35},
                                       {85, 28}, {90, 17},  {93, 12},  {96, 7},  {98, 5},  {99, 3},  {100, 1}};
    const int q = std::clamp(quality, 1, 100);
    int pl = anchors[0][0], pq = anchors[0][1];
    for (int k = 1; k < 14; ++k) {
        const int l = anchors[k][0], qi = anchors[k][1];
        if (q <= l) {
            const int t = q - pl;
            return std::clamp(pq + (qi - pq) * t / std::max(l - pl, 1), 1, 255);
        }
        pl = l;
        pq = qi;
    }
    return 1;
}
