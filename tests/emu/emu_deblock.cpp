// Test harness: mt_deblock_frame / mt_deblock_plane (mt_deblock.hip) run lane by lane on the
// host through the CPU emulation of hip/hip_runtime.h.  Reads the planes and maps from a
// directory, writes the filtered planes and the status word.
#include "hip/hip_runtime.h"
#include "../../maroontree_amd/csrc/mt_deblock.hip"

#include <cstdio>
#include <string>
#include <vector>

thread_local dim3 threadIdx, blockIdx;
dim3 blockDim, gridDim;
namespace emu {
Barrier* g_bar = nullptr;
uint64_t g_xch[1024];
std::atomic<long> g_launches{0};
const char* g_kernel = "";
void launch(const char* name, dim3 g, dim3 b, const std::function<void()>& body) {
    g_kernel = name;
    ++g_launches;
    blockDim = b;
    gridDim = g;
    Barrier bar;
    bar.n = (int)(b.x * b.y * b.z);
    bar.name = name;
    g_bar = &bar;
    Barrier end;
    end.n = bar.n;
    end.name = "block-end";
    std::vector<std::thread> th;
    for (unsigned t = 0; t < b.x; ++t)
        th.emplace_back([&, t] {
            threadIdx = dim3(t, 0, 0);
            for (unsigned by = 0; by < g.y; ++by)
                for (unsigned bx = 0; bx < g.x; ++bx) {
                    blockIdx = dim3(bx, by, 0);
                    body();
                    end.wait();
                }
        });
    for (auto& x : th) x.join();
    g_bar = nullptr;
}
}  // namespace emu

namespace mtlib {
int set_error(int code, const std::string& what) {
    fprintf(stderr, "error %d: %s\n", code, what.c_str());
    return code;
}
int device_ready(int* dev, int* cus) {
    if (dev) *dev = 0;
    if (cus) *cus = 4;
    return MT_OK;
}
}  // namespace mtlib

template <class T> std::vector<T> load(const std::string& p) {
    FILE* f = fopen(p.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", p.c_str());
        exit(2);
    }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<T> v(n / sizeof(T));
    if (fread(v.data(), sizeof(T), v.size(), f) != v.size()) exit(2);
    fclose(f);
    return v;
}
template <class T> void save(const std::string& p, const std::vector<T>& v) {
    FILE* f = fopen(p.c_str(), "wb");
    fwrite(v.data(), sizeof(T), v.size(), f);
    fclose(f);
}

// usage: emu_deblock DIR.  DIR/params.txt: "frame w8 h8 disp_w disp_h sub_x sub_y mono sharp bd
// lvl_y lvl_u lvl_v" or "plane w h vis_w vis_h flags level sharp luma sb_rows4 bd" (flags: 1 =
// explicit edge maps); planes y/u/v.bin (uint16), maps.bin (uint8 planes).
int main(int argc, char** argv) {
    const std::string d = argv[1];
    FILE* f = fopen((d + "/params.txt").c_str(), "r");
    char mode[16];
    if (fscanf(f, "%15s", mode) != 1) return 2;
    uint32_t status = 0;
    int rc = 0;
    if (std::string(mode) == "frame") {
        int w8, h8, dw, dh, sx, sy, mono, sharp, bd, ly, lu, lv;
        if (fscanf(f, "%d %d %d %d %d %d %d %d %d %d %d %d", &w8, &h8, &dw, &dh, &sx, &sy, &mono, &sharp, &bd, &ly,
                   &lu, &lv) != 12)
            return 2;
        auto y = load<uint16_t>(d + "/y.bin");
        auto maps = load<uint8_t>(d + "/maps.bin");
        const size_t n4 = (size_t)(w8 / 4) * (h8 / 4);
        mt_deblock_geom g{};
        g.w8 = w8;
        g.h8 = h8;
        g.disp_w = dw;
        g.disp_h = dh;
        g.blk4 = maps.data();
        g.blk4h = maps.data() + n4;
        g.blk4v = maps.data() + 2 * n4;
        g.blk4t = maps.data() + 3 * n4;
        g.pblk4 = maps.data() + 4 * n4;
        g.pblk4h = maps.data() + 5 * n4;
        g.sub_x = (uint8_t)sx;
        g.sub_y = (uint8_t)sy;
        g.mono = (uint8_t)mono;
        g.sharp = (uint8_t)sharp;
        g.bd = (uint8_t)bd;
        std::vector<uint16_t> oy(y.size()), u, v, ou, ov;
        if (!mono) {
            u = load<uint16_t>(d + "/u.bin");
            v = load<uint16_t>(d + "/v.bin");
            ou.resize(u.size());
            ov.resize(v.size());
        }
        rc = mt_deblock_frame(&g, y.data(), mono ? nullptr : u.data(), mono ? nullptr : v.data(), w8, w8 >> sx,
                              oy.data(), mono ? nullptr : ou.data(), mono ? nullptr : ov.data(), ly, lu, lv, &status,
                              nullptr);
        save(d + "/oy.bin", oy);
        if (!mono) {
            save(d + "/ou.bin", ou);
            save(d + "/ov.bin", ov);
        }
    } else {
        int w, h, vw, vh, flags, level, sharp, luma, sbr, bd;
        if (fscanf(f, "%d %d %d %d %d %d %d %d %d %d", &w, &h, &vw, &vh, &flags, &level, &sharp, &luma, &sbr, &bd) != 10)
            return 2;
        auto y = load<uint16_t>(d + "/y.bin");
        auto maps = load<uint8_t>(d + "/maps.bin");
        const size_t n4 = (size_t)(w / 4) * (h / 4);
        std::vector<uint16_t> oy(y.size());
        mt_lf_plane p{};
        p.src = y.data();
        p.dst = oy.data();
        p.stride = w;
        p.w = w;
        p.h = h;
        p.vis_w = vw;
        p.vis_h = vh;
        p.nc4 = w / 4;
        p.bw4 = maps.data();
        p.bh4 = maps.data() + n4;
        p.vedge4 = (flags & 1) ? maps.data() + 2 * n4 : nullptr;
        p.hedge4 = (flags & 1) ? maps.data() + 3 * n4 : nullptr;
        rc = mt_deblock_plane(&p, level, sharp, luma, sbr, bd, &status, nullptr);
        save(d + "/oy.bin", oy);
    }
    fclose(f);
    FILE* o = fopen((d + "/status.txt").c_str(), "w");
    fprintf(o, "%d %u\n", rc, status);
    fclose(o);
    return 0;
}
