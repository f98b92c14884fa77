// Test harness: mt_trellis (the trellis kernel of mt_rate.hip, with its tables) run lane by
// lane on the host through the CPU emulation of hip/hip_runtime.h.  Reads the descriptors,
// levels and targets from files, writes the rewritten levels and the per-block status.
#include "hip/hip_runtime.h"
#include "../../maroontree_amd/csrc/mt_rate.hip"

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
ProfScope::ProfScope(int slot_, hipStream_t s_) : slot(slot_), s(s_) {}
ProfScope::~ProfScope() {}
}  // namespace mtlib

template <class T> std::vector<T> load(const char* p) {
    FILE* f = fopen(p, "rb");
    if (!f) abort();
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<T> v(n / sizeof(T));
    if (fread(v.data(), sizeof(T), v.size(), f) != v.size()) abort();
    fclose(f);
    return v;
}

int main(int argc, char** argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: emu_trellis blocks levels targets out_levels out_status\n");
        return 2;
    }
    auto blocks = load<mt_trellisblk>(argv[1]);
    auto levels = load<int32_t>(argv[2]);
    auto targets = load<float>(argv[3]);
    if (mtlib::rate_init_device(0)) return 4;
    std::vector<float> work(levels.size());
    std::vector<int32_t> status(blocks.size());
    const int rc = mt_trellis(blocks.data(), (int)blocks.size(), levels.data(), targets.data(), work.data(),
                              status.data(), nullptr);
    if (rc) return 5;
    FILE* f = fopen(argv[4], "wb");
    fwrite(levels.data(), sizeof(int32_t), levels.size(), f);
    fclose(f);
    f = fopen(argv[5], "wb");
    fwrite(status.data(), sizeof(int32_t), status.size(), f);
    fclose(f);
    printf("ok %zu blocks\n", blocks.size());
    return 0;
}
