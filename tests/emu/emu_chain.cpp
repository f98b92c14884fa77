// Test harness: mt_tx_chain (bucket and chain kernels of mtgpu.hip, linked with the rate /
// trellis unit) run lane by lane on the host through the CPU emulation of
// hip/hip_runtime.h, on blocks / pool read from files; writes the per-block results.
#include "hip/hip_runtime.h"
#include "../../include/mtgpu.h"
#include <vector>
#include <cstdio>
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
int pred_init_device(int) { return MT_OK; }
int dist_init_device(int) { return MT_OK; }
int rdo_init_device(int) { return MT_OK; }
}
template <class T> std::vector<T> load(const char* p) {
    FILE* f = fopen(p, "rb");
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<T> v(n / sizeof(T));
    if (fread(v.data(), sizeof(T), v.size(), f) != v.size()) abort();
    fclose(f);
    return v;
}
int main(int argc, char** argv) {
    auto blocks = load<mt_txblk>(argv[1]);
    auto pool = load<int16_t>(argv[2]);
    if (mt_init(0)) { fprintf(stderr, "init: %s\n", mt_last_error()); return 4; }
    std::vector<mt_txres> res(blocks.size());
    const int rc = mt_tx_chain(blocks.data(), (int)blocks.size(), pool.data(), res.data(), nullptr, nullptr, nullptr, nullptr);
    if (rc) { fprintf(stderr, "chain: %s\n", mt_last_error()); return 5; }
    FILE* f = fopen(argv[3], "wb");
    fwrite(res.data(), sizeof(mt_txres), res.size(), f);
    fclose(f);
    printf("ok %zu blocks, %ld launches\n", blocks.size(), emu::g_launches.load());
    return 0;
}
