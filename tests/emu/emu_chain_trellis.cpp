// Test harness: mt_tx_chain_trellis (bucketing, the chain kernels' forward and inverse halves
// of mtgpu.hip and the trellis kernel of mt_rate.hip) run lane by lane on the host through the
// CPU emulation of hip/hip_runtime.h, on blocks / trellis descriptors / pool read from files;
// writes the per-block results.
#include "hip/hip_runtime.h"
#include "../../include/mtgpu.h"
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
int pred_init_device(int) { return MT_OK; }
int dist_init_device(int) { return MT_OK; }
int rdo_init_device(int) { return MT_OK; }
}
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
    if (argc < 5) {
        fprintf(stderr, "usage: emu_chain_trellis blocks trellis pool out\n");
        return 2;
    }
    auto blocks = load<mt_txblk>(argv[1]);
    auto tr = load<mt_trellisblk>(argv[2]);
    auto pool = load<int16_t>(argv[3]);
    if (mt_init(0)) { fprintf(stderr, "init: %s\n", mt_last_error()); return 4; }
    std::vector<mt_txres> res(blocks.size());
    std::vector<int32_t> levels(pool.size()), recon(pool.size());
    std::vector<float> targets(pool.size());
    const int rc = mt_tx_chain_trellis(blocks.data(), tr.data(), (int)blocks.size(), pool.data(), res.data(),
                                       levels.data(), targets.data(), recon.data(), nullptr);
    if (rc) { fprintf(stderr, "chain: %s\n", mt_last_error()); return 5; }
    FILE* f = fopen(argv[4], "wb");
    fwrite(res.data(), sizeof(mt_txres), res.size(), f);
    fclose(f);
    printf("ok %zu blocks, %ld launches\n", blocks.size(), emu::g_launches.load());
    return 0;
}
