// Test harness: mt_intra_rdo with the RDO kernels (mt_rdo.hip), the superblock
// reconstruction (mtgpu.hip) and the rate / trellis unit (mt_rate.hip, linked alongside) run
// lane by lane on the host through the CPU emulation of hip/hip_runtime.h; the batched
// candidate evaluation goes through the oracle's orc_eval_batch (the evaluation kernels have
// their own GPU parity tests).  Prints the
// stats of both and the number of differing samples and node records.
#include "hip/hip_runtime.h"
#include "../../maroontree_amd/csrc/mt_rdo.hip"
#include "../../oracle/oracle.h"

#include <chrono>
#include <string>

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

static_assert(sizeof(mt_cand) == sizeof(orc_cand), "cand");
static_assert(sizeof(mt_candres) == sizeof(orc_candres), "candres");
static_assert(sizeof(mt_plane) == sizeof(orc_plane), "plane");

int eval_launch(const mt_cand* cands_d, int n, const PlaneSet& ps, uint64_t pool_elems, mt_candres* res_d,
                hipStream_t, bool) {
    (void)pool_elems;
    orc_eval_batch((const orc_cand*)cands_d, n, (const orc_plane*)ps.p, ps.n, (orc_candres*)res_d, 8);
    return MT_OK;
}

// the oracle's evaluation keeps no workspace
int eval_reserve(size_t, size_t, hipStream_t, bool) { return MT_OK; }

}  // namespace mtlib

int main(int argc, char** argv) {
    if (argc < 8) {
        fprintf(stderr, "usage: emu_rdo img.raw W H bd qidx qm flags [lambda]\n");
        return 2;
    }
    const int W = atoi(argv[2]), H = atoi(argv[3]), bd = atoi(argv[4]), q = atoi(argv[5]), qm = atoi(argv[6]),
              flags = atoi(argv[7]);
    const float lam = argc > 8 ? (float)atof(argv[8]) : orc_rdo_lambda(q, bd, 2);
    std::vector<uint16_t> img((size_t)W * H);
    FILE* fp = fopen(argv[1], "rb");
    if (!fp || fread(img.data(), 2, img.size(), fp) != img.size()) return 3;
    fclose(fp);
    if (mt_init(0)) {
        fprintf(stderr, "init failed\n");
        return 4;
    }
    const int nsb = ((W + 63) / 64) * ((H + 63) / 64);
    std::vector<uint16_t> rec((size_t)W * H);
    std::vector<mt_rdo_node> nodes((size_t)nsb * 341);
    mt_rdo_frame fr{};
    fr.src = mt_plane{img.data(), W, W, H, W, H, 0};
    fr.recon = rec.data();
    fr.recon_stride = W;
    fr.nodes = nodes.data();
    fr.bd = bd;
    fr.qidx = q;
    fr.qm_level = qm;
    fr.lambda = lam;
    fr.wsse_scale = 1.0f;
    mt_rdo_stats st{};
    std::atomic<bool> done{false};
    std::thread wd([&] {
        int k = 0;
        while (!done) {
            std::this_thread::sleep_for(std::chrono::milliseconds(100));
            if (++k % 50 == 0) fprintf(stderr, "[watch] %ld launches, in %s\n", emu::g_launches.load(), emu::g_kernel);
        }
    });
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = mt_intra_rdo(&fr, flags, &st, nullptr);
    done = true;
    wd.join();
    printf("rc %d launches %ld %.1fs bits %.6f sse %lld cands %lld leaves %d\n", rc, emu::g_launches.load(),
           std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), st.bits, (long long)st.sse,
           (long long)st.candidates, st.leaves);
    orc_rdo_params p{W, H, bd, q, qm, flags, lam, 1.0f, 0, 0, 0, 15};
    std::vector<uint16_t> orec((size_t)W * H);
    std::vector<orc_rdo_node> onodes((size_t)nsb * 341);
    orc_rdo_stats os{};
    orc_intra_rdo(img.data(), W, &p, nullptr, orec.data(), onodes.data(), nullptr, &os, 8);
    printf("oracle bits %.6f sse %lld cands %lld leaves %d\n", os.bits, (long long)os.sse, (long long)os.candidates,
           os.leaves);
    size_t drec = 0, dn = 0;
    for (size_t i = 0; i < rec.size(); ++i) drec += rec[i] != orec[i];
    for (size_t i = 0; i < nodes.size(); ++i) {
        const mt_rdo_node& a = nodes[i];
        const orc_rdo_node& b = onodes[i];
        if (a.split != b.split || a.mode != b.mode || memcmp(&a.cost, &b.cost, 4) ||
            memcmp(&a.none_cost, &b.none_cost, 4)) {
            if (dn < 5)
                printf("node %zu: split %d/%d mode %d/%d cost %g/%g\n", i, a.split, b.split, a.mode, b.mode, a.cost,
                       b.cost);
            ++dn;
        }
    }
    printf("recon diffs %zu node diffs %zu\n", drec, dn);
    return 0;
}
