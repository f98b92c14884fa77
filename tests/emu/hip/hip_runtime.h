// Test harness: a CPU emulation of the HIP subset the library's kernels use, so that the
// device sources compile with g++ and run lane by lane on the host.  One std::thread per
// lane, the blocks of a launch run one after another (__shared__ becomes static storage),
// __syncthreads is a barrier with a deadlock watchdog, shuffles and ballots exchange
// through a shared array, device memory is host memory.  __mul24 checks that its operands
// fit in 24 bits (the range the library's inverse transforms rely on).
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>
#include <cmath>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __constant__
#define __launch_bounds__(...)
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) alignas(n)

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
extern thread_local dim3 threadIdx, blockIdx;
extern dim3 blockDim, gridDim;

typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef int hipError_t;
enum { hipSuccess = 0 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
inline const char* hipGetErrorString(hipError_t) { return "emu error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipMalloc(void* p, size_t n) {
    *(void**)p = calloc(1, n);
    return hipSuccess;
}
template <class T> inline hipError_t hipMalloc(T** p, size_t n) {
    *p = (T*)calloc(1, n);
    return hipSuccess;
}
inline hipError_t hipFree(void* p) {
    free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// Graph capture is not emulated: every graph call fails, so a run with MT_RDO_GRAPH reports an
// error instead of skipping its work.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum { hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeRelaxed = 2 };
constexpr hipError_t kEmuNoGraphs = 801;
inline hipError_t hipStreamCreateWithFlags(hipStream_t*, unsigned) { return kEmuNoGraphs; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return kEmuNoGraphs; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return kEmuNoGraphs; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, char*, size_t) { return kEmuNoGraphs; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return kEmuNoGraphs; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return kEmuNoGraphs; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return kEmuNoGraphs; }
// host execution never captures a stream
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* st) {
    *st = hipStreamCaptureStatusNone;
    return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = (void*)sym;
    return hipSuccess;
}
inline hipError_t hipMemcpyToSymbol(const void* sym, const void* src, size_t n, size_t off = 0) {
    memcpy((char*)sym + off, src, n);
    return hipSuccess;
}

namespace emu {
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int n = 0, count = 0;
    long gen = 0;
    const char* name = "";
    void wait() {
        std::unique_lock<std::mutex> lk(m);
        const long g = gen;
        if (++count == n) {
            count = 0;
            ++gen;
            cv.notify_all();
            return;
        }
        if (!cv.wait_for(lk, std::chrono::seconds(20), [&] { return gen != g; })) {
            fprintf(stderr, "EMU DEADLOCK in %s block (%u,%u) lane %u: %d of %d at barrier\n", name, blockIdx.x,
                    blockIdx.y, threadIdx.x, count, n);
            fflush(stderr);
            abort();
        }
    }
};
extern Barrier* g_bar;
extern uint64_t g_xch[1024];
extern std::atomic<long> g_launches;
extern const char* g_kernel;
void launch(const char* name, dim3 g, dim3 b, const std::function<void()>& body);
}  // namespace emu

#define hipLaunchKernelGGL(k, g, b, shm, st, ...) emu::launch(#k, dim3(g), dim3(b), [&] { k(__VA_ARGS__); })

inline void __syncthreads() { emu::g_bar->wait(); }
// the emulated kernels that use wave-level ordering run one wave per block
inline void __builtin_amdgcn_wave_barrier() { emu::g_bar->wait(); }
#define __builtin_amdgcn_fence(...) ((void)0)

template <class T> inline T emu_shfl(T v, int src_fn(int, int), int arg) {
    const int lane = threadIdx.x;
    uint64_t u = 0;
    memcpy(&u, &v, sizeof(T));
    emu::g_xch[lane] = u;
    emu::g_bar->wait();
    const int base = lane & ~63;
    const int s = src_fn(lane & 63, arg);
    T r;
    uint64_t w = emu::g_xch[base + (s & 63)];
    memcpy(&r, &w, sizeof(T));
    emu::g_bar->wait();
    return r;
}
inline int emu_xor(int l, int o) { return l ^ o; }
inline int emu_up(int l, int o) { return l >= o ? l - o : l; }
inline int emu_idx(int, int s) { return s; }
template <class T> inline T __shfl_xor(T v, int o) { return emu_shfl(v, emu_xor, o); }
template <class T> inline T __shfl_up(T v, unsigned o) { return emu_shfl(v, emu_up, (int)o); }
template <class T> inline T __shfl(T v, int s) { return emu_shfl(v, emu_idx, s); }
inline int __builtin_amdgcn_readfirstlane(int v) { return emu_shfl(v, emu_idx, 0); }
inline uint64_t __ballot(int pred) {
    const int lane = threadIdx.x;
    emu::g_xch[lane] = pred ? 1 : 0;
    emu::g_bar->wait();
    uint64_t m = 0;
    const int base = lane & ~63;
    for (int i = 0; i < 64 && base + i < (int)blockDim.x; ++i)
        if (emu::g_xch[base + i]) m |= 1ull << i;
    emu::g_bar->wait();
    return m;
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
struct int4 { int x, y, z, w; };
struct float2 { float x, y; };
inline float2 make_float2(float x, float y) { return float2{x, y}; }
struct float4 { float x, y, z, w; };
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fmul_rn(float a, float b) { return a * b; }
inline double __dadd_rn(double a, double b) { return a + b; }
inline double __dmul_rn(double a, double b) { return a * b; }
inline uint32_t __float_as_uint(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
inline int __float_as_int(float f) {
    int u;
    memcpy(&u, &f, 4);
    return u;
}
inline float __uint_as_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
inline float __int_as_float(int u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const unsigned long long o = *p;
    if (v < o) *p = v;
    return o;
}
inline unsigned atomicAdd(unsigned* p, unsigned v) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const unsigned o = *p;
    *p = o + v;
    return o;
}
inline int atomicAdd(int* p, int v) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const int o = *p;
    *p = o + v;
    return o;
}
inline unsigned atomicOr(unsigned* p, unsigned v) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const unsigned o = *p;
    *p = o | v;
    return o;
}
struct uint2 { unsigned x, y; };
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline float __fsub_rn(float a, float b) { return a - b; }
inline float __fdiv_rn(float a, float b) { return a / b; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const unsigned long long o = *p;
    *p = o + v;
    return o;
}
struct hipDeviceProp_t {
    int multiProcessorCount = 8;
    char gcnArchName[256] = "gfx950";
};
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    *p = hipDeviceProp_t{};
    return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = nullptr;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
    *ms = 0;
    return hipSuccess;
}
inline int __mul24(int a, int b) {
    if (a < -(1 << 23) || a >= (1 << 23) || b < -(1 << 23) || b >= (1 << 23)) {
        fprintf(stderr, "MUL24 OUT OF RANGE %d * %d\n", a, b);
        fflush(stderr);
        abort();
    }
    return (int)(uint32_t)((int64_t)a * (int64_t)b);
}
inline unsigned __umul24(unsigned a, unsigned b) { return (unsigned)((uint64_t)(a & 0xFFFFFFu) * (b & 0xFFFFFFu)); }
