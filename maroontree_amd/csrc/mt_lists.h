// Work lists by class: a batch's items sorted into per-class index lists on the device, so
// each class runs in its own kernel (one thread per item for the small block shapes, one
// wave per item for the large ones).  The order within a list is not fixed; every result is
// written at its item's own index, so outputs do not depend on it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mtlib {

// One workgroup of 256 threads per contiguous chunk of items.  cls(i) returns the item's
// class in [0, K) or -1 (no list); it may also record side results, and is evaluated twice.
// lists: K lists of `cap` entries; cnt: K counters, zero on entry.
template <int K, typename ClassFn>
__device__ void classify_chunk(int n, ClassFn cls, uint32_t* lists, size_t cap, uint32_t* cnt) {
    __shared__ uint32_t s_cnt[K], s_base[K];
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int k = threadIdx.x; k < K; k += blockDim.x) s_cnt[k] = 0;
    __syncthreads();
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int c = cls(i, true);
        if (c >= 0) atomicAdd(&s_cnt[c], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        s_base[k] = s_cnt[k] ? atomicAdd(&cnt[k], s_cnt[k]) : 0u;
        s_cnt[k] = 0;
    }
    __syncthreads();
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int c = cls(i, false);
        if (c >= 0) lists[(size_t)c * cap + s_base[c] + atomicAdd(&s_cnt[c], 1u)] = (uint32_t)i;
    }
}

template <int K>
__global__ void zero_counts_kernel(uint32_t* cnt) {
    for (int k = threadIdx.x; k < K; k += blockDim.x) cnt[k] = 0;
}

// Shape classes of the thread-per-item paths: 4x4, 4x8, 8x4, 8x8, 4x16, 16x4; else 6.
constexpr int kSmallShapes = 6;
__host__ __device__ inline int small_shape_class(int w, int h) {
    switch (w * 100 + h) {
        case 404: return 0;
        case 408: return 1;
        case 804: return 2;
        case 808: return 3;
        case 416: return 4;
        case 1604: return 5;
        default: return kSmallShapes;
    }
}

}  // namespace mtlib
