// stream_bw — the HBM bandwidth a streaming kernel reaches on this GPU: the
// practical ceiling the update pass is compared with (the pass moves S once
// in and once out per iteration, like a copy).
//
// Four kernels over 1 GiB buffers, each launched with one 16-byte chunk per
// lane and four chunks per thread (one pass over the buffer, no grid-stride
// loop), timed with HIP events over 20 launches after 3 untimed ones:
//   copy_nt     non-temporal loads and stores (the pass's S access pattern)
//   copy        plain loads and stores
//   read_nt     non-temporal loads only (a sum that is never stored)
//   write_nt    non-temporal stores only
// Prints one JSON line of GB/s (bytes read + bytes written per second).

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

typedef float V4 __attribute__((ext_vector_type(4)));
constexpr int kThreads = 256;
constexpr int kPerThread = 4;

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "stream_bw: %s (line %d)\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

template <bool NT>
__global__ void __launch_bounds__(kThreads) copy_kernel(const V4* __restrict__ src, V4* __restrict__ dst,
                                                        size_t n) {
    const size_t base = (size_t)blockIdx.x * kThreads * kPerThread + threadIdx.x;
    V4 v[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        const size_t i = base + (size_t)u * kThreads;
        if (i < n) v[u] = NT ? __builtin_nontemporal_load(&src[i]) : src[i];
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        const size_t i = base + (size_t)u * kThreads;
        if (i < n) {
            if (NT)
                __builtin_nontemporal_store(v[u], &dst[i]);
            else
                dst[i] = v[u];
        }
    }
}

__global__ void __launch_bounds__(kThreads) read_kernel(const V4* __restrict__ src, float* __restrict__ sink,
                                                        size_t n) {
    const size_t base = (size_t)blockIdx.x * kThreads * kPerThread + threadIdx.x;
    float acc = 0.0f;
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        const size_t i = base + (size_t)u * kThreads;
        if (i < n) {
            const V4 v = __builtin_nontemporal_load(&src[i]);
            acc += (v.x + v.y) + (v.z + v.w);
        }
    }
    // the buffers hold small integers: the sum is never negative, so the
    // store never happens, but the loads cannot be dropped
    if (acc < 0.0f) sink[threadIdx.x] = acc;
}

__global__ void __launch_bounds__(kThreads) write_kernel(V4* __restrict__ dst, size_t n) {
    const size_t base = (size_t)blockIdx.x * kThreads * kPerThread + threadIdx.x;
    const V4 v = {1.0f, 2.0f, 3.0f, 4.0f};
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        const size_t i = base + (size_t)u * kThreads;
        if (i < n) __builtin_nontemporal_store(v, &dst[i]);
    }
}

}  // namespace

int main() {
    const size_t bytes = 1ull << 30;
    const size_t n = bytes / sizeof(V4);
    const dim3 grid((unsigned)((n + (size_t)kThreads * kPerThread - 1) / ((size_t)kThreads * kPerThread)));
    V4 *src = nullptr, *dst = nullptr;
    float* sink = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMalloc(&sink, kThreads * sizeof(float)));
    CHECK(hipMemset(src, 0, bytes));
    CHECK(hipMemset(dst, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    auto rate = [&](double moved, auto launch) {
        for (int w = 0; w < 3; ++w) launch();
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, nullptr));
        for (int r = 0; r < 20; ++r) launch();
        CHECK(hipEventRecord(e1, nullptr));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        return moved * 20.0 / (ms * 1e-3) / 1e9;
    };
    const double copy_nt = rate(2.0 * bytes, [&] {
        hipLaunchKernelGGL(copy_kernel<true>, grid, dim3(kThreads), 0, nullptr, src, dst, n);
    });
    const double copy = rate(2.0 * bytes, [&] {
        hipLaunchKernelGGL(copy_kernel<false>, grid, dim3(kThreads), 0, nullptr, src, dst, n);
    });
    const double rd = rate(1.0 * bytes, [&] {
        hipLaunchKernelGGL(read_kernel, grid, dim3(kThreads), 0, nullptr, src, sink, n);
    });
    const double wr = rate(1.0 * bytes, [&] {
        hipLaunchKernelGGL(write_kernel, grid, dim3(kThreads), 0, nullptr, dst, n);
    });
    std::printf("{\"bytes\": %zu, \"copy_nt_GBps\": %.1f, \"copy_GBps\": %.1f, \"read_nt_GBps\": %.1f, "
                "\"write_nt_GBps\": %.1f}\n",
                bytes, copy_nt, copy, rd, wr);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    CHECK(hipFree(sink));
    return 0;
}
