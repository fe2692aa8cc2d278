// noalloc.hip — bld_mult's FP32-bulk solve on caller-owned device buffers
// with no device allocation: the per-instance workspace (digamma rows,
// partial axis sums, descriptors, work lists) lives in one file-scope device
// array, `scratch`, of kScratchBytes, held by each GPU the library runs on
// until the process ends. For callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The kernels are the ones bnmf_bld_mult_f32 runs, on the same tile plan, so
// the two entry points give bitwise identical S. The host never writes the
// scratch: a setup kernel lays out the descriptors and work lists from
// by-value arguments, and the choice between the exact-z pass (S0 >= 0 and
// every beta_k >= 1) and the generic pass is made on the device, by setting
// the item count of one launch to zero.
//
// Every call on a GPU shares that GPU's scratch, which keeps the last
// call's contents; calls that overlap on one GPU (different streams or
// threads) race.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "bnmf_internal.hpp"

namespace bnmf {

// Largest workspace one call may use: 32 MiB (a 2048 x 2048 x 64 instance
// needs about 25 MiB).
constexpr size_t kScratchBytes = 32ull << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

// the device address of `scratch` on each GPU, looked up once per device
void* scratch_ptr(int dev) {
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    BNMF_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (ptr[dev] == nullptr) {
        void* p = nullptr;
        BNMF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = p;
    }
    return ptr[dev];
}

// Byte layout of the scratch for one instance.
struct NoallocLayout {
    size_t desc, counts, items, prep, psib, psia, partA, psiA, pjk, ipk, bytes;
};

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

NoallocLayout noalloc_layout(const InstLayout& L) {
    NoallocLayout o{};
    size_t off = 0;
    auto take = [&](size_t n) {
        const size_t at = off;
        off = align256(off + n);
        return at;
    };
    const size_t x = L.x, y = L.y, z = L.z;
    o.desc = take(2 * sizeof(BldInst));
    o.counts = take(8 * sizeof(uint32_t));
    o.items = take((size_t)L.n_ti * L.n_tj * sizeof(PassItem));
    o.prep = take((size_t)(L.n_cj + L.nA) * sizeof(PrepItem));
    o.psib = take(y * z * 4);
    o.psia = take(x * z * 4);
    o.partA = take((size_t)L.nA * z * 4);
    o.psiA = take(z * 4);
    o.pjk = take((size_t)L.n_ti * y * z * 4);
    o.ipk = take((size_t)L.n_tj * x * z * 4);
    o.bytes = off;
    return o;
}

// counts[0]: work items of the exact-z launch, [1]: of the generic launch,
// [2]: pass items, [3]: prep items, [4]: the nonnegativity flag
enum : int { CNT_EXACT = 0, CNT_GENERIC = 1, CNT_ITEMS = 2, CNT_PREP = 3, CNT_FLAG = 4 };

__global__ void __launch_bounds__(256)
    noalloc_setup_kernel(unsigned char* __restrict__ base, NoallocLayout o, BldInst first, BldInst main,
                         uint32_t n_ti, uint32_t n_tj, uint32_t n_cj, uint32_t nA) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t == 0) {
        BldInst* d = reinterpret_cast<BldInst*>(base + o.desc);
        d[0] = first;
        d[1] = main;
        uint32_t* c = reinterpret_cast<uint32_t*>(base + o.counts);
        c[CNT_ITEMS] = n_ti * n_tj;
        c[CNT_PREP] = n_cj + nA;
        c[CNT_FLAG] = 1u;
    }
    // pass items in row-major order (item order does not change results:
    // each tile writes its own partial-sum slots)
    const uint32_t n_items = n_ti * n_tj;
    if (t < n_items) reinterpret_cast<PassItem*>(base + o.items)[t] = PassItem{0u, t / n_tj, t % n_tj, 0u};
    if (t < n_cj + nA) {
        const PrepItem p = t < n_cj ? PrepItem{0u, 0u, t, 0u} : PrepItem{0u, 1u, t - n_cj, 0u};
        reinterpret_cast<PrepItem*>(base + o.prep)[t] = p;
    }
}

// clears the flag if some beta_k < 1 or some S0 element is negative (or NaN)
__global__ void __launch_bounds__(256)
    noalloc_check_kernel(uint32_t* __restrict__ counts, const double* __restrict__ beta, uint32_t z,
                         const float* __restrict__ S0, size_t n) {
    const size_t t = (size_t)blockIdx.x * 256u + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256u;
    bool ok = true;
    for (size_t e = t; e < z; e += stride) ok = ok && beta[e] >= 1.0;
    if (S0 != nullptr)
        for (size_t e = t; e < n; e += stride) ok = ok && S0[e] >= 0.0f;
    if (!ok) counts[CNT_FLAG] = 0u;  // every writer stores the same value
}

__global__ void noalloc_counts_kernel(uint32_t* __restrict__ counts) {
    const uint32_t n = counts[CNT_ITEMS];
    const bool nonneg = counts[CNT_FLAG] != 0u;
    counts[CNT_EXACT] = nonneg ? n : 0u;
    counts[CNT_GENERIC] = nonneg ? 0u : n;
}

}  // namespace

// Solve; every pointer is device memory the caller owns.
void bld_mult_f32_noalloc(const float* X, uint32_t x, uint32_t y, uint32_t z, const double* alpha,
                          const double* beta, size_t max_iter, double eps, const float* S0, uint64_t seed,
                          float* S, hipStream_t st) {
    const InstLayout L = plan_instance(BNMF_F32, x, y, z, false);
    const NoallocLayout o = noalloc_layout(L);
    BNMF_REQUIRE(o.bytes <= kScratchBytes,
                 "instance needs " + std::to_string(o.bytes) +
                     " bytes of workspace; the no-allocation entry point holds at most 32 MiB");
    ensure_runtime();
    int dev = 0;
    BNMF_HIP(hipGetDevice(&dev));
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(dev));

    // descriptors: the first pass reads S0 (or the draw in S), later ones S
    BldInst d{};
    d.S = S;
    d.S_src = S0 != nullptr ? static_cast<const void*>(S0) : static_cast<const void*>(S);
    d.X = X;
    d.alpha = alpha;
    d.beta = beta;
    d.psib = base + o.psib;
    d.psia = base + o.psia;
    d.partA = base + o.partA;
    d.psiA = base + o.psiA;
    d.part_pjk = base + o.pjk;
    d.part_ipk = base + o.ipk;
    d.eps = eps;
    d.a = d.b = 1.0;
    d.inv_z = 1.0f / (float)z;
    d.x = x;
    d.y = y;
    d.z = z;
    d.rw = L.rw;
    d.hs = L.hs;
    d.hm = L.hm;
    d.ti_rows = L.ti_rows;
    d.tile_cols = L.tile_cols;
    d.n_ti = L.n_ti;
    d.n_tj = L.n_tj;
    d.nA = L.nA;
    d.n_cj = L.n_cj;
    uint32_t zq = 1, lg = 0;
    while (zq < z) {
        zq <<= 1;
        ++lg;
    }
    d.zq = zq;
    d.zq_log2 = lg;
    BldInst dmain = d;
    dmain.S_src = S;

    const uint32_t n_items = L.n_ti * L.n_tj, n_prep = L.n_cj + L.nA;
    const uint32_t n_setup = n_items > n_prep ? n_items : n_prep;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((n_setup + 255) / 256), dim3(256), 0, st, base, o, d, dmain,
                       L.n_ti, L.n_tj, L.n_cj, L.nA);
    BNMF_HIP(hipGetLastError());
    const size_t n_el = (size_t)x * y * z;
    if (S0 == nullptr) BNMF_HIP(launch_init_s(BNMF_F32, S, X, (size_t)x * y, z, seed, st));
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + o.counts);
    {
        const size_t blocks = (n_el + 255) / 256;
        const dim3 grid((unsigned)(blocks < 4096 ? (blocks ? blocks : 1) : 4096));
        hipLaunchKernelGGL(noalloc_check_kernel, grid, dim3(256), 0, st, counts, beta, z, S0, n_el);
        BNMF_HIP(hipGetLastError());
        hipLaunchKernelGGL(noalloc_counts_kernel, dim3(1), dim3(1), 0, st, counts);
        BNMF_HIP(hipGetLastError());
    }
    if (max_iter == 0) {
        if (S0 != nullptr) BNMF_HIP(hipMemcpyAsync(S, S0, n_el * 4, hipMemcpyDeviceToDevice, st));
        return;
    }
    const BldInst* dfirst_dev = reinterpret_cast<const BldInst*>(base + o.desc);
    const BldInst* dmain_dev = dfirst_dev + 1;
    const PassItem* items = reinterpret_cast<const PassItem*>(base + o.items);
    const PrepItem* prep = reinterpret_cast<const PrepItem*>(base + o.prep);
    // the two kernels an instance of this shape can take; a deep instance
    // (or the generic-pass override) has one
    const PassKey k_exact = pass_key(BNMF_F32, MODE_UPDATE, L, true);
    const PassKey k_gen = pass_key(BNMF_F32, MODE_UPDATE, L, false);
    const bool split = k_exact.zexact > 0;
    const PassKey k_sums = pass_key(BNMF_F32, MODE_SUMS, L, false);
    // partial axis sums of the start, then the iterations
    BNMF_HIP(launch_pass_key(k_sums, BNMF_F32, PSIM_F32, MODE_SUMS, dim3(n_items), st, dfirst_dev, items,
                             counts + CNT_ITEMS, z));
    for (size_t t = 0; t < max_iter; ++t) {
        BNMF_HIP(launch_prep(BNMF_F32, PSIM_F32, false, dim3(n_prep), st, dmain_dev, prep, counts + CNT_PREP, 1));
        const BldInst* dp = t == 0 ? dfirst_dev : dmain_dev;
        if (split) {
            BNMF_HIP(launch_pass_key(k_exact, BNMF_F32, PSIM_F32, MODE_UPDATE, dim3(n_items), st, dp, items,
                                     counts + CNT_EXACT, z));
            BNMF_HIP(launch_pass_key(k_gen, BNMF_F32, PSIM_F32, MODE_UPDATE, dim3(n_items), st, dp, items,
                                     counts + CNT_GENERIC, z));
        } else {
            BNMF_HIP(launch_pass_key(k_gen, BNMF_F32, PSIM_F32, MODE_UPDATE, dim3(n_items), st, dp, items,
                                     counts + CNT_ITEMS, z));
        }
    }
}

}  // namespace bnmf
