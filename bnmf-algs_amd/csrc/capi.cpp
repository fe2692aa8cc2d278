// capi.cpp — extern "C" entry points of libbnmf_hip.so (declared in
// include/bnmf_hip.h). Exceptions never cross the boundary: each call maps
// bnmf::Error to its status and records the message for bnmf_last_error().

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bnmf_internal.hpp"

struct bnmf_batch {
    std::unique_ptr<bnmf::Batch> impl;
};

struct bnmf_appr {
    std::unique_ptr<bnmf::ApprBatch> impl;
};

namespace bnmf {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
}  // namespace bnmf

namespace {

template <typename F>
bnmf_status guarded(F&& f) {
    try {
        bnmf::clear_last_error();
        f();
        return BNMF_OK;
    } catch (const bnmf::Error& e) {
        bnmf::set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        bnmf::set_last_error("host allocation failed");
        return BNMF_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        bnmf::set_last_error(e.what());
        return BNMF_ERR_DEVICE;
    }
}

size_t esize(int dtype) { return dtype == BNMF_F32 ? 4 : 8; }

void check_dtype(int dtype) {
    BNMF_REQUIRE(dtype == BNMF_F32 || dtype == BNMF_F64, "dtype must be BNMF_F32 or BNMF_F64");
}

// check_bld_params (src/bld/util_details.hpp:20-28): X nonnegative; the
// alpha/beta lengths are x and z by construction of this ABI. Like the
// reference, no bound on z: z = 0 is an empty instance (S has no elements).
template <typename T>
void check_params(const T* X, size_t x, size_t y, size_t z, const double* alpha,
                  const double* beta) {
    BNMF_REQUIRE(x * y == 0 || X != nullptr, "X is null");
    BNMF_REQUIRE(x == 0 || alpha != nullptr, "alpha is null");
    BNMF_REQUIRE(z == 0 || beta != nullptr, "beta is null");
    BNMF_REQUIRE(x < (1u << 31) && y < (1u << 31) && z < (1u << 31), "x, y and z must be below 2^31");
    for (size_t t = 0; t < x * y; ++t)
        BNMF_REQUIRE(X[t] >= T(0), "X must be nonnegative");
}

int psi_mode(int dtype, int use_psi_appr) {
    if (dtype == BNMF_F32) return bnmf::PSIM_F32;
    return use_psi_appr ? bnmf::PSIM_F64_APPR : bnmf::PSIM_F64_EXACT;
}

template <typename T>
void solve_one(int dtype, const T* X, size_t x, size_t y, size_t z, const double* alpha,
               const double* beta, size_t max_iter, int use_psi_appr, double eps, const T* S0,
               uint64_t seed, T* S_out, int algo = BNMF_ALGO_MULT) {
    check_params(X, x, y, z, alpha, beta);
    BNMF_REQUIRE(x * y * z == 0 || S_out != nullptr, "S_out is null");
    if (x * y * z == 0) return;
    bnmf::InstSpec s;
    s.x = (uint32_t)x;
    s.y = (uint32_t)y;
    s.z = (uint32_t)z;
    s.X = X;
    s.alpha = alpha;
    s.beta = beta;
    s.S0 = S0;
    s.seed = seed;
    bnmf::Batch b(dtype, psi_mode(dtype, use_psi_appr), eps, {s}, algo);
    b.solve_async(max_iter);
    b.synchronize();
    b.get_S(0, S_out);
}

template <typename T>
void appr_one(int dtype, const T* X, size_t x, size_t y, size_t z, const double* alpha,
                     const double* beta, size_t max_iter, double eps, const T* nu0, const T* mu0,
                     uint64_t seed, T* S_out, T* nu_out, T* mu_out) {
    check_params(X, x, y, z, alpha, beta);
    // z = 0: nu, mu and S have no elements
    if (z == 0) return;
    // an empty X (x*y = 0) runs no iteration; the factors come back as
    // initialised, as they do for an empty instance of a batch
    bnmf::ApprSpec s;
    s.x = (uint32_t)x;
    s.y = (uint32_t)y;
    s.z = (uint32_t)z;
    s.X = X;
    s.alpha = alpha;
    s.beta = beta;
    s.nu0 = nu0;
    s.mu0 = mu0;
    s.seed = seed;
    bnmf::ApprBatch a(dtype, eps, {s});
    a.solve_async(max_iter);
    a.get(0, S_out, nu_out, mu_out);
}

}  // namespace

extern "C" {

const char* bnmf_last_error(void) { return bnmf::g_last_error.c_str(); }

int bnmf_abi_version(void) { return 100; }

bnmf_status bnmf_init(int device) {
    return guarded([&] {
        int n = 0;
        BNMF_HIP(hipGetDeviceCount(&n));
        BNMF_REQUIRE(device >= 0 && device < n, "no such HIP device");
        BNMF_HIP(hipSetDevice(device));
        bnmf::ensure_runtime();
    });
}

bnmf_status bnmf_bld_mult_f64(const double* X, size_t x, size_t y, size_t z,
                              const double* alpha, const double* beta, size_t max_iter,
                              int use_psi_appr, double eps, const double* S0, uint64_t seed,
                              double* S_out) {
    return guarded([&] {
        solve_one<double>(BNMF_F64, X, x, y, z, alpha, beta, max_iter, use_psi_appr, eps, S0,
                          seed, S_out);
    });
}

bnmf_status bnmf_bld_mult_f32(const float* X, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, size_t max_iter, int use_psi_appr, double eps,
                              const float* S0, uint64_t seed, float* S_out) {
    return guarded([&] {
        solve_one<float>(BNMF_F32, X, x, y, z, alpha, beta, max_iter, use_psi_appr, eps, S0,
                         seed, S_out);
    });
}

bnmf_status bnmf_bld_mult_f32_noalloc(const float* X, size_t x, size_t y, size_t z, const double* alpha,
                                      const double* beta, size_t max_iter, int use_psi_appr, double eps,
                                      const float* S0, uint64_t seed, float* S_out, void* stream) {
    (void)use_psi_appr;  // the FP32 form has one digamma, as in bnmf_bld_mult_f32
    return guarded([&] {
        BNMF_REQUIRE(x < (1u << 31) && y < (1u << 31) && z < (1u << 31), "x, y and z must be below 2^31");
        if (x * y * z == 0) return;
        BNMF_REQUIRE(X && alpha && beta && S_out, "null argument");
        BNMF_REQUIRE((((uintptr_t)X | (uintptr_t)S0 | (uintptr_t)S_out) & 15u) == 0,
                     "X, S0 and S_out must be 16-byte aligned device buffers");
        BNMF_REQUIRE((((uintptr_t)alpha | (uintptr_t)beta) & 7u) == 0, "alpha and beta must be 8-byte aligned");
        bnmf::bld_mult_f32_noalloc(X, (uint32_t)x, (uint32_t)y, (uint32_t)z, alpha, beta, max_iter, eps, S0, seed,
                                   S_out, static_cast<hipStream_t>(stream));
    });
}

bnmf_status bnmf_bld_add_f64(const double* X, size_t x, size_t y, size_t z, const double* alpha,
                             const double* beta, size_t max_iter, double eps, const double* S0,
                             uint64_t seed, double* S_out) {
    return guarded([&] {
        solve_one<double>(BNMF_F64, X, x, y, z, alpha, beta, max_iter, 1, eps, S0, seed, S_out,
                          BNMF_ALGO_ADD);
    });
}

bnmf_status bnmf_bld_add_f32(const float* X, size_t x, size_t y, size_t z, const double* alpha,
                             const double* beta, size_t max_iter, double eps, const float* S0,
                             uint64_t seed, float* S_out) {
    return guarded([&] {
        solve_one<float>(BNMF_F32, X, x, y, z, alpha, beta, max_iter, 1, eps, S0, seed, S_out,
                         BNMF_ALGO_ADD);
    });
}

bnmf_status bnmf_appr_create(bnmf_appr** out, int dtype, const bnmf_appr_desc* descs, size_t n, double eps) {
    return guarded([&] {
        BNMF_REQUIRE(out != nullptr, "out is null");
        *out = nullptr;
        check_dtype(dtype);
        BNMF_REQUIRE(n == 0 || descs != nullptr, "descs is null");
        std::vector<bnmf::ApprSpec> specs(n);
        for (size_t t = 0; t < n; ++t) {
            const bnmf_appr_desc& d = descs[t];
            if (dtype == BNMF_F32)
                check_params(static_cast<const float*>(d.X), d.x, d.y, d.z, d.alpha, d.beta);
            else
                check_params(static_cast<const double*>(d.X), d.x, d.y, d.z, d.alpha, d.beta);
            bnmf::ApprSpec& s = specs[t];
            s.x = (uint32_t)d.x;
            s.y = (uint32_t)d.y;
            s.z = (uint32_t)d.z;
            s.X = d.X;
            s.alpha = d.alpha;
            s.beta = d.beta;
            s.nu0 = d.nu0;
            s.mu0 = d.mu0;
            s.seed = d.seed;
        }
        auto a = std::make_unique<bnmf_appr>();
        a->impl = std::make_unique<bnmf::ApprBatch>(dtype, eps, specs);
        *out = a.release();
    });
}

bnmf_status bnmf_appr_solve(bnmf_appr* a, size_t max_iter) {
    return guarded([&] {
        BNMF_REQUIRE(a != nullptr, "handle is null");
        a->impl->solve_async(max_iter);
        a->impl->synchronize();
    });
}

bnmf_status bnmf_appr_get(bnmf_appr* a, size_t idx, void* S_out, void* nu_out, void* mu_out) {
    return guarded([&] {
        BNMF_REQUIRE(a != nullptr, "handle is null");
        a->impl->get(idx, S_out, nu_out, mu_out);
    });
}

bnmf_status bnmf_appr_time_iteration(bnmf_appr* a, int reps, double* iter_ms) {
    return guarded([&] {
        BNMF_REQUIRE(a != nullptr && iter_ms != nullptr, "null argument");
        *iter_ms = a->impl->time_iteration(reps);
    });
}

void bnmf_appr_destroy(bnmf_appr* a) { delete a; }

bnmf_status bnmf_bld_appr_f64(const double* X, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, size_t max_iter, double eps, const double* nu0,
                              const double* mu0, uint64_t seed, double* S_out, double* nu_out,
                              double* mu_out) {
    return guarded([&] {
        appr_one<double>(BNMF_F64, X, x, y, z, alpha, beta, max_iter, eps, nu0, mu0, seed, S_out, nu_out,
                         mu_out);
    });
}

bnmf_status bnmf_bld_appr_f32(const float* X, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, size_t max_iter, double eps, const float* nu0,
                              const float* mu0, uint64_t seed, float* S_out, float* nu_out,
                              float* mu_out) {
    return guarded([&] {
        appr_one<float>(BNMF_F32, X, x, y, z, alpha, beta, max_iter, eps, nu0, mu0, seed, S_out, nu_out,
                        mu_out);
    });
}

bnmf_status bnmf_log_marginal_S(int dtype, const void* S, size_t x, size_t y, size_t z,
                                const double* alpha, const double* beta, double a, double b,
                                double* out) {
    return guarded([&] {
        check_dtype(dtype);
        BNMF_REQUIRE(out != nullptr, "out is null");
        BNMF_REQUIRE(x * y * z == 0 || S != nullptr, "S is null");
        BNMF_REQUIRE(x == 0 || alpha != nullptr, "alpha is null");
        BNMF_REQUIRE(z == 0 || beta != nullptr, "beta is null");
        BNMF_REQUIRE(x < (1u << 31) && y < (1u << 31) && z < (1u << 31), "x, y and z must be below 2^31");
        // X is not used by the scoring pass; a zero matrix satisfies the
        // instance layout.
        std::vector<char> Xz(x * y * esize(dtype), 0);
        bnmf::InstSpec s;
        s.x = (uint32_t)x;
        s.y = (uint32_t)y;
        s.z = (uint32_t)z;
        s.X = Xz.data();
        s.alpha = alpha;
        s.beta = beta;
        s.a = a;
        s.b = b;
        s.S0 = S;
        bnmf::Batch bt(dtype, psi_mode(dtype, 0), 1e-50, {s});
        bt.log_marginal(out);
    });
}

bnmf_status bnmf_tensor_sums(int dtype, const void* S, size_t x, size_t y, size_t z, void* S_pjk,
                             void* S_ipk, void* S_ijp) {
    return guarded([&] {
        check_dtype(dtype);
        BNMF_REQUIRE(x * y * z == 0 || S != nullptr, "S is null");
        BNMF_REQUIRE(x < (1u << 31) && y < (1u << 31) && z < (1u << 31), "x, y and z must be below 2^31");
        if (x * y * z == 0) return;
        std::vector<char> Xz(x * y * esize(dtype), 0);
        std::vector<double> ones_x(x, 1.0), ones_z(z, 1.0);
        bnmf::InstSpec s;
        s.x = (uint32_t)x;
        s.y = (uint32_t)y;
        s.z = (uint32_t)z;
        s.X = Xz.data();
        s.alpha = ones_x.data();
        s.beta = ones_z.data();
        s.S0 = S;
        bnmf::Batch bt(dtype, psi_mode(dtype, 0), 1e-50, {s});
        bt.tensor_sums(S_pjk, S_ipk, S_ijp);
    });
}

bnmf_status bnmf_batch_create(bnmf_batch** out, int dtype, const bnmf_bld_desc* descs, size_t n,
                              int use_psi_appr, double eps) {
    return bnmf_batch_create_algo(out, dtype, BNMF_ALGO_MULT, descs, n, use_psi_appr, eps);
}

bnmf_status bnmf_batch_create_algo(bnmf_batch** out, int dtype, int algo, const bnmf_bld_desc* descs,
                                   size_t n, int use_psi_appr, double eps) {
    return guarded([&] {
        BNMF_REQUIRE(out != nullptr, "out is null");
        *out = nullptr;
        check_dtype(dtype);
        BNMF_REQUIRE(n == 0 || descs != nullptr, "descs is null");
        BNMF_REQUIRE(algo == BNMF_ALGO_MULT || algo == BNMF_ALGO_ADD,
                     "algo must be BNMF_ALGO_MULT or BNMF_ALGO_ADD");
        std::vector<bnmf::InstSpec> specs(n);
        for (size_t t = 0; t < n; ++t) {
            const bnmf_bld_desc& d = descs[t];
            if (dtype == BNMF_F32)
                check_params(static_cast<const float*>(d.X), d.x, d.y, d.z, d.alpha, d.beta);
            else
                check_params(static_cast<const double*>(d.X), d.x, d.y, d.z, d.alpha, d.beta);
            bnmf::InstSpec& s = specs[t];
            s.x = (uint32_t)d.x;
            s.y = (uint32_t)d.y;
            s.z = (uint32_t)d.z;
            s.X = d.X;
            s.alpha = d.alpha;
            s.beta = d.beta;
            s.a = d.a;
            s.b = d.b;
            s.S0 = d.S0;
            s.seed = d.seed;
        }
        auto b = std::make_unique<bnmf_batch>();
        b->impl = std::make_unique<bnmf::Batch>(dtype, psi_mode(dtype, use_psi_appr), eps, specs, algo);
        *out = b.release();
    });
}

bnmf_status bnmf_batch_solve(bnmf_batch* b, size_t max_iter) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr, "batch is null");
        b->impl->solve_async(max_iter);
        b->impl->synchronize();
    });
}

bnmf_status bnmf_batch_solve_async(bnmf_batch* b, size_t max_iter) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr, "batch is null");
        b->impl->solve_async(max_iter);
    });
}

bnmf_status bnmf_batch_solve_tol(bnmf_batch* b, size_t max_iter, double tol, size_t check_every,
                                 size_t* iters_run, double* lm_out) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr, "batch is null");
        const size_t done = b->impl->solve_tol(max_iter, tol, check_every, lm_out);
        if (iters_run) *iters_run = done;
    });
}

bnmf_status bnmf_batch_synchronize(bnmf_batch* b) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr, "batch is null");
        b->impl->synchronize();
    });
}

bnmf_status bnmf_batch_get_S(bnmf_batch* b, size_t idx, void* S_out) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr && S_out != nullptr, "null argument");
        b->impl->get_S(idx, S_out);
    });
}

bnmf_status bnmf_batch_log_marginal(bnmf_batch* b, double* out) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr && out != nullptr, "null argument");
        b->impl->log_marginal(out);
    });
}

bnmf_status bnmf_batch_fact(bnmf_batch* b, size_t idx, double eps, double* W, double* H, double* L) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr, "null argument");
        b->impl->fact(idx, eps, W, H, L);
    });
}

bnmf_status bnmf_bld_fact(int dtype, const void* S, size_t x, size_t y, size_t z, const double* alpha,
                          const double* beta, double a, double b, double eps, double* W, double* H,
                          double* L) {
    return guarded([&] {
        check_dtype(dtype);
        BNMF_REQUIRE(x * y * z == 0 || S != nullptr, "S is null");
        BNMF_REQUIRE(x == 0 || alpha != nullptr, "alpha is null");
        BNMF_REQUIRE(z == 0 || beta != nullptr, "beta is null");
        BNMF_REQUIRE(x < (1u << 31) && y < (1u << 31) && z < (1u << 31), "x, y and z must be below 2^31");
        if (x * y * z == 0) return;
        // X is not read by the sums pass; a zero matrix satisfies the layout
        std::vector<char> Xz(x * y * esize(dtype), 0);
        bnmf::InstSpec s;
        s.x = (uint32_t)x;
        s.y = (uint32_t)y;
        s.z = (uint32_t)z;
        s.X = Xz.data();
        s.alpha = alpha;
        s.beta = beta;
        s.a = a;
        s.b = b;
        s.S0 = S;
        bnmf::Batch bt(dtype, dtype == BNMF_F32 ? bnmf::PSIM_F32 : bnmf::PSIM_F64_EXACT, 1e-50, {s});
        bt.fact(0, eps, W, H, L);
    });
}

bnmf_status bnmf_batch_time_pass(bnmf_batch* b, int reps, double* pass_ms, double* iter_ms) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr && pass_ms != nullptr && iter_ms != nullptr, "null argument");
        b->impl->time_pass(reps, pass_ms, iter_ms);
    });
}

bnmf_status bnmf_batch_get_info(bnmf_batch* b, bnmf_batch_info* info) {
    return guarded([&] {
        BNMF_REQUIRE(b != nullptr && info != nullptr, "null argument");
        *info = b->impl->info();
    });
}

void bnmf_batch_destroy(bnmf_batch* b) { delete b; }

bnmf_status bnmf_sample_X(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                          const double* beta, double a, double b, double* X_out) {
    return guarded([&] {
        BNMF_REQUIRE(alpha && beta && X_out, "null argument");
        BNMF_REQUIRE(a > 0 && b > 0, "a and b must be positive");
        bnmf::sample_X(seed, x, y, z, alpha, beta, a, b, X_out);
    });
}

bnmf_status bnmf_sample_S(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                          const double* beta, double a, double b, double* S_out) {
    return guarded([&] {
        BNMF_REQUIRE(alpha && beta && S_out, "null argument");
        BNMF_REQUIRE(a > 0 && b > 0, "a and b must be positive");
        bnmf::sample_S(seed, x, y, z, alpha, beta, a, b, S_out);
    });
}

// sample_X / sample_S with the Poisson draws on the device: the priors
// (O((x+y)z)) on the host, then one device thread per draw of the same
// Philox stream as the host sampler.
static void sample_gpu(int mode, uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                       const double* beta, double a, double b, double* out) {
    BNMF_REQUIRE(alpha && beta && out, "null argument");
    BNMF_REQUIRE(a > 0 && b > 0, "a and b must be positive");
    BNMF_REQUIRE(z >= 1 && z <= 0xffffffffull, "z must be positive");
    const size_t n = mode == 0 ? x * y : x * y * z;
    if (n == 0) return;
    bnmf::ensure_runtime();
    std::vector<double> W, H, L;
    bnmf::sample_priors(seed, x, y, z, alpha, beta, a, b, W, H, L);
    bnmf::DeviceBuffer dW(W.size() * 8), dH(H.size() * 8), dL(L.size() * 8), dout(n * 8);
    BNMF_HIP(hipMemcpy(dW.ptr, W.data(), W.size() * 8, hipMemcpyHostToDevice));
    BNMF_HIP(hipMemcpy(dH.ptr, H.data(), H.size() * 8, hipMemcpyHostToDevice));
    BNMF_HIP(hipMemcpy(dL.ptr, L.data(), L.size() * 8, hipMemcpyHostToDevice));
    BNMF_HIP(bnmf::launch_sample(dout.as<double>(), dW.as<double>(), dH.as<double>(), dL.as<double>(), y,
                                 (uint32_t)z, seed, mode, n, nullptr));
    BNMF_HIP(hipMemcpy(out, dout.ptr, n * 8, hipMemcpyDeviceToHost));
}

bnmf_status bnmf_sample_X_gpu(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, double a, double b, double* X_out) {
    return guarded([&] { sample_gpu(0, seed, x, y, z, alpha, beta, a, b, X_out); });
}

bnmf_status bnmf_sample_S_gpu(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                              const double* beta, double a, double b, double* S_out) {
    return guarded([&] { sample_gpu(1, seed, x, y, z, alpha, beta, a, b, S_out); });
}

bnmf_status bnmf_init_S(int dtype, uint64_t seed, const void* X, size_t x, size_t y, size_t z,
                        void* S0_out) {
    return guarded([&] {
        check_dtype(dtype);
        BNMF_REQUIRE((X && S0_out) || x * y * z == 0, "null argument");
        if (dtype == BNMF_F32)
            bnmf::init_S<float>(seed, static_cast<const float*>(X), x, y, z, static_cast<float*>(S0_out));
        else
            bnmf::init_S<double>(seed, static_cast<const double*>(X), x, y, z, static_cast<double*>(S0_out));
    });
}

bnmf_status bnmf_psi_eval(int kind, const double* x, double* out, size_t n) {
    return guarded([&] {
        BNMF_REQUIRE(kind >= 0 && kind <= 4, "kind must be 0..4");
        BNMF_REQUIRE(n == 0 || (x && out), "null argument");
        if (n == 0) return;
        bnmf::ensure_runtime();
        bnmf::DeviceBuffer dx(n * 8), dout(n * 8);
        BNMF_HIP(hipMemcpy(dx.ptr, x, n * 8, hipMemcpyHostToDevice));
        BNMF_HIP(bnmf::launch_psi_eval(kind, dx.as<double>(), dout.as<double>(), n, nullptr));
        BNMF_HIP(hipDeviceSynchronize());
        BNMF_HIP(hipMemcpy(out, dout.ptr, n * 8, hipMemcpyDeviceToHost));
    });
}

bnmf_status bnmf_tune_set(int dtype, size_t z, size_t fibers, unsigned rows) {
    return guarded([&] {
        check_dtype(dtype);
        BNMF_REQUIRE(z >= 1 && z <= bnmf::kFusedZ, "the tile table covers z in [1, 64] (deeper fibers use fixed tiles)");
        BNMF_REQUIRE(rows == 0 || (rows >= 4 && rows <= 1024), "rows must be 0 or in [4, 1024]");
        bnmf::tune_set(dtype, (uint32_t)z, (uint64_t)fibers, rows);
    });
}

bnmf_status bnmf_tune_get(int dtype, size_t z, size_t fibers, unsigned* rows) {
    return guarded([&] {
        check_dtype(dtype);
        BNMF_REQUIRE(rows != nullptr, "rows is null");
        BNMF_REQUIRE(z >= 1 && z <= bnmf::kFusedZ, "the tile table covers z in [1, 64] (deeper fibers use fixed tiles)");
        *rows = bnmf::tune_ti_cap(dtype, (uint32_t)z, (uint64_t)fibers);
    });
}

bnmf_status bnmf_tune_load(const char* path) {
    return guarded([&] {
        BNMF_REQUIRE(path != nullptr, "path is null");
        bnmf::tune_load(path);
    });
}

bnmf_status bnmf_tune_save(const char* path) {
    return guarded([&] {
        BNMF_REQUIRE(path != nullptr, "path is null");
        bnmf::tune_save(path);
    });
}

unsigned bnmf_tile_rows(int dtype, size_t x, size_t y, size_t z) {
    if ((dtype != BNMF_F32 && dtype != BNMF_F64) || z < 1 || z >= (1u << 31) || x == 0 || y == 0 ||
        x >= (1u << 31) || y >= (1u << 31))
        return 0;
    try {
        return bnmf::choose_ti_rows(dtype, (uint32_t)x, (uint32_t)y, (uint32_t)z);
    } catch (...) {
        return 0;
    }
}

const char* bnmf_tune_source(void) {
    static thread_local std::string src;
    try {
        src = bnmf::tune_source();
    } catch (...) {
        src.clear();
    }
    return src.c_str();
}

}  // extern "C"
