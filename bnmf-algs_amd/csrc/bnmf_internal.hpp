// bnmf_internal.hpp — host-side internals shared by the library sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bnmf_hip.h"
#include "bld_types.hpp"

namespace bnmf {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    bnmf_status status;
    Error(bnmf_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
void set_last_error(const std::string& msg);
void clear_last_error();

#define BNMF_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t bnmf_e_ = (call);                                                    \
        if (bnmf_e_ != hipSuccess)                                                      \
            throw ::bnmf::Error(bnmf_e_ == hipErrorOutOfMemory ? BNMF_ERR_OUT_OF_MEMORY \
                                                               : BNMF_ERR_DEVICE,       \
                                std::string(#call) + ": " + hipGetErrorString(bnmf_e_)); \
    } while (0)

#define BNMF_REQUIRE(cond, msg)                                               \
    do {                                                                      \
        if (!(cond)) throw ::bnmf::Error(BNMF_ERR_INVALID_ARGUMENT, (msg)); \
    } while (0)

// ----------------------------------------------------------------- kernels
enum : int { MODE_UPDATE = 0, MODE_SUMS = 1, MODE_SCORE = 2, MODE_ADD = 3 };
enum : int { PSIM_F32 = 0, PSIM_F64_EXACT = 1, PSIM_F64_APPR = 2 };
hipError_t launch_pass(int dtype, int psi, int mode, int zb, int rw, dim3 grid, hipStream_t st,
                       const BldInst* insts, const PassItem* items, const uint32_t* n_items,
                       uint32_t zmax);
void read_layout(int dtype, uint32_t z, uint32_t* rw, uint32_t* hs, uint32_t* hm);
// raw = false: also the per-instance psi(sum_i alpha_eph) kernel over n_inst
// descriptors
hipError_t launch_prep(int dtype, int psi, bool raw, dim3 grid, hipStream_t st,
                       const BldInst* insts, const PrepItem* items, const uint32_t* n_items,
                       uint32_t n_inst);
hipError_t launch_score(dim3 grid, hipStream_t st, const BldInst* insts, const PrepItem* items,
                        const uint32_t* n_items, uint32_t n_inst);
hipError_t launch_psi_eval(int kind, const double* xin, double* out, size_t n, hipStream_t st);
hipError_t launch_init_s(int dtype, void* S, const void* X, size_t n_fibers, uint32_t z,
                         uint64_t seed, hipStream_t st);
hipError_t init_kernel_attributes();
hipError_t launch_fact(int dtype, const BldInst* inst, uint32_t y, uint32_t z, double eps, double* W,
                       double* H, double* L, hipStream_t st);
// FP64 step pass (MODE_UPDATE or MODE_ADD), two lanes per fiber, z bucket
// zb <= 64 and LDS read width rw (bld_pass_f64.hip)
hipError_t launch_pass_f64(int psi, int mode, int zb, int rw, dim3 grid, hipStream_t st, const BldInst* insts,
                           const PassItem* items, const uint32_t* n_items);
hipError_t init_f64_attributes();
// any mode and dtype for instances deeper than kFusedZ (bld_wide.hip)
hipError_t launch_pass_wide(int dtype, int psi, int mode, dim3 grid, hipStream_t st, const BldInst* insts,
                            const PassItem* items, const uint32_t* n_items);
// FP32 step pass (MODE_UPDATE or MODE_ADD) specialised on the exact fiber
// depth z (bld_pass_f32z.hip)
hipError_t launch_pass_f32z(int z, int mode, dim3 grid, hipStream_t st, const BldInst* insts,
                            const PassItem* items, const uint32_t* n_items);
hipError_t init_f32z_attributes();
// approximate solver (bld_appr.hip): mode 0 sums, 1 nu step, 2 mu step
hipError_t launch_appr_pass(int dtype, int mode, int zb, dim3 grid, hipStream_t st, const ApprInst* insts,
                            const PassItem* items, const uint32_t* n_items, uint32_t zmax);
hipError_t launch_appr_factor(int dtype, int half, dim3 grid, hipStream_t st, const ApprInst* insts,
                              const PassItem* items, const uint32_t* n_items);
hipError_t launch_appr_S(int dtype, const ApprInst* inst, size_t n_fibers, hipStream_t st);
hipError_t init_appr_attributes();
void ensure_runtime();  // per-process kernel attributes, once

// ----------------------------------------------------------------- sampler
int sample_X(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
             const double* beta, double a, double b, double* X);
int sample_S(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
             const double* beta, double a, double b, double* S);
template <typename T>
int init_S(uint64_t seed, const T* X, size_t x, size_t y, size_t z, T* S);
// the priors of bnmf_priors: W (x*z), H (z*y), L (y)
void sample_priors(uint64_t seed, size_t x, size_t y, size_t z, const double* alpha,
                   const double* beta, double a, double b, std::vector<double>& W,
                   std::vector<double>& H, std::vector<double>& L);
// device draws of X (mode 0, n = x*y) or S (mode 1, n = x*y*z) from the priors
hipError_t launch_sample(double* out, const double* W, const double* H, const double* L, size_t y,
                         uint32_t z, uint64_t seed, int mode, size_t n, hipStream_t st);

// ----------------------------------------------------------- tile planner
// z bucket of the fused generic pass (z <= kFusedZ); 0 for the wide pass
uint32_t z_bucket(uint32_t z);
uint32_t choose_ti_rows(int dtype, uint32_t x, uint32_t y, uint32_t z);
// Lays the items of one pass launch out XCD by XCD (L2 sharing of the
// per-column psi staging); results do not depend on item order.
void xcd_order(std::vector<PassItem>& items);

// ------------------------------------------------------ no-allocation solve
// bld_mult's FP32 solve on caller-owned device buffers, workspace in the
// library's fixed device scratch (noalloc.hip); enqueued on st
void bld_mult_f32_noalloc(const float* X, uint32_t x, uint32_t y, uint32_t z, const double* alpha,
                          const double* beta, size_t max_iter, double eps, const float* S0, uint64_t seed,
                          float* S, hipStream_t st);

// ------------------------------------------------------------- tile table
// Rows per pass tile by shape class (tune.cpp); 0 = built-in rule.
int tune_zclass(uint32_t z);
int tune_sclass(uint64_t fibers);
uint32_t tune_ti_cap(int dtype, uint32_t z, uint64_t fibers);
void tune_set(int dtype, uint32_t z, uint64_t fibers, uint32_t cap);
void tune_load(const std::string& path);
void tune_save(const std::string& path);
std::string tune_source();

// --------------------------------------------------------- device buffers
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    template <typename T>
    T* as(size_t byte_off = 0) const {
        return reinterpret_cast<T*>(static_cast<char*>(ptr) + byte_off);
    }
};

// ----------------------------------------------------------------- batches
struct InstSpec {
    uint32_t x = 0, y = 0, z = 0;
    const void* X = nullptr;       // host
    const double* alpha = nullptr; // host
    const double* beta = nullptr;  // host
    double a = 1.0, b = 1.0;
    const void* S0 = nullptr;      // host or null
    uint64_t seed = 0;
};

// Byte layout of one instance inside an arena.
struct InstLayout {
    uint32_t x, y, z, zb, rw, hs, hm, ti_rows, n_ti, n_tj, nA, n_cj;
    uint32_t tile_cols;  // fibers per tile row: kTileCols, or kWideCols (z > kFusedZ)
    size_t off_S, off_S0, off_X, off_alpha, off_beta, off_psib, off_psia, off_partA, off_psiA, off_pjk,
        off_ipk, off_score, off_result;
    size_t bytes;  // total, 256-byte aligned
    size_t elements() const { return (size_t)x * y * z; }
};
// with_S0: reserve a separate buffer for the initial tensor (batches keep S0
// to restart solves; the server solves in place).
InstLayout plan_instance(int dtype, uint32_t x, uint32_t y, uint32_t z, bool with_S0 = true);
BldInst make_desc(const InstLayout& L, void* base, double eps, double a, double b, bool first,
                  bool nonneg = false);

// Kernel of one pass launch: the exact-z FP32 step kernel (zexact > 0), the
// FP64 step kernel (f64, z bucket zb, LDS read width rw), the generic
// z-bucket kernel (zb, rw) or, for z > kFusedZ, the wide kernel (zb = 0).
// Instances with equal keys share a launch. The key depends on the instance
// shape only, so an instance runs the same kernel alone, batched or served.
struct PassKey {
    int zexact = 0, zb = 0, rw = 0, f64 = 0;
    bool operator<(const PassKey& o) const {
        return std::tie(zexact, zb, rw, f64) < std::tie(o.zexact, o.zb, o.rw, o.f64);
    }
    bool wide() const { return zexact == 0 && zb == 0; }
};
// fibers per tile row of an instance: kWideCols for z > kFusedZ, else
// kTileColsF64 (FP64, z <= 32), half that (FP64, z <= 64) or kTileCols (FP32)
uint32_t tile_cols_of(int dtype, uint32_t z);
// nonneg: the instance's S stays >= 0 (S0 >= 0 and beta >= 1)
PassKey pass_key(int dtype, int mode, const InstLayout& L, bool nonneg);
hipError_t launch_pass_key(const PassKey& k, int dtype, int psi, int mode, dim3 grid,
                           hipStream_t st, const BldInst* insts, const PassItem* items,
                           const uint32_t* n_items, uint32_t zmax);

// A fixed set of instances sharing one arena, one stream and one set of
// captured per-iteration graphs.
class Batch {
   public:
    // algo: BNMF_ALGO_MULT (bld_mult) or BNMF_ALGO_ADD (bld_add)
    Batch(int dtype, int psi_mode, double eps, const std::vector<InstSpec>& specs,
          int algo = BNMF_ALGO_MULT);
    ~Batch();
    void solve_async(size_t iters);
    // Restarts from S0 and runs at most max_iter iterations; every
    // check_every iterations the FP64 log_marginal_S of every instance is
    // evaluated, and the solve stops once each relative change is <= tol.
    // Returns the iterations run; lm_out (n values, may be null) receives the
    // final log-marginals. Synchronous.
    size_t solve_tol(size_t max_iter, double tol, size_t check_every, double* lm_out);
    void synchronize();
    void get_S(size_t idx, void* host) const;
    void log_marginal(double* out);
    void time_pass(int reps, double* pass_ms, double* iter_ms);
    void tensor_sums(void* S_pjk, void* S_ipk, void* S_ijp);
    // bld_fact of instance idx's current S: W (x*z), H (z*y), L (y), host
    void fact(size_t idx, double eps, double* W, double* H, double* L);
    bnmf_batch_info info() const;
    size_t size() const { return layouts_.size(); }

   private:
    struct Bucket {
        PassKey key;
        int lane;          // side stream the bucket's pass runs on
        size_t work;       // elements
        uint32_t zmax;
        uint32_t n_items;
        size_t off_items;  // in meta_
        size_t off_count;
    };
    void enqueue_iteration(bool first) { enqueue_iterations(1, first); }
    // first iteration of a solve (step counters reset, S0's partials, the
    // step reading S0), then `rem` more iterations from the captured graphs
    void start_solve();
    void replay_iterations(size_t rem);
    // `iters` iterations; each update group runs prep, psi(A) and its pass
    // on its own stream, iteration after iteration, with one fork and join
    // (with_prep = false: the update passes alone, for timing)
    void enqueue_iterations(int iters, bool first, bool with_prep = true);
    void enqueue_pass(int mode, bool first);
    // pass launches of the update pass (groups_[0]) and of the sums-only and
    // scoring passes (groups_[1])
    std::vector<Bucket>& groups_of(int mode) {
        return groups_[(mode == MODE_UPDATE || mode == MODE_ADD) ? 0 : 1];
    }
    hipGraphExec_t graph_of(int level);

    int dtype_, psi_;
    int algo_ = BNMF_ALGO_MULT;
    int step_mode_ = MODE_UPDATE;  // pass mode of an iteration's step
    int dev_ = 0;
    double eps_;
    size_t es_;
    std::vector<InstLayout> layouts_;
    std::vector<InstSpec> specs_;
    DeviceBuffer arena_;
    DeviceBuffer meta_;  // descriptors, work lists, counts
    size_t off_desc_first_ = 0, off_desc_ = 0, off_prep_items_ = 0, off_prep_count_ = 0;
    size_t off_iter_ = 0;  // additive solver: uint32 iteration counters, one per instance
    uint32_t n_prep_items_ = 0;
    std::vector<Bucket> groups_[2];
    // per stream: descriptor arrays of the instances of its update groups
    // (the groups' pass items index them), its prep work list
    struct Lane {
        size_t off_desc = 0, off_desc_first = 0, off_prep = 0, off_prep_count = 0;
        uint32_t n_prep = 0, n_inst = 0;
    };
    std::vector<Lane> lanes_;
    hipStream_t stream_ = nullptr;
    // z-bucket passes of one iteration are independent: they run on up to
    // four streams forked from and joined back to stream_ (graph branches
    // when captured)
    std::vector<hipStream_t> side_;
    std::vector<hipEvent_t> fork_ev_, join_ev_;
    // captured iteration graphs of 1, 2, 4, ... unroll_ iterations (level l
    // holds 2^l), each instantiated on first use
    static constexpr int kGraphLevels = 7;
    hipGraphExec_t graphs_[kGraphLevels] = {};
    int unroll_ = 32;
    // the partial axis sums in the workspace belong to the current S (true
    // after a solve: its last pass accumulates them)
    bool partials_valid_ = false;
    int cus_ = 256;
    size_t device_bytes_ = 0;
};

// ------------------------------------------------- approximate solver
struct ApprSpec {
    uint32_t x = 0, y = 0, z = 0;
    const void* X = nullptr;        // host, x*y
    const double* alpha = nullptr;  // host, x
    const double* beta = nullptr;   // host, z
    const void* nu0 = nullptr;      // host x*z, or null (drawn from seed)
    const void* mu0 = nullptr;      // host z*y, or null (drawn from seed)
    uint64_t seed = 0;
};

// Instances of bld::bld_appr solved together: the factors stay resident,
// one captured graph per iteration.
class ApprBatch {
   public:
    ApprBatch(int dtype, double eps, const std::vector<ApprSpec>& specs);
    ~ApprBatch();
    // restarts from nu0 / mu0 and runs iters iterations
    void solve_async(size_t iters);
    void synchronize();
    // S (x*y*z, formed from the current factors), nu (x*z), mu (z*y); any null
    void get(size_t idx, void* S, void* nu, void* mu);
    // mean ms of one iteration (graph replays, HIP events); advances the factors
    double time_iteration(int reps);
    size_t device_bytes() const;
    size_t size() const { return lay_.size(); }

   private:
    struct L_ {
        uint32_t x, y, z, ti_rows, n_ti, n_tj, nA, n_cj;
        size_t off_X, off_nu, off_mu, off_nu0, off_mu0, off_alpha, off_beta, off_psib, off_psia, off_partA,
            off_psiA, off_pjk, off_pjk2, off_ipk, off_ipk2, off_S;
    };
    struct Group {
        uint32_t zb, zmax, n_items;
        size_t off_items, off_count;
    };
    void enqueue_iteration();
    void capture();

    int dtype_, psi_ = 0;
    size_t es_;
    std::vector<L_> lay_;
    std::vector<Group> groups_;
    DeviceBuffer arena_, meta_;
    size_t off_ai_ = 0, off_bi_ = 0, off_prep_ = 0, off_prep_count_ = 0, off_fnu_ = 0, off_fnu_count_ = 0,
           off_fmu_ = 0, off_fmu_count_ = 0;
    uint32_t n_prep_ = 0, n_fnu_ = 0, n_fmu_ = 0;
    hipStream_t stream_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
};

}  // namespace bnmf
