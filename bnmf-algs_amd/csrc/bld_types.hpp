// bld_types.hpp — device-side descriptors shared by the host planner and the
// HIP kernels of the multiplicative BLD solver.
//
// HBM layout of one instance (all dense, row-major, k fastest, no per-fiber
// pitch; the reference pads every z-fiber with cudaMalloc3D,
// src/cuda/memory/device_memory_3d.hpp:55-65):
//   S        T[x*y*z]        the solution tensor, updated in place
//   S_src    T[x*y*z]        tensor read by the first pass (S0, or S itself)
//   X        T[x*y]
//   alpha    double[x], beta double[z]    model parameters (Params::alpha/beta)
//   psib     T[y*z]          psi(S_+jk + beta_k)          (beta_eph, :152-163)
//   psia     T[x*z]          psi(S_i+k + alpha_i)         (alpha_eph, :130-141)
//   partA    ACC[nA*z]       column sums of alpha_eph over row chunks of 16
//   psiA     T[z]            psi(sum_i alpha_eph(i,k))    (update_grad_minus)
//   part_pjk ACC[n_ti*y*z]   per-row-tile partial S_+jk
//   part_ipk ACC[n_tj*x*z]   per-column-tile partial S_i+k
//   score    double[n_ti*n_tj + nA + n_cj]  log-marginal partials
#pragma once

#include <cstdint>

namespace bnmf {

// Fibers (columns j) per tile: one wavefront lane per fiber (FP32 passes).
constexpr uint32_t kTileCols = 64;
// FP64 step pass: two lanes per fiber, so a tile row holds 32 fibers (four
// lanes and 16 fibers for 32 < z <= 64).
constexpr uint32_t kTileColsF64 = 32;
// Rows per prep chunk (alpha_eph column-sum granularity) and columns per
// prep chunk: small, so that even one large instance gives the prep launch
// hundreds of blocks.
constexpr uint32_t kPrepRows = 16;
constexpr uint32_t kPrepCols = 16;
// Deepest fiber the fused passes hold in registers; deeper instances
// (z > kFusedZ) run the wide pass: one wavefront per fiber, lanes over k, on
// tiles of kWideRows x kWideCols fibers.
constexpr uint32_t kFusedZ = 64;
constexpr uint32_t kWideRows = 16;
constexpr uint32_t kWideCols = 16;

struct BldInst {
    void* S;              // updated tensor (pass output)
    const void* S_src;    // tensor read by the pass (== S except the first pass)
    const void* X;
    const double* alpha;
    const double* beta;
    void* psib;
    void* psia;
    void* partA;
    void* psiA;           // psi(sum_i alpha_eph(i,k)), z values
    void* part_pjk;
    void* part_ipk;
    double* score;        // score partials (log-marginal), may be null
    double* result;       // log-marginal output (device), may be null
    void* sijp;           // fiber sums S_ij+ (sums-only pass), may be null
    uint32_t* iter;       // additive solver: iterations started (the psi(A)
                          // kernel counts them; the step size depends on it)
    double eps;
    double a, b;          // Gamma prior (log-marginal only)
    float inv_z;          // 1.0f / z (fiber index from element index)
    uint32_t x, y, z;
    uint32_t rw;          // elements per LDS fiber read: 4 (z%4==0), 2 (z even), 1
    uint32_t hs, hm;      // lane l visits its fiber's chunks in the order
                          // m ^ ((l >> hs) & hm): distinct LDS banks per step
    uint32_t ti_rows;     // rows per pass tile
    uint32_t tile_cols;   // fibers per tile row (kTileCols, kTileColsF64 or kWideCols)
    uint32_t n_ti, n_tj;  // tiles along rows / columns
    uint32_t nA;          // prep row chunks
    uint32_t n_cj;        // prep column chunks
    uint32_t zq;          // next power of two >= z
    uint32_t zq_log2;
    uint32_t nonneg;      // 1: S0 >= 0 and beta >= 1, so S stays >= 0 (fast psi(S+1))
};

struct PassItem {
    uint32_t inst, ti, tj, pad;
};

// One instance of the approximate solver (bld::bld_appr): S is never stored;
// S_ijk = X_ij nu_ik mu_kj / (nu mu + eps)_ij is formed per fiber from the
// factors. Prep and psi(A) run on a BldInst that shares psib, psia, partA,
// psiA and the partial buffers.
struct ApprInst {
    const void* X;        // x*y
    void* nu;             // x*z, rows sum to 1
    void* mu;             // z*y, columns sum to 1
    void* psib;           // psi(beta_eph), y*z
    void* psia;           // psi(alpha_eph), x*z
    void* psiA;           // psi(sum_i alpha_eph), z
    void* part_pjk;       // n_ti*y*z: S_+jk partials, then mu's numerator partials
    void* part_ipk;       // n_tj*x*z: S_i+k partials, then nu's numerator partials
    void* part_pjk2;      // n_ti*y*z: mu's denominator partials
    void* part_ipk2;      // n_tj*x*z: nu's denominator partials
    void* S;              // x*y*z output (final materialisation only)
    double eps;
    uint32_t x, y, z;
    uint32_t ti_rows, n_ti, n_tj;
    uint32_t zq, zq_log2;
};

// kind 0: column chunk (psib / column score terms); kind 1: row chunk
// (psia, partA / row score terms).
struct PrepItem {
    uint32_t inst, kind, chunk, pad;
};

}  // namespace bnmf
