// bld_wide.hip — pass kernel for instances whose fibers are deeper than the
// fused passes hold in registers (z > 64). The reference puts no bound on z
// (check_bld_params, src/bld/util_details.hpp:20-28; its kernels loop over
// any z, cuda_src/bld_mult/bld_mult_cuda_kernels.cu:119-255).
//
// Mapping: one wavefront per block; a work item is a tile of kWideRows rows
// x kWideCols columns of one instance, walked row by row. For each fiber
// (i, j) the 64 lanes cover k = lane, lane + 64, ... (a fiber of z > 64
// elements is at least 260 contiguous bytes, so every access is coalesced)
// and the per-fiber sums (S_ij+, nom, denom; S_mult and the exp-row norm of
// the additive step) are butterfly reductions across the wave, which leave
// the same value in every lane (each step adds a commutative pair). Phase B
// re-reads the fiber (it was just read: cache hits) and evaluates the
// element's gradient again rather than keeping z values per lane, so one
// kernel serves every z with no register arrays and no LDS.
//
// The tile owns its partial-sum slots: part_pjk[ti][j][k] for its columns
// (summed over its rows, in row order) and part_ipk[tj][i][k] for its rows
// (summed over its columns, in column order); each slot is written by the
// lane that owns k, first by a store, then by read-add-store, so results are
// deterministic and independent of the rest of the launch. The prep kernel
// reduces them as it reduces the fused passes' partials.
//
// Reference operations: update_grad_plus bld_mult_funcs.hpp:177-192, nom
// :242-259, denom :272-290, update_S :307-327; bld_add's step
// src/bld/bld_add.hpp:137-181; the log-marginal's lgamma(S+1) terms
// alloc_model_funcs.hpp:179-194.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "bld_types.hpp"
#include "digamma.hpp"
#include "pass_common.hpp"

namespace bnmf {

template <typename T, typename ACC, int MODE, int PSI>
__global__ void __launch_bounds__(64)
    bld_wide_kernel(const BldInst* __restrict__ insts, const PassItem* __restrict__ items,
                    const uint32_t* __restrict__ n_items_p) {
    const uint32_t n_items = *n_items_p;
    const uint32_t lane = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const BldInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const uint32_t i0 = item.ti * kWideRows, i1 = min(x, i0 + kWideRows);
        const uint32_t j0 = item.tj * kWideCols, j1 = min(y, j0 + kWideCols);
        // S_src and S are the same tensor after the first pass: no restrict
        const BNMF_GLOBAL T* Ssrc = gconst<T>(I->S_src);
        BNMF_GLOBAL T* Sdst = gmut<T>(I->S);
        BNMF_GLOBAL ACC* __restrict__ pjk = gmut<ACC>(I->part_pjk) + (size_t)item.ti * y * z;
        BNMF_GLOBAL ACC* __restrict__ ipk = gmut<ACC>(I->part_ipk) + (size_t)item.tj * x * z;
        const double eps = I->eps;
        const bool fast_psi = I->nonneg != 0u;
        double eta = 0.0;
        // additive step size 0.1 / (iteration + 1)^0.55 (bld_add.hpp:101-103);
        // *iter counts the iterations started, this one included
        if constexpr (MODE == PASS_ADD) eta = 0.1 / pow((double)*I->iter, 0.55);
        double lg_sum = 0.0;

        for (uint32_t i = i0; i < i1; ++i) {
            const BNMF_GLOBAL T* __restrict__ psia_i = gconst<T>(I->psia) + (size_t)i * z;
            const BNMF_GLOBAL T* __restrict__ psiA = gconst<T>(I->psiA);
            for (uint32_t j = j0; j < j1; ++j) {
                const size_t f = ((size_t)i * y + j) * z;
                const BNMF_GLOBAL T* src = Ssrc + f;
                BNMF_GLOBAL T* dst = Sdst + f;
                const BNMF_GLOBAL T* __restrict__ pb = gconst<T>(I->psib) + (size_t)j * z;
                const bool row0 = i == i0, col0 = j == j0;
                // the new element's contribution to the tile's partial sums
                auto add_partials = [&](uint32_t k, T sn) {
                    const size_t a = (size_t)j * z + k, b = (size_t)i * z + k;
                    pjk[a] = row0 ? ACC(sn) : pjk[a] + ACC(sn);
                    ipk[b] = col0 ? ACC(sn) : ipk[b] + ACC(sn);
                };
                if constexpr (MODE == PASS_SUMS || MODE == PASS_SCORE) {
                    ACC fs = ACC(0);
                    for (uint32_t k = lane; k < z; k += 64) {
                        const T s = src[k];
                        add_partials(k, s);
                        fs += ACC(s);
                        if constexpr (MODE == PASS_SCORE) lg_sum += lgamma((double)s + 1.0);
                    }
                    if constexpr (MODE == PASS_SUMS) {
                        fs = wave_sum(fs);
                        if (lane == 0 && I->sijp != nullptr) gmut<ACC>(I->sijp)[(size_t)i * y + j] = fs;
                    }
                } else if constexpr (MODE == PASS_UPDATE) {
                    const T xv = gconst<T>(I->X)[(size_t)i * y + j];
                    if constexpr (sizeof(T) == 8) {
                        // FP64 plain form: the reference's operations
                        // (X_reciprocal :95, nom :254, denom :284-285,
                        // update :319-323)
                        const T xr = T(1.0) / (xv + eps);
                        auto grad = [&](uint32_t k, T s) { return pb[k] - psi_sp1<PSI, T>(s); };
                        T sumS = T(0), nom = T(0), den = T(0);
                        for (uint32_t k = lane; k < z; k += 64) {
                            const T s = src[k];
                            const T g = grad(k, s);
                            sumS += s;
                            nom += s * xr * (psiA[k] - psia_i[k]);
                            den += s * xr * g;
                        }
                        sumS = wave_sum(sumS);
                        nom = wave_sum(nom);
                        den = wave_sum(den);
                        const T x_over_s = xv / (sumS + eps);
                        for (uint32_t k = lane; k < z; k += 64) {
                            const T s = src[k];
                            const T gm = psiA[k] - psia_i[k];
                            const T sn = as_stored(s * ((grad(k, s) + nom) * x_over_s / (gm + den + eps)));
                            dst[k] = sn;
                            add_partials(k, sn);
                        }
                    } else {
                        // FP32 bulk: eps = 1e-50 flushes to zero in float; the
                        // reference's double-promoted guard makes an X_ij = 0
                        // or all-zero fiber update to exactly 0, taken
                        // explicitly, and an exact-zero denominator (a single
                        // row with beta = 1) updates to 0 rather than 0/0
                        const float epsf = static_cast<float>(eps);
                        const bool live = xv > 0.0f;
                        const float xr = live ? 1.0f / (xv + epsf) : 0.0f;
                        auto grad = [&](uint32_t k, float s) {
                            const float ps = fast_psi ? psi_ge1_f32(s + 1.0f) : psi_sp1<PSI_F32, float>(s);
                            return pb[k] - ps;
                        };
                        float sumS = 0.0f, nom = 0.0f, den = 0.0f;
                        for (uint32_t k = lane; k < z; k += 64) {
                            const float s = src[k];
                            const float g = grad(k, s);
                            sumS += s;
                            const float t = s * xr;
                            nom = fmaf(t, psiA[k] - psia_i[k], nom);
                            den = fmaf(t, g, den);
                        }
                        sumS = wave_sum(sumS);
                        nom = wave_sum(nom);
                        den = wave_sum(den);
                        const bool upd = live && sumS != 0.0f;
                        const float x_over_s = upd ? xv / (sumS + epsf) : 0.0f;
                        float den_e = den + epsf;
                        if (den_e == 0.0f) den_e = FLT_MIN;
                        for (uint32_t k = lane; k < z; k += 64) {
                            const float s = src[k];
                            const float gm = psiA[k] - psia_i[k];
                            const float fct = (grad(k, s) + nom) * x_over_s * recip_fast<float>(gm + den_e);
                            const float sn = as_stored(upd ? s * fct : 0.0f);
                            dst[k] = sn;
                            add_partials(k, sn);
                        }
                    }
                } else {
                    // bld::bld_add's step (src/bld/bld_add.hpp:137-181):
                    // grad_S = psi(beta_eph) - psi(S+1) - psi(A) + psi(alpha_eph),
                    // S_mult = sum_k S grad_S, Z = log S + eps + eta S (X grad_S
                    // - S_mult), S = X exp(Z) / (|exp(Z)|_1 + 1e-50). exp(Z) is
                    // kept in S between the two reductions (each lane rereads
                    // only the elements it wrote).
                    const T xv = gconst<T>(I->X)[(size_t)i * y + j];
                    auto grad = [&](uint32_t k, T s) -> T {
                        if constexpr (sizeof(T) == 8) {
                            const T ps = psi_sp1<PSI, T>(s);
                            return ((pb[k] - ps) - psiA[k]) + psia_i[k];
                        } else {
                            const float ps = fast_psi ? psi_ge1_f32(s + 1.0f) : psi_sp1<PSI_F32, float>(s);
                            return (pb[k] - ps) - (psiA[k] - psia_i[k]);
                        }
                    };
                    T smult = T(0);
                    for (uint32_t k = lane; k < z; k += 64) {
                        const T s = src[k];
                        if constexpr (sizeof(T) == 8)
                            smult += s * grad(k, s);
                        else
                            smult = fmaf(s, grad(k, s), smult);
                    }
                    smult = wave_sum(smult);
                    T norm = T(0);
                    for (uint32_t k = lane; k < z; k += 64) {
                        const T s = src[k];
                        T e;
                        if constexpr (sizeof(T) == 8) {
                            T zk = log(s) + eps;
                            zk += eta * s * (xv * grad(k, s) - smult);
                            e = exp(zk);
                        } else {
                            // exp(log s + t) = s exp(t): one rounding of s
                            // instead of log's (0 stays 0)
                            const float etaf = static_cast<float>(eta);
                            e = s * __expf((etaf * s) * fmaf(xv, grad(k, s), -smult));
                        }
                        dst[k] = e;
                        norm += e;
                    }
                    norm = wave_sum(norm);
                    T scale = T(1);
                    if constexpr (sizeof(T) == 8) {
                        norm += 1e-50;
                    } else {
                        // the reference's 1e-50 guard underflows in float: an
                        // all-zero exp row (X = 0) is taken explicitly
                        scale = norm > 0.0f ? xv / norm : 0.0f;
                    }
                    for (uint32_t k = lane; k < z; k += 64) {
                        const T e = dst[k];
                        T sn;
                        if constexpr (sizeof(T) == 8)
                            sn = as_stored(xv * (e / norm));
                        else
                            sn = as_stored(e * scale);
                        dst[k] = sn;
                        add_partials(k, sn);
                    }
                }
            }
        }
        if constexpr (MODE == PASS_SCORE) {
            lg_sum = wave_sum(lg_sum);
            if (lane == 0) I->score[(size_t)item.ti * I->n_tj + item.tj] = lg_sum;
        }
    }
}

template <typename T, typename ACC, int MODE, int PSI>
static hipError_t launch_wide_t(dim3 grid, hipStream_t st, const BldInst* insts, const PassItem* items,
                                const uint32_t* n_items) {
    hipLaunchKernelGGL((bld_wide_kernel<T, ACC, MODE, PSI>), grid, dim3(64), 0, st, insts, items, n_items);
    return hipGetLastError();
}

// dtype 0 = float, 1 = double; psi and mode as for launch_pass
hipError_t launch_pass_wide(int dtype, int psi, int mode, dim3 grid, hipStream_t st, const BldInst* insts,
                            const PassItem* items, const uint32_t* n_items) {
    if (dtype == 0) {
        switch (mode) {
            case PASS_UPDATE: return launch_wide_t<float, float, PASS_UPDATE, PSI_F32>(grid, st, insts, items, n_items);
            case PASS_SUMS: return launch_wide_t<float, float, PASS_SUMS, PSI_F32>(grid, st, insts, items, n_items);
            case PASS_SCORE: return launch_wide_t<float, double, PASS_SCORE, PSI_F32>(grid, st, insts, items, n_items);
            case PASS_ADD: return launch_wide_t<float, float, PASS_ADD, PSI_F32>(grid, st, insts, items, n_items);
        }
    } else {
        switch (mode) {
            case PASS_UPDATE:
                if (psi == PSI_F64_APPR)
                    return launch_wide_t<double, double, PASS_UPDATE, PSI_F64_APPR>(grid, st, insts, items, n_items);
                return launch_wide_t<double, double, PASS_UPDATE, PSI_F64_EXACT>(grid, st, insts, items, n_items);
            case PASS_SUMS: return launch_wide_t<double, double, PASS_SUMS, PSI_F64_EXACT>(grid, st, insts, items, n_items);
            case PASS_SCORE: return launch_wide_t<double, double, PASS_SCORE, PSI_F64_EXACT>(grid, st, insts, items, n_items);
            case PASS_ADD:
                // the additive solver evaluates every psi with psi_appr
                return launch_wide_t<double, double, PASS_ADD, PSI_F64_APPR>(grid, st, insts, items, n_items);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace bnmf
