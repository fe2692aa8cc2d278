// bld_appr.hip — kernels of the approximate BLD solver bld::bld_appr
// (src/bld/bld_appr.hpp). S is kept in factored form, S_ijk = X_ij nu_ik
// mu_kj / (nu mu + eps)_ij, and never stored during a solve: each pass reads
// X (x*y) and the small factors and forms S per fiber in registers.
//
// One iteration (bld_appr.hpp:320-348) is two half-steps, one per factor:
//   sums pass    X_hat = nu mu, S per fiber, per-tile partial S_+jk / S_i+k
//                (update_X_hat, update_S and the sums of update_alpha_eph /
//                update_beta_eph, :12-71)
//   prep, psi(A) psi(beta_eph), psi(alpha_eph), psi(sum_i alpha_eph) (the
//                multiplicative solver's kernels, shared)
//   step pass    grad_plus (:73-88) and grad_minus (:90-103) per fiber and
//                the numerator and denominator of update_nu (:105-164,
//                reduced over j: row partials) or update_mu (:166-226,
//                reduced over i: column partials)
//   factor       nu *= nom / (den + eps), rows renormalised; or mu, columns
// Tiles are ti_rows x 64 fibers as in the update pass; lane f owns column
// j0 + f (fibers deeper than 64 take a wavefront per fiber instead). The
// double sums over c of update_nu / update_mu factor per fiber:
// sum_c nu_ic mu_cj gm_ic = A_ij and sum_c nu_ic mu_cj gp_ijc = B_ij, so
//   nom_ik = sum_j oa_ij mu_kj gp_ijk + oa2_ij mu_kj A_ij
//   den_ik = sum_j oa_ij mu_kj gm_ik  + oa2_ij mu_kj B_ij
// with oa = X / (X_hat + eps), oa2 = X / (X_hat^2 + eps); update_mu has
// nu_ik in place of mu_kj and sums over i. Every psi is psi_appr (FP64) or
// the FP32 digamma. All reductions have a fixed order (no atomics).

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bld_types.hpp"
#include "digamma.hpp"
#include "pass_common.hpp"

namespace bnmf {

enum ApprMode : int { APPR_SUMS = 0, APPR_NU = 1, APPR_MU = 2 };

// k-major LDS arrays of a tile's 64 columns, padded so that the lanes of a
// row reduction (different k, same column) fall on different banks
constexpr uint32_t kApprStride = 65;

// LDS of a block: [mu: z x 65 T][psib: z x 65 T][psiA: z T]
// [WAVES x (nu row, gm row): 2z T][WAVES x reduction buffer: z x 65 ACC]
template <typename T, typename ACC>
struct ApprLds {
    __host__ __device__ static constexpr size_t psib_off(uint32_t z) { return round16((size_t)z * kApprStride * sizeof(T)); }
    __host__ __device__ static constexpr size_t psiA_off(uint32_t z) {
        return psib_off(z) + round16((size_t)z * kApprStride * sizeof(T));
    }
    __host__ __device__ static constexpr size_t row_off(uint32_t z) { return psiA_off(z) + round16((size_t)z * sizeof(T)); }
    __host__ __device__ static constexpr size_t row_stride(uint32_t z) { return round16((size_t)2 * z * sizeof(T)); }
    __host__ __device__ static constexpr size_t buf_off(uint32_t z, int waves) { return row_off(z) + waves * row_stride(z); }
    __host__ __device__ static constexpr size_t buf_stride(uint32_t z) { return round16((size_t)z * kApprStride * sizeof(ACC)); }
    __host__ __device__ static constexpr size_t bytes(uint32_t z, int waves) { return buf_off(z, waves) + waves * buf_stride(z); }
};

template <typename T, typename ACC, int ZB>
__host__ __device__ constexpr int appr_waves() {
    int w = 4;
    while (w > 1 && ApprLds<T, ACC>::bytes(ZB, w) > kLdsBytes) --w;
    return w;
}

// Row reduction over the tile's columns: buf holds one value per (k, lane)
// at k * 65 + lane; the result for each k < z goes to out[k]. Lanes
// (kq, part) add strided columns in eight independent partial sums, combined
// in a fixed tree and then across parts with xor shuffles.
template <typename ACC>
__device__ __forceinline__ void appr_row_reduce(const ACC* buf, uint32_t z, uint32_t zq, uint32_t zq_log2,
                                                uint32_t nj, BNMF_GLOBAL ACC* out) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t kq = lane & (zq - 1u), part = lane >> zq_log2, L = 64u >> zq_log2;
    ACC r[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) r[v] = ACC(0);
    if (kq < z) {
        const ACC* const rb = buf + kq * kApprStride;
        for (uint32_t f0 = part; f0 < nj; f0 += 8u * L) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const uint32_t f = f0 + (uint32_t)v * L;
                if (f < nj) r[v] += rb[f];
            }
        }
    }
    ACC rs = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (uint32_t off = zq; off < 64u; off <<= 1) rs += __shfl_xor(rs, (int)off);
    if (part == 0 && kq < z) out[kq] = rs;
}

template <typename T, typename ACC, int ZB, int MODE, int PSI, int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
    bld_appr_pass_kernel(const ApprInst* __restrict__ insts, const PassItem* __restrict__ items,
                         const uint32_t* __restrict__ n_items_p) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    using Lds = ApprLds<T, ACC>;
    constexpr uint32_t NT = 64 * WAVES;
    constexpr uint32_t SK = kApprStride;
    const uint32_t n_items = *n_items_p;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));

    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const ApprInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const uint32_t i0 = item.ti * I->ti_rows;
        const uint32_t i1 = min(x, i0 + I->ti_rows);
        const uint32_t j0 = item.tj * kTileCols;
        const uint32_t nj = min(kTileCols, y - j0);
        const T epsT = static_cast<T>(I->eps);

        T* const mu_l = reinterpret_cast<T*>(lds_raw);
        T* const psib_l = reinterpret_cast<T*>(lds_raw + Lds::psib_off(z));
        T* const psiA_l = reinterpret_cast<T*>(lds_raw + Lds::psiA_off(z));
        T* const nu_w = reinterpret_cast<T*>(lds_raw + Lds::row_off(z) + w * Lds::row_stride(z));
        T* const gm_w = nu_w + z;
        unsigned char* const buf_base = lds_raw + Lds::buf_off(z, WAVES);
        ACC* const buf = reinterpret_cast<ACC*>(buf_base + w * Lds::buf_stride(z));

        // the tile's mu columns (and psi(beta_eph) rows), k-major; columns
        // past the instance's edge are zero
        {
            const BNMF_GLOBAL T* __restrict__ mu_g = gconst<T>(I->mu);
            for (uint32_t e = tid; e < z * 64u; e += NT) {
                const uint32_t k = e >> 6, f = e & 63u;
                mu_l[k * SK + f] = f < nj ? mu_g[(size_t)k * y + j0 + f] : T(0);
            }
            if constexpr (MODE != APPR_SUMS) {
                const BNMF_GLOBAL T* __restrict__ pb_g = gconst<T>(I->psib) + (size_t)j0 * z;
                for (uint32_t e = tid; e < z * 64u; e += NT) {
                    const uint32_t f = e / z, k = e - f * z;
                    psib_l[k * SK + f] = f < nj ? pb_g[e] : T(0);
                }
                if (tid < z) psiA_l[tid] = gconst<T>(I->psiA)[tid];
            }
            __syncthreads();
        }

        // column accumulators over the wave's rows: S_+jk (sums pass), or
        // mu's numerator and denominator (mu step)
        ACC c1[ZB], c2[ZB];
#pragma unroll
        for (int k = 0; k < ZB; ++k) c1[k] = c2[k] = ACC(0);

        for (uint32_t i = i0 + w; i < i1; i += WAVES) {
            if (lane < z) {
                nu_w[lane] = gconst<T>(I->nu)[(size_t)i * z + lane];
                if constexpr (MODE != APPR_SUMS)
                    gm_w[lane] = psiA_l[lane] - gconst<T>(I->psia)[(size_t)i * z + lane];
            }
            const T xv = lane < nj ? gconst<T>(I->X)[(size_t)i * y + j0 + lane] : T(0);
            wave_sync();
            // X_hat_ij = sum_k nu_ik mu_kj (update_X_hat :12-15)
            T h = T(0);
#pragma unroll
            for (int k = 0; k < ZB; ++k)
                if (k < (int)z) h += nu_w[k] * mu_l[k * SK + lane];
            // orig_over_appr (:17-23) and orig_over_appr_squared (:25-32). In
            // FP32 eps flushes to zero: an X_ij = 0 fiber contributes 0, as
            // the double guard makes it in the reference
            T oa, oa2;
            if constexpr (sizeof(T) == 8) {
                oa = xv / (h + epsT);
                oa2 = xv / (h * h + epsT);
            } else {
                const bool live = xv > 0.0f;
                oa = live ? xv / (h + epsT) : 0.0f;
                oa2 = live ? xv / (h * h + epsT) : 0.0f;
            }
            if constexpr (MODE == APPR_SUMS) {
                // S_ijk = oa nu_ik mu_kj (update_S :34-48) and its sums
#pragma unroll
                for (int k = 0; k < ZB; ++k) {
                    if (k < (int)z) {
                        const T s = (oa * nu_w[k]) * mu_l[k * SK + lane];
                        c1[k] += ACC(s);
                        buf[k * SK + lane] = ACC(s);
                    }
                }
                wave_sync();
                appr_row_reduce<ACC>(buf, z, I->zq, I->zq_log2, nj,
                                     gmut<ACC>(I->part_ipk) + ((size_t)item.tj * x + i) * z);
            } else {
                // grad_plus = psi(beta_eph) - psi(S+1); A = sum_c nu mu gm,
                // B = sum_c nu mu gp
                T gp[ZB];
                T A = T(0), B = T(0);
#pragma unroll
                for (int k = 0; k < ZB; ++k) {
                    if (k < (int)z) {
                        const T nk = nu_w[k], mk = mu_l[k * SK + lane];
                        const T s = (oa * nk) * mk;
                        T ps;
                        if constexpr (sizeof(T) == 8)
                            ps = psi_sel<PSI, T>(s + T(1));
                        else
                            ps = psi_ge1_f32(s + 1.0f);  // S >= 0
                        const T g = psib_l[k * SK + lane] - ps;
                        gp[k] = g;
                        const T nm = nk * mk;
                        A += nm * gm_w[k];
                        B += nm * g;
                    }
                }
                if constexpr (MODE == APPR_NU) {
                    // numerator, then denominator, summed over the row's columns
#pragma unroll
                    for (int k = 0; k < ZB; ++k) {
                        if (k < (int)z) {
                            const T mk = mu_l[k * SK + lane];
                            buf[k * SK + lane] = ACC((oa * mk) * gp[k] + (oa2 * mk) * A);
                        }
                    }
                    wave_sync();
                    appr_row_reduce<ACC>(buf, z, I->zq, I->zq_log2, nj,
                                         gmut<ACC>(I->part_ipk) + ((size_t)item.tj * x + i) * z);
                    wave_sync();
#pragma unroll
                    for (int k = 0; k < ZB; ++k) {
                        if (k < (int)z) {
                            const T mk = mu_l[k * SK + lane];
                            buf[k * SK + lane] = ACC((oa * mk) * gm_w[k] + (oa2 * mk) * B);
                        }
                    }
                    wave_sync();
                    appr_row_reduce<ACC>(buf, z, I->zq, I->zq_log2, nj,
                                         gmut<ACC>(I->part_ipk2) + ((size_t)item.tj * x + i) * z);
                } else {
#pragma unroll
                    for (int k = 0; k < ZB; ++k) {
                        if (k < (int)z) {
                            const T nk = nu_w[k];
                            c1[k] += ACC((oa * nk) * gp[k] + (oa2 * nk) * A);
                            c2[k] += ACC((oa * nk) * gm_w[k] + (oa2 * nk) * B);
                        }
                    }
                }
            }
            wave_sync();
        }

        // column partials over the tile's rows: the waves' sums meet in LDS
        // and are added in wave order
        if constexpr (MODE != APPR_NU) {
#pragma unroll
            for (int round = 0; round < (MODE == APPR_MU ? 2 : 1); ++round) {
                __syncthreads();
#pragma unroll
                for (int k = 0; k < ZB; ++k)
                    if (k < (int)z) buf[k * SK + lane] = round == 0 ? c1[k] : c2[k];
                __syncthreads();
                BNMF_GLOBAL ACC* __restrict__ pp =
                    gmut<ACC>(round == 0 ? I->part_pjk : I->part_pjk2) + ((size_t)item.ti * y + j0) * z;
                for (uint32_t e = tid; e < nj * z; e += NT) {
                    const uint32_t f = e / z, k = e - f * z;
                    ACC s = ACC(0);
#pragma unroll
                    for (int w2 = 0; w2 < WAVES; ++w2)
                        s += reinterpret_cast<const ACC*>(buf_base + w2 * Lds::buf_stride(z))[k * SK + f];
                    pp[e] = s;
                }
            }
        }
        __syncthreads();
    }
}

// Fibers deeper than 64 (z > 64): one wavefront per block walks a tile of
// ti_rows x kWideCols fibers one fiber at a time, lanes over k in steps of
// 64, and the sums over k (X_hat, A, B) are wavefront sums. The per-tile row
// partials (over the tile's columns) and column partials (over its rows) are
// accumulated in their own slots of the partial arrays by the lane that owns
// k, a store at the tile's first column (row) and read-add-store after, so
// each slot has one writer and a fixed order. The step modes recompute the
// fiber's S and grad_plus in a second walk over k once A and B are known.
template <typename T, typename ACC, int MODE, int PSI>
__global__ void __launch_bounds__(64)
    bld_appr_wide_kernel(const ApprInst* __restrict__ insts, const PassItem* __restrict__ items,
                         const uint32_t* __restrict__ n_items_p) {
    const uint32_t n_items = *n_items_p;
    const uint32_t lane = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const ApprInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const uint32_t i0 = item.ti * I->ti_rows, i1 = min(x, i0 + I->ti_rows);
        const uint32_t j0 = item.tj * kWideCols, j1 = min(y, j0 + kWideCols);
        const T epsT = static_cast<T>(I->eps);
        const BNMF_GLOBAL T* __restrict__ nu = gconst<T>(I->nu);
        const BNMF_GLOBAL T* __restrict__ mu = gconst<T>(I->mu);
        const BNMF_GLOBAL T* __restrict__ psib = gconst<T>(I->psib);
        const BNMF_GLOBAL T* __restrict__ psia = gconst<T>(I->psia);
        const BNMF_GLOBAL T* __restrict__ psiA = gconst<T>(I->psiA);
        // this tile's slots: rows [i*z + k] of tile column tj, columns
        // [j*z + k] of tile row ti
        BNMF_GLOBAL ACC* const rowp = gmut<ACC>(I->part_ipk) + (size_t)item.tj * x * z;
        BNMF_GLOBAL ACC* const rowp2 = gmut<ACC>(I->part_ipk2) + (size_t)item.tj * x * z;
        BNMF_GLOBAL ACC* const colp = gmut<ACC>(I->part_pjk) + (size_t)item.ti * y * z;
        BNMF_GLOBAL ACC* const colp2 = gmut<ACC>(I->part_pjk2) + (size_t)item.ti * y * z;
        for (uint32_t i = i0; i < i1; ++i) {
            const size_t iz = (size_t)i * z;
            for (uint32_t j = j0; j < j1; ++j) {
                const size_t jz = (size_t)j * z;
                const bool first_col = j == j0, first_row = i == i0;
                // X_hat_ij = sum_k nu_ik mu_kj (update_X_hat :12-15)
                T h = T(0);
                for (uint32_t k = lane; k < z; k += 64u) h += nu[iz + k] * mu[(size_t)k * y + j];
                h = wave_sum(h);
                const T xv = gconst<T>(I->X)[(size_t)i * y + j];
                T oa, oa2;
                if constexpr (sizeof(T) == 8) {
                    oa = xv / (h + epsT);
                    oa2 = xv / (h * h + epsT);
                } else {
                    const bool live = xv > 0.0f;
                    oa = live ? xv / (h + epsT) : 0.0f;
                    oa2 = live ? xv / (h * h + epsT) : 0.0f;
                }
                if constexpr (MODE == APPR_SUMS) {
                    // S_ijk (update_S :34-48) into both partials
                    for (uint32_t k = lane; k < z; k += 64u) {
                        const ACC s = ACC((oa * nu[iz + k]) * mu[(size_t)k * y + j]);
                        rowp[iz + k] = first_col ? s : rowp[iz + k] + s;
                        colp[jz + k] = first_row ? s : colp[jz + k] + s;
                    }
                } else {
                    // grad_plus_k = psi(beta_eph_jk) - psi(S_ijk + 1),
                    // A = sum_k nu mu gm_ik, B = sum_k nu mu gp_k
                    auto grad_plus = [&](uint32_t k, T nk, T mk) {
                        const T s = (oa * nk) * mk;
                        T ps;
                        if constexpr (sizeof(T) == 8)
                            ps = psi_sel<PSI, T>(s + T(1));
                        else
                            ps = psi_ge1_f32(s + 1.0f);  // S >= 0
                        return psib[jz + k] - ps;
                    };
                    T A = T(0), B = T(0);
                    for (uint32_t k = lane; k < z; k += 64u) {
                        const T nk = nu[iz + k], mk = mu[(size_t)k * y + j];
                        const T gm = psiA[k] - psia[iz + k];
                        const T nm = nk * mk;
                        A += nm * gm;
                        B += nm * grad_plus(k, nk, mk);
                    }
                    A = wave_sum(A);
                    B = wave_sum(B);
                    for (uint32_t k = lane; k < z; k += 64u) {
                        const T nk = nu[iz + k], mk = mu[(size_t)k * y + j];
                        const T gm = psiA[k] - psia[iz + k];
                        const T g = grad_plus(k, nk, mk);
                        if constexpr (MODE == APPR_NU) {
                            // update_nu's numerator and denominator over j
                            const ACC n = ACC((oa * mk) * g + (oa2 * mk) * A);
                            const ACC d = ACC((oa * mk) * gm + (oa2 * mk) * B);
                            rowp[iz + k] = first_col ? n : rowp[iz + k] + n;
                            rowp2[iz + k] = first_col ? d : rowp2[iz + k] + d;
                        } else {
                            // update_mu's numerator and denominator over i
                            const ACC n = ACC((oa * nk) * g + (oa2 * nk) * A);
                            const ACC d = ACC((oa * nk) * gm + (oa2 * nk) * B);
                            colp[jz + k] = first_row ? n : colp[jz + k] + n;
                            colp2[jz + k] = first_row ? d : colp2[jz + k] + d;
                        }
                    }
                }
            }
        }
    }
}

// update_nu's step and row normalisation (:155-163) or update_mu's step and
// column normalisation (:217-225). One wavefront per row i (HALF 0) or
// column j (HALF 1), lane k: the per-tile partials of the numerator and
// denominator are added in tile order, the step is applied, and the factor
// is renormalised by its sum over k (added in k order, as the reference's
// rowwise / colwise sum). A work item is four rows (columns), one per wave.
template <typename T, int HALF>
__global__ void __launch_bounds__(256)
    bld_appr_factor_kernel(const ApprInst* __restrict__ insts, const PassItem* __restrict__ items,
                           const uint32_t* __restrict__ n_items_p) {
    __shared__ T vbuf[4][64];
    __shared__ T rsum[4];
    const uint32_t n_items = *n_items_p;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const ApprInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const uint32_t n_rows = HALF == 0 ? x : y;
        const uint32_t r = item.ti * 4u + w;
        const T epsT = static_cast<T>(I->eps);
        // HALF 0: nu[r*z + k], partials part_ipk[(t*x + r)*z + k] over n_tj tiles
        // HALF 1: mu[k*y + r], partials part_pjk[(t*y + r)*z + k] over n_ti tiles
        BNMF_GLOBAL T* __restrict__ F = gmut<T>(HALF == 0 ? I->nu : I->mu);
        const BNMF_GLOBAL T* __restrict__ Pn = gconst<T>(HALF == 0 ? I->part_ipk : I->part_pjk);
        const BNMF_GLOBAL T* __restrict__ Pd = gconst<T>(HALF == 0 ? I->part_ipk2 : I->part_pjk2);
        const uint32_t nt = HALF == 0 ? I->n_tj : I->n_ti;
        const size_t tstride = (size_t)n_rows * z;
        if (z > 64u) {
            // deep rows: lanes step k by 64, the new values go to F and their
            // sum over k is a wavefront sum; F is then divided in place (each
            // lane rereads only what it wrote)
            if (r < n_rows) {
                T part = T(0);
                for (uint32_t k = lane; k < z; k += 64u) {
                    const size_t base = (size_t)r * z + k;
                    const size_t fk = HALF == 0 ? base : (size_t)k * y + r;
                    SumAcc<T> n, d;
                    for (uint32_t t = 0; t < nt; ++t) {
                        n.add(Pn[t * tstride + base]);
                        d.add(Pd[t * tstride + base]);
                    }
                    T v;
                    if constexpr (sizeof(T) == 8)
                        v = F[fk] * (n.get() / (d.get() + epsT));
                    else
                        v = n.get() != 0.0f ? F[fk] * (n.get() / (d.get() + epsT)) : 0.0f;
                    F[fk] = v;
                    part += v;
                }
                T rs = wave_sum(part) + epsT;
                if constexpr (sizeof(T) == 4) rs = rs != 0.0f ? rs : 1.0f;
                for (uint32_t k = lane; k < z; k += 64u) {
                    const size_t fk = HALF == 0 ? (size_t)r * z + k : (size_t)k * y + r;
                    F[fk] = F[fk] / rs;
                }
            }
            continue;
        }
        const bool mine = r < n_rows && lane < z;
        const size_t fi = HALF == 0 ? (size_t)r * z + lane : (size_t)lane * y + r;
        if (mine) {
            SumAcc<T> n, d;
            const size_t base = (size_t)r * z + lane;
            for (uint32_t t = 0; t < nt; ++t) {
                n.add(Pn[t * tstride + base]);
                d.add(Pd[t * tstride + base]);
            }
            T v;
            if constexpr (sizeof(T) == 8) {
                v = F[fi] * (n.get() / (d.get() + epsT));
            } else {
                // eps flushes to zero in float: a row with no data
                // (numerator 0) keeps a zero factor, as in double
                v = n.get() != 0.0f ? F[fi] * (n.get() / (d.get() + epsT)) : 0.0f;
            }
            vbuf[w][lane] = v;
        }
        wave_sync();
        if (r < n_rows && lane == 0) {
            T rs = T(0);
            for (uint32_t k = 0; k < z; ++k) rs += vbuf[w][k];
            rs += epsT;
            if constexpr (sizeof(T) == 4) rs = rs != 0.0f ? rs : 1.0f;
            rsum[w] = rs;
        }
        wave_sync();
        if (mine) F[fi] = vbuf[w][lane] / rsum[w];
        wave_sync();
    }
}

// The returned S (bld_appr.hpp:366-369): one thread per fiber.
template <typename T>
__global__ void __launch_bounds__(256) bld_appr_S_kernel(const ApprInst* __restrict__ I) {
    const size_t f = (size_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t y = I->y, z = I->z;
    if (f >= (size_t)I->x * y) return;
    const size_t i = f / y, j = f - i * y;
    const BNMF_GLOBAL T* __restrict__ nu = gconst<T>(I->nu) + i * z;
    const BNMF_GLOBAL T* __restrict__ mu = gconst<T>(I->mu) + j;
    const T xv = gconst<T>(I->X)[f];
    const T epsT = static_cast<T>(I->eps);
    T h = T(0);
    for (uint32_t k = 0; k < z; ++k) h += nu[k] * mu[(size_t)k * y];
    T oa;
    if constexpr (sizeof(T) == 8)
        oa = xv / (h + epsT);
    else
        oa = xv > 0.0f ? xv / (h + epsT) : 0.0f;
    BNMF_GLOBAL T* __restrict__ S = gmut<T>(I->S) + f * z;
    for (uint32_t k = 0; k < z; ++k) S[k] = (oa * nu[k]) * mu[(size_t)k * y];
}

// ----------------------------------------------------------- launchers
template <typename T, typename ACC, int ZB, int MODE, int PSI>
static hipError_t launch_appr_t(dim3 grid, hipStream_t st, const ApprInst* insts, const PassItem* items,
                                const uint32_t* n_items, uint32_t zmax) {
    constexpr int WAVES = appr_waves<T, ACC, ZB>();
    const size_t lds = ApprLds<T, ACC>::bytes(zmax, WAVES);
    hipLaunchKernelGGL((bld_appr_pass_kernel<T, ACC, ZB, MODE, PSI, WAVES>), grid, dim3(64 * WAVES), lds, st,
                       insts, items, n_items);
    return hipGetLastError();
}

template <typename T, typename ACC, int MODE, int PSI>
static hipError_t launch_appr_zb(int zb, dim3 grid, hipStream_t st, const ApprInst* insts,
                                 const PassItem* items, const uint32_t* n, uint32_t zmax) {
    switch (zb) {
        case 0:  // z > 64
            hipLaunchKernelGGL((bld_appr_wide_kernel<T, ACC, MODE, PSI>), grid, dim3(64), 0, st, insts, items, n);
            return hipGetLastError();
        case 8: return launch_appr_t<T, ACC, 8, MODE, PSI>(grid, st, insts, items, n, zmax);
        case 16: return launch_appr_t<T, ACC, 16, MODE, PSI>(grid, st, insts, items, n, zmax);
        case 32: return launch_appr_t<T, ACC, 32, MODE, PSI>(grid, st, insts, items, n, zmax);
        case 64: return launch_appr_t<T, ACC, 64, MODE, PSI>(grid, st, insts, items, n, zmax);
        default: return hipErrorInvalidValue;
    }
}

// mode: 0 sums, 1 nu step, 2 mu step; dtype 0 float, 1 double
hipError_t launch_appr_pass(int dtype, int mode, int zb, dim3 grid, hipStream_t st, const ApprInst* insts,
                            const PassItem* items, const uint32_t* n_items, uint32_t zmax) {
    if (dtype == 0) {
        switch (mode) {
            case APPR_SUMS: return launch_appr_zb<float, float, APPR_SUMS, PSI_F32>(zb, grid, st, insts, items, n_items, zmax);
            case APPR_NU: return launch_appr_zb<float, float, APPR_NU, PSI_F32>(zb, grid, st, insts, items, n_items, zmax);
            case APPR_MU: return launch_appr_zb<float, float, APPR_MU, PSI_F32>(zb, grid, st, insts, items, n_items, zmax);
        }
    } else {
        switch (mode) {
            case APPR_SUMS: return launch_appr_zb<double, double, APPR_SUMS, PSI_F64_APPR>(zb, grid, st, insts, items, n_items, zmax);
            case APPR_NU: return launch_appr_zb<double, double, APPR_NU, PSI_F64_APPR>(zb, grid, st, insts, items, n_items, zmax);
            case APPR_MU: return launch_appr_zb<double, double, APPR_MU, PSI_F64_APPR>(zb, grid, st, insts, items, n_items, zmax);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_appr_factor(int dtype, int half, dim3 grid, hipStream_t st, const ApprInst* insts,
                              const PassItem* items, const uint32_t* n_items) {
    if (dtype == 0) {
        if (half == 0)
            hipLaunchKernelGGL((bld_appr_factor_kernel<float, 0>), grid, dim3(256), 0, st, insts, items, n_items);
        else
            hipLaunchKernelGGL((bld_appr_factor_kernel<float, 1>), grid, dim3(256), 0, st, insts, items, n_items);
    } else {
        if (half == 0)
            hipLaunchKernelGGL((bld_appr_factor_kernel<double, 0>), grid, dim3(256), 0, st, insts, items, n_items);
        else
            hipLaunchKernelGGL((bld_appr_factor_kernel<double, 1>), grid, dim3(256), 0, st, insts, items, n_items);
    }
    return hipGetLastError();
}

hipError_t launch_appr_S(int dtype, const ApprInst* inst, size_t n_fibers, hipStream_t st) {
    const dim3 grid((unsigned)((n_fibers + 255) / 256));
    if (dtype == 0)
        hipLaunchKernelGGL(bld_appr_S_kernel<float>, grid, dim3(256), 0, st, inst);
    else
        hipLaunchKernelGGL(bld_appr_S_kernel<double>, grid, dim3(256), 0, st, inst);
    return hipGetLastError();
}

template <typename T, typename ACC, int MODE, int PSI>
static hipError_t attr_appr_mode() {
    hipError_t e = hipSuccess;
    auto one = [&](auto zbc) {
        constexpr int ZB = decltype(zbc)::value;
        constexpr int WAVES = appr_waves<T, ACC, ZB>();
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(bld_appr_pass_kernel<T, ACC, ZB, MODE, PSI, WAVES>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)ApprLds<T, ACC>::bytes(ZB, WAVES));
    };
    one(std::integral_constant<int, 8>{});
    one(std::integral_constant<int, 16>{});
    one(std::integral_constant<int, 32>{});
    one(std::integral_constant<int, 64>{});
    return e;
}

hipError_t init_appr_attributes() {
    hipError_t e = attr_appr_mode<float, float, APPR_SUMS, PSI_F32>();
    if (e == hipSuccess) e = attr_appr_mode<float, float, APPR_NU, PSI_F32>();
    if (e == hipSuccess) e = attr_appr_mode<float, float, APPR_MU, PSI_F32>();
    if (e == hipSuccess) e = attr_appr_mode<double, double, APPR_SUMS, PSI_F64_APPR>();
    if (e == hipSuccess) e = attr_appr_mode<double, double, APPR_NU, PSI_F64_APPR>();
    if (e == hipSuccess) e = attr_appr_mode<double, double, APPR_MU, PSI_F64_APPR>();
    return e;
}

}  // namespace bnmf
