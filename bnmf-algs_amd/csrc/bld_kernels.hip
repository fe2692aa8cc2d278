// bld_kernels.hip — CDNA4 (gfx950) kernels of the multiplicative BLD solver.
//
// One solver iteration of the reference (src/bld/bld_mult/bld_mult.hpp:123-147)
// is two launches here:
//
//  prep  (small, O((x+y)z)): finishes the axis sums of the current S from
//        per-tile partials, adds the Dirichlet parameters and evaluates
//        psi(beta_eph) (update_beta_eph + the psi half of update_grad_plus,
//        bld_mult_funcs.hpp:152-163,177-192) and psi(alpha_eph) together with
//        column partials of alpha_eph (update_alpha_eph + update_grad_minus,
//        :130-141,205-229).
//  pass  (the hot path, one read and one write of S): per fiber (i,j) it
//        forms grad_plus in registers, the fiber sums S_ij+, nom_ij and
//        denom_ij (update_nom_mult / update_denom_mult, :242-290), applies
//        update_S (:307-327), and accumulates the partial axis sums S_+jk and
//        S_i+k of the NEW tensor for the next iteration's prep.
//
// Pass mapping: a work item is a tile of TI rows x 64 columns of one
// instance. A block of WAVES wavefronts shares the tile's columns; wave w
// walks rows i0+w, i0+w+WAVES, ... For each row the 64 fibers of the tile are
// one contiguous run of 64*z elements of S: lanes load it coalesced (element
// e = c*64 + lane), scatter it into a wave-private LDS buffer with an odd
// fiber stride zp (bank-conflict free), then each lane owns one fiber and
// runs the z-long fiber reductions sequentially in registers, so no
// cross-lane reduction is needed for S_ij+, nom or denom. The updated fiber
// goes back through LDS and leaves with coalesced stores. Row sums over the
// 64 fibers (S_i+k partial) are LDS column sums; column sums over the rows
// (S_+jk partial) stay in registers across the row loop. The next row's
// loads are issued before the current row is computed.
//
// All cross-block reductions are two-stage (per-tile partials, then a
// fixed-order sum in prep), never atomics, so every result is bitwise
// reproducible and independent of which instances share a launch.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "bld_types.hpp"
#include "digamma.hpp"
#include "pass_common.hpp"
#include "philox.hpp"

namespace bnmf {

template <typename T, typename ACC, int ZB, int RW, int MODE, int PSI, int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
    bld_pass_kernel(const BldInst* __restrict__ insts, const PassItem* __restrict__ items,
                    const uint32_t* __restrict__ n_items_p) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    using Lds = PassLds<T, ACC, MODE>;
    constexpr uint32_t NT = 64 * WAVES;
    constexpr uint32_t VEC = 16 / sizeof(T);
    constexpr int CPL = (64 * ZB + 2 * VEC + 64 * VEC - 1) / (64 * VEC);  // 16-byte chunks per lane
    constexpr int NCH = ZB / RW;                                        // fiber chunks
    typedef T V16 __attribute__((ext_vector_type(VEC)));
    typedef typename VecR<T, RW>::type VR;

    const uint32_t n_items = *n_items_p;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = tid >> 6;

    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const BldInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const uint32_t i0 = item.ti * I->ti_rows;
        const uint32_t i1 = min(x, i0 + I->ti_rows);
        const uint32_t j0 = item.tj * I->tile_cols;
        const uint32_t nj = min(I->tile_cols, y - j0);
        const uint32_t n_el = nj * z;
        const uint32_t nchunk = z / RW;                                 // fiber chunks in use
        const uint32_t hsw = RW > 1 ? ((lane >> I->hs) & I->hm) : 0u;  // visit order

        T* const psib_l = reinterpret_cast<T*>(lds_raw);
        T* const psiA_l = reinterpret_cast<T*>(lds_raw + Lds::psiA_off(z));
        T* const gm_l = reinterpret_cast<T*>(lds_raw + Lds::gm_off(z) + w * Lds::gm_stride(z));
        unsigned char* const raw_base = lds_raw + Lds::raw_off(z, WAVES);
        T* const raw = reinterpret_cast<T*>(raw_base + w * Lds::raw_stride(z));

        const BNMF_GLOBAL T* __restrict__ Ssrc = gconst<T>(I->S_src);
        BNMF_GLOBAL T* __restrict__ Sdst = gmut<T>(I->S);
        const double eps = I->eps;
        const bool fast_psi = I->nonneg != 0u;

        if constexpr (pass_steps(MODE)) {
            const BNMF_GLOBAL T* __restrict__ psib_g = gconst<T>(I->psib) + (size_t)j0 * z;
            for (uint32_t e = tid; e < n_el; e += NT) psib_l[e] = psib_g[e];
            if (tid < z) psiA_l[tid] = gconst<T>(I->psiA)[tid];
            __syncthreads();
        }
        // additive step size eta = 0.1 / (iteration + 1)^0.55 (bld_add.hpp:101-103);
        // *iter counts the iterations started, this one included
        double eta = 0.0;
        if constexpr (MODE == PASS_ADD) eta = 0.1 / pow((double)*I->iter, 0.55);

        ACC acc[ZB];
#pragma unroll
        for (int k = 0; k < ZB; ++k) acc[k] = ACC(0);
        double lg_sum = 0.0;

        const uint32_t kq = lane & (I->zq - 1u);
        const uint32_t part = lane >> I->zq_log2;
        const uint32_t L = 64u >> I->zq_log2;

        // prefetch registers: one tile row as 16-byte chunks from the aligned
        // start (delta elements before the row's first element)
        V16 pf[CPL];
        uint32_t i = i0 + w;
        uint32_t dl = 0, nch = 0;
        auto issue_row = [&](uint32_t row, uint32_t& d_out, uint32_t& n_out) {
            const size_t s0 = ((size_t)row * y + j0) * z;
            const uint32_t d = (uint32_t)(s0 % VEC);
            const uint32_t n = (d + n_el + VEC - 1) / VEC;
            const BNMF_GLOBAL V16* src = (const BNMF_GLOBAL V16*)(Ssrc + (s0 - d));
            const uint32_t nfull = n >> 6;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const uint32_t q = (uint32_t)c * 64u + lane;
                if ((uint32_t)c < nfull) {
                    pf[c] = src[q];
                } else if ((uint32_t)c == nfull && q < n) {
                    pf[c] = src[q];
                }
            }
            d_out = d;
            n_out = n;
        };
        if (i < i1) issue_row(i, dl, nch);
        for (; i < i1; i += WAVES) {
            // the current row into the wave's row buffer (natural layout)
            {
                const uint32_t nfull = nch >> 6;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const uint32_t q = (uint32_t)c * 64u + lane;
                    if ((uint32_t)c < nfull || ((uint32_t)c == nfull && q < nch))
                        reinterpret_cast<V16*>(raw)[q] = pf[c];
                }
            }
            const uint32_t cur_dl = dl, cur_nch = nch;
            const size_t cur_s0 = ((size_t)i * y + j0) * z;
            // issue the next row's loads
            const uint32_t inext = i + WAVES;
            if (inext < i1) issue_row(inext, dl, nch);

            if constexpr (MODE == PASS_UPDATE) {
                const BNMF_GLOBAL T* __restrict__ psia_g = gconst<T>(I->psia) + (size_t)i * z;
                if (lane < z) gm_l[lane] = psiA_l[lane] - psia_g[lane];
                T xv = T(0);
                if (lane < nj) xv = gconst<T>(I->X)[(size_t)i * y + j0 + lane];
                wave_sync();
                if (lane < nj) {
                    T* const fb = raw + cur_dl + lane * z;
                    const T* const pbb = psib_l + lane * z;
                    T gp[ZB];
                    T sumS = T(0), nom = T(0), den = T(0);
                    if constexpr (sizeof(T) == 8) {
                        // FP64 plain form: the reference's operations
                        // (X_reciprocal :95, nom :254, denom :284-285,
                        // update :319-323); fiber sums run in visit order.
                        const T xr = T(1.0) / (xv + eps);
#pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            if ((uint32_t)m < nchunk) {
                                const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                                const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                                const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                                const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                                for (int v = 0; v < RW; ++v) {
                                    const T sk = vget<T, RW>(sv, v);
                                    const T g = vget<T, RW>(pb, v) - psi_sp1<PSI, T>(sk);
                                    gp[m * RW + v] = g;
                                    sumS += sk;
                                    nom += sk * xr * vget<T, RW>(gv, v);
                                    den += sk * xr * g;
                                }
                            }
                        }
                        const T x_over_s = xv / (sumS + eps);
#pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            if ((uint32_t)m < nchunk) {
                                const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                                VR sv = *reinterpret_cast<const VR*>(fb + mm);
                                const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                                for (int v = 0; v < RW; ++v) {
                                    const T sk = vget<T, RW>(sv, v);
                                    const T sn = as_stored(sk * ((gp[m * RW + v] + nom) * x_over_s /
                                                                 (vget<T, RW>(gv, v) + den + eps)));
                                    vset<T, RW>(sv, v, sn);
                                    acc[m * RW + v] += ACC(sn);
                                }
                                *reinterpret_cast<VR*>(fb + mm) = sv;
                            }
                        }
                    } else {
                        // FP32 bulk. eps = 1e-50 is not representable in
                        // float; the reference's double-promoted guard makes
                        // an X_ij = 0 or all-zero fiber update to exactly 0,
                        // which is taken explicitly here.
                        const float epsf = static_cast<float>(eps);
                        const bool live = xv > 0.0f;
                        const float xr = live ? 1.0f / (xv + epsf) : 0.0f;
                        // The psi form is chosen once per fiber (a uniform
                        // branch); inside, the chunk loop is straight-line
                        // code the compiler can schedule across elements.
                        auto phase_a = [&](auto psi_of) {
#pragma unroll
                            for (int m = 0; m < NCH; ++m) {
                                if ((uint32_t)m < nchunk) {
                                    const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                                    const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                                    const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                                    const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                                    for (int v = 0; v < RW; ++v) {
                                        const float sk = vget<T, RW>(sv, v);
                                        const float g = vget<T, RW>(pb, v) - psi_of(sk);
                                        gp[m * RW + v] = g;
                                        sumS += sk;
                                        const float t = sk * xr;
                                        nom = fmaf(t, vget<T, RW>(gv, v), nom);
                                        den = fmaf(t, g, den);
                                    }
                                }
                            }
                        };
                        if (fast_psi) {
                            // S >= 0 throughout when S0 >= 0 and beta >= 1,
                            // so s + 1 >= 1: no guard for negative S
                            phase_a([](float sk) { return psi_ge1_f32(sk + 1.0f); });
                        } else {
                            phase_a([](float sk) { return psi_sp1<PSI_F32, float>(sk); });
                        }
                        const bool upd = live && sumS != 0.0f;
                        const float x_over_s = upd ? xv / (sumS + epsf) : 0.0f;
                        // an exact-zero denominator (grad_plus = grad_minus = 0,
                        // e.g. a single row with beta = 1) would make 0/0; the
                        // reference's double eps guard makes that update 0
                        float den_e = den + epsf;
                        if (den_e == 0.0f) den_e = FLT_MIN;
#pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            if ((uint32_t)m < nchunk) {
                                const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                                VR sv = *reinterpret_cast<const VR*>(fb + mm);
                                const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                                for (int v = 0; v < RW; ++v) {
                                    const float sk = vget<T, RW>(sv, v);
                                    const float fct = (gp[m * RW + v] + nom) * x_over_s *
                                                      recip_fast<float>(vget<T, RW>(gv, v) + den_e);
                                    const float sn = as_stored(upd ? sk * fct : 0.0f);
                                    vset<T, RW>(sv, v, sn);
                                    acc[m * RW + v] += ACC(sn);
                                }
                                *reinterpret_cast<VR*>(fb + mm) = sv;
                            }
                        }
                    }
                }
            } else if constexpr (MODE == PASS_ADD) {
                // bld::bld_add's iteration (src/bld/bld_add.hpp:137-181) per
                // fiber. The LDS row holds psi(alpha_eph(i,k)) in FP64 (the
                // reference's operation order) and psi(A) - psi(alpha_eph)
                // in FP32 (the exact-z kernel's order)
                const BNMF_GLOBAL T* __restrict__ psia_g = gconst<T>(I->psia) + (size_t)i * z;
                if (lane < z) gm_l[lane] = sizeof(T) == 8 ? psia_g[lane] : psiA_l[lane] - psia_g[lane];
                T xv = T(0);
                if (lane < nj) xv = gconst<T>(I->X)[(size_t)i * y + j0 + lane];
                wave_sync();
                if (lane < nj) {
                    T* const fb = raw + cur_dl + lane * z;
                    const T* const pbb = psib_l + lane * z;
                    T gz[ZB];  // grad_S, then exp(Z)
                    T smult = T(0);
                    // grad_S = psi(beta_eph) - psi(S+1) - psi(A) + psi(alpha_eph)
                    // (:137-146) and S_mult = sum_k S grad_S (:148-155)
#pragma unroll
                    for (int m = 0; m < NCH; ++m) {
                        if ((uint32_t)m < nchunk) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                            const VR pA = *reinterpret_cast<const VR*>(psiA_l + mm);
                            const VR pa = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const T sk = vget<T, RW>(sv, v);
                                if constexpr (sizeof(T) == 8) {
                                    const T ps = psi_sp1<PSI, T>(sk);
                                    const T g = ((vget<T, RW>(pb, v) - ps) - vget<T, RW>(pA, v)) + vget<T, RW>(pa, v);
                                    gz[m * RW + v] = g;
                                    smult += sk * g;
                                } else {
                                    const float ps = fast_psi ? psi_ge1_f32(sk + 1.0f) : psi_sp1<PSI_F32, float>(sk);
                                    const float g = (vget<T, RW>(pb, v) - ps) - vget<T, RW>(pa, v);
                                    gz[m * RW + v] = g;
                                    smult = fmaf(sk, g, smult);
                                }
                            }
                        }
                    }
                    // Z = log S + eps + eta S (X grad_S - S_mult) (:150,158-166),
                    // exp(Z) and its L1 norm along k with util::normalize's
                    // default 1e-50 guard (:167-173, util.hpp:321-383)
                    T norm = T(0);
#pragma unroll
                    for (int m = 0; m < NCH; ++m) {
                        if ((uint32_t)m < nchunk) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const T sk = vget<T, RW>(sv, v);
                                T e;
                                if constexpr (sizeof(T) == 8) {
                                    T zk = log(sk) + eps;
                                    zk += eta * sk * (xv * gz[m * RW + v] - smult);
                                    e = exp(zk);
                                } else {
                                    // exp(log s + t) = s exp(t): one rounding of
                                    // s instead of log's (0 stays 0)
                                    const float etaf = static_cast<float>(eta);
                                    e = sk * __expf((etaf * sk) * fmaf(xv, gz[m * RW + v], -smult));
                                }
                                gz[m * RW + v] = e;
                                norm += e;
                            }
                        }
                    }
                    // S = X exp(Z) / (|exp(Z)|_1 + 1e-50) (:175-181). In FP32
                    // the guard underflows to 0: an all-zero row (X = 0)
                    // is taken explicitly
                    T scale;
                    if constexpr (sizeof(T) == 8) {
                        norm += 1e-50;
                        scale = T(1);
                    } else {
                        scale = norm > 0.0f ? xv / norm : 0.0f;
                    }
#pragma unroll
                    for (int m = 0; m < NCH; ++m) {
                        if ((uint32_t)m < nchunk) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            VR sv;
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                T sn;
                                if constexpr (sizeof(T) == 8)
                                    sn = as_stored(xv * (gz[m * RW + v] / norm));
                                else
                                    sn = as_stored(gz[m * RW + v] * scale);
                                vset<T, RW>(sv, v, sn);
                                acc[m * RW + v] += ACC(sn);
                            }
                            *reinterpret_cast<VR*>(fb + mm) = sv;
                        }
                    }
                }
            } else {
                wave_sync();
                if (lane < nj) {
                    const T* const fb = raw + cur_dl + lane * z;
                    ACC fs = ACC(0);
#pragma unroll
                    for (int k = 0; k < ZB; ++k) {
                        if (k < (int)z) {
                            const T sk = fb[k];
                            acc[k] += ACC(sk);
                            fs += ACC(sk);
                            if constexpr (MODE == PASS_SCORE) lg_sum += lgamma((double)sk + 1.0);
                        }
                    }
                    if constexpr (MODE == PASS_SUMS)
                        if (I->sijp != nullptr) gmut<ACC>(I->sijp)[(size_t)i * y + j0 + lane] = fs;
                }
            }
            wave_sync();

            // S_i+k partial over the tile's columns: eight independent
            // partial sums combined in a fixed tree (deterministic)
            {
                ACC r[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) r[v] = ACC(0);
                if (kq < z) {
                    const T* const rb = raw + cur_dl + kq;
                    for (uint32_t f0 = part; f0 < nj; f0 += 8u * L) {
#pragma unroll
                        for (int v = 0; v < 8; ++v) {
                            const uint32_t f = f0 + (uint32_t)v * L;
                            if (f < nj) r[v] += ACC(rb[f * z]);
                        }
                    }
                }
                ACC rs = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
                for (uint32_t off = I->zq; off < 64u; off <<= 1) rs += __shfl_xor(rs, (int)off);
                if (part == 0 && kq < z)
                    gmut<ACC>(I->part_ipk)[((size_t)item.tj * x + i) * z + kq] = rs;
            }

            if constexpr (pass_steps(MODE)) {
                // 16-byte stores of every chunk that lies inside the row
                // segment; the (at most two) chunks straddling its ends are
                // written element by element so neighbours are not touched.
                BNMF_GLOBAL T* __restrict__ dst = Sdst + (cur_s0 - cur_dl);
                const uint32_t lo = cur_dl, hi = cur_dl + n_el;
                const uint32_t nfull = cur_nch >> 6;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const uint32_t q = (uint32_t)c * 64u + lane;
                    if ((uint32_t)c < nfull || ((uint32_t)c == nfull && q < cur_nch)) {
                        const uint32_t e0 = q * VEC;
                        if (e0 >= lo && e0 + VEC <= hi) {
                            reinterpret_cast<BNMF_GLOBAL V16*>(dst)[q] =
                                reinterpret_cast<const V16*>(raw)[q];
                        } else {
                            for (uint32_t v = 0; v < VEC; ++v)
                                if (e0 + v >= lo && e0 + v < hi) dst[e0 + v] = raw[e0 + v];
                        }
                    }
                }
            }
            wave_sync();
        }

        // S_+jk partial over the tile's rows: the waves' column sums meet in
        // LDS (natural fiber layout) and are added in wave order
        __syncthreads();
        ACC* const comb_self = reinterpret_cast<ACC*>(raw_base + w * Lds::raw_stride(z));
        if (lane < nj) {
            if constexpr (pass_steps(MODE)) {
#pragma unroll
                for (int m = 0; m < NCH; ++m) {
                    if ((uint32_t)m < nchunk) {
                        const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                        // RW-wide writes in the chunk visit order: lanes of
                        // one access cycle on different banks
                        VR cv;
#pragma unroll
                        for (int v = 0; v < RW; ++v) vset<T, RW>(cv, v, acc[m * RW + v]);
                        *reinterpret_cast<VR*>(comb_self + lane * z + mm) = cv;
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < ZB; ++k)
                    if (k < (int)z) comb_self[lane * z + k] = acc[k];
            }
        }
        __syncthreads();
        {
            BNMF_GLOBAL ACC* __restrict__ pp = gmut<ACC>(I->part_pjk) + ((size_t)item.ti * y + j0) * z;
            for (uint32_t e = tid; e < n_el; e += NT) {
                ACC sacc = ACC(0);
#pragma unroll
                for (int w2 = 0; w2 < WAVES; ++w2)
                    sacc += reinterpret_cast<const ACC*>(raw_base + w2 * Lds::raw_stride(z))[e];
                pp[e] = sacc;
            }
        }
        if constexpr (MODE == PASS_SCORE) {
            // block sum of the lgamma(S+1) lane sums, fixed tree order
            __syncthreads();
            double* red = reinterpret_cast<double*>(raw_base);
            for (uint32_t off = 32; off > 0; off >>= 1) lg_sum += __shfl_xor(lg_sum, (int)off);
            if (lane == 0) red[w] = lg_sum;
            __syncthreads();
            if (tid == 0) {
                double sacc = 0.0;
                for (int w2 = 0; w2 < WAVES; ++w2) sacc += red[w2];
                I->score[(size_t)item.ti * I->n_tj + item.tj] = sacc;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// prep kernel
// ---------------------------------------------------------------------------
// s += base[t * stride] for t = 0..n-1 in order; the loads are issued eight
// at a time ahead of their (sequential) additions.
template <typename T>
__device__ __forceinline__ void add_strided(SumAcc<T>& s, const T* __restrict__ base,
                                            size_t stride, uint32_t n) {
    uint32_t t = 0;
    for (; t + 8 <= n; t += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = base[(size_t)(t + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) s.add(v[u]);
    }
    for (; t < n; ++t) s.add(base[(size_t)t * stride]);
}

template <typename T, int PSI, bool RAW>
__global__ void __launch_bounds__(256)
    bld_prep_kernel(const BldInst* __restrict__ insts, const PrepItem* __restrict__ items,
                    const uint32_t* __restrict__ n_items_p) {
    __shared__ T aeph[kPrepRows * 64];
    const uint32_t n_items = *n_items_p;
    const uint32_t tid = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PrepItem item = items[it];
        const BldInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const float inv_z = I->inv_z;
        if (item.kind == 0) {
            const uint32_t c0 = item.chunk * kPrepCols;
            const uint32_t nc = min(kPrepCols, y - c0);
            const T* __restrict__ pp = static_cast<const T*>(I->part_pjk);
            T* __restrict__ psib = static_cast<T*>(I->psib);
            const size_t stride = (size_t)y * z;
            for (uint32_t e = tid; e < nc * z; e += 256) {
                const size_t idx = (size_t)c0 * z + e;
                const uint32_t k = z <= kFusedZ ? e - fiber_of(e, inv_z) * z : e % z;
                SumAcc<T> s;
                add_strided(s, pp + idx, stride, I->n_ti);
                if constexpr (RAW) {
                    psib[idx] = s.get();
                } else {
                    const T beta_eph = s.get() + static_cast<T>(I->beta[k]);
                    psib[idx] = psi_sel<PSI, T>(beta_eph);
                }
            }
        } else if (z > kFusedZ) {
            // deep fibers: one thread per k walks the chunk's rows in order,
            // so the alpha_eph column sums need no LDS staging
            const uint32_t r0 = item.chunk * kPrepRows;
            const uint32_t nr = min(kPrepRows, x - r0);
            const T* __restrict__ pp = static_cast<const T*>(I->part_ipk);
            T* __restrict__ psia = static_cast<T*>(I->psia);
            const size_t stride = (size_t)x * z;
            for (uint32_t k = tid; k < z; k += 256) {
                SumAcc<T> A;
                for (uint32_t r = 0; r < nr; ++r) {
                    const size_t idx = (size_t)(r0 + r) * z + k;
                    SumAcc<T> s;
                    add_strided(s, pp + idx, stride, I->n_tj);
                    if constexpr (RAW) {
                        psia[idx] = s.get();
                    } else {
                        const T alpha_eph = s.get() + static_cast<T>(I->alpha[r0 + r]);
                        psia[idx] = psi_sel<PSI, T>(alpha_eph);
                        A.add(alpha_eph);
                    }
                }
                if constexpr (!RAW) static_cast<T*>(I->partA)[(size_t)item.chunk * z + k] = A.get();
            }
        } else {
            const uint32_t r0 = item.chunk * kPrepRows;
            const uint32_t nr = min(kPrepRows, x - r0);
            const T* __restrict__ pp = static_cast<const T*>(I->part_ipk);
            T* __restrict__ psia = static_cast<T*>(I->psia);
            const size_t stride = (size_t)x * z;
            for (uint32_t e = tid; e < nr * z; e += 256) {
                const size_t idx = (size_t)r0 * z + e;
                const uint32_t r = fiber_of(e, inv_z);
                SumAcc<T> s;
                add_strided(s, pp + idx, stride, I->n_tj);
                if constexpr (RAW) {
                    psia[idx] = s.get();
                } else {
                    const T alpha_eph = s.get() + static_cast<T>(I->alpha[r0 + r]);
                    psia[idx] = psi_sel<PSI, T>(alpha_eph);
                    aeph[e] = alpha_eph;
                }
            }
            __syncthreads();
            if (!RAW && tid < z) {
                SumAcc<T> A;
                for (uint32_t r = 0; r < nr; ++r) A.add(aeph[r * z + tid]);
                static_cast<T*>(I->partA)[(size_t)item.chunk * z + tid] = A.get();
            }
            __syncthreads();
        }
    }
}

// psi(A_k), A_k = sum_i alpha_eph(i,k) (the first term of update_grad_minus,
// bld_mult_funcs.hpp:213-220), once per instance after prep: the prep row
// chunks' column sums (partA) are reduced by P threads per k over strided
// chunks, then the P partials are added in thread order (fixed tree).
template <typename T, int PSI>
__global__ void __launch_bounds__(256)
    bld_psiA_kernel(const BldInst* __restrict__ insts, uint32_t n_inst) {
    __shared__ T part[256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t n = blockIdx.x; n < n_inst; n += gridDim.x) {
        const BldInst* __restrict__ I = insts + n;
        const uint32_t z = I->z, nA = I->nA;
        if ((size_t)I->x * I->y * z == 0) continue;
        if (z > kFusedZ) {
            // deep fibers: one thread per k sums the chunk sums in order
            for (uint32_t k = tid; k < z; k += 256) {
                SumAcc<T> B;
                add_strided(B, static_cast<const T*>(I->partA) + k, z, nA);
                gmut<T>(I->psiA)[k] = psi_sel<PSI, T>(B.get());
            }
            if (tid == 0 && I->iter != nullptr) I->iter[0] += 1u;
            __syncthreads();
            continue;
        }
        const uint32_t P = 256u / z;  // partial sums per k (z <= 64: P >= 4)
        const uint32_t k = tid % z, p = tid / z;
        const BNMF_GLOBAL T* __restrict__ pA = gconst<T>(I->partA);
        SumAcc<T> A;
        if (p < P && p < nA)
            add_strided(A, (const T*)(pA + (size_t)p * z + k), (size_t)P * z, (nA - p + P - 1) / P);
        part[tid] = A.get();
        __syncthreads();
        if (tid < z) {
            SumAcc<T> B;
            for (uint32_t q = 0; q < P; ++q) B.add(part[q * z + tid]);
            gmut<T>(I->psiA)[tid] = psi_sel<PSI, T>(B.get());
        }
        // the additive solver's iteration count: this kernel runs once per
        // instance and iteration, before the iteration's pass
        if (tid == 0 && I->iter != nullptr) I->iter[0] += 1u;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// log-marginal scoring (alloc_model_funcs.hpp:574-586), FP64 throughout.
// Needs a PASS_SCORE pass first: double partials of S_+jk / S_i+k and the
// per-tile sums of lgamma(S_ijk + 1) (fourth term, :179-194).
// score buffer layout: [n_ti*n_tj tile lgamma sums | nA row-chunk values |
//                       n_cj column-chunk values | nA*z row-chunk column sums]
// ---------------------------------------------------------------------------
__device__ inline double block_sum_256(double v, double* red) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t off = 32; off > 0; off >>= 1) v += __shfl_xor(v, (int)off);
    __syncthreads();
    if ((tid & 63u) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0) s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

__global__ void __launch_bounds__(256)
    bld_score_terms_kernel(const BldInst* __restrict__ insts, const PrepItem* __restrict__ items,
                           const uint32_t* __restrict__ n_items_p) {
    __shared__ double red[4];
    __shared__ double sik[kPrepRows * 64];
    const uint32_t n_items = *n_items_p;
    const uint32_t tid = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PrepItem item = items[it];
        const BldInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const float inv_z = I->inv_z;
        const size_t ntile = (size_t)I->n_ti * I->n_tj;
        if (item.kind == 0) {
            // second and third terms for a chunk of columns j
            const uint32_t c0 = item.chunk * kPrepCols;
            const uint32_t nc = min(kPrepCols, y - c0);
            const double* __restrict__ pp = static_cast<const double*>(I->part_pjk);
            const size_t stride = (size_t)y * z;
            double beta_sum = 0.0;
            for (uint32_t k = 0; k < z; ++k) beta_sum += I->beta[k];
            double v = 0.0;
            if (tid < nc) {
                const uint32_t j = c0 + tid;
                double spj = 0.0;
                for (uint32_t k = 0; k < z; ++k) {
                    double s = 0.0;
                    for (uint32_t t = 0; t < I->n_ti; ++t) s += pp[t * stride + (size_t)j * z + k];
                    spj += s;
                    v += lgamma(I->beta[k] + s);
                }
                v -= lgamma(beta_sum + spj);
                v += lgamma(I->a + spj);
                v -= (I->a + spj) * log(I->b + 1.0);
            }
            const double s = block_sum_256(v, red);
            if (tid == 0) I->score[ntile + I->nA + item.chunk] = s;
        } else if (z > kFusedZ) {
            // the same terms for deep fibers: one thread per k over the
            // chunk's rows (no LDS staging of S_i+k)
            const uint32_t r0 = item.chunk * kPrepRows;
            const uint32_t nr = min(kPrepRows, x - r0);
            const double* __restrict__ pp = static_cast<const double*>(I->part_ipk);
            const size_t stride = (size_t)x * z;
            double v = 0.0;
            for (uint32_t k = tid; k < z; k += 256) {
                double cs = 0.0;
                for (uint32_t r = 0; r < nr; ++r) {
                    const size_t idx = (size_t)(r0 + r) * z + k;
                    double s = 0.0;
                    for (uint32_t t = 0; t < I->n_tj; ++t) s += pp[t * stride + idx];
                    v += lgamma(I->alpha[r0 + r] + s);
                    cs += s;
                }
                I->score[ntile + I->nA + I->n_cj + (size_t)item.chunk * z + k] = cs;
            }
            const double s = block_sum_256(v, red);
            if (tid == 0) I->score[ntile + item.chunk] = s;
        } else {
            // first term for a chunk of rows i: sum lgamma(alpha_i + S_i+k)
            // and the chunk's column sums of S_i+k (for S_++k)
            const uint32_t r0 = item.chunk * kPrepRows;
            const uint32_t nr = min(kPrepRows, x - r0);
            const double* __restrict__ pp = static_cast<const double*>(I->part_ipk);
            const size_t stride = (size_t)x * z;
            double v = 0.0;
            for (uint32_t e = tid; e < nr * z; e += 256) {
                const size_t idx = (size_t)r0 * z + e;
                const uint32_t r = fiber_of(e, inv_z);
                double s = 0.0;
                for (uint32_t t = 0; t < I->n_tj; ++t) s += pp[t * stride + idx];
                sik[e] = s;
                v += lgamma(I->alpha[r0 + r] + s);
            }
            const double s = block_sum_256(v, red);
            if (tid == 0) I->score[ntile + item.chunk] = s;
            if (tid < z) {
                double cs = 0.0;
                for (uint32_t r = 0; r < nr; ++r) cs += sik[r * z + tid];
                I->score[ntile + I->nA + I->n_cj + (size_t)item.chunk * z + tid] = cs;
            }
            __syncthreads();
        }
    }
}

// One block per instance: combines the partial terms in a fixed order.
__global__ void __launch_bounds__(256)
    bld_score_final_kernel(const BldInst* __restrict__ insts, uint32_t n_inst) {
    __shared__ double red[4];
    __shared__ double s_alpha_sum;
    const uint32_t tid = threadIdx.x;
    for (uint32_t in = blockIdx.x; in < n_inst; in += gridDim.x) {
        const BldInst* __restrict__ I = insts + in;
        if (I->result == nullptr) continue;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const size_t ntile = (size_t)I->n_ti * I->n_tj;
        const double* sc = I->score;
        // alpha sums
        double asum = 0.0, alg = 0.0;
        for (uint32_t i = tid; i < x; i += 256) {
            asum += I->alpha[i];
            alg += lgamma(I->alpha[i]);
        }
        const double alpha_sum = block_sum_256(asum, red);
        const double lg_alpha = block_sum_256(alg, red);
        // per-k: S_++k from row-chunk column sums
        double tk = 0.0;
        if (tid == 0) s_alpha_sum = alpha_sum;
        __syncthreads();
        for (uint32_t k = tid; k < z; k += 256) {
            double spk = 0.0;
            for (uint32_t c = 0; c < I->nA; ++c) spk += sc[ntile + I->nA + I->n_cj + (size_t)c * z + k];
            tk -= lgamma(s_alpha_sum + spk);
        }
        const double t_k = block_sum_256(tk, red);
        double tl = 0.0;
        for (size_t t = tid; t < ntile; t += 256) tl += sc[t];
        const double fourth = block_sum_256(tl, red);
        double tr = 0.0;
        for (uint32_t c = tid; c < I->nA; c += 256) tr += sc[ntile + c];
        const double rows = block_sum_256(tr, red);
        double tc = 0.0;
        for (uint32_t c = tid; c < I->n_cj; c += 256) tc += sc[ntile + I->nA + c];
        const double cols = block_sum_256(tc, red);
        if (tid == 0) {
            double beta_sum = 0.0, lg_beta = 0.0;
            for (uint32_t k = 0; k < z; ++k) {
                beta_sum += I->beta[k];
                lg_beta += lgamma(I->beta[k]);
            }
            const double first = (lgamma(alpha_sum) * z - lg_alpha * z) + (t_k + rows);
            const double second_third = (lgamma(beta_sum) * y - lg_beta * y) +
                                        (-lgamma(I->a) * y + I->a * log(I->b) * y) + cols;
            *I->result = first + second_third - fourth;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host-side launchers (explicit instantiation table)
// ---------------------------------------------------------------------------
template <typename T, typename ACC, int ZB, int RW, int MODE, int PSI>
static hipError_t launch_pass_t(dim3 grid, hipStream_t st, const BldInst* insts,
                                const PassItem* items, const uint32_t* n_items, uint32_t zmax) {
    constexpr int WAVES = pass_waves<T, ACC, ZB, MODE>();
    const size_t lds = PassLds<T, ACC, MODE>::bytes(zmax, WAVES);
    auto* fn = bld_pass_kernel<T, ACC, ZB, RW, MODE, PSI, WAVES>;
    hipLaunchKernelGGL(fn, grid, dim3(64 * WAVES), lds, st, insts, items, n_items);
    return hipGetLastError();
}

template <typename T, typename ACC, int RW, int MODE, int PSI>
static hipError_t launch_pass_zb(int zb, dim3 grid, hipStream_t st, const BldInst* insts,
                                 const PassItem* items, const uint32_t* n, uint32_t zmax) {
    switch (zb) {
        case 8: return launch_pass_t<T, ACC, 8, RW, MODE, PSI>(grid, st, insts, items, n, zmax);
        case 16: return launch_pass_t<T, ACC, 16, RW, MODE, PSI>(grid, st, insts, items, n, zmax);
        case 32: return launch_pass_t<T, ACC, 32, RW, MODE, PSI>(grid, st, insts, items, n, zmax);
        case 64: return launch_pass_t<T, ACC, 64, RW, MODE, PSI>(grid, st, insts, items, n, zmax);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, typename ACC, int MODE, int PSI>
static hipError_t launch_pass_rw(int rw, int zb, dim3 grid, hipStream_t st, const BldInst* insts,
                                 const PassItem* items, const uint32_t* n, uint32_t zmax) {
    if constexpr (!pass_steps(MODE)) {
        // sums-only and scoring passes read fibers element by element
        (void)rw;
        return launch_pass_zb<T, ACC, 1, MODE, PSI>(zb, grid, st, insts, items, n, zmax);
    } else {
        if (rw == 1) return launch_pass_zb<T, ACC, 1, MODE, PSI>(zb, grid, st, insts, items, n, zmax);
        if (rw == 2) return launch_pass_zb<T, ACC, 2, MODE, PSI>(zb, grid, st, insts, items, n, zmax);
        if constexpr (sizeof(T) == 4) {
            if (rw == 4) return launch_pass_zb<T, ACC, 4, MODE, PSI>(zb, grid, st, insts, items, n, zmax);
        }
        return hipErrorInvalidValue;
    }
}

// dtype: 0 = float, 1 = double. rw: LDS read width of the instances in the
// launch (update mode). zmax: the largest z among them (sizes the LDS).
hipError_t launch_pass(int dtype, int psi, int mode, int zb, int rw, dim3 grid, hipStream_t st,
                       const BldInst* insts, const PassItem* items, const uint32_t* n_items,
                       uint32_t zmax) {
    if (dtype == 0) {
        switch (mode) {
            case PASS_UPDATE:
                return launch_pass_rw<float, float, PASS_UPDATE, PSI_F32>(rw, zb, grid, st, insts, items, n_items, zmax);
            case PASS_SUMS:
                return launch_pass_rw<float, float, PASS_SUMS, PSI_F32>(rw, zb, grid, st, insts, items, n_items, zmax);
            case PASS_SCORE:
                return launch_pass_rw<float, double, PASS_SCORE, PSI_F32>(rw, zb, grid, st, insts, items, n_items, zmax);
            case PASS_ADD:
                return launch_pass_rw<float, float, PASS_ADD, PSI_F32>(rw, zb, grid, st, insts, items, n_items, zmax);
        }
    } else {
        switch (mode) {
            case PASS_UPDATE:
                if (psi == PSI_F64_APPR)
                    return launch_pass_rw<double, double, PASS_UPDATE, PSI_F64_APPR>(rw, zb, grid, st, insts, items, n_items, zmax);
                return launch_pass_rw<double, double, PASS_UPDATE, PSI_F64_EXACT>(rw, zb, grid, st, insts, items, n_items, zmax);
            case PASS_SUMS:
                return launch_pass_rw<double, double, PASS_SUMS, PSI_F64_EXACT>(rw, zb, grid, st, insts, items, n_items, zmax);
            case PASS_SCORE:
                return launch_pass_rw<double, double, PASS_SCORE, PSI_F64_EXACT>(rw, zb, grid, st, insts, items, n_items, zmax);
            case PASS_ADD:
                // the additive solver evaluates every psi with psi_appr
                return launch_pass_rw<double, double, PASS_ADD, PSI_F64_APPR>(rw, zb, grid, st, insts, items, n_items, zmax);
        }
    }
    return hipErrorInvalidValue;
}

// LDS read width of an instance of depth z (update pass) and the XOR mask of
// its chunk visit order. A read of RW elements serves
// kLdsBankBytes / (RW*sizeof(T)) lanes per LDS cycle; fibers of lanes l and l' start on the same bank group
// when l*z and l'*z agree modulo the group count, so lanes that repeat a
// group (period lanes_per_cycle / P) get distinct XOR masks below P, where P
// is the largest power of two dividing the chunk count (so m ^ mask stays a
// permutation of the chunks).
void read_layout(int dtype, uint32_t z, uint32_t* rw, uint32_t* hs, uint32_t* hm) {
    uint32_t r = 1;
    if (dtype == 0)
        r = (z % 4 == 0) ? 4 : (z % 2 == 0 ? 2 : 1);
    else
        r = (z % 2 == 0) ? 2 : 1;
    const uint32_t es = dtype == 0 ? 4 : 8;
    const uint32_t lanes_per_cycle = kLdsBankBytes / (r * es);
    const uint32_t C = z / r;
    uint32_t P = 1;
    while (P < lanes_per_cycle && (C % (P * 2)) == 0) P *= 2;
    if (r == 1) P = 1;
    const uint32_t period = lanes_per_cycle / P;
    uint32_t sh = 0;
    while ((1u << sh) < period) ++sh;
    *rw = r;
    *hs = sh;
    *hm = P - 1;
}

hipError_t launch_prep(int dtype, int psi, bool raw, dim3 grid, hipStream_t st,
                       const BldInst* insts, const PrepItem* items, const uint32_t* n_items,
                       uint32_t n_inst) {
    const dim3 gA(n_inst < 65535u ? (n_inst ? n_inst : 1u) : 65535u);
    if (dtype == 0) {
        if (raw) {
            hipLaunchKernelGGL((bld_prep_kernel<float, PSI_F32, true>), grid, dim3(256), 0, st, insts, items, n_items);
        } else {
            hipLaunchKernelGGL((bld_prep_kernel<float, PSI_F32, false>), grid, dim3(256), 0, st, insts, items, n_items);
            hipLaunchKernelGGL((bld_psiA_kernel<float, PSI_F32>), gA, dim3(256), 0, st, insts, n_inst);
        }
    } else if (raw) {
        hipLaunchKernelGGL((bld_prep_kernel<double, PSI_F64_EXACT, true>), grid, dim3(256), 0, st, insts, items, n_items);
    } else if (psi == PSI_F64_APPR) {
        hipLaunchKernelGGL((bld_prep_kernel<double, PSI_F64_APPR, false>), grid, dim3(256), 0, st, insts, items, n_items);
        hipLaunchKernelGGL((bld_psiA_kernel<double, PSI_F64_APPR>), gA, dim3(256), 0, st, insts, n_inst);
    } else {
        hipLaunchKernelGGL((bld_prep_kernel<double, PSI_F64_EXACT, false>), grid, dim3(256), 0, st, insts, items, n_items);
        hipLaunchKernelGGL((bld_psiA_kernel<double, PSI_F64_EXACT>), gA, dim3(256), 0, st, insts, n_inst);
    }
    return hipGetLastError();
}

// bld::bld_fact (src/bld/bld_fact.hpp:36-80) from the axis sums of S, which
// a raw prep has left in the descriptor's psib (S_+jk, y x z) and psia
// (S_i+k, x x z) buffers. Blocks 0..z-1: column k of W, normalised by its
// sum over i; blocks z..: 256 columns j of H (normalised over k) and L_j.
// Sums in double, per-thread strided partials added in a fixed tree.
template <typename T>
__global__ void __launch_bounds__(256)
    bld_fact_kernel(const BldInst* __restrict__ I, double eps, double* __restrict__ W,
                    double* __restrict__ H, double* __restrict__ L) {
    __shared__ double red[4];
    const uint32_t x = I->x, y = I->y, z = I->z, tid = threadIdx.x;
    const T* __restrict__ S_pjk = static_cast<const T*>(I->psib);
    const T* __restrict__ S_ipk = static_cast<const T*>(I->psia);
    if (blockIdx.x < z) {
        const uint32_t k = blockIdx.x;
        double part = 0.0;
        for (uint32_t i = tid; i < x; i += 256) part += I->alpha[i] + (double)S_ipk[(size_t)i * z + k] - 1.0;
        const double cs = block_sum_256(part, red);
        __shared__ double colsum;
        if (tid == 0) colsum = cs + eps;
        __syncthreads();
        for (uint32_t i = tid; i < x; i += 256)
            W[(size_t)i * z + k] = (I->alpha[i] + (double)S_ipk[(size_t)i * z + k] - 1.0) / colsum;
    } else {
        const uint32_t j = (blockIdx.x - z) * 256 + tid;
        if (j >= y) return;
        double hs = 0.0, sp = 0.0;
        for (uint32_t k = 0; k < z; ++k) {
            const double s = (double)S_pjk[(size_t)j * z + k];
            hs += I->beta[k] + s - 1.0;
            sp += s;
        }
        L[j] = (I->a + sp - 1.0) / (I->b + 1.0 + eps);
        hs += eps;
        for (uint32_t k = 0; k < z; ++k)
            H[(size_t)k * y + j] = (I->beta[k] + (double)S_pjk[(size_t)j * z + k] - 1.0) / hs;
    }
}

hipError_t launch_fact(int dtype, const BldInst* inst, uint32_t y, uint32_t z, double eps, double* W,
                       double* H, double* L, hipStream_t st) {
    const dim3 grid(z + (y + 255) / 256);
    if (dtype == 0)
        hipLaunchKernelGGL(bld_fact_kernel<float>, grid, dim3(256), 0, st, inst, eps, W, H, L);
    else
        hipLaunchKernelGGL(bld_fact_kernel<double>, grid, dim3(256), 0, st, inst, eps, W, H, L);
    return hipGetLastError();
}

// Element-wise evaluation of the device digamma forms (test hook).
__global__ void __launch_bounds__(256)
    bld_psi_eval_kernel(int kind, const double* __restrict__ xin, double* __restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double v = xin[t];
    double r;
    if (kind == 0) {
        const float f = (float)v;
        r = (double)psi_f32(f);
    } else if (kind == 3) {
        r = (double)psi_sp1<PSI_F32, float>((float)v - 1.0f);
    } else if (kind == 1) {
        r = psi_f64(v);
    } else if (kind == 4) {
        r = psi_sp1_f64(v - 1.0);
    } else {
        r = psi_appr_ref<double>(v);
    }
    out[t] = r;
}

hipError_t launch_psi_eval(int kind, const double* xin, double* out, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(bld_psi_eval_kernel, dim3(blocks), dim3(256), 0, st, kind, xin, out, n);
    return hipGetLastError();
}

// Sets the dynamic-LDS limit of every pass instantiation once per process.
template <typename T, typename ACC, int ZB, int RW, int MODE, int PSI>
static hipError_t attr_one() {
    constexpr int WAVES = pass_waves<T, ACC, ZB, MODE>();
    return hipFuncSetAttribute(
        reinterpret_cast<const void*>(bld_pass_kernel<T, ACC, ZB, RW, MODE, PSI, WAVES>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)PassLds<T, ACC, MODE>::bytes(ZB, WAVES));
}

template <typename T, typename ACC, int RW, int MODE, int PSI>
static hipError_t attr_zb() {
    hipError_t e = attr_one<T, ACC, 8, RW, MODE, PSI>();
    if (e == hipSuccess) e = attr_one<T, ACC, 16, RW, MODE, PSI>();
    if (e == hipSuccess) e = attr_one<T, ACC, 32, RW, MODE, PSI>();
    if (e == hipSuccess) e = attr_one<T, ACC, 64, RW, MODE, PSI>();
    return e;
}

hipError_t init_kernel_attributes() {
    hipError_t e = attr_zb<float, float, 1, PASS_UPDATE, PSI_F32>();
    if (e == hipSuccess) e = attr_zb<float, float, 2, PASS_UPDATE, PSI_F32>();
    if (e == hipSuccess) e = attr_zb<float, float, 4, PASS_UPDATE, PSI_F32>();
    if (e == hipSuccess) e = attr_zb<float, float, 1, PASS_SUMS, PSI_F32>();
    if (e == hipSuccess) e = attr_zb<float, double, 1, PASS_SCORE, PSI_F32>();
    if (e == hipSuccess) e = attr_zb<double, double, 1, PASS_UPDATE, PSI_F64_EXACT>();
    if (e == hipSuccess) e = attr_zb<double, double, 2, PASS_UPDATE, PSI_F64_EXACT>();
    if (e == hipSuccess) e = attr_zb<double, double, 1, PASS_UPDATE, PSI_F64_APPR>();
    if (e == hipSuccess) e = attr_zb<double, double, 2, PASS_UPDATE, PSI_F64_APPR>();
    if (e == hipSuccess) e = attr_zb<double, double, 1, PASS_SUMS, PSI_F64_EXACT>();
    if (e == hipSuccess) e = attr_zb<double, double, 1, PASS_SCORE, PSI_F64_EXACT>();
    return e;
}

hipError_t launch_score(dim3 grid, hipStream_t st, const BldInst* insts, const PrepItem* items,
                        const uint32_t* n_items, uint32_t n_inst) {
    hipLaunchKernelGGL(bld_score_terms_kernel, grid, dim3(256), 0, st, insts, items, n_items);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bld_score_final_kernel, dim3(n_inst < 1024 ? n_inst : 1024), dim3(256), 0,
                       st, insts, n_inst);
    return hipGetLastError();
}

// init_S on the device (bld_mult_funcs.hpp:36-74): one thread per fiber,
// the same Philox stream as the host init_S. The Exp(1) draws are made twice
// (once for their sum, once for the output) instead of being stored, so no
// per-thread array is needed; the values agree with the host's to the last
// bit of the double log.
template <typename T>
__global__ void __launch_bounds__(256)
    bld_init_s_kernel(T* __restrict__ S, const T* __restrict__ X, size_t n_fibers, uint32_t z,
                      uint64_t seed) {
    const size_t f = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= n_fibers) return;
    const double xv = (double)X[f];
    T* out = S + f * z;
    if (xv == 0.0) {
        for (uint32_t k = 0; k < z; ++k) out[k] = T(0);
        return;
    }
    double tot = 0.0;
    {
        Philox g(seed, ST_INIT, f);
        for (uint32_t k = 0; k < z; ++k) tot += -log(g.uniform());
    }
    Philox g(seed, ST_INIT, f);
    for (uint32_t k = 0; k < z; ++k) out[k] = (T)(xv * (-log(g.uniform()) / tot));
}

// Allocation-model draws on the device from the host sampler's Philox
// streams (alloc_model_funcs.hpp:528-552): mode 0 X_ij ~ Poisson(L_j
// sum_k W_ik H_kj) (stream ST_X, index i*y + j); mode 1 S_ijk ~ Poisson(W_ik
// H_kj L_j) (stream ST_S, index (i*y + j)*z + k). One thread per draw.
__global__ void __launch_bounds__(256)
    bld_sample_kernel(double* __restrict__ out, const double* __restrict__ W,
                      const double* __restrict__ H, const double* __restrict__ L, size_t y,
                      uint32_t z, uint64_t seed, int mode, size_t n) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    if (mode == 0) {
        const size_t i = idx / y, j = idx - i * y;
        double lam = 0.0;
        for (uint32_t k = 0; k < z; ++k) lam += W[i * z + k] * H[(size_t)k * y + j];
        Philox g(seed, ST_X, idx);
        out[idx] = poisson_draw(g, lam * L[j]);
    } else {
        const size_t f = idx / z, k = idx - f * z;
        const size_t i = f / y, j = f - i * y;
        Philox g(seed, ST_S, idx);
        out[idx] = poisson_draw(g, W[i * z + k] * H[k * y + j] * L[j]);
    }
}

hipError_t launch_sample(double* out, const double* W, const double* H, const double* L, size_t y,
                         uint32_t z, uint64_t seed, int mode, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(bld_sample_kernel, grid, dim3(256), 0, st, out, W, H, L, y, z, seed, mode, n);
    return hipGetLastError();
}

hipError_t launch_init_s(int dtype, void* S, const void* X, size_t n_fibers, uint32_t z,
                         uint64_t seed, hipStream_t st) {
    if (n_fibers == 0) return hipSuccess;
    const dim3 grid((unsigned)((n_fibers + 255) / 256));
    if (dtype == 0)
        hipLaunchKernelGGL(bld_init_s_kernel<float>, grid, dim3(256), 0, st, (float*)S,
                           (const float*)X, n_fibers, z, seed);
    else
        hipLaunchKernelGGL(bld_init_s_kernel<double>, grid, dim3(256), 0, st, (double*)S,
                           (const double*)X, n_fibers, z, seed);
    return hipGetLastError();
}

}  // namespace bnmf

