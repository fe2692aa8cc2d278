// bld_pass_f64.hip — the FP64 step passes (bld_mult's multiplicative update,
// exact or approximate digamma, and bld_add's additive step) for fibers of
// depth z <= 64.
//
// The update needs, per element, the gradient g_ijk and, per fiber, three
// sums over k before any element can be rewritten; in double a lane that
// owned a whole fiber would hold the tile's S_+jk accumulators for z
// elements (2z registers) beside the next row in flight and the digamma
// temporaries, which at z = 32 leaves one wave per SIMD. Here LPF lanes share
// a fiber (2 for z <= 32, 4 for z <= 64): lane LPF*f + p holds part p of
// fiber f (a contiguous run of ceil(c/LPF) of its c = z / RW chunks of RW
// elements), the fiber sums are joined across the LPF lanes with xor
// shuffles (every lane then holds the same sum), and a tile row is 64 / LPF
// fibers. The gradients of the current row live in a per-wave LDS row, not
// in registers. Per row a wave moves (64 / LPF) * z doubles, in flight as
// 16-byte streaming loads while the previous row is computed (fixed
// instruction count: clamped chunk indices, unconditional range-checked
// buffer stores, as in the FP32 exact-z pass).
//
// Digamma of the per-element term psi(S + 1) comes from psi_sp1_f64 (a
// rational fit on [1, 2) reached by recurrence, one division) for the exact
// form, and from util::psi_appr's own formula for the approximate one.
// Everything else follows the reference's operations in double
// (src/bld/bld_mult/bld_mult_funcs.hpp: X_reciprocal :95, grad_plus
// :177-192, nom :254, denom :284-285, update_S :319-323;
// src/bld/bld_add.hpp:137-181); fiber sums run in the lane's visit order and
// join the two halves, so they differ from the reference's sequential sums
// only by rounding.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bld_types.hpp"
#include "digamma.hpp"
#include "pass_common.hpp"

namespace bnmf {

namespace {
constexpr int kBufWord3F64 = 0x00020000;   // raw buffer resource, gfx9 family
constexpr uint32_t kBufDropF64 = 0x80000000u;  // offset past any resource
constexpr int kStreamAuxF64 = 2;           // nt: streaming store of S
}  // namespace

// lanes per fiber of a z bucket, and the tile row width that follows
__host__ __device__ constexpr int f64_lpf(int zb) { return zb <= 32 ? 2 : 4; }

// LDS of a block: [psib: TW fibers x ZB][psiA: ZB][WAVES x grad_minus row:
// ZB][WAVES x 32 X_ij][WAVES x gradient row: TW x ZB][WAVES x row buffer:
// CPL x 64 16-byte chunks], TW = 64 / LPF. A row segment starts 0 or 1
// double past a 16-byte boundary. The gradient row holds g_ijk of the wave's
// current row (fiber f at f * z).
template <int ZB>
struct F64Lds {
    static constexpr uint32_t TW = 64u / f64_lpf(ZB);
    static constexpr uint32_t CPL = (TW * ZB * 8u + 8u + 1023u) / 1024u;  // chunks per lane
    static constexpr size_t psiA_off = round16((size_t)TW * ZB * 8);
    static constexpr size_t gm_off = psiA_off + round16((size_t)ZB * 8);
    static constexpr size_t gm_stride = round16((size_t)ZB * 8);
    static constexpr size_t xs_off(int waves) { return gm_off + waves * gm_stride; }
    static constexpr size_t gb_off(int waves) { return xs_off(waves) + waves * 32 * 8; }
    static constexpr size_t gb_stride = (size_t)TW * ZB * 8;
    static constexpr size_t raw_off(int waves) { return gb_off(waves) + waves * gb_stride; }
    static constexpr size_t raw_stride = (size_t)CPL * 64 * 16;
    static constexpr size_t bytes(int waves) { return raw_off(waves) + waves * raw_stride; }
};

// four waves per block, two where four would not let two blocks share a CU
template <int ZB>
constexpr int f64_waves() {
    return F64Lds<ZB>::bytes(4) <= kLdsBytes / 2 ? 4 : 2;
}

// waves per SIMD the register allocation must leave room for (two: about
// 256 registers; the kernels need about 170 to 260 without a budget)
#define BNMF_F64_WPE(ZB) 2

template <int ZB, int RW, int MODE, int PSI, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(BNMF_F64_WPE(ZB))))
    bld_pass_f64_kernel(const BldInst* __restrict__ insts, const PassItem* __restrict__ items,
                        const uint32_t* __restrict__ n_items_p) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    using Lds = F64Lds<ZB>;
    constexpr uint32_t NT = 64 * WAVES;
    constexpr int CPL = (int)Lds::CPL;
    constexpr int LPF = f64_lpf(ZB);
    constexpr uint32_t TW = Lds::TW;
    constexpr int NH = (ZB / RW + LPF - 1) / LPF;  // chunks a lane holds at most
    typedef double V16 __attribute__((ext_vector_type(2)));
    typedef uint32_t U16 __attribute__((ext_vector_type(4)));
    typedef uint32_t U8 __attribute__((ext_vector_type(2)));
    typedef typename VecR<double, RW>::type VR;

    const uint32_t n_items = *n_items_p;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));
    const uint32_t f = lane / LPF, p = lane % LPF;  // fiber of the tile row, part of it

    double* const psib_l = reinterpret_cast<double*>(lds_raw);
    double* const psiA_l = reinterpret_cast<double*>(lds_raw + Lds::psiA_off);
    double* const gm_l = reinterpret_cast<double*>(lds_raw + Lds::gm_off + w * Lds::gm_stride);
    double* const xs_l = reinterpret_cast<double*>(lds_raw + Lds::xs_off(WAVES) + w * 32 * 8);
    double* const gbuf = reinterpret_cast<double*>(lds_raw + Lds::gb_off(WAVES) + w * Lds::gb_stride);
    unsigned char* const raw_base = lds_raw + Lds::raw_off(WAVES);
    double* const raw = reinterpret_cast<double*>(raw_base + w * Lds::raw_stride);

    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const BldInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y, z = I->z;
        const uint32_t i0 = item.ti * I->ti_rows;
        const uint32_t i1 = min(x, i0 + I->ti_rows);
        const uint32_t j0 = item.tj * TW;
        const uint32_t nj = min(TW, y - j0);
        const uint32_t n_el = nj * z;

        const BNMF_GLOBAL double* Ssrc = gconst<double>(I->S_src);
        BNMF_GLOBAL double* Sdst = gmut<double>(I->S);
        const BNMF_GLOBAL double* __restrict__ Xg = gconst<double>(I->X);
        const BNMF_GLOBAL double* __restrict__ psia_g = gconst<double>(I->psia);
        const double eps = I->eps;
        const uint32_t lane_x = min(f, nj - 1u);  // clamped fiber of this lane
        const uint32_t lane_z = min(lane, z - 1u);
        // this lane's chunks of its fiber and their visit order: chunk
        // m0 + (u ^ xm) for u < mcount; when the chunk count divides evenly
        // the parts are equal and divisible by the xor range, so lanes of
        // neighbouring fibers read different banks
        const uint32_t nchunk = z / RW;
        const uint32_t cp = (nchunk + LPF - 1u) / LPF;
        const uint32_t m0 = min(p * cp, nchunk);
        const uint32_t mcount = min(cp, nchunk - m0);
        uint32_t xm = 0;
        if (nchunk % LPF == 0u) {
            const uint32_t P = min(cp & (0u - cp), 4u);
            xm = f & (P - 1u);
        }

        // prefetch registers of one row: S chunks from the aligned start
        // (indices clamped to the row's last chunk), X_ij of the lane's
        // fiber and psi(alpha_eph)(i, lane)
        V16 pf[CPL];
        double pf_x = 0.0, pf_pa = 0.0;
        uint32_t i = i0 + w;
        uint32_t dl = 0, nch = 0;
        // real = false: there is no next row; every lane loads the row's
        // first chunk so the instruction count stays fixed
        auto issue_row = [&](uint32_t row, bool real, uint32_t& d_out, uint32_t& n_out) {
            const size_t s0 = ((size_t)row * y + j0) * z;
            const uint32_t d = (uint32_t)(s0 & 1u);
            const uint32_t n = (d + n_el + 1u) / 2u;
            const BNMF_GLOBAL V16* src = (const BNMF_GLOBAL V16*)(Ssrc + (s0 - d));
            const uint32_t last = real ? n - 1u : 0u;
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                pf[c] = __builtin_nontemporal_load(&src[min((uint32_t)c * 64u + lane, last)]);
            pf_x = Xg[(size_t)row * y + j0 + lane_x];
            pf_pa = psia_g[(size_t)row * z + lane_z];
            d_out = d;
            n_out = n;
        };
        if (i < i1) issue_row(i, true, dl, nch);
        {
            const BNMF_GLOBAL double* __restrict__ psib_g = gconst<double>(I->psib) + (size_t)j0 * z;
            for (uint32_t e = tid; e < n_el; e += NT) psib_l[e] = psib_g[e];
            if (tid < z) psiA_l[tid] = gconst<double>(I->psiA)[tid];
            __syncthreads();
        }
        // additive step size 0.1 / (iteration + 1)^0.55 (bld_add.hpp:101-103);
        // *iter counts the iterations started, this one included
        double eta = 0.0;
        if constexpr (MODE == PASS_ADD) eta = 0.1 / pow((double)*I->iter, 0.55);

        double acc[NH * RW];
#pragma unroll
        for (int k = 0; k < NH * RW; ++k) acc[k] = 0.0;

        const uint32_t kq = lane & (I->zq - 1u);
        const uint32_t part = lane >> I->zq_log2;
        const uint32_t L = 64u >> I->zq_log2;

        auto row_step = [&](uint32_t i) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) reinterpret_cast<V16*>(raw)[(uint32_t)c * 64u + lane] = pf[c];
            // grad_minus row psi(A) - psi(alpha_eph) for the update; the
            // additive step keeps psi(alpha_eph) (its operation order)
            if (lane < z) gm_l[lane] = MODE == PASS_UPDATE ? psiA_l[lane] - pf_pa : pf_pa;
            if (p == 0u) xs_l[f] = pf_x;
            const uint32_t cur_dl = dl, cur_nch = nch;
            const size_t cur_s0 = ((size_t)i * y + j0) * z;
            const uint32_t inext = i + WAVES;
            const bool more = inext < i1;
            issue_row(more ? inext : i, more, dl, nch);
            wave_sync();
            const double xv = f < nj ? xs_l[f] : 0.0;
            if (f < nj) {
                double* const fb = raw + cur_dl + f * z;
                double* const gb = gbuf + f * z;
                const double* const pbb = psib_l + f * z;
                if constexpr (MODE == PASS_UPDATE) {
                    const double xr = 1.0 / (xv + eps);
                    // Phase A: g = psi(beta_eph) - psi(S+1) into the gradient
                    // row, S_ij+, nom and denom (sums in the lane's visit
                    // order; an element whose S+1 is outside [1, 10) adds its
                    // denom term in the rare-path loop below instead)
                    double sumS = 0.0, nom = 0.0, den = 0.0;
                    bool slow = false;
#pragma unroll
                    for (int u = 0; u < NH; ++u) {
                        if ((uint32_t)u < mcount) {
                            const uint32_t mm = (m0 + ((uint32_t)u ^ xm)) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                            const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
                            VR gw;
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const double sk = vget<double, RW>(sv, v);
                                double ps;
                                bool fast = true;
                                if constexpr (PSI == PSI_F64_EXACT) {
                                    // 1 <= S+1 < 10 inline; other arguments
                                    // are redone below (a rare path)
                                    const double xa = sk + 1.0;
                                    fast = xa >= 1.0 && xa < 10.0;
                                    slow = slow || !fast;
                                    ps = psi_1to10(fast ? xa : 1.0);
                                } else {
                                    ps = psi_sp1<PSI, double>(sk);
                                }
                                const double g = vget<double, RW>(pb, v) - ps;
                                vset<double, RW>(gw, v, g);
                                sumS += sk;
                                nom += sk * xr * vget<double, RW>(gv, v);
                                if (fast) den += sk * xr * g;
                            }
                            *reinterpret_cast<VR*>(gb + mm) = gw;
                        }
                    }
                    if constexpr (PSI == PSI_F64_EXACT) {
                        if (__builtin_amdgcn_ballot_w64(slow) != 0) {
                            for (uint32_t u = 0; u < mcount; ++u) {
                                const uint32_t mm = (m0 + (u ^ xm)) * RW;
                                for (int v = 0; v < RW; ++v) {
                                    const double xa = fb[mm + v] + 1.0;
                                    if (!(xa >= 1.0 && xa < 10.0)) {
                                        const double g = pbb[mm + v] - psi_f64_general(xa);
                                        gb[mm + v] = g;
                                        den += fb[mm + v] * xr * g;
                                    }
                                }
                            }
                        }
                    }
                    // the other parts of the fiber (each butterfly step adds a
                    // pair in both orders: every lane gets the same sum)
#pragma unroll
                    for (int o = 1; o < LPF; o <<= 1) {
                        sumS += __shfl_xor(sumS, o);
                        nom += __shfl_xor(nom, o);
                        den += __shfl_xor(den, o);
                    }
                    const double x_over_s = xv / (sumS + eps);
                    // Phase B: the multiplicative update and the S_+jk partials
#pragma unroll
                    for (int u = 0; u < NH; ++u) {
                        if ((uint32_t)u < mcount) {
                            const uint32_t mm = (m0 + ((uint32_t)u ^ xm)) * RW;
                            VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            const VR gw = *reinterpret_cast<const VR*>(gb + mm);
                            const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const double sk = vget<double, RW>(sv, v);
                                // gm + den + eps is in div_moderate's range
                                // unless gm + den cancels eps below 2^-1000
                                const double sn = as_stored(sk * div_moderate((vget<double, RW>(gw, v) + nom) * x_over_s,
                                                                               vget<double, RW>(gv, v) + den + eps));
                                vset<double, RW>(sv, v, sn);
                                acc[u * RW + v] += sn;
                            }
                            *reinterpret_cast<VR*>(fb + mm) = sv;
                        }
                    }
                } else {
                    // bld_add (src/bld/bld_add.hpp:137-181): grad_S =
                    // psi(beta_eph) - psi(S+1) - psi(A) + psi(alpha_eph),
                    // S_mult = sum_k S grad_S
                    double smult = 0.0;
#pragma unroll
                    for (int u = 0; u < NH; ++u) {
                        if ((uint32_t)u < mcount) {
                            const uint32_t mm = (m0 + ((uint32_t)u ^ xm)) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                            const VR pA = *reinterpret_cast<const VR*>(psiA_l + mm);
                            const VR pa = *reinterpret_cast<const VR*>(gm_l + mm);
                            VR gw;
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const double sk = vget<double, RW>(sv, v);
                                const double ps = psi_sp1<PSI, double>(sk);
                                const double g = ((vget<double, RW>(pb, v) - ps) - vget<double, RW>(pA, v)) +
                                                 vget<double, RW>(pa, v);
                                vset<double, RW>(gw, v, g);
                                smult += sk * g;
                            }
                            *reinterpret_cast<VR*>(gb + mm) = gw;
                        }
                    }
#pragma unroll
                    for (int o = 1; o < LPF; o <<= 1) smult += __shfl_xor(smult, o);
                    // Z = log S + eps + eta S (X grad_S - S_mult) (:150,158-166),
                    // exp(Z) (into the gradient row) and its L1 norm with
                    // util::normalize's 1e-50 guard (:167-173)
                    double norm = 0.0;
#pragma unroll
                    for (int u = 0; u < NH; ++u) {
                        if ((uint32_t)u < mcount) {
                            const uint32_t mm = (m0 + ((uint32_t)u ^ xm)) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            VR gw = *reinterpret_cast<const VR*>(gb + mm);
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const double sk = vget<double, RW>(sv, v);
                                double zk = log(sk) + eps;
                                zk += eta * sk * (xv * vget<double, RW>(gw, v) - smult);
                                const double e = exp(zk);
                                vset<double, RW>(gw, v, e);
                                norm += e;
                            }
                            *reinterpret_cast<VR*>(gb + mm) = gw;
                        }
                    }
#pragma unroll
                    for (int o = 1; o < LPF; o <<= 1) norm += __shfl_xor(norm, o);
                    norm += 1e-50;
#pragma unroll
                    for (int u = 0; u < NH; ++u) {
                        if ((uint32_t)u < mcount) {
                            const uint32_t mm = (m0 + ((uint32_t)u ^ xm)) * RW;
                            const VR gw = *reinterpret_cast<const VR*>(gb + mm);
                            VR sv;
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const double sn = as_stored(xv * (vget<double, RW>(gw, v) / norm));
                                vset<double, RW>(sv, v, sn);
                                acc[u * RW + v] += sn;
                            }
                            *reinterpret_cast<VR*>(fb + mm) = sv;
                        }
                    }
                }
            }
            wave_sync();

            // S_i+k partial over the tile's columns: eight independent
            // partial sums combined in a fixed tree (deterministic)
            {
                double r[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) r[v] = 0.0;
                if (kq < z) {
                    const double* const rb = raw + cur_dl + kq;
                    for (uint32_t f0 = part; f0 < nj; f0 += 8u * L) {
#pragma unroll
                        for (int v = 0; v < 8; ++v) {
                            const uint32_t fc = f0 + (uint32_t)v * L;
                            if (fc < nj) r[v] += rb[fc * z];
                        }
                    }
                }
                double rs = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
                for (uint32_t off = I->zq; off < 64u; off <<= 1) rs += __shfl_xor(rs, (int)off);
                if (part == 0 && kq < z) gmut<double>(I->part_ipk)[((size_t)item.tj * x + i) * z + kq] = rs;
            }

            // the row back to S through a buffer resource spanning its
            // chunks; a lane with nothing to write gives an offset past the
            // resource (dropped by the hardware). Whole 16-byte chunks inside
            // the segment go as one store; the (at most two) straddling the
            // segment ends go element by element, lanes 0-1 the head chunk
            // and lanes 2-3 the tail
            {
                const __amdgpu_buffer_rsrc_t dst = __builtin_amdgcn_make_buffer_rsrc(
                    (void*)(Sdst + (cur_s0 - cur_dl)), (short)0, (int)(cur_nch * 16u), kBufWord3F64);
                const uint32_t lo = cur_dl, hi = cur_dl + n_el;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const uint32_t q = (uint32_t)c * 64u + lane;
                    const uint32_t e0 = q * 2u;
                    const bool inside = q < cur_nch && e0 >= lo && e0 + 2u <= hi;
                    const V16 v = reinterpret_cast<const V16*>(raw)[q];
                    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U16, v), dst,
                                                           (int)(inside ? q * 16u : kBufDropF64), 0, kStreamAuxF64);
                }
                const uint32_t qe = lane < 2u ? 0u : cur_nch - 1u;
                const uint32_t e = qe * 2u + (lane & 1u);
                const bool mine = lane < 4u && e >= lo && e < hi && ((qe * 2u < lo) || (qe * 2u + 2u > hi));
                const double v = raw[min(e, (uint32_t)(Lds::raw_stride / 8u) - 1u)];
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U8, v), dst,
                                                      (int)(mine ? e * 8u : kBufDropF64), 0, 0);
            }
            wave_sync();
        };
        if (i < i1) {
            row_step(i);
            for (i += WAVES; i < i1; i += WAVES) row_step(i);
        }

        // S_+jk partial over the tile's rows: the waves' column sums meet in
        // LDS (natural fiber layout) and are added in wave order
        __syncthreads();
        double* const comb_self = reinterpret_cast<double*>(raw_base + w * Lds::raw_stride);
        if (f < nj) {
            // RW-wide writes in the chunk visit order (bank-spread, as the
            // reads of the row are)
#pragma unroll
            for (int u = 0; u < NH; ++u) {
                if ((uint32_t)u < mcount) {
                    const uint32_t mm = (m0 + ((uint32_t)u ^ xm)) * RW;
                    VR cv;
#pragma unroll
                    for (int v = 0; v < RW; ++v) vset<double, RW>(cv, v, acc[u * RW + v]);
                    *reinterpret_cast<VR*>(comb_self + f * z + mm) = cv;
                }
            }
        }
        __syncthreads();
        {
            BNMF_GLOBAL double* __restrict__ pp = gmut<double>(I->part_pjk) + ((size_t)item.ti * y + j0) * z;
            for (uint32_t e = tid; e < n_el; e += NT) {
                double sacc = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < WAVES; ++w2)
                    sacc += reinterpret_cast<const double*>(raw_base + w2 * Lds::raw_stride)[e];
                pp[e] = sacc;
            }
        }
        __syncthreads();
    }
}

template <int ZB, int RW, int MODE, int PSI>
static hipError_t launch_f64_t(dim3 grid, hipStream_t st, const BldInst* insts, const PassItem* items,
                               const uint32_t* n_items) {
    constexpr int WAVES = f64_waves<ZB>();
    hipLaunchKernelGGL((bld_pass_f64_kernel<ZB, RW, MODE, PSI, WAVES>), grid, dim3(64 * WAVES),
                       F64Lds<ZB>::bytes(WAVES), st, insts, items, n_items);
    return hipGetLastError();
}

template <int ZB, int RW, int MODE, int PSI>
static hipError_t attr_f64_t() {
    constexpr int WAVES = f64_waves<ZB>();
    return hipFuncSetAttribute(reinterpret_cast<const void*>(bld_pass_f64_kernel<ZB, RW, MODE, PSI, WAVES>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)F64Lds<ZB>::bytes(WAVES));
}

template <int RW, int MODE, int PSI>
static hipError_t launch_f64_zb(int zb, dim3 grid, hipStream_t st, const BldInst* insts, const PassItem* items,
                                const uint32_t* n) {
    switch (zb) {
        case 8: return launch_f64_t<8, RW, MODE, PSI>(grid, st, insts, items, n);
        case 16: return launch_f64_t<16, RW, MODE, PSI>(grid, st, insts, items, n);
        case 32: return launch_f64_t<32, RW, MODE, PSI>(grid, st, insts, items, n);
        case 64: return launch_f64_t<64, RW, MODE, PSI>(grid, st, insts, items, n);
        default: return hipErrorInvalidValue;
    }
}

template <int MODE, int PSI>
static hipError_t launch_f64_rw(int rw, int zb, dim3 grid, hipStream_t st, const BldInst* insts,
                                const PassItem* items, const uint32_t* n) {
    if (rw == 2) return launch_f64_zb<2, MODE, PSI>(zb, grid, st, insts, items, n);
    if (rw == 1) return launch_f64_zb<1, MODE, PSI>(zb, grid, st, insts, items, n);
    return hipErrorInvalidValue;
}

// FP64 step pass of instances in z bucket zb (8, 16, 32, 64) with LDS read
// width rw (2 when z is even, else 1): mode PASS_UPDATE with psi
// PSI_F64_EXACT or PSI_F64_APPR, or PASS_ADD (psi_appr, as bld_add).
hipError_t launch_pass_f64(int psi, int mode, int zb, int rw, dim3 grid, hipStream_t st, const BldInst* insts,
                           const PassItem* items, const uint32_t* n_items) {
    if (mode == PASS_ADD) return launch_f64_rw<PASS_ADD, PSI_F64_APPR>(rw, zb, grid, st, insts, items, n_items);
    if (mode != PASS_UPDATE) return hipErrorInvalidValue;
    if (psi == PSI_F64_APPR)
        return launch_f64_rw<PASS_UPDATE, PSI_F64_APPR>(rw, zb, grid, st, insts, items, n_items);
    return launch_f64_rw<PASS_UPDATE, PSI_F64_EXACT>(rw, zb, grid, st, insts, items, n_items);
}

template <int MODE, int PSI>
static hipError_t attr_f64_mode() {
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = attr_f64_t<8, 1, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<8, 2, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<16, 1, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<16, 2, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<32, 1, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<32, 2, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<64, 1, MODE, PSI>();
    if (e == hipSuccess) e = attr_f64_t<64, 2, MODE, PSI>();
    return e;
}

hipError_t init_f64_attributes() {
    hipError_t e = attr_f64_mode<PASS_UPDATE, PSI_F64_EXACT>();
    if (e == hipSuccess) e = attr_f64_mode<PASS_UPDATE, PSI_F64_APPR>();
    if (e == hipSuccess) e = attr_f64_mode<PASS_ADD, PSI_F64_APPR>();
    return e;
}

}  // namespace bnmf
