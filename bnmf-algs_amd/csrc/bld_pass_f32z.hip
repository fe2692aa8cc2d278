// bld_pass_f32z.hip — the FP32-bulk step passes (bld_mult's multiplicative
// update and bld_add's additive one) specialised on the exact fiber depth z
// (1..64).
//
// Same work decomposition, LDS layout and arithmetic as the generic update
// pass in bld_kernels.hip (a block of WAVES wavefronts per tile of ti_rows x
// 64 fibers; each lane owns one fiber of the current row, held in a
// wave-private LDS row buffer), but with z a compile-time constant: the
// chunk loops have no per-chunk guards, the LDS read width and the XOR visit
// order are constants, the per-fiber arrays have exactly z entries (fewer
// registers, so more waves per SIMD for small z) and, for z <= 24, the fiber
// stays in registers between the two phases instead of being read again.
// It serves instances whose S stays nonnegative (S0 >= 0 and beta >= 1, the
// descriptor's nonneg flag), so psi(S+1) always takes its x >= 1 form; other
// instances run the generic pass. The update is the generic pass's formula
// with X_reciprocal and X / (S_ij+ + eps) folded into per-fiber constants
// (nom and den scaled once per fiber, one reciprocal per element), so the
// two kernels agree to FP32 rounding rather than bit for bit.
//
// Reference operations (src/bld/bld_mult/bld_mult_funcs.hpp): grad_plus
// :177-192, nom :242-259, denom :272-290, update_S :307-327; the axis sums of
// the new S (bld_mult.hpp:124-133) leave as per-tile partials.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "bld_types.hpp"
#include "digamma.hpp"
#include "pass_common.hpp"

namespace bnmf {

// Buffer resource word 3 for raw (untyped) buffer access on gfx9-family
// GPUs, and an offset past any resource of this kernel: a store given it is
// dropped by the range check.
constexpr int kBufWord3 = 0x00020000;
constexpr uint32_t kBufDrop = 0x80000000u;
// Cache policy of the S stores: nt (bit 1 on gfx94x/gfx950), a streaming
// write. S is read once and written once per pass, so both its loads
// (__builtin_nontemporal_load) and its whole-chunk stores are streaming
// accesses, which leaves L2 and the last-level cache to the partial sums, X
// and the psi staging. Measured on the 256-instance workload: loads and
// stores both streaming, pass 578 us against 608 us; either alone was slower
// than neither.
constexpr int kStreamAux = 2;

// Compile-time form of read_layout() for float: LDS read width RW and the XOR
// mask of the chunk visit order (lane l visits chunk m as m ^ ((l >> HS) & HM)).
template <int Z>
struct F32Layout {
    static constexpr int RW = (Z % 4 == 0) ? 4 : ((Z % 2 == 0) ? 2 : 1);
    static constexpr int NCH = Z / RW;
    static constexpr int LPC = (int)kLdsBankBytes / (RW * 4);
    static constexpr int p_of() {
        int p = 1;
        if (RW == 1) return 1;
        while (p < LPC && (NCH % (p * 2)) == 0) p *= 2;
        return p;
    }
    static constexpr int P = p_of();
    static constexpr int hs_of() {
        int s = 0;
        while ((1 << s) < LPC / P) ++s;
        return s;
    }
    static constexpr int HS = hs_of();
    static constexpr int HM = P - 1;
    static constexpr int zq_of() {
        int q = 1;
        while (q < Z) q *= 2;
        return q;
    }
    static constexpr int ZQ = zq_of();
    static constexpr int ZQL = (ZQ == 1 ? 0 : ZQ == 2 ? 1 : ZQ == 4 ? 2 : ZQ == 8 ? 3 : ZQ == 16 ? 4 : ZQ == 32 ? 5 : 6);
};

// LDS of a block: [psib 64*Z][psiA Z][WAVES x grad_minus Z][WAVES x 64 X_ij]
// [WAVES x row buffer]. A row buffer holds CPL*64 16-byte chunks: every lane writes all of
// its chunks each row (the ones past the row end hold repeated data), so the
// prefetch has a fixed instruction count.
template <int Z>
struct F32zLds {
    // A row segment starts d floats past a 16-byte boundary: d = 0 when
    // z % 4 == 0, d in {0, 2} when z is even, else d <= 3. It spans
    // ceil((d + 64 z) / 4) 16-byte chunks.
    static constexpr uint32_t DMAX = (Z % 4 == 0) ? 0u : ((Z % 2 == 0) ? 2u : 3u);
    static constexpr uint32_t CPL = (16u * Z + (DMAX ? 1u : 0u) + 63u) / 64u;  // chunks per lane
    // psi(beta_eph) rows of the tile: stride Z when z % 4 == 0 (read with
    // the fiber's own chunk order); otherwise padded to a multiple of 4
    // floats with an odd number of 16-byte groups, so each lane reads it
    // 16 bytes at a time with the lanes of an access cycle on distinct
    // banks, whatever z is
    static constexpr uint32_t PBS = (Z % 4 == 0) ? (uint32_t)Z
                                    : (((Z + 3) / 4) % 2 == 1 ? (Z + 3) / 4 * 4 : (Z + 3) / 4 * 4 + 4);
    static constexpr size_t psiA_off = round16((size_t)64 * PBS * 4);
    static constexpr size_t gm_off = psiA_off + round16((size_t)Z * 4);
    static constexpr size_t gm_stride = round16((size_t)Z * 4);
    static constexpr size_t raw_stride = (size_t)CPL * 64 * 16;
    static constexpr size_t xs_off(int waves) { return gm_off + waves * gm_stride; }
    static constexpr size_t raw_off(int waves) { return xs_off(waves) + waves * 256; }
    static constexpr size_t bytes(int waves) { return raw_off(waves) + waves * raw_stride; }
};

template <int Z>
constexpr int f32z_waves() {
    int w = 4;
    while (w > 1 && F32zLds<Z>::bytes(w) > kLdsBytes) --w;
    return w;
}

// Waves per SIMD the register allocation must leave room for.
#ifndef BNMF_F32Z_WPE
#define BNMF_F32Z_WPE(Z) ((Z) <= 12 ? 4 : (Z) <= 20 ? 3 : ((Z) <= 32 || (Z) == 36) ? 2 : 1)
#endif
// The additive step keeps more live values per element (exp(Z) and the
// row's scale): one wave fewer where the multiplicative budget would spill.
#define BNMF_F32Z_WPE_ADD(Z) ((Z) <= 7 ? 4 : (Z) <= 16 ? 3 : (Z) <= 32 ? 2 : 1)

// Memory pipelining: a wave issues the next row's loads (S chunks, X, psia)
// right after moving the current row into LDS, and every global store of a
// row is an unconditional buffer store (lanes with nothing to write pass an
// out-of-range offset, which the hardware drops). The instruction count between a row's loads and their
// first use is then fixed, so the wait for them leaves the previous row's
// stores in flight instead of draining the whole memory queue.
// MODE: PASS_UPDATE (bld_mult's multiplicative step) or PASS_ADD (bld_add's
// additive step, src/bld/bld_add.hpp:137-181).
template <int Z, int WAVES, int MODE>
__global__ void __launch_bounds__(64 * WAVES)
    __attribute__((amdgpu_waves_per_eu(MODE == PASS_ADD ? BNMF_F32Z_WPE_ADD(Z) : BNMF_F32Z_WPE(Z))))
    bld_pass_f32z_kernel(const BldInst* __restrict__ insts, const PassItem* __restrict__ items,
                         const uint32_t* __restrict__ n_items_p) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    using Lds = F32zLds<Z>;
    using Lay = F32Layout<Z>;
    constexpr int RW = Lay::RW;
    constexpr int NCH = Lay::NCH;
    constexpr uint32_t NT = 64 * WAVES;
    constexpr uint32_t VEC = 4;
    constexpr int CPL = (int)Lds::CPL;
    constexpr bool KEEP = Z <= 24;  // fiber kept in registers across the phases
    constexpr uint32_t ZQ = Lay::ZQ;
    constexpr uint32_t L = 64u / ZQ;
    typedef float V16 __attribute__((ext_vector_type(4)));
    typedef uint32_t U16 __attribute__((ext_vector_type(4)));
    typedef typename VecR<float, RW>::type VR;
    typedef float V8 __attribute__((ext_vector_type(2)));
    constexpr bool PAD = (Z % 4) != 0;
    constexpr int NG = (Z + 3) / 4;  // groups of 4 elements of a fiber
    // elements 4*g4 .. 4*g4+3 (those below Z) of a fiber in RW-wide LDS
    // accesses; fb is RW-aligned (z even: RW = 2, d even)
    auto load4 = [](const float* fb, int g4, float* s4) {
#pragma unroll
        for (int v = 0; v < 4; v += RW) {
            if (4 * g4 + v < Z) {
                if constexpr (RW == 2) {
                    const V8 t = *reinterpret_cast<const V8*>(fb + 4 * g4 + v);
                    s4[v] = t[0];
                    s4[v + 1] = t[1];
                } else {
                    s4[v] = fb[4 * g4 + v];
                }
            } else {
                s4[v] = 0.0f;
                if constexpr (RW == 2) s4[v + 1] = 0.0f;
            }
        }
    };
    auto store4 = [](float* fb, int g4, const float* s4) {
#pragma unroll
        for (int v = 0; v < 4; v += RW) {
            if (4 * g4 + v < Z) {
                if constexpr (RW == 2) {
                    V8 t;
                    t[0] = s4[v];
                    t[1] = s4[v + 1];
                    *reinterpret_cast<V8*>(fb + 4 * g4 + v) = t;
                } else {
                    fb[4 * g4 + v] = s4[v];
                }
            }
        }
    };

    const uint32_t n_items = *n_items_p;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    // wave index as a scalar: row loops and their branches stay uniform
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));
    const uint32_t hsw = RW > 1 ? ((lane >> Lay::HS) & (uint32_t)Lay::HM) : 0u;

    float* const psib_l = reinterpret_cast<float*>(lds_raw);
    float* const psiA_l = reinterpret_cast<float*>(lds_raw + Lds::psiA_off);
    float* const gm_l = reinterpret_cast<float*>(lds_raw + Lds::gm_off + w * Lds::gm_stride);
    float* const xs_l = reinterpret_cast<float*>(lds_raw + Lds::xs_off(WAVES) + w * 256);
    unsigned char* const raw_base = lds_raw + Lds::raw_off(WAVES);
    float* const raw = reinterpret_cast<float*>(raw_base + w * Lds::raw_stride);

    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PassItem item = items[it];
        const BldInst* __restrict__ I = insts + item.inst;
        const uint32_t x = I->x, y = I->y;
        const uint32_t i0 = item.ti * I->ti_rows;
        const uint32_t i1 = min(x, i0 + I->ti_rows);
        const uint32_t j0 = item.tj * kTileCols;
        const uint32_t nj = min(kTileCols, y - j0);
        const uint32_t n_el = nj * Z;

        const BNMF_GLOBAL float* __restrict__ Ssrc = gconst<float>(I->S_src);
        BNMF_GLOBAL float* __restrict__ Sdst = gmut<float>(I->S);
        const BNMF_GLOBAL float* __restrict__ Xg = gconst<float>(I->X);
        const BNMF_GLOBAL float* __restrict__ psia_g = gconst<float>(I->psia);
        const float epsf = static_cast<float>(I->eps);
        const uint32_t lane_x = min(lane, nj - 1u);  // clamped fiber of this lane
        const uint32_t lane_z = min(lane, (uint32_t)Z - 1u);

        // prefetch registers of one row: S chunks from the aligned start
        // (chunk indices clamped to the row's last chunk), X_ij of the
        // lane's fiber and psi(alpha_eph)(i, lane)
        V16 pf[CPL];
        float pf_x = 0.0f, pf_pa = 0.0f;
        uint32_t i = i0 + w;
        uint32_t dl = 0, nch = 0;
        // real = false: there is no next row; every lane loads the row's
        // first chunk (one cache line) so the instruction count stays fixed
        auto issue_row = [&](uint32_t row, bool real, uint32_t& d_out, uint32_t& n_out) {
            const size_t s0 = ((size_t)row * y + j0) * Z;
            const uint32_t d = (uint32_t)(s0 % VEC);
            const uint32_t n = (d + n_el + VEC - 1) / VEC;
            const BNMF_GLOBAL V16* src = (const BNMF_GLOBAL V16*)(Ssrc + (s0 - d));
            const uint32_t last = real ? n - 1u : 0u;
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                pf[c] = __builtin_nontemporal_load(&src[min((uint32_t)c * 64u + lane, last)]);
            pf_x = Xg[(size_t)row * y + j0 + lane_x];
            pf_pa = psia_g[(size_t)row * Z + lane_z];
            d_out = d;
            n_out = n;
        };
        // the tile's first row is in flight while psi(beta_eph) is staged
        if (i < i1) issue_row(i, true, dl, nch);

        {
            const BNMF_GLOBAL float* __restrict__ psib_g = gconst<float>(I->psib) + (size_t)j0 * Z;
            constexpr uint32_t PB = (64u * Z + NT - 1) / NT;  // elements per thread
            float pbv[PB];
#pragma unroll
            for (uint32_t u = 0; u < PB; ++u) pbv[u] = psib_g[min(tid + u * NT, n_el - 1u)];
#pragma unroll
            for (uint32_t u = 0; u < PB; ++u) {
                const uint32_t e = tid + u * NT;
                if (e < n_el) {
                    const uint32_t f = e / (uint32_t)Z, k = e - f * (uint32_t)Z;
                    psib_l[f * Lds::PBS + k] = pbv[u];
                }
            }
            // psi(sum_i alpha_eph(i,k)) (update_grad_minus,
            // bld_mult_funcs.hpp:213-220), from the prep launch
            if (tid < Z) psiA_l[tid] = gconst<float>(I->psiA)[tid];
            __syncthreads();
        }

        // additive step size 0.1 / (iteration + 1)^0.55 (bld_add.hpp:101-103);
        // *iter counts the iterations started, this one included
        float etaf = 0.0f;
        if constexpr (MODE == PASS_ADD) etaf = static_cast<float>(0.1 / pow((double)*I->iter, 0.55));

        float acc[Z];
#pragma unroll
        for (int k = 0; k < Z; ++k) acc[k] = 0.0f;

        const uint32_t kq = lane & (ZQ - 1u);
        const uint32_t part = lane >> Lay::ZQL;

        // One row of the tile. The first row is peeled off the loop so that
        // at the loop head the previous row's stores always follow the
        // current row's loads in issue order.
        auto row_step = [&](uint32_t i) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) reinterpret_cast<V16*>(raw)[(uint32_t)c * 64u + lane] = pf[c];
            if (lane < (uint32_t)Z) gm_l[lane] = psiA_l[lane] - pf_pa;
            // X_ij goes through LDS too: the prefetch register is then free
            // before the next row's load, and the loop carries no copy that
            // would have to wait for that load
            xs_l[lane] = pf_x;
            const uint32_t cur_dl = dl, cur_nch = nch;
            const size_t cur_s0 = ((size_t)i * y + j0) * Z;
            const uint32_t inext = i + WAVES;
            const bool more = inext < i1;
            issue_row(more ? inext : i, more, dl, nch);
            wave_sync();
            const float xv = lane < nj ? xs_l[lane] : 0.0f;
            if (lane < nj) {
                float* const fb = raw + cur_dl + lane * Z;
                const float* const pbb = psib_l + lane * Lds::PBS;
                float gp[Z];
                float sv_keep[KEEP ? Z : 1];
                if constexpr (MODE == PASS_ADD) {
                    // bld_add (src/bld/bld_add.hpp:137-181). Phase A:
                    // grad_S = psi(beta_eph) - psi(S+1) - (psi(A) - psi(alpha_eph))
                    // and S_mult = sum_k S grad_S
                    float smult = 0.0f;
                    auto elem_a = [&](int k, float sk, float pbk, float gmk) {
                        if constexpr (KEEP) sv_keep[k] = sk;
                        const float g = (pbk - psi_ge1_f32(sk + 1.0f)) - gmk;
                        gp[k] = g;
                        smult = fmaf(sk, g, smult);
                    };
                    if constexpr (!PAD) {
#pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                            const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
#pragma unroll
                            for (int v = 0; v < RW; ++v)
                                elem_a(m * RW + v, vget<float, RW>(sv, v), vget<float, RW>(pb, v),
                                       vget<float, RW>(gv, v));
                        }
                    } else {
#pragma unroll
                        for (int g4 = 0; g4 < NG; ++g4) {
                            const V16 pb = *reinterpret_cast<const V16*>(pbb + 4 * g4);
                            const V16 gv = *reinterpret_cast<const V16*>(gm_l + 4 * g4);
                            float s4[4];
                            load4(fb, g4, s4);
#pragma unroll
                            for (int v = 0; v < 4; ++v)
                                if (4 * g4 + v < Z) elem_a(4 * g4 + v, s4[v], pb[v], gv[v]);
                        }
                    }
                    // Phase B: exp(Z) = exp(log S + eta S (X grad_S - S_mult))
                    // = S exp(eta S (X grad_S - S_mult)) (one rounding of S
                    // instead of log's; S = 0 stays 0) and its L1 norm
                    float norm = 0.0f;
                    auto elem_e = [&](int k, float sk) {
                        const float e = sk * __expf((etaf * sk) * fmaf(xv, gp[k], -smult));
                        gp[k] = e;
                        norm += e;
                    };
                    if constexpr (!PAD) {
#pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            VR sv;
                            if constexpr (KEEP) {
#pragma unroll
                                for (int v = 0; v < RW; ++v) vset<float, RW>(sv, v, sv_keep[m * RW + v]);
                            } else {
                                sv = *reinterpret_cast<const VR*>(fb + mm);
                            }
#pragma unroll
                            for (int v = 0; v < RW; ++v) elem_e(m * RW + v, vget<float, RW>(sv, v));
                        }
                    } else {
#pragma unroll
                        for (int g4 = 0; g4 < NG; ++g4) {
                            float s4[4];
                            if constexpr (KEEP) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) s4[v] = 4 * g4 + v < Z ? sv_keep[4 * g4 + v] : 0.0f;
                            } else {
                                load4(fb, g4, s4);
                            }
#pragma unroll
                            for (int v = 0; v < 4; ++v)
                                if (4 * g4 + v < Z) elem_e(4 * g4 + v, s4[v]);
                        }
                    }
                    // Phase C: S = X exp(Z) / |exp(Z)|_1 (the reference's
                    // 1e-50 guard underflows in float: an all-zero row, X = 0,
                    // is taken explicitly) and the S_+jk partials
                    const float scale = norm > 0.0f ? xv / norm : 0.0f;
                    if constexpr (!PAD) {
#pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            VR sv;
#pragma unroll
                            for (int v = 0; v < RW; ++v) {
                                const float sn = as_stored(gp[m * RW + v] * scale);
                                acc[m * RW + v] += sn;
                                vset<float, RW>(sv, v, sn);
                            }
                            *reinterpret_cast<VR*>(fb + mm) = sv;
                        }
                    } else {
#pragma unroll
                        for (int g4 = 0; g4 < NG; ++g4) {
                            float s4[4];
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int k = 4 * g4 + v;
                                s4[v] = 0.0f;
                                if (k < Z) {
                                    s4[v] = as_stored(gp[k] * scale);
                                    acc[k] += s4[v];
                                }
                            }
                            store4(fb, g4, s4);
                        }
                    }
                } else {
                    float sumS = 0.0f, nom = 0.0f, den = 0.0f;
                    // Phase A: per element g = psi(beta_eph) - psi(S+1) and the
                    // fiber sums S_ij+, sum_k S*gm and sum_k S*g.
                    // S >= 0 throughout (S0 >= 0 and beta >= 1), so S + 1 >= 1
                    // and psi takes its x >= 1 form directly.
                    auto elem_a = [&](int k, float sk, float pbk, float gmk) {
                        if constexpr (KEEP) sv_keep[k] = sk;
                        // psi(S+1) rounded at its own magnitude first: forming
                        // pbk - ln2*log2(S+1) before adding the small polynomial
                        // term would round at the (larger) magnitude of
                        // psi(beta_eph), which ill-conditioned instances amplify
                        const float g = pbk - psi_ge1_f32(sk + 1.0f);
                        gp[k] = g;
                        sumS += sk;
                        nom = fmaf(sk, gmk, nom);
                        den = fmaf(sk, g, den);
                    };
                    if constexpr (!PAD) {
    #pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            const VR sv = *reinterpret_cast<const VR*>(fb + mm);
                            const VR pb = *reinterpret_cast<const VR*>(pbb + mm);
                            const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
    #pragma unroll
                            for (int v = 0; v < RW; ++v)
                                elem_a(m * RW + v, vget<float, RW>(sv, v), vget<float, RW>(pb, v),
                                       vget<float, RW>(gv, v));
                        }
                    } else {
                        // z % 4 != 0: the fiber in RW-wide reads, psi(beta_eph)
                        // and gm 16 bytes at a time
    #pragma unroll
                        for (int g4 = 0; g4 < NG; ++g4) {
                            const V16 pb = *reinterpret_cast<const V16*>(pbb + 4 * g4);
                            const V16 gv = *reinterpret_cast<const V16*>(gm_l + 4 * g4);
                            float s4[4];
                            load4(fb, g4, s4);
    #pragma unroll
                            for (int v = 0; v < 4; ++v)
                                if (4 * g4 + v < Z) elem_a(4 * g4 + v, s4[v], pb[v], gv[v]);
                        }
                    }
                    // Per-fiber constants. With Xh = 1/(X + eps) (X_reciprocal)
                    // the reference's nom and denom are Xh * sum, and its update
                    //   S *= (g + nom) * X / (S_ij+ + eps) / (gm + den + eps)
                    // is S *= (g + nom) / (gm * q + (den + eps) * q) with
                    // q = (S_ij+ + eps) / X. eps = 1e-50 flushes to zero in
                    // float; the reference's double-promoted guard makes an
                    // X_ij = 0 or all-zero fiber update to exactly 0, taken here
                    // as q = 0, (den + eps) * q = inf: the factor is then 0.
                    const bool upd = xv > 0.0f && sumS != 0.0f;
                    const float xh = upd ? 1.0f / (xv + epsf) : 0.0f;
                    nom *= xh;
                    den *= xh;
                    const float q = upd ? (sumS + epsf) / xv : 0.0f;
                    // den >= 0 here; kept off exact zero so that an all-zero
                    // numerator (grad_plus = grad_minus = 0, e.g. a single row
                    // with beta = 1) updates to 0, as the reference's double
                    // eps guard makes it, instead of 0/0
                    const float dq = upd ? fmaxf((den + epsf) * q, FLT_MIN) : __builtin_inff();
                    // Phase B: the multiplicative update and the S_+jk partials
                    auto elem_b = [&](int k, float sk, float gmk) {
                        const float r = recip_fast<float>(fmaf(gmk, q, dq));
                        const float sn = as_stored(sk * ((gp[k] + nom) * r));
                        acc[k] += sn;
                        return sn;
                    };
                    if constexpr (!PAD) {
    #pragma unroll
                        for (int m = 0; m < NCH; ++m) {
                            const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                            VR sv;
                            if constexpr (KEEP) {
    #pragma unroll
                                for (int v = 0; v < RW; ++v) vset<float, RW>(sv, v, sv_keep[m * RW + v]);
                            } else {
                                sv = *reinterpret_cast<const VR*>(fb + mm);
                            }
                            const VR gv = *reinterpret_cast<const VR*>(gm_l + mm);
    #pragma unroll
                            for (int v = 0; v < RW; ++v)
                                vset<float, RW>(sv, v, elem_b(m * RW + v, vget<float, RW>(sv, v),
                                                              vget<float, RW>(gv, v)));
                            *reinterpret_cast<VR*>(fb + mm) = sv;
                        }
                    } else {
    #pragma unroll
                        for (int g4 = 0; g4 < NG; ++g4) {
                            const V16 gv = *reinterpret_cast<const V16*>(gm_l + 4 * g4);
                            float s4[4];
                            if constexpr (KEEP) {
    #pragma unroll
                                for (int v = 0; v < 4; ++v) s4[v] = 4 * g4 + v < Z ? sv_keep[4 * g4 + v] : 0.0f;
                            } else {
                                load4(fb, g4, s4);
                            }
    #pragma unroll
                            for (int v = 0; v < 4; ++v)
                                if (4 * g4 + v < Z) s4[v] = elem_b(4 * g4 + v, s4[v], gv[v]);
                            store4(fb, g4, s4);
                        }
                    }
                }
            }
            wave_sync();

            // S_i+k partial over the tile's columns: eight independent
            // partial sums combined in a fixed tree (deterministic)
            {
                float r[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) r[v] = 0.0f;
                if (kq < (uint32_t)Z) {
                    const float* const rb = raw + cur_dl + kq;
                    for (uint32_t f0 = part; f0 < nj; f0 += 8u * L) {
#pragma unroll
                        for (int v = 0; v < 8; ++v) {
                            const uint32_t f = f0 + (uint32_t)v * L;
                            if (f < nj) r[v] += rb[f * Z];
                        }
                    }
                }
                float rs = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
#pragma unroll
                for (uint32_t off = ZQ; off < 64u; off <<= 1) rs += __shfl_xor(rs, (int)off);
                if (part == 0 && kq < (uint32_t)Z)
                    gmut<float>(I->part_ipk)[((size_t)item.tj * x + i) * Z + kq] = rs;
            }

            // the row back to S through a buffer resource spanning the row's
            // chunks: every lane issues its stores, and a lane with nothing
            // to write gives an offset past the resource, which the hardware
            // drops. 16-byte chunks that lie inside the row segment go whole;
            // the (at most two) chunks straddling the segment ends go element
            // by element, lanes 0-3 the head chunk and lanes 4-7 the tail.
            {
                const __amdgpu_buffer_rsrc_t dst = __builtin_amdgcn_make_buffer_rsrc(
                    (void*)(Sdst + (cur_s0 - cur_dl)), (short)0, (int)(cur_nch * 16u), kBufWord3);
                const uint32_t lo = cur_dl, hi = cur_dl + n_el;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const uint32_t q = (uint32_t)c * 64u + lane;
                    const uint32_t e0 = q * VEC;
                    const bool inside = q < cur_nch && e0 >= lo && e0 + VEC <= hi;
                    const V16 v = reinterpret_cast<const V16*>(raw)[q];
                    __builtin_amdgcn_raw_buffer_store_b128(
                        __builtin_bit_cast(U16, v), dst, (int)(inside ? q * 16u : kBufDrop), 0, kStreamAux);
                }
                const uint32_t qe = lane < 4u ? 0u : cur_nch - 1u;
                const uint32_t e = qe * VEC + (lane & 3u);
                const bool mine = lane < 8u && e >= lo && e < hi &&
                                  ((qe * VEC < lo) || (qe * VEC + VEC > hi));
                const float v = raw[min(e, (uint32_t)Lds::raw_stride / 4u - 1u)];
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), dst,
                                                      (int)(mine ? e * 4u : kBufDrop), 0, 0);
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
        float* const comb_self = reinterpret_cast<float*>(raw_base + w * Lds::raw_stride);
        if (lane < nj) {
            // RW-wide writes in the fiber's chunk visit order: lanes of one
            // access cycle land on different banks (scalar writes at stride
            // Z put every lane on one bank when Z is a multiple of 8)
#pragma unroll
            for (int m = 0; m < NCH; ++m) {
                const uint32_t mm = ((uint32_t)m ^ hsw) * RW;
                VR cv;
#pragma unroll
                for (int v = 0; v < RW; ++v) vset<float, RW>(cv, v, acc[m * RW + v]);
                *reinterpret_cast<VR*>(comb_self + lane * Z + mm) = cv;
            }
        }
        __syncthreads();
        {
            BNMF_GLOBAL float* __restrict__ pp = gmut<float>(I->part_pjk) + ((size_t)item.ti * y + j0) * Z;
            for (uint32_t e = tid; e < n_el; e += NT) {
                float sacc = 0.0f;
#pragma unroll
                for (int w2 = 0; w2 < WAVES; ++w2)
                    sacc += reinterpret_cast<const float*>(raw_base + w2 * Lds::raw_stride)[e];
                pp[e] = sacc;
            }
        }
        __syncthreads();
    }
}

template <int Z, int MODE>
static hipError_t launch_f32z_t(dim3 grid, hipStream_t st, const BldInst* insts,
                                const PassItem* items, const uint32_t* n_items) {
    constexpr int WAVES = f32z_waves<Z>();
    const size_t lds = F32zLds<Z>::bytes(WAVES);
    hipLaunchKernelGGL((bld_pass_f32z_kernel<Z, WAVES, MODE>), grid, dim3(64 * WAVES), lds, st, insts,
                       items, n_items);
    return hipGetLastError();
}

template <int Z>
static hipError_t attr_f32z_t() {
    constexpr int WAVES = f32z_waves<Z>();
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(bld_pass_f32z_kernel<Z, WAVES, PASS_UPDATE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)F32zLds<Z>::bytes(WAVES));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(bld_pass_f32z_kernel<Z, WAVES, PASS_ADD>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)F32zLds<Z>::bytes(WAVES));
}

template <int... Zs>
struct F32zTable {
    template <int MODE>
    static hipError_t launch(int z, dim3 grid, hipStream_t st, const BldInst* insts,
                             const PassItem* items, const uint32_t* n_items) {
        hipError_t e = hipErrorInvalidValue;
        (void)((Zs == z ? (e = launch_f32z_t<Zs, MODE>(grid, st, insts, items, n_items), true) : false) || ...);
        return e;
    }
    static hipError_t attrs() {
        hipError_t e = hipSuccess;
        (void)((e = attr_f32z_t<Zs>(), e != hipSuccess) || ...);
        return e;
    }
};

using F32zAll = F32zTable<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21,
                          22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40,
                          41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59,
                          60, 61, 62, 63, 64>;

// FP32 step pass (mode PASS_UPDATE or PASS_ADD) of instances that all have
// depth z (1..64).
hipError_t launch_pass_f32z(int z, int mode, dim3 grid, hipStream_t st, const BldInst* insts,
                            const PassItem* items, const uint32_t* n_items) {
    if (mode == PASS_ADD) return F32zAll::launch<PASS_ADD>(z, grid, st, insts, items, n_items);
    if (mode == PASS_UPDATE) return F32zAll::launch<PASS_UPDATE>(z, grid, st, insts, items, n_items);
    return hipErrorInvalidValue;
}

hipError_t init_f32z_attributes() { return F32zAll::attrs(); }

}  // namespace bnmf
