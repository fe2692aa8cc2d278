// pass_common.hpp — device helpers shared by the fused pass kernels
// (bld_kernels.hip: generic z-bucket forms; bld_pass_f32z.hip: exact-z FP32
// update forms): LDS layout of a pass block, compensated sums, vector LDS
// reads and global-address-space views.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bld_types.hpp"

namespace bnmf {

// UPDATE: multiplicative step (bld_mult); SUMS: axis sums only; SCORE: sums
// and lgamma(S+1) for the log-marginal; ADD: additive step (bld_add).
enum PassMode : int { PASS_UPDATE = 0, PASS_SUMS = 1, PASS_SCORE = 2, PASS_ADD = 3 };

// passes that read psi(beta_eph), psi(A) and a psi(alpha_eph) row and
// rewrite S
__host__ __device__ constexpr bool pass_steps(int mode) { return mode == PASS_UPDATE || mode == PASS_ADD; }

// Element index within a run of whole fibers -> fiber index. Exact for
// e < 2^16 and z <= 64: (e + 0.5)/z lies at least 0.5/z from an integer and
// the float error is below 1e-5.
__device__ __forceinline__ uint32_t fiber_of(uint32_t e, float inv_z) {
    return static_cast<uint32_t>((static_cast<float>(e) + 0.5f) * inv_z);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum over the wavefront, the same value in every lane (a butterfly: every
// lane adds the same pairs in the same order)
template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <typename A>
struct Kahan {
    A s = A(0), c = A(0);
    __device__ __forceinline__ void add(A v) {
        const A yv = v - c;
        const A t = s + yv;
        c = (t - s) - yv;
        s = t;
    }
};

// Compensated sums only where the summand type is float; double sums stay
// plain (the FP64 form follows the reference's plain sums).
template <typename A>
struct SumAcc;
template <>
struct SumAcc<float> {
    Kahan<float> k;
    __device__ __forceinline__ void add(float v) { k.add(v); }
    __device__ __forceinline__ float get() const { return k.s; }
};
template <>
struct SumAcc<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double v) { s += v; }
    __device__ __forceinline__ double get() const { return s; }
};

// A new S element as it is stored: an opaque copy, so that adding it to a
// partial axis sum is not contracted with the product that produced it (an
// FMA would add the unrounded product). The partials of a step pass then
// equal the sums of S as stored, which a sums-only pass reproduces bit for bit.
template <typename T>
__device__ __forceinline__ T as_stored(T v) {
    asm("" : "+v"(v));
    return v;
}

template <typename T>
__device__ __forceinline__ T recip_fast(T v);
template <>
__device__ __forceinline__ float recip_fast<float>(float v) { return __builtin_amdgcn_rcpf(v); }

// ---------------------------------------------------------------------------
// pass kernel
// ---------------------------------------------------------------------------
// Global-address-space views: pointers read out of the descriptor are generic
// to the compiler; casting them keeps S traffic on global_load/global_store
// (flat accesses also count against the LDS wait counter, which would
// serialise the row prefetch with the LDS work).
#define BNMF_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const BNMF_GLOBAL T* gconst(const void* p) {
    return (const BNMF_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ BNMF_GLOBAL T* gmut(void* p) {
    return (BNMF_GLOBAL T*)p;
}

// RW-element vector of T (one LDS read of a fiber chunk); RW = 1 is T.
template <typename T, int RW>
struct VecR {
    typedef T type __attribute__((ext_vector_type(RW)));
};
template <typename T>
struct VecR<T, 1> {
    typedef T type;
};
template <typename T, int RW>
__device__ __forceinline__ T vget(const typename VecR<T, RW>::type& v, int i) {
    if constexpr (RW == 1) {
        (void)i;
        return v;
    } else {
        return v[i];
    }
}
template <typename T, int RW>
__device__ __forceinline__ void vset(typename VecR<T, RW>::type& v, int i, T val) {
    if constexpr (RW == 1) {
        (void)i;
        v = val;
    } else {
        v[i] = val;
    }
}

__host__ __device__ constexpr size_t round16(size_t b) { return (b + 15) / 16 * 16; }

// LDS layout of a pass block, for an instance of depth z:
//   [psib: 64*z T][psiA: z T][WAVES x grad_minus row: z T]   (step passes)
//   [WAVES x row buffer: 64*z + 2*VEC slots of max(T, ACC)]
// A row buffer holds one tile row of S in its natural layout: fiber f,
// element k at delta + f*z + k, delta = misalignment of the row start.
template <typename T, typename ACC, int MODE>
struct PassLds {
    static constexpr size_t VEC = 16 / sizeof(T);
    static constexpr size_t SLOT = sizeof(ACC) > sizeof(T) ? sizeof(ACC) : sizeof(T);
    __host__ __device__ static constexpr size_t psiA_off(uint32_t z) {
        return pass_steps(MODE) ? round16((size_t)64 * z * sizeof(T)) : 0;
    }
    __host__ __device__ static constexpr size_t gm_off(uint32_t z) {
        return pass_steps(MODE) ? psiA_off(z) + round16((size_t)z * sizeof(T)) : 0;
    }
    __host__ __device__ static constexpr size_t gm_stride(uint32_t z) {
        return pass_steps(MODE) ? round16((size_t)z * sizeof(T)) : 0;
    }
    __host__ __device__ static constexpr size_t raw_stride(uint32_t z) {
        return round16(((size_t)64 * z + 2 * VEC) * SLOT);
    }
    __host__ __device__ static constexpr size_t raw_off(uint32_t z, int waves) {
        return gm_off(z) + (size_t)waves * gm_stride(z);
    }
    __host__ __device__ static constexpr size_t bytes(uint32_t z, int waves) {
        return raw_off(z, waves) + (size_t)waves * raw_stride(z);
    }
};

constexpr size_t kLdsBytes = 160 * 1024;
// Bytes of LDS bank width the chunk visit order spreads a cycle over: lanes
// whose addresses agree modulo this in one access cycle conflict.
constexpr uint32_t kLdsBankBytes = 128;

// Wavefronts per pass block: four (one per SIMD), so that two blocks can
// share a CU and hide each other's tile prologue and epilogue; fewer when
// the LDS of a CU cannot hold four row buffers.
template <typename T, typename ACC, int ZB, int MODE>
__host__ __device__ constexpr int pass_waves() {
    using Lds = PassLds<T, ACC, MODE>;
    int w = 4;
    while (w > 1 && Lds::bytes(ZB, w) > kLdsBytes) --w;
    return w;
}

}  // namespace bnmf
