// philox.hpp — Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable on the
// host and on the GPU. A (seed, stream, index) triple names an independent
// sequence, so a variate's value does not depend on which thread or device
// draws it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bnmf {

// stream ids: one per kind of variate
enum : uint32_t { ST_L = 1, ST_W = 2, ST_H = 3, ST_X = 4, ST_S = 5, ST_INIT = 6 };

struct Philox {
    uint32_t k0, k1;
    uint32_t c0, c1, c2, c3;
    uint32_t buf[4];
    int pos;

    __host__ __device__ Philox(uint64_t seed, uint32_t stream, uint64_t index)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), c0(0), c1(stream),
          c2((uint32_t)index), c3((uint32_t)(index >> 32)), buf{0, 0, 0, 0}, pos(4) {}

    __host__ __device__ void block() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, a = k0, b = k1;
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
            const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
            const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
            const uint32_t n0 = hi1 ^ x1 ^ a, n2 = hi0 ^ x3 ^ b;
            x0 = n0;
            x1 = lo1;
            x2 = n2;
            x3 = lo0;
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
        buf[0] = x0;
        buf[1] = x1;
        buf[2] = x2;
        buf[3] = x3;
        ++c0;  // next block of this variate's stream
        pos = 0;
    }
    __host__ __device__ uint32_t next32() {
        if (pos == 4) block();
        return buf[pos++];
    }
    // uniform on the open interval (0, 1), 53 random bits
    __host__ __device__ double uniform() {
        const uint64_t hi = next32(), lo = next32();
        const uint64_t m = ((hi << 21) ^ (lo >> 11)) & ((1ull << 53) - 1);
        return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
    }
};

// Poisson(mu): Knuth's product method for mu < 12, otherwise the transformed
// rejection method PTRS (W. Hormann, "The transformed rejection method for
// generating Poisson random variables", 1993). Host and device run the same
// operations, so they draw the same value unless their exp / log / lgamma
// differ in the last bit at a comparison's boundary.
__host__ __device__ inline double poisson_draw(Philox& g, double mu) {
    if (!(mu > 0.0)) return 0.0;
    if (mu < 12.0) {
        const double lim = exp(-mu);
        double prod = g.uniform();
        double k = 0;
        while (prod > lim) {
            prod *= g.uniform();
            k += 1;
        }
        return k;
    }
    const double smu = sqrt(mu);
    const double b = 0.931 + 2.53 * smu;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double log_mu = log(mu);
    for (;;) {
        const double U = g.uniform() - 0.5;
        const double V = g.uniform();
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -mu + k * log_mu - lgamma(k + 1.0))
            return k;
    }
}

// init_S for one fiber (bld_mult_funcs.hpp:36-74): S0_k = X * D_k with
// D ~ Dirichlet(1_z), i.e. normalised Exp(1) draws. w must hold z doubles.
template <typename T>
__host__ __device__ inline void init_fiber(uint64_t seed, uint64_t fiber, double xv, uint32_t z,
                                           T* out, double* w) {
    if (xv == 0.0) {
        for (uint32_t k = 0; k < z; ++k) out[k] = T(0);
        return;
    }
    Philox g(seed, ST_INIT, fiber);
    double tot = 0.0;
    for (uint32_t k = 0; k < z; ++k) {
        w[k] = -log(g.uniform());  // Gamma(1, 1)
        tot += w[k];
    }
    for (uint32_t k = 0; k < z; ++k) out[k] = (T)(xv * (w[k] / tot));
}

}  // namespace bnmf
