// rng.hpp -- counter-based random numbers and Poisson/Normal samplers.
//
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC11): a keyed bijection of a 128-bit counter, so every
// (seed, bin, frequency, realization, attempt) has its own independent draw and
// results do not depend on launch shape or thread order.
#pragma once
#include "common.hpp"

namespace hd {

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// xoshiro128** (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", ACM TOMS 47, 2021): a 128-bit xorshift state with the "**"
// output scrambler, 32 bits per step, all full-rate integer ops.  Used as a
// short sequential stream seeded from one Philox block (a keyed hash of the
// stream's coordinates), so streams stay independent and reproducible while a
// draw costs a few instructions instead of a Philox evaluation.
struct Xo128 {
    uint32_t s0, s1, s2, s3;
    __host__ __device__ __forceinline__ void seed(U4 w) {
        s0 = w.x; s1 = w.y; s2 = w.z; s3 = w.w;
        if ((s0 | s1 | s2 | s3) == 0u) s0 = 0x9E3779B9u;   // the all-zero state is a fixed point
    }
    __host__ __device__ __forceinline__ uint32_t next() {
        const uint32_t m = s1 * 5u;
        const uint32_t r = ((m << 7) | (m >> 25)) * 9u;
        const uint32_t t = s1 << 9;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = (s3 << 11) | (s3 >> 21);
        return r;
    }
};

// uniform in (0,1) with 53 random bits
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    uint64_t x = ((uint64_t)hi << 21) | (lo >> 11);
    return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

// A per-draw stream: counter (c0, c1, c2, attempt), 4 words per block.
struct Stream {
    uint32_t c0, c1, c2, att;
    uint32_t k0, k1;
    U4 buf;
    int used;   // doubles consumed from buf (0..2)
    __device__ __forceinline__ void init(uint32_t a, uint32_t b, uint32_t c, uint32_t tag, uint64_t seed) {
        c0 = a; c1 = b; c2 = c; att = tag << 16;
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        used = 2;
    }
    __device__ __forceinline__ double next() {
        if (used == 2) {
            U4 ctr = {c0, c1, c2, att};
            buf = philox4x32_10(ctr, k0, k1);
            att++;
            used = 0;
        }
        double u = (used == 0) ? u53(buf.x, buf.y) : u53(buf.z, buf.w);
        used++;
        return u;
    }
};

// Poisson by inversion of the upper tail, for lam < 10.  S_k = P(N >= k);
// n = #{k >= 1 : u < S_k}; S_1 = -expm1(-lam) is exact for tiny lam.
__device__ __forceinline__ double poisson_inversion(double u, double lam, double q1, double elam) {
    double n = 0.0, S = q1, p = elam;
    int k = 0;
    while (u < S && k < 1000) {
        k++;
        p = p * lam / k;
        S = S - p;
    }
    n = (double)k;
    return n;
}

// PTRS transformed rejection (W. Hoermann, "The transformed rejection method
// for generating Poisson random variables", Insurance: Mathematics and
// Economics 12, 1993), for lam >= 10.
__device__ __forceinline__ double poisson_ptrs(Stream& st, double lam) {
    double slam = sqrt(lam), loglam = log(lam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2);
    for (int it = 0; it < 100000; it++) {
        double U = st.next() - 0.5;
        double V = st.next();
        double us = 0.5 - fabs(U);
        double k = floor((2 * a / us + b) * U + lam + 0.43);
        if ((us >= 0.07) && (V <= vr)) return k;
        if ((k < 0) || ((us < 0.013) && (V > us))) continue;
        if ((log(V) + log(invalpha) - log(a / (us * us) + b)) <= (-lam + k * loglam - lgamma(k + 1.0))) return k;
    }
    return floor(lam);   // unreachable in practice; keeps the loop bounded
}

// Normal(lam, sqrt(lam)) by Box-Muller, not floored (cyutils.pyx:1328-1332).
__device__ __forceinline__ double normal_count(Stream& st, double lam) {
    double u1 = st.next(), u2 = st.next();
    double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    return lam + sqrt(lam) * z;
}

}  // namespace hd
