// gen.h -- seeded, counter-based R-MAT used identically by the host
// generator (gen.cpp) and the device generator (gen_device.hip).
//
// Line e (0 <= e < n_draws) of R-MAT(scale, a, b, c): for each level l from
// the top bit down, a 32-bit draw r picks the quadrant
//   r < t1 -> (0,0), r < t2 -> (0,1), r < t3 -> (1,0), else (1,1)
// with t1 = a*2^32, t2 = (a+b)*2^32, t3 = (a+b+c)*2^32.  Draws for levels
// (2j, 2j+1) are the two halves of splitmix64(base ^ (32 e + j)).  Node ids
// are then passed through a seeded bijection of [0, 2^scale).  Clones: node x
// is a clone iff q(x) < n_clones for the affine bijection q; its lines are
// replaced by the lines of y = q^-1(n_clones + h(x) mod (N - n_clones)).
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RBSC_HD __host__ __device__ __forceinline__
#else
#define RBSC_HD static inline
#endif

struct RmatSpec {
    uint32_t scale;
    uint64_t n_draws;
    uint32_t t1, t2, t3;
    uint64_t base;          // per-seed stream base
    uint64_t mask;          // 2^scale - 1
    uint64_t m1, c1, m2, c2;  // node-id bijection
    uint64_t qm, qc, qm_inv;  // clone-selection bijection
    uint64_t n_clones;
    uint64_t clone_seed;
};

RBSC_HD uint64_t gen_mix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

RBSC_HD uint64_t rmat_perm(const RmatSpec& s, uint64_t x) {
    const unsigned sh = (s.scale + 1) / 2;
    x = (x * s.m1 + s.c1) & s.mask;
    x ^= x >> sh;
    x = (x * s.m2 + s.c2) & s.mask;
    x ^= x >> sh;
    return x;
}

// (src << 32) | dst of draw e, after the node-id bijection.
RBSC_HD uint64_t rmat_draw(const RmatSpec& s, uint64_t e) {
    uint64_t u = 0, v = 0;
    for (uint32_t j = 0; 2 * j < s.scale; ++j) {
        const uint64_t h = gen_mix64(s.base ^ (e * 32 + j));
        for (uint32_t half = 0; half < 2; ++half) {
            const uint32_t l = 2 * j + half;
            if (l >= s.scale) break;
            const uint32_t r = half ? (uint32_t)(h >> 32) : (uint32_t)h;
            const uint32_t bu = r >= s.t2;                       // quadrants c, d
            const uint32_t bv = (r >= s.t1 && r < s.t2) || r >= s.t3;  // quadrants b, d
            const uint32_t bit = s.scale - 1 - l;
            u |= (uint64_t)bu << bit;
            v |= (uint64_t)bv << bit;
        }
    }
    return (rmat_perm(s, u) << 32) | rmat_perm(s, v);
}

RBSC_HD bool rmat_is_clone(const RmatSpec& s, uint64_t x) {
    return ((x * s.qm + s.qc) & s.mask) < s.n_clones;
}

// Node whose lines node x carries (x itself unless x is a clone).
RBSC_HD uint64_t rmat_source(const RmatSpec& s, uint64_t x) {
    if (!rmat_is_clone(s, x)) return x;
    const uint64_t rest = (s.mask + 1) - s.n_clones;
    const uint64_t j = gen_mix64(s.clone_seed ^ x) % rest;
    return ((s.n_clones + j - s.qc) * s.qm_inv) & s.mask;
}

static inline uint64_t gen_inv_odd(uint64_t m) {
    uint64_t x = m;  // Newton iteration for m^-1 mod 2^64
    for (int i = 0; i < 6; ++i) x *= 2 - m * x;
    return x;
}

static inline uint32_t gen_threshold(double p) {
    double t = floor(p * 4294967296.0 + 0.5);
    if (t < 0) t = 0;
    if (t > 4294967295.0) t = 4294967295.0;
    return (uint32_t)t;
}

static inline RmatSpec rmat_spec(uint32_t scale, uint64_t n_draws, double a, double b, double c,
                                 double clone_frac, uint64_t seed) {
    RmatSpec s;
    s.scale = scale;
    s.n_draws = n_draws;
    s.t1 = gen_threshold(a);
    s.t2 = gen_threshold(a + b);
    s.t3 = gen_threshold(a + b + c);
    s.base = gen_mix64(seed ^ 0x5851f42d4c957f2dULL);
    s.mask = (scale >= 64) ? ~0ULL : ((1ULL << scale) - 1);
    s.m1 = gen_mix64(seed + 1) | 1;
    s.c1 = gen_mix64(seed + 2);
    s.m2 = gen_mix64(seed + 3) | 1;
    s.c2 = gen_mix64(seed + 4);
    s.qm = gen_mix64(seed + 5) | 1;
    s.qc = gen_mix64(seed + 6) & s.mask;
    s.qm_inv = gen_inv_odd(s.qm);
    const double nn = (double)(s.mask + 1);
    uint64_t nc = (uint64_t)floor(clone_frac * nn);
    if (nc >= s.mask + 1) nc = s.mask;  // keep at least one source
    s.n_clones = nc;
    s.clone_seed = gen_mix64(seed ^ 0xc10e5eedULL);
    return s;
}
