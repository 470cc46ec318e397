// oracle_numerics.h -- numerics of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: nothing in the product may include this file.
//
// The reference is built with nvcc --use_fast_math (CMakeLists.txt:26), so its
// exp/sqrt/divide results are NVIDIA-specific and unpinnable.  The oracle and the
// HIP path therefore agree on one explicit numerics contract instead, each
// restating it in its own code:
//   * IEEE fp32 arithmetic, correctly rounded divide and sqrt, no
//     contraction (-ffp-contract=off); fused multiply-adds only where written
//     as fmaf (NCC tap loop, homography projection, bilinear lerp, polynomials).
//   * exp: Cody-Waite reduction by ln2 with a 16-bit high part, Taylor
//     polynomial (degree 7 in float, 13 in double), exact power-of-two scaling.
//   * sin/cos: reduction by pi/2 (three-part Cody-Waite) and Taylor
//     polynomials to degree 9/10.
//   * rsqrtf(x) (DPE.cu:275) is restated as 1.0f / sqrtf(x).
//   * curand XORWOW (DPE.cu:1036) is replaced by stateless Philox4x32-10
//     (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
//     SC'11), keyed by the 64-bit seed; the counter is (draw block, pixel, phase,
//     call).  curand_uniform -> ((x >> 8) + 1) * 2^-24 in (0, 1].
#ifndef DPE_ORACLE_NUMERICS_H_
#define DPE_ORACLE_NUMERICS_H_

#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

inline float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline float o_expf(float x) {
    if (x != x) return x;
    if (!(x > -87.0f)) return 0.0f;
    if (x > 88.0f) return INFINITY;
    const float k = std::rint(x * 1.44269502162933349609375f);
    float r = std::fma(-k, 0.693145751953125f, x);
    r = std::fma(-k, 1.428606820309417232e-06f, r);
    float p = 1.0f / 5040.0f;
    p = std::fma(p, r, 1.0f / 720.0f);
    p = std::fma(p, r, 1.0f / 120.0f);
    p = std::fma(p, r, 1.0f / 24.0f);
    p = std::fma(p, r, 1.0f / 6.0f);
    p = std::fma(p, r, 0.5f);
    p = std::fma(p, r, 1.0f);
    p = std::fma(p, r, 1.0f);
    const int ki = (int)k;
    return p * bits_to_float((uint32_t)(ki + 127) << 23);
}

inline double o_exp(double x) {
    if (x != x) return x;
    if (!(x > -708.0)) return 0.0;
    if (x > 709.0) return INFINITY;
    const double k = std::rint(x * 1.4426950408889634074);
    double r = std::fma(-k, 6.93147180369123816490e-01, x);
    r = std::fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.0 / 6227020800.0;  // 1/13!
    const double inv_fact[13] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                 1.0 / 362880.0,    1.0 / 40320.0,    1.0 / 5040.0,
                                 1.0 / 720.0,       1.0 / 120.0,      1.0 / 24.0,
                                 1.0 / 6.0,         0.5,              1.0, 1.0};
    for (int i = 0; i < 13; ++i) p = std::fma(p, r, inv_fact[i]);
    const int64_t ki = (int64_t)k;
    uint64_t bits = (uint64_t)(ki + 1023) << 52;
    double scale;
    std::memcpy(&scale, &bits, 8);
    return p * scale;
}

// sin and cos of x in float.  Reduction x = q*(pi/2) + r with pi/2 split in three
// float parts (each exact times a small integer q), then Taylor polynomials in r.
inline void o_sincosf(float x, float *s, float *c) {
    const float q = std::rint(x * 0.636619746685028076171875f);
    float r = std::fma(-q, 1.5703125f, x);
    r = std::fma(-q, 4.83751296997070312500e-04f, r);
    r = std::fma(-q, 7.549790126404332e-08f, r);
    const float r2 = r * r;
    float ps = 1.0f / 362880.0f;
    ps = std::fma(ps, r2, -1.0f / 5040.0f);
    ps = std::fma(ps, r2, 1.0f / 120.0f);
    ps = std::fma(ps, r2, -1.0f / 6.0f);
    ps = std::fma(ps * r2, r, r);
    float pc = -1.0f / 3628800.0f;
    pc = std::fma(pc, r2, 1.0f / 40320.0f);
    pc = std::fma(pc, r2, -1.0f / 720.0f);
    pc = std::fma(pc, r2, 1.0f / 24.0f);
    pc = std::fma(pc, r2, -0.5f);
    pc = std::fma(pc, r2, 1.0f);
    const int qi = ((int)q) & 3;
    float sv, cv;
    switch (qi) {
        case 0: sv = ps; cv = pc; break;
        case 1: sv = pc; cv = -ps; break;
        case 2: sv = -ps; cv = -pc; break;
        default: sv = -pc; cv = ps; break;
    }
    *s = sv;
    *c = cv;
}

// acos in float: acos(x) = 2 asin(sqrt((1-x)/2)) for x >= 0.5, pi - (the same
// for -x) for x <= -0.5, pi/2 - asin(x) otherwise; asin by its Taylor series
// sum_k (2k)! / (4^k (k!)^2 (2k+1)) z^(2k+1) to z^23 (|z| <= 0.5).
inline float o_asinf_small(float z) {
    static const float c[12] = {1.0f, 1.0f / 6.0f, 3.0f / 40.0f, 15.0f / 336.0f, 105.0f / 3456.0f,
                                945.0f / 42240.0f, 10395.0f / 599040.0f, 135135.0f / 9676800.0f,
                                2027025.0f / 175472640.0f, 34459425.0f / 3530096640.0f,
                                654729075.0f / 78033715200.0f, 13749310575.0f / 1880240947200.0f};
    const float z2 = z * z;
    float p = c[11];
    for (int i = 10; i >= 0; --i) p = std::fma(p, z2, c[i]);
    return p * z;
}
inline float o_acosf(float x) {
    if (x != x || x > 1.0f || x < -1.0f) return NAN;
    if (x >= 0.5f) return 2.0f * o_asinf_small(std::sqrt((1.0f - x) * 0.5f));
    if (x <= -0.5f) return 3.14159274101257324f - 2.0f * o_asinf_small(std::sqrt((1.0f + x) * 0.5f));
    return 1.57079637050628662f - o_asinf_small(x);
}

// Philox4x32-10 block function.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Random stream of one pixel in one phase of one process call.
struct Rng {
    uint32_t key[2];
    uint32_t pixel, phase, call;
    uint32_t n = 0;
    uint32_t blk[4];
    uint32_t cur = 0xFFFFFFFFu;
    Rng(uint64_t seed, uint32_t pixel_, uint32_t phase_, uint32_t call_)
        : pixel(pixel_), phase(phase_), call(call_) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
    }
    // test access only: when set, draws come from this array (cyclically)
    // instead of the Philox stream
    const uint32_t *replay = nullptr;
    uint32_t replay_n = 0;
    uint32_t next() {
        if (replay) return replay[(n++) % replay_n];
        const uint32_t b = n >> 2;
        if (b != cur) {
            const uint32_t ctr[4] = {b, pixel, phase, call};
            philox4x32_10(ctr, key, blk);
            cur = b;
        }
        return blk[(n++) & 3];
    }
    float uniform() { return (float)((next() >> 8) + 1u) * 5.9604644775390625e-08f; }
};

// Phase identifiers of the random streams within one process call.
enum : uint32_t {
    PH_GEN_NEIGHBOURS = 1,
    PH_RANDOM_INIT = 2,
};
inline uint32_t ph_strong(int it) { return 8u + 4u * (uint32_t)it; }
inline uint32_t ph_ransac(int it) { return 9u + 4u * (uint32_t)it; }
inline uint32_t ph_weak(int it) { return 10u + 4u * (uint32_t)it; }

}  // namespace oracle
#endif
