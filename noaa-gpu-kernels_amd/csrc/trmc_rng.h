// Counter-based random numbers and deterministic log/exp from basic IEEE
// operations, shared by the host generators (trmc_gen.cpp) and the device
// forcing generator (trmc_kernels.hip): both sides produce the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#ifndef TRMC_HD
#define TRMC_HD __host__ __device__ __forceinline__
#endif

namespace trmc {
namespace rng {

TRMC_HD uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

TRMC_HD uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    return mix64(seed * 0x2545F4914F6CDD1Dull ^ mix64(stream + 0x1234567ull));
}

// uniform in (0, 1): 53 random bits, never exactly 0 or 1
TRMC_HD double uniform_at(uint64_t key, uint64_t i)
{
    const uint64_t x = mix64(key ^ mix64(i));
    return ((double)(x >> 11) + 0.5) * 0x1.0p-53;
}

// natural log from basic operations: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^21.
TRMC_HD double det_log(double x)
{
    int e = 0;
    double m = frexp(x, &e);                 // m in [0.5, 1)
    if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double term = s, sum = 0.0;
    for (int k = 1; k <= 21; k += 2) { sum += term / (double)k; term *= s2; }
    return 2.0 * sum + (double)e * 0.69314718055994530942;
}

// exp from basic operations: x = k ln2 + r, |r| <= ln2/2, Taylor to r^17.
TRMC_HD double det_exp(double x)
{
    if (x < -700.0) return 0.0;
    if (x > 700.0) return HUGE_VAL;
    const double kf = floor(x / 0.69314718055994530942 + 0.5);
    const double r = (x - kf * 0.693145751953125) - kf * 1.42860682030941723212e-6;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k <= 17; ++k) { term *= r / (double)k; sum += term; }
    return ldexp(sum, (int)kf);
}

// standard normal, Marsaglia polar method on sub-draws of one counter
TRMC_HD double normal_at(uint64_t key, uint64_t i)
{
    const uint64_t sub = mix64(key + 0x51ED27ull) ^ mix64(i);
    for (uint64_t a = 0;; ++a) {
        const double u = 2.0 * uniform_at(sub, 2 * a) - 1.0;
        const double v = 2.0 * uniform_at(sub, 2 * a + 1) - 1.0;
        const double s = u * u + v * v;
        if (s > 0.0 && s < 1.0) return u * sqrt(-2.0 * det_log(s) / s);
    }
}

TRMC_HD double f32r(double x) { return (double)(float)x; }

// Lateral inflow draw j (= t * n_segments + segment) of member m:
// lognormal with mean 0.1 m3/s and sigma_ln 1 (mu = ln 0.1 - 1/2), rounded
// to float.
TRMC_HD double forcing_at(uint64_t seed, int m, uint64_t j)
{
    const double mu = det_log(0.1) - 0.5;
    return f32r(det_exp(mu + normal_at(stream_key(seed, 1000 + (uint64_t)m), j)));
}

}  // namespace rng
}  // namespace trmc
