// pa_numeric.hpp -- numerics shared by the host table builder and the gfx950 kernels.
//
// Everything here is `__host__ __device__` so the host (table construction, pose checks) and
// the device compute identical FP64 bits: each function disables multiply-add contraction
// and uses fma only where written, so IEEE-754 correctly rounded +, -, *, /, sqrt and fma
// give the same result on the CPU and on gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pa {

// src/constants.jl:3-4 (CODATA 2014 values used by kUPS).
constexpr double KB = 1.38064852e-23 / 1.6021766208e-19;
constexpr double KE = 14.399645351950548;
constexpr double PI = 3.141592653589793;
constexpr double TWO_PI = 6.283185307179586;
constexpr int MAX_SITES = 16;
// src/constants.jl:10: one pascal in eV/Å^3.
constexpr double PASCAL_EV = 1.0 / (1.0e30 * 1.6021766208e-19);

// ---------------------------------------------------------------------------------------
// Counter-based RNG (src/state.jl:225-230, src/rng.jl:14-81): a key folded by SplitMix64,
// word i = splitmix64(key, i).
// ---------------------------------------------------------------------------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + index * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Chain {
    uint64_t key;
    uint64_t draw;
    __host__ __device__ Chain(uint64_t chain_seed, uint64_t cycle, uint64_t move)
        : key(splitmix64(splitmix64(chain_seed, cycle), move)), draw(0) {}
    __host__ __device__ uint64_t next() { return splitmix64(key, draw++); }
};

// unit_from_bits(raw, Float64): top 53 bits scaled into [0, 1) (src/rng.jl:70).
__host__ __device__ inline double unit_f64(uint64_t raw) { return (double)(raw >> 11) * 1.1102230246251565e-16; }

// sin(pi x), cos(pi x): exact reduction by the nearest half-integer, then Taylor series of
// sin(pi r), cos(pi r) for |r| <= 1/4 in Horner form with fused multiply-adds. The
// coefficients are (-1)^k pi^(2k+1)/(2k+1)! and (-1)^k pi^(2k)/(2k)! rounded to double.
__host__ __device__ inline void sincospi(double x, double *s_out, double *c_out)
{
#pragma clang fp contract(off)
    const double S[9] = {0x1.921fb54442d18p+1, -0x1.4abbce625be53p+2, 0x1.466bc6775aae2p+1,
                         -0x1.32d2cce62bd86p-1, 0x1.50783487ee782p-4, -0x1.e3074fde8871fp-8,
                         0x1.e8f434d018d63p-12, -0x1.6fadb9f155744p-16, 0x1.aaec32af93359p-21};
    const double Cc[10] = {0x1.0000000000000p+0, -0x1.3bd3cc9be45dep+2, 0x1.03c1f081b5ac4p+2,
                           -0x1.55d3c7e3cbffap+0, 0x1.e1f506891babbp-3, -0x1.a6d1f2a204a8cp-6,
                           0x1.f9d38a3763cc3p-10, -0x1.b6e24f44b128fp-14, 0x1.20c62c2f2d7f5p-18,
                           -0x1.2a0c591af8314p-23};
    double n = __builtin_rint(2.0 * x);
    double r = x - 0.5 * n;
    double r2 = r * r;
    double ps = S[8];
    for (int k = 7; k >= 0; --k) ps = __builtin_fma(ps, r2, S[k]);
    double s = r * ps;
    double pc = Cc[9];
    for (int k = 8; k >= 0; --k) pc = __builtin_fma(pc, r2, Cc[k]);
    double c = pc;
    int q = (int)((int64_t)n & 3);
    double so, co;
    if (q == 0) { so = s; co = c; }
    else if (q == 1) { so = c; co = -s; }
    else if (q == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *s_out = so;
    *c_out = co;
}

// Shoemake (1992) uniform unit quaternion, (x, y, z, w) order (src/rng.jl:126-129).
__host__ __device__ inline void shoemake(double u1, double u2, double u3, double q[4])
{
#pragma clang fp contract(off)
    double a = __builtin_sqrt(1.0 - u1), b = __builtin_sqrt(u1);
    double s2, c2, s3, c3;
    sincospi(2.0 * u2, &s2, &c2);
    sincospi(2.0 * u3, &s3, &c3);
    q[0] = a * c2;
    q[1] = b * s3;
    q[2] = b * c3;
    q[3] = a * s2;
}

// Column-major 3x3 times a vector, summed left to right.
__host__ __device__ inline void matvec(const double A[9], const double v[3], double out[3])
{
#pragma clang fp contract(off)
    for (int i = 0; i < 3; ++i) out[i] = A[i] * v[0] + A[i + 3] * v[1] + A[i + 6] * v[2];
}

// ---------------------------------------------------------------------------------------
// Draws of the move path (src/rng.jl:14-24,93-141; src/moves.jl:81-161)
// ---------------------------------------------------------------------------------------
// High 64 bits of the 128-bit product, from 32-bit halves (exact).
__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b)
{
    const uint64_t a0 = a & 0xffffffffULL, a1 = a >> 32, b0 = b & 0xffffffffULL, b1 = b >> 32;
    const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffULL) + (p10 & 0xffffffffULL);
    return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// rand_range: uniform integer in 1..n by Lemire's multiply-shift map (src/rng.jl:93-97).
__host__ __device__ inline int64_t rand_range(Chain &c, int64_t n) { return (int64_t)mulhi64(c.next(), (uint64_t)n) + 1; }

// rand_normal: Box-Muller cosine branch on u1, u2 (src/rng.jl:109-115); cos(2 pi u2) as
// cospi(2 u2).
__host__ __device__ inline double rand_normal(Chain &c)
{
#pragma clang fp contract(off)
    const double u1 = unit_f64(c.next());
    const double u2 = unit_f64(c.next());
    const double r = __builtin_sqrt(-2.0 * log(1.0 - u1));
    double s, co;
    sincospi(2.0 * u2, &s, &co);
    return r * co;
}

// rand_quaternion (src/rng.jl:136-141).
__host__ __device__ inline void rand_quaternion(Chain &c, double q[4])
{
    const double u1 = unit_f64(c.next());
    const double u2 = unit_f64(c.next());
    const double u3 = unit_f64(c.next());
    shoemake(u1, u2, u3, q);
}

// qmul(q2, q1): apply q1, then q2, (x, y, z, w) order (src/moves.jl:81-89).
__host__ __device__ inline void qmul(const double q2[4], const double q1[4], double out[4])
{
#pragma clang fp contract(off)
    const double x1 = q1[0], y1 = q1[1], z1 = q1[2], w1 = q1[3];
    const double x2 = q2[0], y2 = q2[1], z2 = q2[2], w2 = q2[3];
    out[3] = w2 * w1 - (x2 * x1 + y2 * y1 + z2 * z1);
    out[0] = w2 * x1 + w1 * x2 + (y2 * z1 - z2 * y1);
    out[1] = w2 * y1 + w1 * y2 + (z2 * x1 - x2 * z1);
    out[2] = w2 * z1 + w1 * z2 + (x2 * y1 - y2 * x1);
}

// quaternion_pow: the same axis with the rotation angle scaled by t (src/moves.jl:102-110).
__host__ __device__ inline void quaternion_pow(const double q[4], double t, double out[4])
{
#pragma clang fp contract(off)
    const double w = q[3] < -1.0 ? -1.0 : (q[3] > 1.0 ? 1.0 : q[3]);
    const double theta = 2.0 * acos(w);
    const double s2 = 1.0 - w * w;
    const double s = __builtin_sqrt(s2 > 0.0 ? s2 : 0.0);
    double ax[3] = {1.0, 0.0, 0.0};
    if (s > 2.220446049250313e-16) {
        ax[0] = q[0] / s;
        ax[1] = q[1] / s;
        ax[2] = q[2] / s;
    }
    const double half = t * theta / 2.0;
    const double sh = sin(half);
    out[0] = ax[0] * sh;
    out[1] = ax[1] * sh;
    out[2] = ax[2] * sh;
    out[3] = cos(half);
}

__host__ __device__ inline void qnormalize(double q[4])
{
#pragma clang fp contract(off)
    const double nrm = __builtin_sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] = q[i] / nrm;
}

// Widom pose: fractional position f (3 uniforms) and quaternion (3 uniforms) drawn from
// ChainRNG(splitmix64(seed, sys+1), j+1, 0) -- the reference's `propose_reinsertion`
// placement (src/moves.jl:154-161) under SystemState's chain seeds (src/state.jl:434).
__host__ __device__ inline void pose_uniforms(uint64_t seed, int64_t sys, int64_t j, double f[3], double q[4])
{
#pragma clang fp contract(off)
    Chain c(splitmix64(seed, (uint64_t)(sys + 1)), (uint64_t)(j + 1), 0);
    f[0] = unit_f64(c.next());
    f[1] = unit_f64(c.next());
    f[2] = unit_f64(c.next());
    double u1 = unit_f64(c.next());
    double u2 = unit_f64(c.next());
    double u3 = unit_f64(c.next());
    shoemake(u1, u2, u3, q);
}

// src/energy.jl:13-16: v + 2 u x (u x v + w v).
__host__ __device__ inline void rotate(const double q[4], const double v[3], double out[3])
{
#pragma clang fp contract(off)
    double u0 = q[0], u1 = q[1], u2 = q[2], w = q[3];
    double t0 = (u1 * v[2] - u2 * v[1]) + w * v[0];
    double t1 = (u2 * v[0] - u0 * v[2]) + w * v[1];
    double t2 = (u0 * v[1] - u1 * v[0]) + w * v[2];
    out[0] = v[0] + 2.0 * (u1 * t2 - u2 * t1);
    out[1] = v[1] + 2.0 * (u2 * t0 - u0 * t2);
    out[2] = v[2] + 2.0 * (u0 * t1 - u1 * t0);
}

// src/cell.jl:34-37 minimum image, round = nearest-even (rint).
__host__ __device__ inline void minimum_image(const double A[9], const double invA[9], const double d[3], double out[3])
{
#pragma clang fp contract(off)
    double f[3], r[3], s[3];
    matvec(invA, d, f);
    for (int i = 0; i < 3; ++i) r[i] = __builtin_rint(f[i]);
    matvec(A, r, s);
    for (int i = 0; i < 3; ++i) out[i] = d[i] - s[i];
}

// 32-bit fixed-point fractional coordinate: round(frac(f) * 2^32) mod 2^32. The difference of
// two such values, read as a signed 32-bit integer, is the nearest-image fractional offset
// (the minimum image of src/cell.jl:34-37 up to exact ties).
__host__ __device__ inline uint32_t to_fixed(double f)
{
#pragma clang fp contract(off)
    double t = f - __builtin_floor(f);
    double v = __builtin_rint(t * 4294967296.0);
    return (uint32_t)(uint64_t)(int64_t)v;
}

// pair_erfc_dev (src/ewald.jl:88-130): erfc on z = alpha*r in [0, 4] by a Chebyshev series
// of log(erfc(z)/t) + z^2 in the rescaled variable of t = 2/(2+z). The coefficient values
// are the reference's own literals (src/ewald.jl:103-113), regenerated and checked by the
// test suite from the construction the reference documents.
constexpr double PAIR_ERFC_XMAX = 4.0;
constexpr double PAIR_ERFC_XLO = 2.0 * (2.0 / (2.0 + PAIR_ERFC_XMAX)) - 1.0;

// FUSED = true evaluates the recurrence with fma (one rounding per step instead of two); the
// move path and the Widom FP64 path use FUSED = false.
template <bool FUSED>
__host__ __device__ inline double pair_erfc_f64_t(double z)
{
#pragma clang fp contract(off)
    const double c[17] = {-0.888894912038317, 0.4479385778277769, -0.00021608983881297707,
                          -0.003410058708799526, 8.555152981629436e-5, 5.532830637302869e-5,
                          -5.2098383667087075e-6, -8.380063246741613e-7, 1.9764068422519322e-7,
                          1.7838566371819855e-9, -5.26989861621132e-9, 5.378177355603348e-10,
                          7.748359588558286e-11, -2.3770397709020423e-11, 8.934868696420547e-13,
                          5.207576546902621e-13, -9.42604840757663e-14};
    double t = 2.0 / (2.0 + z);
    double xi = 2.0 * (2.0 * t - 1.0 - PAIR_ERFC_XLO) / (1.0 - PAIR_ERFC_XLO) - 1.0;
    double d = 0.0, dd = 0.0;
    if (FUSED) {
        const double x2 = 2.0 * xi;
#pragma unroll
        for (int j = 16; j >= 1; --j) {
            double tmp = d;
            d = __builtin_fma(x2, d, c[j] - dd);
            dd = tmp;
        }
        return t * exp(__builtin_fma(xi, d, c[0] / 2.0 - dd) - z * z);
    }
    for (int j = 16; j >= 1; --j) {
        double tmp = d;
        d = 2.0 * xi * d - dd + c[j];
        dd = tmp;
    }
    return t * exp(xi * d - dd + c[0] / 2.0 - z * z);
}

__host__ __device__ inline double pair_erfc_f64(double z) { return pair_erfc_f64_t<false>(z); }

// The reference's Float32 coefficient set (src/ewald.jl:120-123).
__host__ __device__ constexpr float pair_erfc_c32(int j)
{
    return j == 0 ? -0.8888949f
         : j == 1 ? 0.4479386f
         : j == 2 ? -0.00021608984f
         : j == 3 ? -0.0034100588f
         : j == 4 ? 8.555145e-5f
         : j == 5 ? 5.5327768e-5f
         : j == 6 ? -5.2045684e-6f
                  : -8.397902e-7f;
}

// The same Float32 series as an ordinary polynomial in xi, coefficient k of
// c0/2 + sum_{j=1..7} c_j T_j(xi) (T_j the Chebyshev polynomials, expanded by
// T_{j+1} = 2 xi T_j - T_{j-1} in FP64 at compile time). Evaluated by Horner in FP32 it takes
// 7 FMAs instead of Clenshaw's 14 operations; over xi in [-1, 1] both forms stay within 1e-7 of
// the FP64 value of the series.
__host__ __device__ constexpr double pair_erfc_mono32_d(int k)
{
    double T[8][8] = {};
    T[0][0] = 1.0;
    T[1][1] = 1.0;
    for (int j = 2; j < 8; ++j)
        for (int i = 0; i < 8; ++i) T[j][i] = (i > 0 ? 2.0 * T[j - 1][i - 1] : 0.0) - T[j - 2][i];
    double v = 0.5 * (double)pair_erfc_c32(0) * T[0][k];
    for (int j = 1; j < 8; ++j) v += (double)pair_erfc_c32(j) * T[j][k];
    return v;
}

__host__ __device__ constexpr float pair_erfc_mono32(int k) { return (float)pair_erfc_mono32_d(k); }

// log(erfc(z)/t) + z^2 of the Float32 series at xi (Horner form of pair_erfc_c32).
__host__ __device__ inline float pair_erfc_poly32(float xi)
{
    float p = pair_erfc_mono32(7);
#pragma unroll
    for (int k = 6; k >= 0; --k) p = p * xi + pair_erfc_mono32(k);
    return p;
}

// The same polynomial times log2(e), for an exp2 of the result.
__host__ __device__ inline float pair_erfc_poly32_log2e(float xi)
{
    constexpr double L2E = 1.4426950408889634;
    float p = (float)(pair_erfc_mono32_d(7) * L2E);
#pragma unroll
    for (int k = 6; k >= 0; --k) p = p * xi + (float)(pair_erfc_mono32_d(k) * L2E);
    return p;
}

}  // namespace pa
