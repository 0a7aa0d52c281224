/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 CPU restatement of PureAdsorb.jl's Widom insertion path, written from the
 * reference's Julia source (file:line citations below point into the reference tree).
 * It is the checker for the HIP product path: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and never as the thing measured or shipped.
 * The product library never links, loads or calls anything in this directory.
 *
 * Pinning (what fixes each function independently of this restatement) is listed in
 * tests/test_oracle_*.py; functions without an independent check are marked
 * "parity unpinned" below.
 *
 * Build: oracle/Makefile -> oracle/_build/liboracle.so (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/constants.jl:3-4 */
#define OR_KB (1.38064852e-23 / 1.6021766208e-19)
#define OR_KE 14.399645351950548
#define OR_PI 3.141592653589793
#define OR_TWO_PI 6.283185307179586
#define OR_MAXSITES 16

/* ------------------------------------------------------------------------------------------
 * Counter-based RNG: src/state.jl:225-230 (splitmix64), src/rng.jl:14-24 (mulhi64),
 * src/rng.jl:43-81 (ChainRNG, next_bits, unit_from_bits, rand_uniform),
 * src/rng.jl:93-97 (rand_range).
 * ---------------------------------------------------------------------------------------- */
uint64_t or_splitmix64(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + index * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t or_mulhi64(uint64_t a, uint64_t b)
{
    uint64_t a_lo = a & 0xffffffffULL, a_hi = a >> 32;
    uint64_t b_lo = b & 0xffffffffULL, b_hi = b >> 32;
    uint64_t lo_lo = a_lo * b_lo, hi_lo = a_hi * b_lo, lo_hi = a_lo * b_hi, hi_hi = a_hi * b_hi;
    uint64_t cross = (lo_lo >> 32) + (hi_lo & 0xffffffffULL) + (lo_hi & 0xffffffffULL);
    return hi_hi + (hi_lo >> 32) + (lo_hi >> 32) + (cross >> 32);
}

typedef struct { uint64_t key, draw; } or_chain;

static or_chain or_chain_new(uint64_t chain_seed, uint64_t cycle, uint64_t move)
{
    or_chain c;
    c.key = or_splitmix64(or_splitmix64(chain_seed, cycle), move);
    c.draw = 0;
    return c;
}

static uint64_t or_next_bits(or_chain *c)
{
    uint64_t raw = or_splitmix64(c->key, c->draw);
    c->draw += 1;
    return raw;
}

double or_unit_f64(uint64_t raw) { return (double)(raw >> 11) * 1.1102230246251565e-16; }
float or_unit_f32(uint64_t raw) { return (float)(raw >> 40) * 5.9604645e-8f; }

/* Draws words 0..nwords-1 of ChainRNG(chain_seed, cycle, move) (test hook). */
void or_chain_words(uint64_t chain_seed, uint64_t cycle, uint64_t move, int nwords, uint64_t *out)
{
    or_chain c = or_chain_new(chain_seed, cycle, move);
    for (int i = 0; i < nwords; ++i) out[i] = or_next_bits(&c);
}

/* rand_range: 1-based index in 1:n (src/rng.jl:93-97). */
int64_t or_rand_range(uint64_t raw, int64_t n) { return (int64_t)or_mulhi64(raw, (uint64_t)n) + 1; }

/* ------------------------------------------------------------------------------------------
 * sin(pi x), cos(pi x). The reference uses Julia Base `sinpi`/`cospi` (src/rng.jl:126-129);
 * this build fixes its own contract: exact reduction to r in [-1/4, 1/4] by the nearest
 * half-integer, then Taylor series of sin(pi r)/cos(pi r) in r^2 (coefficients
 * (-1)^k pi^(2k+1)/(2k+1)!, (-1)^k pi^(2k)/(2k)!, rounded from 40-digit values), Horner with
 * fused multiply-add. Pinned against mpmath to <= 1 ulp (tests/test_oracle_rng.py).
 * ULP-level parity with Julia Base is unpinned.
 * ---------------------------------------------------------------------------------------- */
static const double OR_SINPI_C[9] = {
    0x1.921fb54442d18p+1, -0x1.4abbce625be53p+2, 0x1.466bc6775aae2p+1,
    -0x1.32d2cce62bd86p-1, 0x1.50783487ee782p-4, -0x1.e3074fde8871fp-8,
    0x1.e8f434d018d63p-12, -0x1.6fadb9f155744p-16, 0x1.aaec32af93359p-21,
};
static const double OR_COSPI_C[10] = {
    0x1.0000000000000p+0, -0x1.3bd3cc9be45dep+2, 0x1.03c1f081b5ac4p+2,
    -0x1.55d3c7e3cbffap+0, 0x1.e1f506891babbp-3, -0x1.a6d1f2a204a8cp-6,
    0x1.f9d38a3763cc3p-10, -0x1.b6e24f44b128fp-14, 0x1.20c62c2f2d7f5p-18,
    -0x1.2a0c591af8314p-23,
};

void or_sincospi(double x, double *s_out, double *c_out)
{
    double n = rint(2.0 * x);          /* nearest multiple of 1/2, exact */
    double r = x - 0.5 * n;            /* exact: |r| <= 1/4 */
    double r2 = r * r;
    double ps = OR_SINPI_C[8];
    for (int k = 7; k >= 0; --k) ps = fma(ps, r2, OR_SINPI_C[k]);
    double s = r * ps;
    double pc = OR_COSPI_C[9];
    for (int k = 8; k >= 0; --k) pc = fma(pc, r2, OR_COSPI_C[k]);
    double c = pc;
    int64_t q = (int64_t)n & 3;
    double so, co;
    switch (q) {
    case 0: so = s; co = c; break;
    case 1: so = c; co = -s; break;
    case 2: so = -s; co = -c; break;
    default: so = -c; co = s; break;
    }
    *s_out = so;
    *c_out = co;
}

/* Shoemake (1992) quaternion in (x, y, z, w) order: src/rng.jl:126-129. */
void or_shoemake(double u1, double u2, double u3, double q[4])
{
    double a = sqrt(1.0 - u1), b = sqrt(u1);
    double s2, c2, s3, c3;
    or_sincospi(2.0 * u2, &s2, &c2);
    or_sincospi(2.0 * u3, &s3, &c3);
    q[0] = a * c2;
    q[1] = b * s3;
    q[2] = b * c3;
    q[3] = a * s2;
}

/* Column-major 3x3 times vector, summed left to right (StaticArrays order). */
static inline void or_matvec(const double A[9], const double v[3], double out[3])
{
    for (int i = 0; i < 3; ++i) out[i] = A[i] * v[0] + A[i + 3] * v[1] + A[i + 6] * v[2];
}

/* Widom pose of system `sys` (0-based), per-system insertion `j` (0-based): the reference's
 * counter-based rigid-guest placement `propose_reinsertion` (src/moves.jl:154-161) keyed as
 * ChainRNG(splitmix64(seed, sys+1), j+1, 0); chain seeds follow SystemState's
 * rng_seed[n] = splitmix64(seed, n) (src/state.jl:434). */
void or_pose(uint64_t seed, int64_t sys, int64_t j, const double A[9], double pos[3], double q[4])
{
    or_chain c = or_chain_new(or_splitmix64(seed, (uint64_t)(sys + 1)), (uint64_t)(j + 1), 0);
    double f[3];
    for (int i = 0; i < 3; ++i) f[i] = or_unit_f64(or_next_bits(&c));
    or_matvec(A, f, pos);
    double u1 = or_unit_f64(or_next_bits(&c));
    double u2 = or_unit_f64(or_next_bits(&c));
    double u3 = or_unit_f64(or_next_bits(&c));
    or_shoemake(u1, u2, u3, q);
}

/* ------------------------------------------------------------------------------------------
 * Cell geometry: src/cell.jl:2-37.
 * ---------------------------------------------------------------------------------------- */
/* sin/cos of an angle in degrees with exact zeros at multiples of 90 (Julia `sind`/`cosd`
 * semantics at those points; other values are unpinned at the ULP level). */
static void or_sincosd(double deg, double *s, double *c)
{
    double n = rint(deg / 90.0);
    double r = deg - 90.0 * n;   /* exact for the angles a cell uses */
    double rs, rc;
    or_sincospi(r / 180.0, &rs, &rc);
    int64_t q = ((int64_t)n) & 3;
    switch (q) {
    case 0: *s = rs; *c = rc; break;
    case 1: *s = rc; *c = -rs; break;
    case 2: *s = -rs; *c = -rc; break;
    default: *s = -rc; *c = rs; break;
    }
}

int or_cell_matrix(double a, double b, double c, double al, double be, double ga, double A[9])
{
    if (!(a > 0 && b > 0 && c > 0)) return -1;
    double sa, ca, sb, cb, sg, cg;
    or_sincosd(al, &sa, &ca);
    or_sincosd(be, &sb, &cb);
    or_sincosd(ga, &sg, &cg);
    (void)sa; (void)sb;
    if (sg == 0.0) return -2;
    double cy = (ca - cb * cg) / sg;
    double rad = 1.0 - cb * cb - cy * cy;
    if (!(rad > 0)) return -3;
    double cz = sqrt(rad);
    A[0] = a; A[1] = 0; A[2] = 0;
    A[3] = b * cg; A[4] = b * sg; A[5] = 0;
    A[6] = c * cb; A[7] = c * cy; A[8] = c * cz;
    return 0;
}

double or_det3(const double A[9])
{
    /* A[r + 3c] */
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
           A[6] * (A[1] * A[5] - A[4] * A[2]);
}

/* Closed-form 3x3 inverse by cofactors (StaticArrays' `inv` is the reference; ULP-level
 * agreement with it is unpinned, energies are checked at rtol 1e-12). */
void or_inv3(const double A[9], double inv[9])
{
    double d = or_det3(A);
    inv[0] = (A[4] * A[8] - A[7] * A[5]) / d;
    inv[1] = -(A[1] * A[8] - A[7] * A[2]) / d;
    inv[2] = (A[1] * A[5] - A[4] * A[2]) / d;
    inv[3] = -(A[3] * A[8] - A[6] * A[5]) / d;
    inv[4] = (A[0] * A[8] - A[6] * A[2]) / d;
    inv[5] = -(A[0] * A[5] - A[3] * A[2]) / d;
    inv[6] = (A[3] * A[7] - A[6] * A[4]) / d;
    inv[7] = -(A[0] * A[7] - A[6] * A[1]) / d;
    inv[8] = (A[0] * A[4] - A[3] * A[1]) / d;
}

static void or_cross(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static double or_norm3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* src/cell.jl:14-20 */
void or_perpendicular_lengths(const double A[9], double L[3])
{
    double V = fabs(or_det3(A));
    double x[3];
    or_cross(A + 3, A + 6, x); L[0] = V / or_norm3(x);
    or_cross(A + 0, A + 6, x); L[1] = V / or_norm3(x);
    or_cross(A + 0, A + 3, x); L[2] = V / or_norm3(x);
}

/* src/cell.jl:34-37: Delta - A*round(invA*Delta), round = nearest-even. */
void or_minimum_image(const double A[9], const double invA[9], const double d[3], double out[3])
{
    double f[3], r[3], s[3];
    or_matvec(invA, d, f);
    for (int i = 0; i < 3; ++i) r[i] = rint(f[i]);
    or_matvec(A, r, s);
    for (int i = 0; i < 3; ++i) out[i] = d[i] - s[i];
}

/* src/energy.jl:13-16: v + 2 u x (u x v + w v). */
void or_rotate(const double q[4], const double v[3], double out[3])
{
    double u[3] = {q[0], q[1], q[2]}, w = q[3];
    double t[3], uv[3];
    or_cross(u, v, uv);
    for (int i = 0; i < 3; ++i) t[i] = uv[i] + w * v[i];
    double ut[3];
    or_cross(u, t, ut);
    for (int i = 0; i < 3; ++i) out[i] = v[i] + 2.0 * ut[i];
}

/* ------------------------------------------------------------------------------------------
 * erfc series: src/ewald.jl:47-74 (erfc_dev, 28 terms) and src/ewald.jl:88-130
 * (pair_erfc_dev, 17 terms in Float64, 8 in Float32). The coefficient literals below are
 * the reference's (src/ewald.jl:47-56,103-113); tests/test_oracle_erfc.py regenerates them
 * with mpmath from the construction the reference documents and checks them against scipy.
 * ---------------------------------------------------------------------------------------- */
static const double OR_ERFC_COF[28] = {
    -1.3026537197817094, 6.419697923564902e-1, 1.9476473204185836e-2,
    -9.561514786808632e-3, -9.465953444820369e-4, 3.6683949785276145e-4, 4.252332480690777e-5,
    -2.0278578112534242e-5, -1.6242900046470256e-6, 1.3036558355805232e-6, 1.5626441722066142e-8,
    -8.523809591492654e-8, 6.5290544390988515e-9, 5.0593434955514685e-9, -9.913641564930322e-10,
    -2.2736512229318417e-10, 9.646791102014962e-11, 2.394038083065756e-12, -6.886027526536277e-12,
    8.944879271730368e-13, 3.130921408127142e-13, -1.1270822525221733e-13, 3.8108713588528985e-16,
    7.106125536922576e-15, -1.5230878975373303e-15, -9.464871412184462e-17, 1.218683266906788e-16,
    -3.0494735135424696e-17,
};

static double or_erfccheb(double z)
{
    double t = 2.0 / (2.0 + z);
    double ty = 4.0 * t - 2.0;
    double d = 0.0, dd = 0.0;
    for (int j = 27; j >= 1; --j) {
        double tmp = d;
        d = ty * d - dd + OR_ERFC_COF[j];
        dd = tmp;
    }
    return t * exp(-z * z + (ty / 2.0) * d - dd + OR_ERFC_COF[0] / 2.0);
}

double or_erfc_dev(double x) { return x >= 0 ? or_erfccheb(x) : 2.0 - or_erfccheb(-x); }

#define OR_PAIR_ERFC_XMAX 4.0
static const double OR_PAIR_COF64[17] = {
    -0.888894912038317, 0.4479385778277769, -0.00021608983881297707, -0.003410058708799526,
    8.555152981629436e-5, 5.532830637302869e-5, -5.2098383667087075e-6, -8.380063246741613e-7,
    1.9764068422519322e-7, 1.7838566371819855e-9, -5.26989861621132e-9, 5.378177355603348e-10,
    7.748359588558286e-11, -2.3770397709020423e-11, 8.934868696420547e-13, 5.207576546902621e-13,
    -9.42604840757663e-14,
};
static const float OR_PAIR_COF32[8] = {
    -0.8888949f, 0.4479386f, -0.00021608984f, -0.0034100588f,
    8.555145e-5f, 5.5327768e-5f, -5.2045684e-6f, -8.397902e-7f,
};

/* Coefficient literals for the regeneration test: which = 0 (erfc_dev, 28), 1 (pair F64, 17),
 * 2 (pair F32, 8). */
double or_erfc_coef(int which, int j)
{
    if (which == 0 && j >= 0 && j < 28) return OR_ERFC_COF[j];
    if (which == 1 && j >= 0 && j < 17) return OR_PAIR_COF64[j];
    if (which == 2 && j >= 0 && j < 8) return (double)OR_PAIR_COF32[j];
    return NAN;
}

double or_pair_erfc_f64(double z)
{
    const double xlo = 2.0 * (2.0 / (2.0 + OR_PAIR_ERFC_XMAX)) - 1.0;
    double t = 2.0 / (2.0 + z);
    double xi = 2.0 * (2.0 * t - 1.0 - xlo) / (1.0 - xlo) - 1.0;
    double d = 0.0, dd = 0.0;
    for (int j = 16; j >= 1; --j) {
        double tmp = d;
        d = 2.0 * xi * d - dd + OR_PAIR_COF64[j];
        dd = tmp;
    }
    return t * exp(xi * d - dd + OR_PAIR_COF64[0] / 2.0 - z * z);
}

/* Float32 arithmetic throughout (the reference's F32 method, src/ewald.jl:120-130). */
float or_pair_erfc_f32(float z)
{
    const float xlo = (float)(2.0 * (2.0 / (2.0 + OR_PAIR_ERFC_XMAX)) - 1.0);
    float t = 2.0f / (2.0f + z);
    float xi = 2.0f * (2.0f * t - 1.0f - xlo) / (1.0f - xlo) - 1.0f;
    float d = 0.0f, dd = 0.0f;
    for (int j = 7; j >= 1; --j) {
        float tmp = d;
        d = 2.0f * xi * d - dd + OR_PAIR_COF32[j];
        dd = tmp;
    }
    return t * expf(xi * d - dd + OR_PAIR_COF32[0] / 2.0f - z * z);
}

/* ------------------------------------------------------------------------------------------
 * Ewald parameters and tables: src/ewald.jl:21-39,142-170; src/reference.jl:10-16.
 * ---------------------------------------------------------------------------------------- */
/* Bisection on erfc(z) = rc*prec/2 (src/ewald.jl:21-32). The reference uses
 * SpecialFunctions.erfc; this uses C99 erfc. Returns -1 when the precision is unreachable. */
double or_ewald_alpha(double cutoff, double precision)
{
    double target = cutoff * precision / 2.0;
    double lo = 0.0, hi = 20.0;
    if (!(erfc(hi) - target < 0)) return -1.0;
    for (int it = 0; it < 200; ++it) {
        double mid = (lo + hi) / 2.0;
        if (erfc(mid) - target > 0) lo = mid; else hi = mid;
        if (hi - lo < 1.0e-12) break;
    }
    return (lo + hi) / 2.0 / cutoff;
}

double or_ewald_kmax(double alpha, double precision) { return 2.0 * alpha * sqrt(-log(precision / 2.0)); }

/* src/ewald.jl:36-39 */
double or_pk(double k2, double alpha, double V) { return 2.0 * OR_PI / V * exp(-k2 / (4.0 * alpha * alpha)) / k2; }

/* Half-space enumeration (src/ewald.jl:142-157): n1 in 0..n1max outer, n2 then n3 in
 * -n..n, skip 0, keep |k| <= kmax, weight 1 if n1 == 0 else 2. Writes up to `cap` entries
 * and returns the total count (call with cap = 0 to size). */
int or_kvectors(const double A[9], double kmax, int cap, double *ks, double *w, int *coeffs)
{
    double invA[9], B[9];
    or_inv3(A, invA);
    /* B = 2pi * inv(A)', B[i + 3j] = 2pi * invA[j + 3i] */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i + 3 * j] = OR_TWO_PI * invA[j + 3 * i];
    int nmax[3];
    for (int i = 0; i < 3; ++i) nmax[i] = (int)ceil(kmax * or_norm3(A + 3 * i) / OR_TWO_PI);
    int count = 0;
    for (int n1 = 0; n1 <= nmax[0]; ++n1)
        for (int n2 = -nmax[1]; n2 <= nmax[1]; ++n2)
            for (int n3 = -nmax[2]; n3 <= nmax[2]; ++n3) {
                if (n1 == 0 && n2 == 0 && n3 == 0) continue;
                double k[3];
                for (int i = 0; i < 3; ++i) k[i] = B[i] * n1 + B[i + 3] * n2 + B[i + 6] * n3;
                if (!(or_norm3(k) <= kmax)) continue;
                if (count < cap) {
                    ks[3 * count] = k[0]; ks[3 * count + 1] = k[1]; ks[3 * count + 2] = k[2];
                    w[count] = n1 == 0 ? 1.0 : 2.0;
                    coeffs[3 * count] = n1; coeffs[3 * count + 1] = n2; coeffs[3 * count + 2] = n3;
                }
                ++count;
            }
    return count;
}

static double or_dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* src/ewald.jl:159-170 */
void or_structure_factor(int nk, const double *ks, int n, const double *pos, const double *q,
                         double *Sre, double *Sim)
{
    for (int i = 0; i < nk; ++i) {
        double re = 0.0, im = 0.0;
        for (int j = 0; j < n; ++j) {
            double ph = or_dot3(ks + 3 * i, pos + 3 * j);
            re += q[j] * cos(ph);
            im += q[j] * sin(ph);
        }
        Sre[i] = re;
        Sim[i] = im;
    }
}

/* Full Ewald energy of one periodic system (src/ewald.jl:177-203): the oracle of the oracle,
 * pinned by the NaCl Madelung constant and alpha independence. */
double or_ewald_energy(const double A[9], int n, const double *pos, const double *q, const int *mol,
                       double alpha, double cutoff, int nk, const double *ks, const double *w)
{
    double invA[9];
    or_inv3(A, invA);
    double V = fabs(or_det3(A));
    double E_sr = 0.0, E_excl = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
            if (mol[i] == mol[j]) {
                double r = or_norm3(d);
                E_excl -= q[i] * q[j] * (1.0 - or_erfc_dev(alpha * r)) / r;
            } else {
                double m[3];
                or_minimum_image(A, invA, d, m);
                double r = or_norm3(m);
                if (r < cutoff) E_sr += q[i] * q[j] * or_erfc_dev(alpha * r) / r;
            }
        }
    double *Sre = malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
    double *Sim = malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
    or_structure_factor(nk, ks, n, pos, q, Sre, Sim);
    double E_lr = 0.0;
    for (int i = 0; i < nk; ++i)
        E_lr += w[i] * or_pk(or_dot3(ks + 3 * i, ks + 3 * i), alpha, V) * (Sre[i] * Sre[i] + Sim[i] * Sim[i]);
    free(Sre);
    free(Sim);
    double q2 = 0.0, Q = 0.0;
    for (int i = 0; i < n; ++i) { q2 += q[i] * q[i]; Q += q[i]; }
    double E_self = -alpha / sqrt(OR_PI) * q2;
    double E_net = -OR_PI / (2.0 * V * alpha * alpha) * Q * Q;
    return OR_KE * (E_sr + E_lr + E_self + E_excl + E_net);
}

/* ------------------------------------------------------------------------------------------
 * Force-field tail correction: src/forcefield.jl:103-124. sigma/eps are ntypes x ntypes,
 * column-major (index a + ntypes*b).
 * ---------------------------------------------------------------------------------------- */
static double or_tail_coefficient(int ntypes, const double *sigma, const double *eps, double cutoff, int i, int j)
{
    double e = eps[i + ntypes * j];
    if (e == 0.0) return 0.0;
    double s = sigma[i + ntypes * j];
    double x = s / cutoff;
    double x3 = x * x * x;
    return e * (s * s * s) * (x3 * x3 * x3 / 3.0 - x3);
}

double or_tail_delta(int ntypes, const double *sigma, const double *eps, double cutoff, int tail,
                     const int64_t *counts, const int64_t *gcounts, double V)
{
    if (!tail) return 0.0;
    double acc = 0.0;
    for (int a = 0; a < ntypes; ++a)
        for (int b = 0; b < ntypes; ++b) {
            double c = or_tail_coefficient(ntypes, sigma, eps, cutoff, a, b);
            acc += 2.0 * (double)counts[a] * (double)gcounts[b] * c + (double)gcounts[a] * (double)gcounts[b] * c;
        }
    return (8.0 * OR_PI / 3.0) / V * acc;
}

/* ------------------------------------------------------------------------------------------
 * One framework system with the per-framework tables of FrameworkBatch
 * (src/batch.jl:237-432): coupled-k filter, replication check, 64-orientation guest self
 * term, constant_offset. Atoms stay in input (CIF/replicate) order; the reference's cell sort
 * (src/batch.jl:332-335) changes only the summation order.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    double A[9], invA[9], V;
    int natoms;
    double *pos, *q;
    int *type;
    int rep[3];
    int ntypes;
    double *sigma, *eps;
    double rc_lj;
    int tail;
    int nsites;
    double gsite[3 * OR_MAXSITES], gq[OR_MAXSITES];
    int gtype[OR_MAXSITES];
    int charged;
    double rc_ew, prec, alpha, kmax;
    int nkfull;
    double *kfull, *kpref_full, *Sre_full, *Sim_full;
    int *coef_full;
    int nk;
    int *kidx;
    double self_mean, self_halfrange, constant_offset, tail_term, self_term, excl_term, net_term;
    int status;
    char err[512];
} or_system;

/* Self-term orientations: ChainRNG(0x5e1f, 0, o), o = 1..64, three F64 uniforms each.
 * The reference draws its 64 orientations from Xoshiro(0x5e1f) (src/batch.jl:302-303),
 * which cannot be reproduced here; energies then differ from the reference by at most
 * 2*self_term_halfrange (the reference's own tolerance for this approximation). */
void or_self_orientation(int o, double q[4])
{
    or_chain c = or_chain_new(0x5e1fULL, 0, (uint64_t)o);
    double u1 = or_unit_f64(or_next_bits(&c));
    double u2 = or_unit_f64(or_next_bits(&c));
    double u3 = or_unit_f64(or_next_bits(&c));
    or_shoemake(u1, u2, u3, q);
}

static void or_fail(or_system *s, const char *msg)
{
    s->status = -1;
    snprintf(s->err, sizeof s->err, "%s", msg);
}

void or_system_free(or_system *s)
{
    if (!s) return;
    free(s->pos); free(s->q); free(s->type); free(s->sigma); free(s->eps);
    free(s->kfull); free(s->kpref_full); free(s->Sre_full); free(s->Sim_full); free(s->coef_full);
    free(s->kidx);
    free(s);
}

or_system *or_system_new(const double *A, int natoms, const double *frac, const int *type, const double *q,
                         const int *rep, int ntypes, const double *sigma, const double *eps, double rc_lj,
                         int tail, int nsites, const double *gsites, const int *gtypes, const double *gq,
                         double rc_ew, double prec)
{
    or_system *s = calloc(1, sizeof(or_system));
    if (nsites < 1 || nsites > OR_MAXSITES) { or_fail(s, "guest site count out of range"); return s; }
    memcpy(s->A, A, sizeof s->A);
    or_inv3(s->A, s->invA);
    s->V = fabs(or_det3(s->A));
    s->natoms = natoms;
    s->pos = malloc(sizeof(double) * 3 * (size_t)(natoms > 0 ? natoms : 1));
    s->q = malloc(sizeof(double) * (size_t)(natoms > 0 ? natoms : 1));
    s->type = malloc(sizeof(int) * (size_t)(natoms > 0 ? natoms : 1));
    for (int j = 0; j < natoms; ++j) {
        or_matvec(s->A, frac + 3 * j, s->pos + 3 * j);
        s->q[j] = q[j];
        s->type[j] = type[j];
    }
    memcpy(s->rep, rep, sizeof s->rep);
    s->ntypes = ntypes;
    s->sigma = malloc(sizeof(double) * (size_t)ntypes * ntypes);
    s->eps = malloc(sizeof(double) * (size_t)ntypes * ntypes);
    memcpy(s->sigma, sigma, sizeof(double) * (size_t)ntypes * ntypes);
    memcpy(s->eps, eps, sizeof(double) * (size_t)ntypes * ntypes);
    s->rc_lj = rc_lj;
    s->tail = tail;
    s->nsites = nsites;
    s->charged = 0;
    double rguest = 0.0;
    for (int a = 0; a < nsites; ++a) {
        for (int i = 0; i < 3; ++i) s->gsite[3 * a + i] = gsites[3 * a + i];
        s->gtype[a] = gtypes[a];
        s->gq[a] = gq[a];
        if (gq[a] != 0.0) s->charged = 1;
        double r = or_norm3(gsites + 3 * a);
        if (r > rguest) rguest = r;
    }
    s->rc_ew = rc_ew;
    s->prec = prec;
    (void)rguest;
    /* src/batch.jl:289-297 */
    s->alpha = or_ewald_alpha(rc_ew, prec);
    if (s->alpha < 0) { or_fail(s, "cannot reach precision with cutoff"); return s; }
    if (s->alpha * rc_ew > OR_PAIR_ERFC_XMAX) { or_fail(s, "alpha*ewald_cutoff exceeds PAIR_ERFC_XMAX"); return s; }
    s->kmax = or_ewald_kmax(s->alpha, prec);
    int need_k = s->charged || rep[0] != 1 || rep[1] != 1 || rep[2] != 1;
    if (need_k) {
        int nk = or_kvectors(s->A, s->kmax, 0, NULL, NULL, NULL);
        s->nkfull = nk;
        s->kfull = malloc(sizeof(double) * 3 * (size_t)(nk > 0 ? nk : 1));
        double *w = malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
        s->coef_full = malloc(sizeof(int) * 3 * (size_t)(nk > 0 ? nk : 1));
        or_kvectors(s->A, s->kmax, nk, s->kfull, w, s->coef_full);
        s->kpref_full = malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
        for (int i = 0; i < nk; ++i) s->kpref_full[i] = w[i] * or_pk(or_dot3(s->kfull + 3 * i, s->kfull + 3 * i), s->alpha, s->V);
        free(w);
        s->Sre_full = malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
        s->Sim_full = malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
        or_structure_factor(nk, s->kfull, natoms, s->pos, s->q, s->Sre_full, s->Sim_full);
        s->kidx = malloc(sizeof(int) * (size_t)(nk > 0 ? nk : 1));
        s->nk = 0;
        /* src/batch.jl:355-359: coupled iff every coefficient is divisible by the replication */
        for (int i = 0; i < nk; ++i) {
            int c = 1;
            for (int d = 0; d < 3; ++d) if (s->coef_full[3 * i + d] % rep[d] != 0) c = 0;
            if (c && s->charged) s->kidx[s->nk++] = i;
        }
        /* replication check, src/batch.jl:173-195 */
        if (rep[0] != 1 || rep[1] != 1 || rep[2] != 1) {
            int *unc = malloc(sizeof(int) * (size_t)(nk > 0 ? nk : 1));
            int nu = 0;
            for (int i = 0; i < nk; ++i) {
                int c = 1;
                for (int d = 0; d < 3; ++d) if (s->coef_full[3 * i + d] % rep[d] != 0) c = 0;
                if (!c) unc[nu++] = i;
            }
            if (nu > 0) {
                int stride = (nu + 31) / 32;
                if (stride < 1) stride = 1;
                double maxS = 0.0, sabs = 0.0;
                int taken = 0;
                for (int t = 0; t < nu && taken < 32; t += stride, ++taken) {
                    int i = unc[t];
                    double m = hypot(s->Sre_full[i], s->Sim_full[i]);
                    if (m > maxS) maxS = m;
                }
                for (int j = 0; j < natoms; ++j) sabs += fabs(s->q[j]);
                double thr = fmax(1.0e-8, 100.0 * 2.220446049250313e-16) * sabs;
                if (maxS > thr) { free(unc); or_fail(s, "replication does not describe the atoms"); return s; }
            }
            free(unc);
        }
    }
    /* 64-orientation self term, src/batch.jl:360-389 */
    if (s->charged) {
        double samples[64];
        for (int o = 0; o < 64; ++o) {
            double qo[4];
            or_self_orientation(o + 1, qo);
            double gs[3 * OR_MAXSITES];
            for (int a = 0; a < nsites; ++a) or_rotate(qo, s->gsite + 3 * a, gs + 3 * a);
            double acc = 0.0;
            for (int i = 0; i < s->nkfull; ++i) {
                double re = 0.0, im = 0.0;
                for (int a = 0; a < nsites; ++a) {
                    double ph = or_dot3(s->kfull + 3 * i, gs + 3 * a);
                    re += s->gq[a] * cos(ph);
                    im += s->gq[a] * sin(ph);
                }
                acc += s->kpref_full[i] * (re * re + im * im);
            }
            samples[o] = OR_KE * acc;
        }
        double sum = 0.0, mx = samples[0], mn = samples[0];
        for (int o = 0; o < 64; ++o) {
            sum += samples[o];
            if (samples[o] > mx) mx = samples[o];
            if (samples[o] < mn) mn = samples[o];
        }
        s->self_mean = sum / 64.0;
        s->self_halfrange = (mx - mn) / 2.0;
        const double KT_REF = OR_KB * 300.0;
        if (2.0 * s->self_halfrange / KT_REF > 1.0e-3) { or_fail(s, "guest self-term half-range exceeds tolerance"); return s; }
    }
    /* constant_offset, src/batch.jl:396-407 */
    int64_t *counts = calloc((size_t)ntypes, sizeof(int64_t));
    int64_t *gcounts = calloc((size_t)ntypes, sizeof(int64_t));
    for (int j = 0; j < natoms; ++j) counts[type[j]] += 1;
    for (int a = 0; a < nsites; ++a) gcounts[gtypes[a]] += 1;
    s->tail_term = or_tail_delta(ntypes, sigma, eps, rc_lj, tail, counts, gcounts, s->V);
    free(counts);
    free(gcounts);
    double q2 = 0.0;
    for (int a = 0; a < nsites; ++a) q2 += s->gq[a] * s->gq[a];
    s->self_term = -s->alpha / sqrt(OR_PI) * q2;
    double excl = 0.0;
    for (int a = 0; a < nsites; ++a)
        for (int c = a + 1; c < nsites; ++c) {
            double d[3] = {s->gsite[3 * a] - s->gsite[3 * c], s->gsite[3 * a + 1] - s->gsite[3 * c + 1],
                           s->gsite[3 * a + 2] - s->gsite[3 * c + 2]};
            double r = or_norm3(d);
            excl -= s->gq[a] * s->gq[c] * (1.0 - or_erfc_dev(s->alpha * r)) / r;
        }
    s->excl_term = excl;
    double Qh = 0.0, Qg = 0.0;
    for (int j = 0; j < natoms; ++j) Qh += s->q[j];
    for (int a = 0; a < nsites; ++a) Qg += s->gq[a];
    s->net_term = -OR_PI / (2.0 * s->V * s->alpha * s->alpha) * ((Qh + Qg) * (Qh + Qg) - Qh * Qh);
    s->constant_offset = s->tail_term + OR_KE * (s->self_term + s->excl_term + s->net_term) + s->self_mean;
    s->status = 0;
    return s;
}

int or_system_status(const or_system *s, char *buf, int n)
{
    if (buf && n > 0) snprintf(buf, (size_t)n, "%s", s->err);
    return s->status;
}

/* out: V, alpha, kmax, constant_offset, self_mean, self_halfrange, tail, self, excl, net,
 * nk (coupled), nkfull */
void or_system_scalars(const or_system *s, double out[12])
{
    out[0] = s->V; out[1] = s->alpha; out[2] = s->kmax; out[3] = s->constant_offset;
    out[4] = s->self_mean; out[5] = s->self_halfrange; out[6] = s->tail_term; out[7] = s->self_term;
    out[8] = s->excl_term; out[9] = s->net_term; out[10] = s->nk; out[11] = s->nkfull;
}

/* Copies the full (which = 1) or coupled (which = 0) k table. */
void or_system_ktable(const or_system *s, int which, double *ks, double *kpref, double *Sre, double *Sim, int *coef)
{
    int n = which ? s->nkfull : s->nk;
    for (int t = 0; t < n; ++t) {
        int i = which ? t : s->kidx[t];
        for (int d = 0; d < 3; ++d) { ks[3 * t + d] = s->kfull[3 * i + d]; coef[3 * t + d] = s->coef_full[3 * i + d]; }
        kpref[t] = s->kpref_full[i];
        Sre[t] = s->Sre_full[i];
        Sim[t] = s->Sim_full[i];
    }
}

/* ------------------------------------------------------------------------------------------
 * Insertion energy. `or_insertion_energy` is the production contract of FrameworkBatch +
 * insertion_energy (src/energy.jl:42-82): LJ + screened Coulomb with pair_erfc_dev, the
 * reciprocal cross term over the coupled k subset only, the orientation-averaged self term
 * left to constant_offset. Minimum image per (site, atom) pair, the oracle's own semantics
 * (src/reference.jl:39-56), which equals the reference's single image per atom whenever its
 * construction guard holds and stays exact down to perpendicular widths of 2*cutoff.
 * Returns E without constant_offset; *npairs receives the in-cutoff (site, atom) count.
 * ---------------------------------------------------------------------------------------- */
double or_insertion_energy(const or_system *s, const double pos[3], const double q[4], int64_t *npairs)
{
    double rc_lj2 = s->rc_lj * s->rc_lj, rc_ew2 = s->rc_ew * s->rc_ew;
    double gp[3 * OR_MAXSITES];
    for (int a = 0; a < s->nsites; ++a) {
        double r[3];
        or_rotate(q, s->gsite + 3 * a, r);
        for (int i = 0; i < 3; ++i) gp[3 * a + i] = pos[i] + r[i];
    }
    double E_lj = 0.0, E_sr = 0.0;
    int64_t np = 0;
    for (int j = 0; j < s->natoms; ++j) {
        const double *hp = s->pos + 3 * j;
        int ht = s->type[j];
        double hq = s->q[j];
        for (int a = 0; a < s->nsites; ++a) {
            double d[3] = {gp[3 * a] - hp[0], gp[3 * a + 1] - hp[1], gp[3 * a + 2] - hp[2]};
            double m[3];
            or_minimum_image(s->A, s->invA, d, m);
            double r2 = or_dot3(m, m);
            if (!(r2 < rc_lj2 || r2 < rc_ew2)) continue;
            ++np;
            int gt = s->gtype[a];
            if (r2 < rc_lj2) {
                double sg = s->sigma[gt + s->ntypes * ht], ep = s->eps[gt + s->ntypes * ht];
                double x = sg * sg / r2;
                x = x * x * x;
                E_lj += 4.0 * ep * (x * x - x);
            }
            if (r2 < rc_ew2 && s->charged) {
                double r = sqrt(r2);
                E_sr += s->gq[a] * hq * or_pair_erfc_f64(s->alpha * r) / r;
            }
        }
    }
    double E_lr = 0.0;
    for (int t = 0; t < s->nk; ++t) {
        int i = s->kidx[t];
        double re = 0.0, im = 0.0;
        for (int a = 0; a < s->nsites; ++a) {
            double ph = or_dot3(s->kfull + 3 * i, gp + 3 * a);
            re += s->gq[a] * cos(ph);
            im += s->gq[a] * sin(ph);
        }
        E_lr += s->kpref_full[i] * 2.0 * (s->Sre_full[i] * re + s->Sim_full[i] * im);
    }
    if (npairs) *npairs = np;
    return E_lj + OR_KE * (E_sr + E_lr);
}

/* insertion_energy_reference (src/reference.jl:31-67): per-site minimum image, erfc_dev,
 * FULL k table with the cross AND the guest self term. Without constant_offset pieces. */
double or_insertion_energy_reference(const or_system *s, const double pos[3], const double q[4])
{
    double rc_lj2 = s->rc_lj * s->rc_lj, rc_ew2 = s->rc_ew * s->rc_ew;
    double gp[3 * OR_MAXSITES];
    for (int a = 0; a < s->nsites; ++a) {
        double r[3];
        or_rotate(q, s->gsite + 3 * a, r);
        for (int i = 0; i < 3; ++i) gp[3 * a + i] = pos[i] + r[i];
    }
    double E_lj = 0.0, E_sr = 0.0;
    for (int a = 0; a < s->nsites; ++a) {
        int gt = s->gtype[a];
        for (int j = 0; j < s->natoms; ++j) {
            const double *hp = s->pos + 3 * j;
            double d[3] = {gp[3 * a] - hp[0], gp[3 * a + 1] - hp[1], gp[3 * a + 2] - hp[2]};
            double m[3];
            or_minimum_image(s->A, s->invA, d, m);
            double r2 = or_dot3(m, m);
            if (!(r2 < rc_lj2 || r2 < rc_ew2)) continue;
            if (r2 < rc_lj2) {
                int ht = s->type[j];
                double sg = s->sigma[gt + s->ntypes * ht], ep = s->eps[gt + s->ntypes * ht];
                double x = sg * sg / r2;
                x = x * x * x;
                E_lj += 4.0 * ep * (x * x - x);
            }
            if (r2 < rc_ew2) {
                double r = sqrt(r2);
                E_sr += s->gq[a] * s->q[j] * or_erfc_dev(s->alpha * r) / r;
            }
        }
    }
    double E_lr = 0.0;
    for (int i = 0; i < s->nkfull; ++i) {
        double re = 0.0, im = 0.0;
        for (int a = 0; a < s->nsites; ++a) {
            double ph = or_dot3(s->kfull + 3 * i, gp + 3 * a);
            re += s->gq[a] * cos(ph);
            im += s->gq[a] * sin(ph);
        }
        E_lr += s->kpref_full[i] * (2.0 * (s->Sre_full[i] * re + s->Sim_full[i] * im) + (re * re + im * im));
    }
    return E_lj + OR_KE * (E_sr + E_lr);
}

/* Per-insertion Delta U = insertion_energy + constant_offset for insertions j0..j0+n-1 of
 * system `sysidx` under `seed` (the widom_kernel! contract, src/widom.jl:157-175). Optional
 * outputs: poses (7 doubles each: pos, quat) and the in-cutoff pair count per insertion. */
void or_energies(const or_system *s, uint64_t seed, int64_t sysidx, int64_t j0, int64_t n, double *dU,
                 double *poses, int64_t *npairs, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < n; ++t) {
        double pos[3], q[4];
        or_pose(seed, sysidx, j0 + t, s->A, pos, q);
        int64_t np = 0;
        double e = or_insertion_energy(s, pos, q, &np);
        dU[t] = e + s->constant_offset;
        if (poses) {
            for (int i = 0; i < 3; ++i) poses[7 * t + i] = pos[i];
            for (int i = 0; i < 4; ++i) poses[7 * t + 3 + i] = q[i];
        }
        if (npairs) npairs[t] = np;
    }
}

/* ------------------------------------------------------------------------------------------
 * Widom bookkeeping: src/widom.jl:34-53 (system_counts, default_run), :80-84
 * (boltzmann_weight), :394-441 (per-system blocks), :451-492 (_reduce).
 * ---------------------------------------------------------------------------------------- */
int64_t or_default_run(int64_t ninsert, int64_t nsys)
{
    int64_t r = (ninsert / nsys) / 4;
    if (r < 1) r = 1;
    if (r > 256) r = 256;
    return r;
}

void or_system_counts(int64_t ninsert, int64_t nsys, int64_t run, int64_t *ns)
{
    int64_t q = ninsert / run, rem = ninsert % run;
    int64_t full = q / nsys, hi = q % nsys;
    for (int64_t s = 0; s < nsys; ++s) ns[s] = full * run;
    for (int64_t r = 0; r < hi; ++r) ns[r] += run;
    if (rem > 0) ns[q % nsys] += rem;
}

void or_boltzmann_weight(double dU, double kT, double *w, double *uw)
{
    *w = exp(-dU / kT);
    *uw = (*w == 0.0) ? 0.0 : dU * *w;
}

/* Returns 0, or -1 (zero-sample block), -2 (W/UW overflow), -3 (mu_ex or q_st non-finite).
 * out: mu_ex, mu_ex_err, K_H, K_H_err, q_st, q_st_err, nsamples, nblocks. */
int or_reduce(int nb, const double *sW, const double *sUW, const int64_t *n, double kT, double V, double out[8])
{
    int64_t ntot = 0;
    double tW = 0.0, tUW = 0.0;
    for (int b = 0; b < nb; ++b) {
        if (n[b] <= 0) return -1;
        ntot += n[b];
        tW += sW[b];
        tUW += sUW[b];
    }
    double W = tW / (double)ntot, UW = tUW / (double)ntot;
    if (!(isfinite(W) && isfinite(UW))) return -2;
    double varW = 0.0, varUW = 0.0, cov = 0.0;
    for (int b = 0; b < nb; ++b) {
        double mW = sW[b] / (double)n[b], mUW = sUW[b] / (double)n[b];
        varW += (mW - W) * (mW - W);
        varUW += (mUW - UW) * (mUW - UW);
        cov += (mW - W) * (mUW - UW);
    }
    varW /= (nb - 1);
    varUW /= (nb - 1);
    cov /= (nb - 1);
    double semW = sqrt(varW / nb);
    double ratio = UW / W;
    double var_ratio = UW == 0.0 ? 0.0 : ratio * ratio * (varUW / (UW * UW) + varW / (W * W) - 2.0 * cov / (UW * W)) / nb;
    double mu_ex = -kT * log(W);
    double q_st = kT - ratio;
    if (!isfinite(mu_ex) || !isfinite(q_st)) return -3;
    out[0] = mu_ex;
    out[1] = kT * semW / W;
    out[2] = V * W / kT;
    out[3] = V * semW / kT;
    out[4] = q_st;
    out[5] = sqrt(var_ratio > 0.0 ? var_ratio : 0.0);
    out[6] = (double)ntot;
    out[7] = (double)nb;
    return 0;
}

/* Widom for one system of a batch: ns insertions keyed (seed, sysidx, j), blocks of
 * ns/nblocks with the remainder in the last block, FP64 Boltzmann sums in sample order.
 * block_sums: 3*nblocks (sW, sUW, n). Returns or_reduce's status. */
int or_widom(const or_system *s, double T, int64_t ns, uint64_t seed, int64_t sysidx, int nblocks,
             double *block_sums, double out[8], int nthreads)
{
    double kT = OR_KB * T;
    double *dU = malloc(sizeof(double) * (size_t)(ns > 0 ? ns : 1));
    or_energies(s, seed, sysidx, 0, ns, dU, NULL, NULL, nthreads);
    double *sW = calloc((size_t)nblocks, sizeof(double));
    double *sUW = calloc((size_t)nblocks, sizeof(double));
    int64_t *n = calloc((size_t)nblocks, sizeof(int64_t));
    int64_t block_len = ns / nblocks;
    for (int64_t j = 0; j < ns; ++j) {
        int64_t blk = block_len > 0 ? j / block_len : nblocks - 1;
        if (blk > nblocks - 1) blk = nblocks - 1;
        double w, uw;
        or_boltzmann_weight(dU[j], kT, &w, &uw);
        sW[blk] += w;
        sUW[blk] += uw;
        n[blk] += 1;
    }
    for (int b = 0; b < nblocks; ++b) {
        block_sums[3 * b] = sW[b];
        block_sums[3 * b + 1] = sUW[b];
        block_sums[3 * b + 2] = (double)n[b];
    }
    int st = or_reduce(nblocks, sW, sUW, n, kT, s->V, out);
    free(dU); free(sW); free(sUW); free(n);
    return st;
}

/* ------------------------------------------------------------------------------------------
 * Reweighting gradients of ln K_H and q_st of one system (no counterpart in the reference,
 * which has no derivative path). Over insertions 0..ns-1 with w_i = exp(-beta dU_i - m):
 *   d ln K_H / d theta = -beta G / W,
 *   d q_st / d theta   = -(G - beta H) / W - beta (R / W) (G / W),
 * W = sum w, R = sum w dU, G = sum w d dU/d theta, H = sum w dU d dU/d theta.
 * theta runs over
 *  - the per-type LJ parameters eps_t, sigma_t under Lorentz-Berthelot mixing
 *    (eps_xy = sqrt(eps_x eps_y), sigma_xy = (sigma_x + sigma_y)/2, src/forcefield.jl:30-45),
 *    through every pair term and the tail correction (src/forcefield.jl:103-124);
 *  - each framework charge q_j: real-space pairs, the reciprocal cross term over the FULL k
 *    table, and the net-charge term of constant_offset. With a replication claim the entry of
 *    atom j is the mean over its images (atoms in replicate order, image-major), i.e. the
 *    derivative along "all images of j change together" divided by the image count.
 * Per-pair derivatives: d u_LJ/d eps_xy = 4 (x^12 - x^6), d u_LJ/d sigma_xy =
 * 4 eps (12 x^12 - 6 x^6) / sigma with x = sigma/r; d E/d q_j = KE q_a erfc(alpha r)/r (pair
 * series) + KE sum_k kpref 2 (cos(k.r_j) Re S_g + sin(k.r_j) Im S_g).
 * Pinned by central finite differences of or_widom's ln K_H and q_st (tests).
 * out: ln K_H, q_st. Returns 0, or -1 when no insertion has a finite energy.
 * ---------------------------------------------------------------------------------------- */
int or_widom_grad(const or_system *s, double T, int64_t ns, uint64_t seed, int64_t sysidx, int nthreads,
                  double out[2], double *lnkh_eps, double *lnkh_sig, double *qst_eps, double *qst_sig,
                  double *lnkh_q, double *qst_q)
{
    const double kT = OR_KB * T, beta = 1.0 / kT;
    const int NT = s->ntypes, na = s->natoms, nkf = s->nkfull;
    double *dU = malloc(sizeof(double) * (size_t)(ns > 0 ? ns : 1));
    or_energies(s, seed, sysidx, 0, ns, dU, NULL, NULL, nthreads);
    double m = -INFINITY;
    for (int64_t i = 0; i < ns; ++i) if (-beta * dU[i] > m) m = -beta * dU[i];
    if (!(m > -INFINITY)) { free(dU); return -1; }
    const size_t nlj = (size_t)NT * NT;
    /* accumulators: [0..3] per mixed pair (G_eps, H_eps, G_sig, H_sig), then Gq, Hq, then
     * A_re, A_im, B_re, B_im per full k, then W, R */
    const size_t nacc = 4 * nlj + 2 * (size_t)na + 4 * (size_t)nkf + 2;
    double *acc = calloc(nacc, sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *loc = calloc(nacc, sizeof(double));
        double *Ge = loc, *He = loc + nlj, *Gs = loc + 2 * nlj, *Hs = loc + 3 * nlj;
        double *Gq = loc + 4 * nlj, *Hq = Gq + na, *Ak = Hq + na, *Wl = Ak + 4 * (size_t)nkf;
        const double rc_lj2 = s->rc_lj * s->rc_lj, rc_ew2 = s->rc_ew * s->rc_ew;
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < ns; ++i) {
            const double w = exp(-beta * dU[i] - m);
            if (w == 0.0) continue;
            const double wu = w * dU[i];
            Wl[0] += w;
            Wl[1] += wu;
            double pos[3], q[4], gp[3 * OR_MAXSITES];
            or_pose(seed, sysidx, i, s->A, pos, q);
            for (int a = 0; a < s->nsites; ++a) {
                double r[3];
                or_rotate(q, s->gsite + 3 * a, r);
                for (int d = 0; d < 3; ++d) gp[3 * a + d] = pos[d] + r[d];
            }
            for (int j = 0; j < na; ++j) {
                const double *hp = s->pos + 3 * j;
                const int ht = s->type[j];
                for (int a = 0; a < s->nsites; ++a) {
                    double d[3] = {gp[3 * a] - hp[0], gp[3 * a + 1] - hp[1], gp[3 * a + 2] - hp[2]}, mi[3];
                    or_minimum_image(s->A, s->invA, d, mi);
                    const double r2 = or_dot3(mi, mi);
                    const int gt = s->gtype[a];
                    const size_t p = (size_t)gt + (size_t)NT * ht;
                    if (r2 < rc_lj2) {
                        const double sg = s->sigma[p], ep = s->eps[p];
                        double x6 = sg * sg / r2;
                        x6 = x6 * x6 * x6;
                        const double de = 4.0 * (x6 * x6 - x6);
                        const double ds = sg > 0.0 ? 4.0 * ep * (12.0 * x6 * x6 - 6.0 * x6) / sg : 0.0;
                        Ge[p] += w * de; He[p] += wu * de;
                        Gs[p] += w * ds; Hs[p] += wu * ds;
                    }
                    if (s->charged && r2 < rc_ew2) {
                        const double r = sqrt(r2);
                        const double f = OR_KE * s->gq[a] * or_pair_erfc_f64(s->alpha * r) / r;
                        Gq[j] += w * f;
                        Hq[j] += wu * f;
                    }
                }
            }
            if (s->charged)
                for (int k = 0; k < nkf; ++k) {
                    double re = 0.0, im = 0.0;
                    for (int a = 0; a < s->nsites; ++a) {
                        const double ph = or_dot3(s->kfull + 3 * k, gp + 3 * a);
                        re += s->gq[a] * cos(ph);
                        im += s->gq[a] * sin(ph);
                    }
                    Ak[4 * k] += w * re; Ak[4 * k + 1] += w * im;
                    Ak[4 * k + 2] += wu * re; Ak[4 * k + 3] += wu * im;
                }
        }
#pragma omp critical
        for (size_t t = 0; t < nacc; ++t) acc[t] += loc[t];
        free(loc);
    }
    double *Ge = acc, *He = acc + nlj, *Gs = acc + 2 * nlj, *Hs = acc + 3 * nlj;
    double *Gq = acc + 4 * nlj, *Hq = Gq + na, *Ak = Hq + na;
    const double W = Ak[4 * (size_t)nkf], R = Ak[4 * (size_t)nkf + 1];
    /* tail correction: every insertion carries d tail / d (eps_xy, sigma_xy) */
    if (s->tail) {
        int64_t *counts = calloc((size_t)NT, sizeof(int64_t)), *gcounts = calloc((size_t)NT, sizeof(int64_t));
        for (int j = 0; j < na; ++j) counts[s->type[j]] += 1;
        for (int a = 0; a < s->nsites; ++a) gcounts[s->gtype[a]] += 1;
        const double pre = (8.0 * OR_PI / 3.0) / s->V, rc = s->rc_lj;
        for (int x = 0; x < NT; ++x)
            for (int y = 0; y < NT; ++y) {
                const double mult = 2.0 * (double)counts[x] * (double)gcounts[y] + (double)gcounts[x] * (double)gcounts[y];
                const size_t p = (size_t)x + (size_t)NT * y;
                const double sg = s->sigma[p], ep = s->eps[p], u = sg / rc, u3 = u * u * u;
                const double dce = sg * sg * sg * (u3 * u3 * u3 / 3.0 - u3);
                const double dcs = ep * (12.0 * pow(sg, 11) / (3.0 * pow(rc, 9)) - 6.0 * pow(sg, 5) / pow(rc, 3));
                Ge[p] += W * pre * mult * dce; He[p] += R * pre * mult * dce;
                Gs[p] += W * pre * mult * dcs; Hs[p] += R * pre * mult * dcs;
            }
        free(counts); free(gcounts);
    }
    /* Lorentz-Berthelot chain rule */
    for (int t = 0; t < NT; ++t) {
        double ge = 0.0, he = 0.0, gs = 0.0, hs = 0.0;
        const double et = s->eps[t + (size_t)NT * t];
        for (int x = 0; x < NT; ++x)
            for (int y = 0; y < NT; ++y) {
                const int c = (x == t) + (y == t);
                if (!c) continue;
                const size_t p = (size_t)x + (size_t)NT * y;
                const double fe = s->eps[p] / (2.0 * et) * c, fs = 0.5 * c;
                if (Ge[p] != 0.0 || He[p] != 0.0) { ge += fe * Ge[p]; he += fe * He[p]; }
                gs += fs * Gs[p]; hs += fs * Hs[p];
            }
        lnkh_eps[t] = -beta * ge / W;
        lnkh_sig[t] = -beta * gs / W;
        qst_eps[t] = -(ge - beta * he) / W - beta * (R / W) * (ge / W);
        qst_sig[t] = -(gs - beta * hs) / W - beta * (R / W) * (gs / W);
    }
    /* charges: full-k reciprocal per atom, net-charge term, mean over replication images */
    double Qg = 0.0;
    for (int a = 0; a < s->nsites; ++a) Qg += s->gq[a];
    const double cnet = s->charged ? -OR_KE * OR_PI * Qg / (s->V * s->alpha * s->alpha) : 0.0;
    for (int j = 0; j < na; ++j) {
        double gw = 0.0, gu = 0.0;
        if (s->charged)
            for (int k = 0; k < nkf; ++k) {
                const double ph = or_dot3(s->kfull + 3 * k, s->pos + 3 * j), c = cos(ph), sn = sin(ph);
                gw += 2.0 * s->kpref_full[k] * (c * Ak[4 * k] + sn * Ak[4 * k + 1]);
                gu += 2.0 * s->kpref_full[k] * (c * Ak[4 * k + 2] + sn * Ak[4 * k + 3]);
            }
        Gq[j] += OR_KE * gw + W * cnet;
        Hq[j] += OR_KE * gu + R * cnet;
    }
    const int nimg = s->rep[0] * s->rep[1] * s->rep[2];
    if (nimg > 1 && na % nimg == 0) {
        const int nu = na / nimg;
        for (int a = 0; a < nu; ++a) {
            double sg = 0.0, sh = 0.0;
            for (int im = 0; im < nimg; ++im) { sg += Gq[im * nu + a]; sh += Hq[im * nu + a]; }
            for (int im = 0; im < nimg; ++im) { Gq[im * nu + a] = sg / nimg; Hq[im * nu + a] = sh / nimg; }
        }
    }
    for (int j = 0; j < na; ++j) {
        lnkh_q[j] = -beta * Gq[j] / W;
        qst_q[j] = -(Gq[j] - beta * Hq[j]) / W - beta * (R / W) * (Gq[j] / W);
    }
    out[0] = log(s->V) - log(kT) + log(W / (double)ns) + m;
    out[1] = kT - R / W;
    free(acc);
    free(dU);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Hard-core rejection (src/reject.jl, src/widom.jl:99-151,202-253), FP64 throughout.
 * ---------------------------------------------------------------------------------------- */
/* theta_F(Float64), src/reject.jl:11-19 */
double or_theta_f64(void)
{
    double lo = 0.0, hi = 2000.0;
    while (nextafter(lo, hi) != hi) {
        double mid = lo + (hi - lo) / 2.0;
        if (exp(-mid) == 0.0) hi = mid; else lo = mid;
    }
    return hi;
}

static double or_lj(double r2, double sg, double ep)
{
    double x = sg * sg / r2;
    x = x * x * x;
    return 4.0 * ep * (x * x - x);
}

/* src/reject.jl:22 */
double or_pair_energy(double r, double sg, double ep, double K, double alpha)
{
    return or_lj(r * r, sg, ep) + K * or_pair_erfc_f64(alpha * r) / r;
}

/* src/reject.jl:45-69; NaN when no bracket */
double or_find_r0(double sg, double ep, double K, double alpha, double r_lj)
{
    double hi = sg * pow(2.0, 1.0 / 6.0);
    if (!(or_pair_energy(hi, sg, ep, K, alpha) < 0.0)) return NAN;
    double lo = hi;
    for (int it = 0; or_pair_energy(lo, sg, ep, K, alpha) <= 0.0; ++it) {
        lo /= 2.0;
        if (it > 200) return NAN;
    }
    for (int it = 0; hi - lo > 4.0 * 2.220446049250313e-16 * hi; ++it) {
        double mid = (lo + hi) / 2.0;
        if (or_pair_energy(mid, sg, ep, K, alpha) > 0.0) lo = mid; else hi = mid;
        if (it > 200) return NAN;
    }
    double r = (lo + hi) / 2.0;
    return r < r_lj ? r : r_lj;
}

/* src/reject.jl:191-209 */
double or_find_rho2(double sg, double ep, double kmin, double margin, double r_lj)
{
    if (isinf(margin) || ep == 0.0) return 0.0;
#define OR_F(r) (or_lj((r) * (r), sg, ep) - fabs(kmin) / (r) - margin)
    double hi = sg;
    for (int it = 0; OR_F(hi) >= 0.0; ++it) { hi *= 2.0; if (it > 200) return NAN; }
    double lo = hi;
    for (int it = 0; OR_F(lo) < 0.0; ++it) { lo /= 2.0; if (it > 200) return NAN; }
    for (int it = 0; hi - lo > 4.0 * 2.220446049250313e-16 * hi; ++it) {
        double mid = (lo + hi) / 2.0;
        if (OR_F(mid) > 0.0) lo = mid; else hi = mid;
        if (it > 200) return NAN;
    }
#undef OR_F
    double r = (lo + hi) / 2.0;
    if (r > r_lj) r = r_lj;
    return r * r;
}

/* hardcore_bound (src/reject.jl:101-131) over every (site, atom) pair of the system, with the
 * reciprocal bound over the coupled k subset (the k-vectors the energy sums). */
double or_hardcore_bound(const or_system *s)
{
    double acc = 0.0, qg = 0.0;
    for (int a = 0; a < s->nsites; ++a) {
        qg += fabs(s->gq[a]);
        for (int j = 0; j < s->natoms; ++j) {
            size_t p = (size_t)s->gtype[a] + (size_t)s->ntypes * s->type[j];
            double sg = s->sigma[p], ep = s->eps[p], K = OR_KE * s->gq[a] * s->q[j];
            if (K >= 0.0) { acc += ep; continue; }
            if (ep == 0.0) return INFINITY;
            double r0 = or_find_r0(sg, ep, K, s->alpha, s->rc_lj);
            if (isnan(r0)) return INFINITY;
            acc += ep + fabs(K) * or_pair_erfc_f64(s->alpha * r0) / r0;
        }
    }
    double Rs = 0.0;
    for (int t = 0; t < s->nk; ++t) {
        int i = s->kidx[t];
        Rs += s->kpref_full[i] * hypot(s->Sre_full[i], s->Sim_full[i]);
    }
    return acc + 2.0 * OR_KE * qg * Rs;
}

/* build_rejection_tables for one system (src/widom.jl:222-253) with the rounding allowance of
 * arithmetic at unit roundoff eps_f/2 (2.220446049250313e-16 for FP64 energies, 1.1920929e-7
 * for FP32). rho2: nsites x ntypes (force-field types). Returns the margin (eV). */
double or_rejection_tables(const or_system *s, double kT, double eps_f, double *rho2)
{
    double theta = or_theta_f64();
    double Bs = or_hardcore_bound(s), cs = s->constant_offset;
    double n = (double)s->nsites * s->natoms + s->nk + 8;
    double safety = 2.0 * n * eps_f * (Bs + fabs(cs)) + 4.0e-6 * Bs;
    double margin = isinf(Bs) ? INFINITY : (theta + 2.0) * kT + safety + Bs - cs;
    for (int a = 0; a < s->nsites; ++a)
        for (int t = 0; t < s->ntypes; ++t) {
            double kmin = 0.0;
            for (int j = 0; j < s->natoms; ++j)
                if (s->type[j] == t) {
                    double K = OR_KE * s->gq[a] * s->q[j];
                    if (K < kmin) kmin = K;
                }
            size_t p = (size_t)s->gtype[a] + (size_t)s->ntypes * t;
            rho2[a * s->ntypes + t] = margin > 0.0 ? or_find_rho2(s->sigma[p], s->eps[p], kmin, margin, s->rc_lj) : NAN;
        }
    return margin;
}

/* Phase-0 flags of insertions j0..j0+n-1 by brute force over every atom, per-site minimum
 * image: flag = some site within sqrt(rho2[site][type]) of an atom of that type. */
void or_hardcore_flags(const or_system *s, uint64_t seed, int64_t sysidx, int64_t j0, int64_t n, const double *rho2,
                       uint8_t *flags, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < n; ++t) {
        double pos[3], q[4];
        or_pose(seed, sysidx, j0 + t, s->A, pos, q);
        uint8_t f = 0;
        for (int a = 0; a < s->nsites && !f; ++a) {
            double r[3], gp[3];
            or_rotate(q, s->gsite + 3 * a, r);
            for (int d = 0; d < 3; ++d) gp[d] = pos[d] + r[d];
            for (int j = 0; j < s->natoms; ++j) {
                double dd[3] = {gp[0] - s->pos[3 * j], gp[1] - s->pos[3 * j + 1], gp[2] - s->pos[3 * j + 2]}, m[3];
                or_minimum_image(s->A, s->invA, dd, m);
                if (or_dot3(m, m) < rho2[a * s->ntypes + s->type[j]]) { f = 1; break; }
            }
        }
        flags[t] = f;
    }
}

/* The reference's two-phase CPU path (src/widom.jl:390-420): phase-0 flags by the brute-force
 * check, then energies of the survivors only; rejected insertions get dU = +inf (weight 0). */
void or_energies_twophase(const or_system *s, uint64_t seed, int64_t sysidx, int64_t j0, int64_t n, const double *rho2,
                          double *dU, int64_t *npairs, int nthreads)
{
    uint8_t *flags = malloc((size_t)(n > 0 ? n : 1));
    or_hardcore_flags(s, seed, sysidx, j0, n, rho2, flags, nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < n; ++t) {
        if (flags[t]) {
            dU[t] = INFINITY;
            if (npairs) npairs[t] = 0;
            continue;
        }
        double pos[3], q[4];
        or_pose(seed, sysidx, j0 + t, s->A, pos, q);
        int64_t np = 0;
        dU[t] = or_insertion_energy(s, pos, q, &np) + s->constant_offset;
        if (npairs) npairs[t] = np;
    }
    free(flags);
}

/* ------------------------------------------------------------------------------------------
 * Move-energy path (src/guest.jl, src/moves.jl, src/state.jl, src/gcmc.jl): one Monte Carlo
 * chain over a system built with every k-vector (fullk), FP64, serial. The move logic and
 * the per-attempt random keying follow the reference's select_and_propose /
 * evaluate_*_kernel! / decide_*_kernel! (src/moves.jl:275-288, 379-553, 827-1128): attempt c
 * of a chain with seed S draws from ChainRNG(S, c, 0) -- guest index (Lemire), proposal,
 * acceptance uniform.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    const or_system *s;
    int cap, occ;
    double *pos, *quat, *host_e;  /* cap x 3, cap x 4, cap */
    double *Sre, *Sim;            /* running structure factor over the full k table */
    double energy, energy_abs;
    uint64_t seed, counter;
    double const_p, const_q;
    int acc[5], att[5];
} or_mc;

static void or_sites_at(const or_system *s, const double p[3], const double q[4], double out[][3])
{
    for (int a = 0; a < s->nsites; ++a) {
        double r[3];
        or_rotate(q, s->gsite + 3 * a, r);
        for (int d = 0; d < 3; ++d) out[a][d] = p[d] + r[d];
    }
}

static double or_mi_r2(const or_system *s, const double a[3], const double b[3])
{
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]}, m[3];
    or_minimum_image(s->A, s->invA, d, m);
    return or_dot3(m, m);
}

/* LJ inside rc_lj plus KE q_a q_b erfc(alpha r)/r inside rc_ew, one site pair. */
static double or_site_pair(const or_system *s, double r2, int ta, int tb, double qa, double qb)
{
    double e = 0.0;
    size_t p = (size_t)ta + (size_t)s->ntypes * tb;
    if (r2 < s->rc_lj * s->rc_lj) {
        double x = s->sigma[p] * s->sigma[p] / r2;
        x = x * x * x;
        e += 4.0 * s->eps[p] * (x * x - x);
    }
    if (s->charged && r2 < s->rc_ew * s->rc_ew) {
        double r = sqrt(r2);
        e += OR_KE * qa * qb * or_pair_erfc_f64(s->alpha * r) / r;
    }
    return e;
}

/* host_guest_realspace_energy (src/guest.jl:176-222), per-site minimum image. */
double or_host_guest(const or_system *s, const double p[3], const double q[4])
{
    double sites[OR_MAXSITES][3], e = 0.0;
    or_sites_at(s, p, q, sites);
    for (int j = 0; j < s->natoms; ++j)
        for (int a = 0; a < s->nsites; ++a)
            e += or_site_pair(s, or_mi_r2(s, sites[a], s->pos + 3 * j), s->gtype[a], s->type[j], s->gq[a], s->q[j]);
    return e;
}

/* guest_pair_realspace_energy (src/guest.jl:49-71), LJ + KE E_sr. */
static double or_guest_pair(const or_system *s, double sa[][3], double sb[][3])
{
    double e = 0.0;
    for (int a = 0; a < s->nsites; ++a)
        for (int b = 0; b < s->nsites; ++b)
            e += or_site_pair(s, or_mi_r2(s, sa[a], sb[b]), s->gtype[a], s->gtype[b], s->gq[a], s->gq[b]);
    return e;
}

/* KE (self + exclusion) of one guest (guest_self_terms, src/guest.jl:26-42). */
double or_guest_self(const or_system *s)
{
    double q2 = 0.0, excl = 0.0;
    for (int a = 0; a < s->nsites; ++a) q2 += s->gq[a] * s->gq[a];
    for (int a = 0; a < s->nsites; ++a)
        for (int c = a + 1; c < s->nsites; ++c) {
            double d[3] = {s->gsite[3 * a] - s->gsite[3 * c], s->gsite[3 * a + 1] - s->gsite[3 * c + 1],
                           s->gsite[3 * a + 2] - s->gsite[3 * c + 2]};
            double r = or_norm3(d);
            excl -= s->gq[a] * s->gq[c] * (1.0 - or_erfc_dev(s->alpha * r)) / r;
        }
    return OR_KE * (-s->alpha / sqrt(OR_PI) * q2 + excl);
}

/* tail and net-charge terms at N guests (src/guest.jl:9-14, 332-341). */
double or_guest_tail(const or_system *s, int64_t N)
{
    int64_t *counts = calloc((size_t)s->ntypes, sizeof(int64_t)), *g = calloc((size_t)s->ntypes, sizeof(int64_t));
    for (int j = 0; j < s->natoms; ++j) counts[s->type[j]] += 1;
    for (int a = 0; a < s->nsites; ++a) g[s->gtype[a]] += N;
    double t = or_tail_delta(s->ntypes, s->sigma, s->eps, s->rc_lj, s->tail, counts, g, s->V);
    free(counts); free(g);
    return t;
}

double or_net(const or_system *s, int64_t N)
{
    double Qh = 0.0, Qg = 0.0;
    for (int j = 0; j < s->natoms; ++j) Qh += s->q[j];
    for (int a = 0; a < s->nsites; ++a) Qg += s->gq[a];
    return -OR_KE * OR_PI / (2.0 * s->V * s->alpha * s->alpha) * ((Qh + N * Qg) * (Qh + N * Qg) - Qh * Qh);
}

/* exchange_constant_term (src/moves.jl:728-746) */
double or_exchange_constant(const or_system *s, int64_t N)
{
    return or_guest_self(s) + or_guest_tail(s, N + 1) - or_guest_tail(s, N) + or_net(s, N + 1) - or_net(s, N);
}

static void or_mc_rebuild_sk(const or_mc *m, double *Sre, double *Sim)
{
    const or_system *s = m->s;
    for (int k = 0; k < s->nkfull; ++k) { Sre[k] = s->Sre_full[k]; Sim[k] = s->Sim_full[k]; }
    if (!s->charged) return;
    double sites[OR_MAXSITES][3];
    for (int i = 0; i < m->occ; ++i) {
        or_sites_at(s, m->pos + 3 * i, m->quat + 4 * i, sites);
        for (int k = 0; k < s->nkfull; ++k)
            for (int a = 0; a < s->nsites; ++a) {
                double ph = or_dot3(s->kfull + 3 * k, sites[a]);
                Sre[k] += s->gq[a] * cos(ph);
                Sim[k] += s->gq[a] * sin(ph);
            }
    }
}

/* total_energy (src/guest.jl:293-343) from the poses, with the structure factor rebuilt. */
double or_mc_total_energy(const or_mc *m)
{
    const or_system *s = m->s;
    double e = 0.0;
    for (int i = 0; i < m->occ; ++i) e += or_host_guest(s, m->pos + 3 * i, m->quat + 4 * i);
    double si[OR_MAXSITES][3], sj[OR_MAXSITES][3];
    for (int i = 0; i < m->occ; ++i) {
        or_sites_at(s, m->pos + 3 * i, m->quat + 4 * i, si);
        for (int j = i + 1; j < m->occ; ++j) {
            or_sites_at(s, m->pos + 3 * j, m->quat + 4 * j, sj);
            e += or_guest_pair(s, si, sj);
        }
    }
    double *Sre = malloc(sizeof(double) * (size_t)(s->nkfull + 1)), *Sim = malloc(sizeof(double) * (size_t)(s->nkfull + 1));
    or_mc_rebuild_sk(m, Sre, Sim);
    double rec = 0.0;
    for (int k = 0; k < s->nkfull; ++k)
        rec += s->kpref_full[k] * ((Sre[k] * Sre[k] + Sim[k] * Sim[k]) -
                                   (s->Sre_full[k] * s->Sre_full[k] + s->Sim_full[k] * s->Sim_full[k]));
    free(Sre); free(Sim);
    e += OR_KE * rec;
    e += m->occ * or_guest_self(s) + or_guest_tail(s, m->occ) + or_net(s, m->occ);
    return e;
}

or_mc *or_mc_new(const or_system *s, int cap, uint64_t chain_seed, int occ, const double *pos, const double *quat)
{
    or_mc *m = calloc(1, sizeof(or_mc));
    m->s = s;
    m->cap = cap;
    m->occ = occ;
    m->pos = calloc((size_t)(cap + 1) * 3, sizeof(double));
    m->quat = calloc((size_t)(cap + 1) * 4, sizeof(double));
    m->host_e = calloc((size_t)(cap + 1), sizeof(double));
    m->Sre = calloc((size_t)(s->nkfull + 1), sizeof(double));
    m->Sim = calloc((size_t)(s->nkfull + 1), sizeof(double));
    for (int i = 0; i < cap; ++i) m->quat[4 * i + 3] = 1.0;
    for (int i = 0; i < occ; ++i) {
        memcpy(m->pos + 3 * i, pos + 3 * i, 3 * sizeof(double));
        memcpy(m->quat + 4 * i, quat + 4 * i, 4 * sizeof(double));
        m->host_e[i] = or_host_guest(s, m->pos + 3 * i, m->quat + 4 * i);
    }
    or_mc_rebuild_sk(m, m->Sre, m->Sim);
    m->energy = or_mc_total_energy(m);
    m->seed = chain_seed;
    m->const_p = or_exchange_constant(s, 0);
    m->const_q = or_exchange_constant(s, 1) - m->const_p;
    return m;
}

void or_mc_free(or_mc *m)
{
    if (!m) return;
    free(m->pos); free(m->quat); free(m->host_e); free(m->Sre); free(m->Sim);
    free(m);
}

static double or_rand_normal(or_chain *c)
{
    double u1 = or_unit_f64(or_next_bits(c)), u2 = or_unit_f64(or_next_bits(c));
    double r = sqrt(-2.0 * log(1.0 - u1)), sn, cs;
    or_sincospi(2.0 * u2, &sn, &cs);
    return r * cs;
}

static void or_rand_quat(or_chain *c, double q[4])
{
    double u1 = or_unit_f64(or_next_bits(c)), u2 = or_unit_f64(or_next_bits(c)), u3 = or_unit_f64(or_next_bits(c));
    or_shoemake(u1, u2, u3, q);
}

/* qmul, quaternion_pow (src/moves.jl:81-110) */
static void or_qmul(const double q2[4], const double q1[4], double o[4])
{
    o[3] = q2[3] * q1[3] - (q2[0] * q1[0] + q2[1] * q1[1] + q2[2] * q1[2]);
    o[0] = q2[3] * q1[0] + q1[3] * q2[0] + (q2[1] * q1[2] - q2[2] * q1[1]);
    o[1] = q2[3] * q1[1] + q1[3] * q2[1] + (q2[2] * q1[0] - q2[0] * q1[2]);
    o[2] = q2[3] * q1[2] + q1[3] * q2[2] + (q2[0] * q1[1] - q2[1] * q1[0]);
}

static void or_qpow(const double q[4], double t, double o[4])
{
    double w = q[3] < -1.0 ? -1.0 : (q[3] > 1.0 ? 1.0 : q[3]);
    double th = 2.0 * acos(w), s2 = 1.0 - w * w, s = sqrt(s2 > 0.0 ? s2 : 0.0);
    double ax[3] = {1.0, 0.0, 0.0};
    if (s > 2.220446049250313e-16) { ax[0] = q[0] / s; ax[1] = q[1] / s; ax[2] = q[2] / s; }
    double half = t * th / 2.0, sh = sin(half);
    o[0] = ax[0] * sh; o[1] = ax[1] * sh; o[2] = ax[2] * sh; o[3] = cos(half);
}

/* Sampler pieces exported for the distribution tests: n successive rand_normal draws
 * (src/rng.jl:109-141) and n rand_quaternion draws from one ChainRNG(chain_seed, cycle, move)
 * stream, quaternion_pow and qmul (src/moves.jl:81-110). */
void or_normals(uint64_t chain_seed, uint64_t cycle, uint64_t move, int64_t n, double *out)
{
    or_chain c = or_chain_new(chain_seed, cycle, move);
    for (int64_t i = 0; i < n; ++i) out[i] = or_rand_normal(&c);
}

void or_quats(uint64_t chain_seed, uint64_t cycle, uint64_t move, int64_t n, double *out)
{
    or_chain c = or_chain_new(chain_seed, cycle, move);
    for (int64_t i = 0; i < n; ++i) or_rand_quat(&c, out + 4 * i);
}

void or_quat_pow(const double q[4], double t, double out[4]) { or_qpow(q, t, out); }

void or_quat_mul(const double q2[4], const double q1[4], double out[4]) { or_qmul(q2, q1, out); }

/* One attempt of `type` (1 translation, 2 rotation, 3 reinsertion, 4 insertion, 5 deletion).
 * fug: eV/Å^3. Returns accepted (0/1); *dU receives the energy change. Returns -1 when an
 * insertion the Metropolis test accepts finds the chain at capacity (nothing is changed). */
int or_mc_attempt(or_mc *m, int type, double step_t, double step_r, double fug, double kT, double *dU_out,
                  int *slot_out)
{
    const or_system *s = m->s;
    or_chain rng = or_chain_new(m->seed, m->counter, 0);
    int slot = 0;
    double op[3] = {0, 0, 0}, oq[4] = {0, 0, 0, 1}, np[3] = {0, 0, 0}, nq[4] = {0, 0, 0, 1};
    if (type != 4) {
        int64_t idx = or_rand_range(or_next_bits(&rng), m->occ > 0 ? m->occ : 1);
        slot = (int)(idx - 1);
        if (slot > m->cap - 1) slot = m->cap > 0 ? m->cap - 1 : 0;
        memcpy(op, m->pos + 3 * slot, sizeof op);
        memcpy(oq, m->quat + 4 * slot, sizeof oq);
    }
    if (type == 1) {
        double dx = or_rand_normal(&rng), dy = or_rand_normal(&rng), dz = or_rand_normal(&rng);
        np[0] = op[0] + step_t * dx; np[1] = op[1] + step_t * dy; np[2] = op[2] + step_t * dz;
        memcpy(nq, oq, sizeof nq);
    } else if (type == 2) {
        double qr[4], qp[4];
        or_rand_quat(&rng, qr);
        or_qpow(qr, step_r, qp);
        or_qmul(qp, oq, nq);
        double nrm = sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
        for (int i = 0; i < 4; ++i) nq[i] /= nrm;
        memcpy(np, op, sizeof np);
    } else if (type == 3 || type == 4) {
        double f[3];
        for (int i = 0; i < 3; ++i) f[i] = or_unit_f64(or_next_bits(&rng));
        or_matvec(s->A, f, np);
        or_rand_quat(&rng, nq);
    }
    double u = or_unit_f64(or_next_bits(&rng));
    int has_old = type != 4, has_new = type != 5;
    double os[OR_MAXSITES][3], ns[OR_MAXSITES][3], sj[OR_MAXSITES][3];
    or_sites_at(s, op, oq, os);
    or_sites_at(s, np, nq, ns);
    double e_new = has_new ? or_host_guest(s, np, nq) : 0.0;
    double gg = 0.0;
    for (int j = 0; j < m->occ; ++j) {
        if (has_old && j == slot) continue;
        or_sites_at(s, m->pos + 3 * j, m->quat + 4 * j, sj);
        if (has_new) gg += or_guest_pair(s, ns, sj);
        if (type <= 3) gg -= or_guest_pair(s, os, sj);
        else if (type == 5) gg += or_guest_pair(s, os, sj);
    }
    double rec = 0.0;
    double *dre = malloc(sizeof(double) * (size_t)(s->nkfull + 1)), *dim = malloc(sizeof(double) * (size_t)(s->nkfull + 1));
    for (int k = 0; k < s->nkfull; ++k) {
        double re = 0.0, im = 0.0;
        if (s->charged)
            for (int a = 0; a < s->nsites; ++a) {
                if (has_new) {
                    double ph = or_dot3(s->kfull + 3 * k, ns[a]);
                    re += s->gq[a] * cos(ph); im += s->gq[a] * sin(ph);
                }
                if (has_old) {
                    double ph = or_dot3(s->kfull + 3 * k, os[a]);
                    re -= s->gq[a] * cos(ph); im -= s->gq[a] * sin(ph);
                }
            }
        dre[k] = re; dim[k] = im;
        rec += s->kpref_full[k] * (2.0 * (m->Sre[k] * re + m->Sim[k] * im) + (re * re + im * im));
    }
    rec *= OR_KE;
    double dU;
    int accept, status = 0;
    if (type <= 3) {
        dU = (e_new - m->host_e[slot]) + gg + rec;
        accept = isfinite(dU) && (-dU / kT > log(u)) && m->occ > 0;
    } else if (type == 4) {
        dU = e_new + gg + rec + (m->const_p + m->const_q * m->occ);
        double lp = log(fug * s->V) - log((double)m->occ + 1.0) - log(kT);
        int would = isfinite(dU) && (lp - dU / kT > log(u));
        accept = would && m->occ < m->cap;
        if (would && !accept) status = -1;
    } else {
        int occ_eff = m->occ > 0 ? m->occ : 1;
        dU = -(m->host_e[slot] + gg + (m->const_p + m->const_q * (m->occ - 1))) + rec;
        double lp = -(log(fug * s->V) - log((double)(occ_eff - 1) + 1.0) - log(kT));
        accept = isfinite(dU) && (lp - dU / kT > log(u)) && m->occ > 0;
    }
    if (accept) {
        if (type <= 3) {
            memcpy(m->pos + 3 * slot, np, sizeof np);
            memcpy(m->quat + 4 * slot, nq, sizeof nq);
            m->host_e[slot] = e_new;
        } else if (type == 4) {
            memcpy(m->pos + 3 * m->occ, np, sizeof np);
            memcpy(m->quat + 4 * m->occ, nq, sizeof nq);
            m->host_e[m->occ] = e_new;
            m->occ += 1;
        } else {
            int hi = m->occ - 1;
            if (slot != hi) {
                memcpy(m->pos + 3 * slot, m->pos + 3 * hi, 3 * sizeof(double));
                memcpy(m->quat + 4 * slot, m->quat + 4 * hi, 4 * sizeof(double));
                m->host_e[slot] = m->host_e[hi];
            }
            m->occ -= 1;
        }
        m->energy += dU;
        m->energy_abs += fabs(dU);
        for (int k = 0; k < s->nkfull; ++k) { m->Sre[k] += dre[k]; m->Sim[k] += dim[k]; }
        m->acc[type - 1] += 1;
    }
    free(dre); free(dim);
    m->att[type - 1] += 1;
    m->counter += 1;
    if (dU_out) *dU_out = dU;
    if (slot_out) *slot_out = slot;
    return status < 0 ? status : accept;
}

/* One GCMC cycle of the chain: max(occupancy, min_cycle) attempts, the move type of attempt c
 * from ChainRNG(seed, c, 1) word 0 against the cumulative probabilities pcum[5]. Returns the
 * number of attempts, or -1 on a capacity hit. */
int64_t or_mc_cycle(or_mc *m, const double pcum[5], int64_t min_cycle, double step_t, double step_r, double fug,
                    double kT)
{
    int64_t len = m->occ > min_cycle ? m->occ : min_cycle;
    for (int64_t i = 0; i < len; ++i) {
        or_chain r1 = or_chain_new(m->seed, m->counter, 1);
        double v = or_unit_f64(or_next_bits(&r1));
        int t = 5;
        for (int k = 0; k < 5; ++k)
            if (v < pcum[k]) { t = k + 1; break; }
        if (or_mc_attempt(m, t, step_t, step_r, fug, kT, NULL, NULL) < 0) return -1;
    }
    return len;
}

/* accessors */
void or_mc_scalars(const or_mc *m, double out[6])
{
    out[0] = m->occ; out[1] = m->energy; out[2] = m->energy_abs; out[3] = (double)m->counter;
    out[4] = m->const_p; out[5] = m->const_q;
}

void or_mc_counts(const or_mc *m, int *acc, int *att)
{
    for (int t = 0; t < 5; ++t) { acc[t] = m->acc[t]; att[t] = m->att[t]; }
}

void or_mc_guests(const or_mc *m, double *pos, double *quat, double *host_e)
{
    for (int i = 0; i < m->occ; ++i) {
        if (pos) memcpy(pos + 3 * i, m->pos + 3 * i, 3 * sizeof(double));
        if (quat) memcpy(quat + 4 * i, m->quat + 4 * i, 4 * sizeof(double));
        if (host_e) host_e[i] = m->host_e[i];
    }
}

void or_mc_sk(const or_mc *m, double *re, double *im)
{
    for (int k = 0; k < m->s->nkfull; ++k) { re[k] = m->Sre[k]; im[k] = m->Sim[k]; }
}

/* ------------------------------------------------------------------------------------------
 * NVT driver pieces (src/nvt.jl): test insertions into a chain's current configuration and
 * the energy audit's reset, for the serial restatement of run_nvt!.
 * ---------------------------------------------------------------------------------------- */

/* insertion_constant_term (src/nvt.jl:44-66): constant_offset at Ng = 0; otherwise
 * constant_offset with its own 0 -> 1 tail and net-charge terms taken out and the Ng -> Ng+1
 * ones put in. */
double or_nvt_const_term(const or_system *s, int64_t Ng)
{
    if (Ng == 0) return s->constant_offset;
    double base = s->constant_offset - or_guest_tail(s, 1) - or_net(s, 1);
    return base + or_guest_tail(s, Ng + 1) - or_guest_tail(s, Ng) + or_net(s, Ng + 1) - or_net(s, Ng);
}

/* widom_chain_kernel! (src/nvt.jl:85-119) for one pose: host-guest real space, the reciprocal
 * cross term 2 Re(conj(S) S_g) against the chain's RUNNING structure factor over the full k
 * table (insertion_energy, src/energy.jl:69-80), guest-guest real space against every live
 * guest (guest_pair_realspace_energy), plus const_term. Minimum image per site pair. */
double or_chain_insertion(const or_mc *m, const double pos[3], const double q[4], double const_term)
{
    const or_system *s = m->s;
    double ns[OR_MAXSITES][3], sj[OR_MAXSITES][3];
    or_sites_at(s, pos, q, ns);
    double e = or_host_guest(s, pos, q);
    double lr = 0.0;
    if (s->charged)
        for (int k = 0; k < s->nkfull; ++k) {
            double re = 0.0, im = 0.0;
            for (int a = 0; a < s->nsites; ++a) {
                double ph = or_dot3(s->kfull + 3 * k, ns[a]);
                re += s->gq[a] * cos(ph);
                im += s->gq[a] * sin(ph);
            }
            lr += s->kpref_full[k] * 2.0 * (m->Sre[k] * re + m->Sim[k] * im);
        }
    double gg = 0.0;
    for (int j = 0; j < m->occ; ++j) {
        or_sites_at(s, m->pos + 3 * j, m->quat + 4 * j, sj);
        gg += or_guest_pair(s, ns, sj);
    }
    return e + OR_KE * lr + gg + const_term;
}

/* Test insertions j0 .. j0+n-1 of system `sysidx` into the chain as it stands, poses from the
 * Widom stream or_pose(widom_seed, sysidx, j). */
void or_chain_energies(const or_mc *m, uint64_t widom_seed, int64_t sysidx, int64_t j0, int64_t n, double const_term,
                       double *dU, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < n; ++t) {
        double pos[3], q[4];
        or_pose(widom_seed, sysidx, j0 + t, m->s->A, pos, q);
        dU[t] = or_chain_insertion(m, pos, q, const_term);
    }
}

/* audit_energy!'s reset on success (src/audit.jl:97-165): the running structure factor and
 * energy become the values rebuilt from the poses, the |dU| accumulator restarts. Returns the
 * discrepancy recomputed - running. */
double or_mc_audit(or_mc *m)
{
    double E = or_mc_total_energy(m);
    double d = E - m->energy;
    or_mc_rebuild_sk(m, m->Sre, m->Sim);
    m->energy = E;
    m->energy_abs = 0.0;
    return d;
}
