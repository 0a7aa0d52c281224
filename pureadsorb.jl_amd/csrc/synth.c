/*
 * synth.c -- random sequential addition of atoms in a periodic cell, the position rule of the
 * synthetic benchmark frameworks (pureadsorb_amd/synth.py documents the whole generator).
 *
 * Atom i (0-based) draws attempt k (1, 2, ...) from ChainRNG(seed, i+1, k) words 0..2
 * (SplitMix64 counter keys, the reference's src/rng.jl scheme) as fractional coordinates in
 * [0, 1)^3 and is accepted when its minimum-image distance to every atom placed before it is
 * at least min_sep. A cell grid with perpendicular cell widths >= min_sep limits the check
 * to the neighbouring cells.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t sm64(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + index * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double unit(uint64_t raw) { return (double)(raw >> 11) * 1.1102230246251565e-16; }

static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

static void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* A: column-major cell matrix. Returns the number of atoms placed (natoms), or -(i+1) when
 * atom i could not be placed within max_attempts. */
int pa_synth_positions(const double *A, int natoms, uint64_t seed, double min_sep, int max_attempts, double *frac)
{
    double V, L[3], x[3];
    V = fabs(A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
             A[6] * (A[1] * A[5] - A[4] * A[2]));
    cross3(A + 3, A + 6, x); L[0] = V / norm3(x);
    cross3(A + 0, A + 6, x); L[1] = V / norm3(x);
    cross3(A + 0, A + 3, x); L[2] = V / norm3(x);
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = (int)floor(L[d] / min_sep);
        if (n[d] < 1) n[d] = 1;
        if (n[d] > 256) n[d] = 256;
    }
    const int ncell = n[0] * n[1] * n[2];
    int *head = malloc(sizeof(int) * (size_t)ncell);
    int *next = malloc(sizeof(int) * (size_t)(natoms > 0 ? natoms : 1));
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    const double min2 = min_sep * min_sep;
    int status = natoms;
    for (int i = 0; i < natoms && status == natoms; ++i) {
        int placed = 0;
        for (int k = 1; k <= max_attempts && !placed; ++k) {
            uint64_t key = sm64(sm64(seed, (uint64_t)(i + 1)), (uint64_t)k);
            double f[3] = {unit(sm64(key, 0)), unit(sm64(key, 1)), unit(sm64(key, 2))};
            int h[3];
            for (int d = 0; d < 3; ++d) {
                h[d] = (int)(f[d] * n[d]);
                if (h[d] >= n[d]) h[d] = n[d] - 1;
            }
            int ok = 1;
            /* neighbouring cells; an axis with fewer than 3 cells is scanned once */
            int lo[3], cnt[3];
            for (int d = 0; d < 3; ++d) {
                if (n[d] < 3) { lo[d] = 0; cnt[d] = n[d]; }
                else { lo[d] = h[d] - 1; cnt[d] = 3; }
            }
            for (int a = 0; a < cnt[0] && ok; ++a)
                for (int b = 0; b < cnt[1] && ok; ++b)
                    for (int c = 0; c < cnt[2] && ok; ++c) {
                        int cx = (lo[0] + a + n[0]) % n[0], cy = (lo[1] + b + n[1]) % n[1], cz = (lo[2] + c + n[2]) % n[2];
                        for (int j = head[cx + n[0] * (cy + n[1] * cz)]; j >= 0 && ok; j = next[j]) {
                            double df[3], v[3];
                            for (int d = 0; d < 3; ++d) {
                                df[d] = f[d] - frac[3 * j + d];
                                df[d] -= rint(df[d]);
                            }
                            for (int r = 0; r < 3; ++r) v[r] = A[r] * df[0] + A[r + 3] * df[1] + A[r + 6] * df[2];
                            if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < min2) ok = 0;
                        }
                    }
            if (ok) {
                for (int d = 0; d < 3; ++d) frac[3 * i + d] = f[d];
                int cell = h[0] + n[0] * (h[1] + n[1] * h[2]);
                next[i] = head[cell];
                head[cell] = i;
                placed = 1;
            }
        }
        if (!placed) status = -(i + 1);
    }
    free(head);
    free(next);
    return status;
}
