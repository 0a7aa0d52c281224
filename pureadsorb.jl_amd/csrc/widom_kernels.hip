// widom_kernels.hip -- gfx950 kernels of the Widom insertion path.
//
//  * widom_f64_kernel      FP64 plain path: one work-item per insertion, every (site, atom)
//                          pair in FP64 with the reference's 17-term pair erfc and libm exp,
//                          minimum image per pair (src/energy.jl:42-82, src/reference.jl:39-56).
//  (the mixed-precision pipeline is in widom_binned.hip)
//  * boltzmann_tiles_kernel per-tile log-sum-exp Boltzmann sums {max, sum w, sum U w} in FP64,
//                          deterministic (fixed per-thread order, wavefront butterfly,
//                          wavefronts merged in order).
//  * widom_pose_kernel     device-drawn poses, for bit-exact placement checks.
#include "pa_device.hpp"

namespace pa {

__device__ inline int find_system(const int64_t *seg, int nsys, int64_t g)
{
    // largest s with seg[s] <= g
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------------------------------
// FP64 plain path
// ------------------------------------------------------------------------------------------
template <int NS, bool CHARGED>
__global__ __launch_bounds__(256) void widom_f64_kernel(WidomArgs a)
{
#pragma clang fp contract(off)
    const int64_t g = a.g0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.g1) return;
    const int s = find_system(a.seg, a.nsys, g);
    const int64_t j = g - a.seg[s];
    const FwDesc &F = a.fw[a.fw_of[s]];
    double f[3], q[4], pos[3];
    pose_uniforms(a.seed, s, j, f, q);
    matvec(F.A, f, pos);
    double gp[NS][3];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        double r[3];
        rotate(q, a.guest.site[t], r);
        for (int i = 0; i < 3; ++i) gp[t][i] = pos[i] + r[i];
    }
    double E_lj = 0.0, E_sr = 0.0;
    const double4 *atoms = a.atom64 + F.atom_off;
    const int *types = a.atype + F.atom_off;
    for (int n = 0; n < F.natoms; ++n) {
        const double4 h = atoms[n];
        const int ht = types[n];
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            double d[3] = {gp[t][0] - h.x, gp[t][1] - h.y, gp[t][2] - h.z};
            double m[3];
            minimum_image(F.A, F.invA, d, m);
            double r2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
            if (!(r2 < a.rc_lj2 || r2 < a.rc_ew2)) continue;
            if (r2 < a.rc_lj2) {
                const double2 p = a.pair64[t * a.ntypes_c + ht];
                double x = p.x * p.x / r2;
                x = x * x * x;
                E_lj += 4.0 * p.y * (x * x - x);
            }
            if (CHARGED && r2 < a.rc_ew2) {
                double r = __builtin_sqrt(r2);
                E_sr += a.guest.q[t] * h.w * pair_erfc_f64(F.alpha * r) / r;
            }
        }
    }
    double E_lr = 0.0;
    if (CHARGED) {
        const double4 *ks = a.k64 + F.k_off;
        const double2 *S = a.S64 + F.k_off;
        for (int i = 0; i < F.nk; ++i) {
            const double4 k = ks[i];
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                double ph = k.x * gp[t][0] + k.y * gp[t][1] + k.z * gp[t][2];
                double sn, cs;
                sincos(ph, &sn, &cs);
                re += a.guest.q[t] * cs;
                im += a.guest.q[t] * sn;
            }
            E_lr += k.w * 2.0 * (S[i].x * re + S[i].y * im);
        }
    }
    a.dU[g - a.out_base] = (E_lj + KE * (E_sr + E_lr)) + F.const_off;
}

// ------------------------------------------------------------------------------------------
// Boltzmann sums
// ------------------------------------------------------------------------------------------
struct Lse {
    double m, w, u;
};

__device__ inline void lse_add(Lse &acc, double lw, double du)
{
    if (lw == -INFINITY) return;  // w == 0 exactly: contributes 0 to both sums (src/widom.jl:80-84)
    if (lw > acc.m) {
        const double sc = exp(acc.m - lw);
        acc.w = acc.w * sc + 1.0;
        acc.u = acc.u * sc + du;
        acc.m = lw;
    } else {
        const double e = exp(lw - acc.m);
        acc.w += e;
        acc.u += du * e;
    }
}

__device__ inline Lse lse_merge(Lse x, Lse y)
{
    if (y.m == -INFINITY) return x;
    if (x.m == -INFINITY) return y;
    const double m = x.m > y.m ? x.m : y.m;
    const double ex = exp(x.m - m), ey = exp(y.m - m);
    Lse r;
    r.m = m;
    r.w = x.w * ex + y.w * ey;
    r.u = x.u * ex + y.u * ey;
    return r;
}

constexpr int TILE_THREADS = 256;

// One workgroup per tile of consecutive insertions (tiles never straddle a block). out[t] =
// {max log w, sum w e^-max, sum dU w e^-max, NaN count}. Each work-item accumulates its
// strided insertions with a running maximum, each wavefront merges its 64 partial sums with
// shuffles (butterfly, fixed order), and work-item 0 merges the wavefronts' results in order.
__global__ __launch_bounds__(TILE_THREADS) void boltzmann_tiles_kernel(const double *dU, const int64_t *tile_g0,
                                                                       const int *tile_n, double kT, double4 *out)
{
    constexpr int NW = TILE_THREADS / 64;
    __shared__ double4 part[NW];
    const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t g0 = tile_g0[t];
    const int n = tile_n[t];
    Lse acc = {-INFINITY, 0.0, 0.0};
    int nan = 0;
    for (int i = tid; i < n; i += TILE_THREADS) {
        const double du = dU[g0 + i];
        const double lw = -du / kT;
        if (lw != lw) { ++nan; continue; }
        lse_add(acc, lw, du);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const Lse y = {__shfl_xor(acc.m, off, 64), __shfl_xor(acc.w, off, 64), __shfl_xor(acc.u, off, 64)};
        nan += __shfl_xor(nan, off, 64);
        // both lanes of a pair merge (lower lane's value first), so they hold the same result
        acc = (lane & off) ? lse_merge(y, acc) : lse_merge(acc, y);
    }
    if (lane == 0) part[w] = make_double4(acc.m, acc.w, acc.u, (double)nan);
    __syncthreads();
    if (tid == 0) {
        Lse r = {part[0].x, part[0].y, part[0].z};
        double nn = part[0].w;
        for (int i = 1; i < NW; ++i) {
            r = lse_merge(r, Lse{part[i].x, part[i].y, part[i].z});
            nn += part[i].w;
        }
        out[t] = make_double4(r.m, r.w, r.u, nn);
    }
}

// ------------------------------------------------------------------------------------------
// Poses
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void widom_pose_kernel(const FwDesc *fw, const int *fw_of, int sys, uint64_t seed,
                                                         int64_t j0, int64_t n, double *pos, double *quat)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const FwDesc &F = fw[fw_of[sys]];
    double f[3], q[4], p[3];
    pose_uniforms(seed, sys, j0 + i, f, q);
    matvec(F.A, f, p);
    for (int d = 0; d < 3; ++d) pos[3 * i + d] = p[d];
    for (int d = 0; d < 4; ++d) quat[4 * i + d] = q[d];
}

// ------------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------------
template <int NS>
static void launch_ns(const WidomArgs &a, hipStream_t st)
{
    const int64_t n = a.g1 - a.g0;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (a.guest.charged) hipLaunchKernelGGL((widom_f64_kernel<NS, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((widom_f64_kernel<NS, false>), grid, block, 0, st, a);
}

// FP64 plain-path energy launch over g in [a.g0, a.g1).
hipError_t launch_energy(const WidomArgs &a, int precision, bool upper, hipStream_t st)
{
    (void)precision;
    (void)upper;
    if (a.g1 <= a.g0) return hipSuccess;
    switch (a.guest.nsites) {
    case 1: launch_ns<1>(a, st); break;
    case 2: launch_ns<2>(a, st); break;
    case 3: launch_ns<3>(a, st); break;
    case 4: launch_ns<4>(a, st); break;
    case 5: launch_ns<5>(a, st); break;
    case 6: launch_ns<6>(a, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tiles(const double *dU, const int64_t *tile_g0, const int *tile_n, int ntiles, double kT, double4 *out,
                        hipStream_t st)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(boltzmann_tiles_kernel, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, st, dU, tile_g0, tile_n, kT,
                       out);
    return hipGetLastError();
}

hipError_t launch_poses(const FwDesc *fw, const int *fw_of, int sys, uint64_t seed, int64_t j0, int64_t n, double *pos,
                        double *quat, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(widom_pose_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, fw, fw_of, sys, seed, j0, n,
                       pos, quat);
    return hipGetLastError();
}

}  // namespace pa
