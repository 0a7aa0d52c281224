// widom_kernels.hip -- gfx950 kernels of the Widom insertion path.
//
//  * widom_f64_kernel      FP64 plain path: one work-item per insertion, every (site, atom)
//                          pair in FP64 with the reference's 17-term pair erfc and libm exp,
//                          minimum image per pair (src/energy.jl:42-82, src/reference.jl:39-56).
//  * widom_mixed_kernel    one wavefront per insertion at a time, lanes over framework atoms:
//                          32-bit fixed-point fractional coordinates make the minimum image an
//                          integer subtraction; in-cutoff (site, atom) pairs are compacted into
//                          a per-wave LDS ring so the FP32 LJ + screened-Coulomb math runs on
//                          full wavefronts; the reciprocal cross term uses exact fixed-point
//                          phases (n . f mod 1) and the hardware sin/cos (revolution input).
//                          Pair terms are FP32, every accumulator is FP64.
//  * boltzmann_tiles_kernel per-tile log-sum-exp Boltzmann sums {max, sum w, sum U w} in FP64,
//                          deterministic (fixed per-thread order, fixed tree).
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
// Mixed path
// ------------------------------------------------------------------------------------------
constexpr int MIX_WAVE = 64;
constexpr int MIX_INS_PER_WAVE = 64;
// Ring capacity: fewer than 64 entries stay pending after a drain, and one sweep over 64
// atoms appends at most 64 * NS more.
template <int NS>
constexpr int ring_size() { return NS <= 1 ? 128 : NS <= 3 ? 256 : NS <= 7 ? 512 : 1024; }
constexpr int MAX_TYPES_MIXED = 64;

template <bool CHARGED>
__device__ inline float pair_energy_f32(float r2, float qj, int info, const float2 *pair, int ntc,
                                        const float *qke, float rc_lj2, float rc_ew2, float alpha,
                                        float alpha2)
{
    const int site = info >> 16, t = info & 0xffff;
    r2 = fmaxf(r2, 1.0e-30f);
    float e = 0.0f;
    if (r2 < rc_lj2) {
        const float2 p = pair[site * ntc + t];
        float x = p.x * __builtin_amdgcn_rcpf(r2);
        float x3 = x * x * x;
        e = p.y * x3 * (x3 - 1.0f);
    }
    if (CHARGED && r2 < rc_ew2) {
        // pair_erfc_dev in Float32 (src/ewald.jl:242-254): the reference's 8-term series
        constexpr float XLO = (float)PAIR_ERFC_XLO;
        constexpr float XA = 4.0f / (1.0f - XLO);
        constexpr float XB = -2.0f * (1.0f + XLO) / (1.0f - XLO) - 1.0f;
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float z = alpha * (r2 * rinv);
        const float tt = 2.0f * __builtin_amdgcn_rcpf(2.0f + z);
        const float xi = XA * tt + XB;
        const float xi2 = 2.0f * xi;
        float d = 0.0f, dd = 0.0f;
#pragma unroll
        for (int jj = 7; jj >= 1; --jj) {
            float tmp = d;
            d = xi2 * d - dd + pair_erfc_c32(jj);
            dd = tmp;
        }
        const float arg = xi * d - dd + 0.5f * pair_erfc_c32(0) - alpha2 * r2;
        const float erfcz = tt * __builtin_amdgcn_exp2f(arg * 1.4426950408889634f);
        e += qke[site] * qj * erfcz * rinv;
    }
    return e;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NS, bool CHARGED, bool UPPER>
__global__ __launch_bounds__(MIX_WAVE) void widom_mixed_kernel(WidomArgs a)
{
    constexpr int RING = ring_size<NS>();
    __shared__ float ring_r2[RING];
    __shared__ float ring_q[RING];
    __shared__ int ring_info[RING];
    __shared__ float2 pair_lds[NS * MAX_TYPES_MIXED];
    __shared__ float qke_lds[NS];

    const int lane = threadIdx.x;
    const int ntc = a.ntypes_c;
    for (int i = lane; i < NS * ntc; i += MIX_WAVE) pair_lds[i] = a.pair32[i];
    if (lane < NS) qke_lds[lane] = a.guest.qke[lane];
    __syncthreads();

    const float rc_lj2 = a.rc_lj2f, rc_ew2 = a.rc_ew2f, rc_max2 = a.rc_max2f;
    const int64_t gwave = a.g0 + (int64_t)blockIdx.x * MIX_INS_PER_WAVE;
    const uint64_t lanebit = 1ull << lane;
    const uint64_t below = lanebit - 1ull;

    for (int ii = 0; ii < MIX_INS_PER_WAVE; ++ii) {
        const int64_t g = gwave + ii;
        if (g >= a.g1) break;
        // ---- pose (uniform across the wave; every lane computes the same FP64 values)
        const int s = find_system(a.seg, a.nsys, g);
        const int64_t j = g - a.seg[s];
        const int fwi = __builtin_amdgcn_readfirstlane(a.fw_of[s]);
        const FwDesc &F = a.fw[fwi];
        double f[3], q[4];
        pose_uniforms(a.seed, s, j, f, q);
        uint32_t sx[NS], sy[NS], sz[NS];
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            double r[3], df[3];
            rotate(q, a.guest.site[t], r);
            matvec(F.invA, r, df);
            sx[t] = __builtin_amdgcn_readfirstlane(to_fixed(f[0] + df[0]));
            sy[t] = __builtin_amdgcn_readfirstlane(to_fixed(f[1] + df[1]));
            sz[t] = __builtin_amdgcn_readfirstlane(to_fixed(f[2] + df[2]));
        }
        const float A0 = F.Af[0], A3 = F.Af[3], A6 = F.Af[6], A4 = F.Af[4], A7 = F.Af[7], A8 = F.Af[8];
        const float A1 = F.Af[1], A2 = F.Af[2], A5 = F.Af[5];
        const float alpha = F.alphaf, alpha2 = F.alphaf * F.alphaf;
        const uint4 *atoms = a.atomfx + F.atom_off;
        const int *types = a.atype + F.atom_off;
        const int natoms = F.natoms;

        double acc = 0.0;
        uint32_t head = 0, tail = 0;
        for (int base = 0; base < natoms; base += MIX_WAVE) {
            const int n = base + lane;
            const bool valid = n < natoms;
            uint4 at = make_uint4(0u, 0u, 0u, 0u);
            int ht = 0;
            if (valid) {
                at = atoms[n];
                ht = types[n];
            }
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                const float dx = (float)(int32_t)(sx[t] - at.x);
                const float dy = (float)(int32_t)(sy[t] - at.y);
                const float dz = (float)(int32_t)(sz[t] - at.z);
                float ex, ey, ez;
                if (UPPER) {
                    ex = A0 * dx + A3 * dy + A6 * dz;
                    ey = A4 * dy + A7 * dz;
                    ez = A8 * dz;
                } else {
                    ex = A0 * dx + A3 * dy + A6 * dz;
                    ey = A1 * dx + A4 * dy + A7 * dz;
                    ez = A2 * dx + A5 * dy + A8 * dz;
                }
                const float r2 = ex * ex + ey * ey + ez * ez;
                const bool in = valid && (r2 < rc_max2);
                const uint64_t mask = __ballot(in);
                if (in) {
                    const uint32_t slot = (tail + (uint32_t)__popcll(mask & below)) & (RING - 1);
                    ring_r2[slot] = r2;
                    ring_q[slot] = __uint_as_float(at.w);
                    ring_info[slot] = (t << 16) | ht;
                }
                tail += (uint32_t)__popcll(mask);
            }
            __syncthreads();
            while (tail - head >= (uint32_t)MIX_WAVE) {
                const uint32_t slot = (head + lane) & (RING - 1);
                acc += (double)pair_energy_f32<CHARGED>(ring_r2[slot], ring_q[slot], ring_info[slot], pair_lds, ntc,
                                                        qke_lds, rc_lj2, rc_ew2, alpha, alpha2);
                head += MIX_WAVE;
            }
            __syncthreads();
        }
        if ((uint32_t)lane < tail - head) {
            const uint32_t slot = (head + lane) & (RING - 1);
            acc += (double)pair_energy_f32<CHARGED>(ring_r2[slot], ring_q[slot], ring_info[slot], pair_lds, ntc,
                                                    qke_lds, rc_lj2, rc_ew2, alpha, alpha2);
        }
        __syncthreads();
        if (CHARGED) {
            const int4 *kn = a.kn + F.k_off;
            const float2 *kc = a.kc + F.k_off;
            for (int kb = 0; kb < F.nk; kb += MIX_WAVE) {
                const int k = kb + lane;
                if (k < F.nk) {
                    const int4 nn = kn[k];
                    const float2 c = kc[k];
                    float re = 0.0f, im = 0.0f;
#pragma unroll
                    for (int t = 0; t < NS; ++t) {
                        const uint32_t ph = (uint32_t)nn.x * sx[t] + (uint32_t)nn.y * sy[t] + (uint32_t)nn.z * sz[t];
                        const float rev = (float)(int32_t)ph * 2.3283064365386963e-10f;
                        const float qs = qke_lds[t];
                        re += qs * __builtin_amdgcn_cosf(rev);
                        im += qs * __builtin_amdgcn_sinf(rev);
                    }
                    acc += (double)(c.x * re + c.y * im);
                }
            }
        }
        const double E = wave_sum(acc);
        if (lane == 0) a.dU[g - a.out_base] = E + F.const_off;
    }
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
// {max log w, sum w e^-max, sum dU w e^-max, NaN count}.
__global__ __launch_bounds__(TILE_THREADS) void boltzmann_tiles_kernel(const double *dU, const int64_t *tile_g0,
                                                                       const int *tile_n, double kT, double4 *out)
{
    __shared__ double sm[TILE_THREADS], sw[TILE_THREADS], su[TILE_THREADS];
    __shared__ int snan[TILE_THREADS];
    const int t = blockIdx.x, tid = threadIdx.x;
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
    sm[tid] = acc.m; sw[tid] = acc.w; su[tid] = acc.u; snan[tid] = nan;
    __syncthreads();
    for (int stride = TILE_THREADS / 2; stride >= 1; stride >>= 1) {
        if (tid < stride) {
            Lse x = {sm[tid], sw[tid], su[tid]}, y = {sm[tid + stride], sw[tid + stride], su[tid + stride]};
            Lse r = lse_merge(x, y);
            sm[tid] = r.m; sw[tid] = r.w; su[tid] = r.u;
            snan[tid] += snan[tid + stride];
        }
        __syncthreads();
    }
    if (tid == 0) out[t] = make_double4(sm[0], sw[0], su[0], (double)snan[0]);
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
static void launch_ns(const WidomArgs &a, int precision, bool upper, hipStream_t st)
{
    const int64_t n = a.g1 - a.g0;
    if (precision == 0) {
        const dim3 grid((unsigned)((n + 255) / 256)), block(256);
        if (a.guest.charged) hipLaunchKernelGGL((widom_f64_kernel<NS, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((widom_f64_kernel<NS, false>), grid, block, 0, st, a);
        return;
    }
    const dim3 grid((unsigned)((n + MIX_INS_PER_WAVE - 1) / MIX_INS_PER_WAVE)), block(MIX_WAVE);
    if (a.guest.charged) {
        if (upper) hipLaunchKernelGGL((widom_mixed_kernel<NS, true, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((widom_mixed_kernel<NS, true, false>), grid, block, 0, st, a);
    } else {
        if (upper) hipLaunchKernelGGL((widom_mixed_kernel<NS, false, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((widom_mixed_kernel<NS, false, false>), grid, block, 0, st, a);
    }
}

// Energy launch over g in [a.g0, a.g1). `upper`: every framework cell is upper triangular.
hipError_t launch_energy(const WidomArgs &a, int precision, bool upper, hipStream_t st)
{
    if (a.g1 <= a.g0) return hipSuccess;
    switch (a.guest.nsites) {
    case 1: launch_ns<1>(a, precision, upper, st); break;
    case 2: launch_ns<2>(a, precision, upper, st); break;
    case 3: launch_ns<3>(a, precision, upper, st); break;
    case 4: launch_ns<4>(a, precision, upper, st); break;
    case 5: launch_ns<5>(a, precision, upper, st); break;
    case 6: launch_ns<6>(a, precision, upper, st); break;
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
