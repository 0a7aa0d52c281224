// widom_grad.hip -- adjoint kernel of the Widom estimators (gradients of ln K_H and q_st).
//
// The host runs the energy pipeline first, so every insertion's Delta U is known, and lists the
// insertions whose Boltzmann weight relative to the system's best insertion matters in FP64
// (w = exp(-beta U - m_s) > e^-36). This kernel takes one such insertion per wavefront:
//  * lanes sweep the framework's atoms with the exact 32-bit fixed-point minimum image; pairs
//    inside the cutoff are compacted into an LDS ring (ballot + mbcnt) and evaluated on full
//    wavefronts;
//  * each pair adds its Lennard-Jones derivatives dU/d eps_sh = 4 (x^12 - x^6) and
//    dU/d sigma_sh = 4 eps (12 x^12 - 6 x^6) / sigma into per-(site, host type) sums in LDS,
//    and its Coulomb factor dU/dq_j = KE q_s erfc(alpha r)/r, times w and w U, into the atom's
//    FP64 accumulators (global atomics);
//  * lanes over the coupled k-vectors add w S_g(k) and w U S_g(k) (the guest structure factor)
//    into per-system FP64 accumulators, from which the host forms the reciprocal part
//    KE sum_k kprefactor 2 Re(conj(e^{i k r_j}) A_k) of every atom's charge derivative.
// Pair and k-space terms are FP32; every sum is FP64.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "pa_device.hpp"

namespace pa {

constexpr int GRAD_RING = 512;   // holds a partial wave (< 64) plus one 64-atom sweep of 6 sites

__device__ inline int find_sys_g(const int64_t *seg, int nsys, int64_t g)
{
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int NS, bool CHARGED, bool UPPER>
__global__ __launch_bounds__(64) void widom_grad_kernel(GradArgs a)
{
    __shared__ float ring_r2[GRAD_RING];
    __shared__ int ring_atom[GRAD_RING];
    __shared__ int ring_site[GRAD_RING];
    __shared__ double lj_sum[NS * 64 * 2];   // per (site, type): dU/deps, dU/dsig for this insertion
    __shared__ float4 pair_lds[NS * 64];

    static_assert(63 + 64 * NS <= GRAD_RING, "ring too small for one sweep");
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const int ntc = a.ntypes_c;
    for (int i = lane; i < NS * ntc; i += 64) pair_lds[i] = a.pairg[i];
    for (int i = lane; i < NS * ntc * 2; i += 64) lj_sum[i] = 0.0;
    __syncthreads();

    const int64_t idx = blockIdx.x;
    if (idx >= a.nlist) return;
    const int64_t g = a.list_g[idx];
    const int s = find_sys_g(a.seg, a.nsys, g);
    const double du = a.dU[g];
    const double w = exp(-a.beta * du - a.msys[s]), wu = w * du;
    const int64_t j = g - a.seg[s];
    const FwDesc &F = a.fw[a.fw_of[s]];
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
    const float A0 = F.Af[0], A1 = F.Af[1], A2 = F.Af[2], A3 = F.Af[3], A4 = F.Af[4], A5 = F.Af[5], A6 = F.Af[6],
                A7 = F.Af[7], A8 = F.Af[8];
    const float alpha = F.alphaf;
    const uint4 *atoms = a.atomfx + F.atom_off;
    const int *types = a.atype + F.atom_off;
    double *accq = a.acc_q + 2 * a.qoff[s];

    auto drain = [&](uint32_t slot, bool active) {
        if (!active) return;
        float r2 = fmaxf(ring_r2[slot], 1.0e-30f);
        const int n = ring_atom[slot], site = ring_site[slot];
        const int ht = types[n];
        if (r2 < a.rc_lj2f) {
            const float4 p = pair_lds[site * ntc + ht];
            const float x2 = p.x * __builtin_amdgcn_rcpf(r2);
            const float x6 = x2 * x2 * x2;
            const float dde = 4.0f * x6 * (x6 - 1.0f);                       // dU/d eps_sh
            const float dds = p.y * (12.0f * x6 * x6 - 6.0f * x6) * p.z;     // dU/d sigma_sh
            unsafeAtomicAdd(&lj_sum[2 * (site * ntc + ht)], (double)dde);
            unsafeAtomicAdd(&lj_sum[2 * (site * ntc + ht) + 1], (double)dds);
        }
        if (CHARGED && r2 < a.rc_ew2f) {
            constexpr float XLO = (float)PAIR_ERFC_XLO;
            constexpr float XA = 4.0f / (1.0f - XLO);
            constexpr float XB = -2.0f * (1.0f + XLO) / (1.0f - XLO) - 1.0f;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float z = alpha * (r2 * rinv);
            const float tt = 2.0f * __builtin_amdgcn_rcpf(2.0f + z);
            const float xi = XA * tt + XB, xi2 = 2.0f * xi;
            float d = 0.0f, dd = 0.0f;
#pragma unroll
            for (int jj = 7; jj >= 1; --jj) {
                const float tmp = d;
                d = xi2 * d - dd + pair_erfc_c32(jj);
                dd = tmp;
            }
            const float arg = xi * d - dd + 0.5f * pair_erfc_c32(0) - alpha * alpha * r2;
            const double fq = (double)(a.guest.qke[site] * (tt * __builtin_amdgcn_exp2f(arg * 1.4426950408889634f)) * rinv);
            unsafeAtomicAdd(&accq[2 * n], w * fq);
            unsafeAtomicAdd(&accq[2 * n + 1], wu * fq);
        }
    };

    uint32_t head = 0, tail = 0;
    for (int base = 0; base < F.natoms; base += 64) {
        const int n = base + lane;
        const bool valid = n < F.natoms;
        uint4 at = make_uint4(0u, 0u, 0u, 0u);
        if (valid) at = atoms[n];
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const float dx = (float)(int32_t)(sx[t] - at.x);
            const float dy = (float)(int32_t)(sy[t] - at.y);
            const float dz = (float)(int32_t)(sz[t] - at.z);
            const float ex = A0 * dx + A3 * dy + A6 * dz;
            const float ey = UPPER ? A4 * dy + A7 * dz : A1 * dx + A4 * dy + A7 * dz;
            const float ez = UPPER ? A8 * dz : A2 * dx + A5 * dy + A8 * dz;
            const float r2 = ex * ex + ey * ey + ez * ez;
            const bool in = valid && r2 < a.rc_max2f;
            const uint64_t mask = __ballot(in);
            if (in) {
                const uint32_t slot = (tail + (uint32_t)__popcll(mask & below)) & (GRAD_RING - 1);
                ring_r2[slot] = r2;
                ring_atom[slot] = n;
                ring_site[slot] = t;
            }
            tail += (uint32_t)__popcll(mask);
        }
        __syncthreads();
        while (tail - head >= 64u) {
            drain((head + lane) & (GRAD_RING - 1), true);
            head += 64u;
        }
        __syncthreads();
    }
    drain((head + lane) & (GRAD_RING - 1), (uint32_t)lane < tail - head);
    __syncthreads();
    // per-(site, type) LJ sums of this insertion, weighted into the system's accumulators
    double *acclj = a.acc_lj + (size_t)s * NS * ntc * 4;
    for (int c = lane; c < NS * ntc; c += 64) {
        const double de = lj_sum[2 * c], ds = lj_sum[2 * c + 1];
        if (de != 0.0 || ds != 0.0) {
            unsafeAtomicAdd(&acclj[4 * c + 0], w * de);
            unsafeAtomicAdd(&acclj[4 * c + 1], wu * de);
            unsafeAtomicAdd(&acclj[4 * c + 2], w * ds);
            unsafeAtomicAdd(&acclj[4 * c + 3], wu * ds);
        }
    }
    if (CHARGED) {
        double *acck = a.acc_k + 4 * a.koff[s];
        const int4 *kn = a.kn + F.k_off;
        for (int kb = 0; kb < F.nk; kb += 64) {
            const int k = kb + lane;
            if (k < F.nk) {
                const int4 nn = kn[k];
                float re = 0.0f, im = 0.0f;
#pragma unroll
                for (int t = 0; t < NS; ++t) {
                    const uint32_t ph = (uint32_t)nn.x * sx[t] + (uint32_t)nn.y * sy[t] + (uint32_t)nn.z * sz[t];
                    const float rev = (float)(int32_t)ph * 2.3283064365386963e-10f;
                    const float qs = (float)a.guest.q[t];
                    re += qs * __builtin_amdgcn_cosf(rev);
                    im += qs * __builtin_amdgcn_sinf(rev);
                }
                unsafeAtomicAdd(&acck[4 * k + 0], w * re);
                unsafeAtomicAdd(&acck[4 * k + 1], w * im);
                unsafeAtomicAdd(&acck[4 * k + 2], wu * re);
                unsafeAtomicAdd(&acck[4 * k + 3], wu * im);
            }
        }
    }
}

// Flags the insertions whose weight relative to their system's best insertion exceeds e^-40.
__global__ __launch_bounds__(256) void grad_flag_kernel(const double *dU, const int64_t *seg, int nsys, const double *msys,
                                                        double beta, int64_t G, unsigned char *flag)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= G) return;
    const int s = find_sys_g(seg, nsys, g);
    const double x = -beta * dU[g] - msys[s];
    flag[g] = x > -40.0 ? 1 : 0;
}

// Reciprocal part of every framework atom's charge derivative, one work-item per (system, atom):
// KE sum_k kprefactor 2 (cos(k.r_j) A_re(k) + sin(k.r_j) A_im(k)), with k.r_j = 2 pi n.f_j taken
// from the fixed-point coordinates (an exact wrapping integer phase), FP64 sin/cos of pi x.
__global__ __launch_bounds__(256) void grad_recip_kernel(GradArgs a, const double4 *k64, int64_t natoms_total)
{
#pragma clang fp contract(off)
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x >= natoms_total) return;
    int lo = 0, hi = a.nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.qoff[mid] <= x) lo = mid; else hi = mid - 1;
    }
    const int s = lo;
    const int j = (int)(x - a.qoff[s]);
    const FwDesc &F = a.fw[a.fw_of[s]];
    const uint4 at = a.atomfx[F.atom_off + j];
    const int4 *kn = a.kn + F.k_off;
    const double4 *kp = k64 + F.k_off;
    const double *acck = a.acc_k + 4 * a.koff[s];
    double gw = 0.0, gu = 0.0;
    for (int k = 0; k < F.nk; ++k) {
        const int4 nn = kn[k];
        const uint32_t ph = (uint32_t)nn.x * at.x + (uint32_t)nn.y * at.y + (uint32_t)nn.z * at.z;
        const double rev = (double)(int32_t)ph * 2.3283064365386963e-10;
        double sn, cs;
        sincospi(2.0 * rev, &sn, &cs);
        const double f = 2.0 * kp[k].w;
        gw += f * (cs * acck[4 * k] + sn * acck[4 * k + 1]);
        gu += f * (cs * acck[4 * k + 2] + sn * acck[4 * k + 3]);
    }
    double *accq = a.acc_q + 2 * a.qoff[s];
    accq[2 * j] += KE * gw;
    accq[2 * j + 1] += KE * gu;
}

hipError_t grad_select(const double *dU, const int64_t *seg, int nsys, const double *msys, double beta, int64_t G,
                       unsigned char *flag, int64_t *list, int64_t *count, void *tmp, size_t *tmp_bytes, hipStream_t st)
{
    if (!tmp) {
        return hipcub::DeviceSelect::Flagged(nullptr, *tmp_bytes, hipcub::CountingInputIterator<int64_t>(0), flag, list,
                                             count, G, st);
    }
    hipLaunchKernelGGL(grad_flag_kernel, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, st, dU, seg, nsys, msys, beta,
                       G, flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceSelect::Flagged(tmp, *tmp_bytes, hipcub::CountingInputIterator<int64_t>(0), flag, list, count,
                                         G, st);
}

hipError_t launch_grad_recip(const GradArgs &a, const double4 *k64, int64_t natoms_total, hipStream_t st)
{
    if (natoms_total <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_recip_kernel, dim3((unsigned)((natoms_total + 255) / 256)), dim3(256), 0, st, a, k64,
                       natoms_total);
    return hipGetLastError();
}

template <int NS>
static void launch_grad_ns(const GradArgs &a, bool upper, hipStream_t st)
{
    const dim3 grid((unsigned)a.nlist), block(64);
    if (a.guest.charged) {
        if (upper) hipLaunchKernelGGL((widom_grad_kernel<NS, true, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((widom_grad_kernel<NS, true, false>), grid, block, 0, st, a);
    } else {
        if (upper) hipLaunchKernelGGL((widom_grad_kernel<NS, false, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((widom_grad_kernel<NS, false, false>), grid, block, 0, st, a);
    }
}

hipError_t launch_grad(const GradArgs &a, bool upper, hipStream_t st)
{
    if (a.guest.nsites < 1 || a.guest.nsites > 6) return hipErrorInvalidValue;
    // one block per weighted insertion, at most 2^24 blocks per launch
    const int64_t CH = (int64_t)1 << 24;
    for (int64_t off = 0; off < a.nlist; off += CH) {
        GradArgs c = a;
        c.list_g = a.list_g + off;
        c.nlist = std::min(CH, a.nlist - off);
        switch (a.guest.nsites) {
        case 1: launch_grad_ns<1>(c, upper, st); break;
        case 2: launch_grad_ns<2>(c, upper, st); break;
        case 3: launch_grad_ns<3>(c, upper, st); break;
        case 4: launch_grad_ns<4>(c, upper, st); break;
        case 5: launch_grad_ns<5>(c, upper, st); break;
        default: launch_grad_ns<6>(c, upper, st); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace pa
