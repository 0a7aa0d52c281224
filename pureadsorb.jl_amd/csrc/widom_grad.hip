// widom_grad.hip -- adjoint kernels of the Widom estimators (gradients of ln K_H and q_st).
//
// The Widom pipeline runs first, so every insertion's Delta U is known. grad_select lists, on
// the device, the insertions whose weight relative to their system's best insertion matters in
// FP64 (w = exp(-beta U - m_s) > e^-40). Then, in gather form (no per-pair atomics):
//  * grad_prep_kernel stores every listed insertion's sites as 32-bit fixed-point fractional
//    coordinates and its (w, w dU);
//  * widom_grad_gather_kernel: one workgroup per (block of 256 framework atoms of one system,
//    chunk of that system's listed insertions), atoms in hard-core grid-cell order, insertions
//    streamed through LDS in tiles; each work-item owns one atom and
//    sums over every (insertion, site) pair inside the cutoff, with the exact fixed-point minimum
//    image: the Coulomb factor KE q_s erfc(alpha r)/r times w and w dU into the atom's own FP64
//    sums, and the Lennard-Jones derivatives dU/d eps_sh = 4 (x^12 - x^6) and
//    dU/d sigma_sh = 4 eps (12 x^12 - 6 x^6)/sigma times w and w dU into per-site sums of the
//    atom's type, reduced over the workgroup per (site, type);
//  * grad_kspace_kernel: one workgroup per (256 coupled k-vectors of one system, chunk of that
//    system's listed insertions) sums w S_g(k) and w dU S_g(k), one work-item per k-vector;
//  * grad_recip_kernel turns those into every atom's reciprocal charge derivative.
// Pair and k-space terms are FP32; every sum is FP64.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "pa_device.hpp"

namespace pa {

constexpr int GRAD_WG = 256;
constexpr int GRAD_TILE = 256;

__device__ inline int find_sys_g(const int64_t *seg, int nsys, int64_t g)
{
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int NS>
__global__ __launch_bounds__(256) void grad_prep_kernel(GradArgs a)
{
#pragma clang fp contract(off)
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.nlist) return;
    const int64_t g = a.list_g[idx];
    const int s = find_sys_g(a.seg, a.nsys, g);
    const int64_t j = g - a.seg[s];
    const FwDesc &F = a.fw[a.fw_of[s]];
    double f[3], q[4];
    pose_uniforms(a.seed, s, j, f, q);
    uint32_t *out = a.lsite + (size_t)idx * NS * 3;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        double r[3], df[3];
        rotate(q, a.guest.site[t], r);
        matvec(F.invA, r, df);
        for (int d = 0; d < 3; ++d) out[3 * t + d] = to_fixed(f[d] + df[d]);
    }
    const double du = a.dU[g];
    const double w = exp(-a.beta * du - a.msys[s]);
    a.lw[idx] = make_double2(w, w * du);
}

template <int NS, bool CHARGED, bool UPPER>
__global__ __launch_bounds__(GRAD_WG) void widom_grad_gather_kernel(GradArgs a)
{
    __shared__ uint32_t t_site[GRAD_TILE * NS * 3];
    __shared__ double2 t_w[GRAD_TILE];
    __shared__ double lj_red[NS * 64 * 4];
    const int tid = threadIdx.x;
    const int4 bk = a.blk[blockIdx.x];   // system, first sorted atom, list range
    const int s = bk.x;
    const FwDesc &F = a.fw[a.fw_of[s]];
    const int ntc = a.ntypes_c;
    for (int i = tid; i < NS * ntc * 4; i += GRAD_WG) lj_red[i] = 0.0;
    const int j = bk.y + tid;
    const bool valid = j < F.natoms;
    uint4 at = make_uint4(0u, 0u, 0u, 0u);
    if (valid) at = a.corefx[F.atom_off + j];
    const int ht = valid ? (int)at.w : 0;
    float4 pr[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) pr[t] = a.pairg[t * ntc + ht];
    const float A0 = F.Af[0], A1 = F.Af[1], A2 = F.Af[2], A3 = F.Af[3], A4 = F.Af[4], A5 = F.Af[5], A6 = F.Af[6],
                A7 = F.Af[7], A8 = F.Af[8];
    const float alpha = F.alphaf;
    double ge[NS], he[NS], gs[NS], hs[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) ge[t] = he[t] = gs[t] = hs[t] = 0.0;
    double gq = 0.0, hq = 0.0;
    const int64_t l0 = a.loff[s] + bk.z, l1 = a.loff[s] + bk.w;
    for (int64_t base = l0; base < l1; base += GRAD_TILE) {
        const int nt = (int)((l1 - base) < GRAD_TILE ? (l1 - base) : GRAD_TILE);
        __syncthreads();
        for (int i = tid; i < nt * NS * 3; i += GRAD_WG) t_site[i] = a.lsite[(size_t)base * NS * 3 + i];
        for (int i = tid; i < nt; i += GRAD_WG) t_w[i] = a.lw[base + i];
        __syncthreads();
        if (!valid) continue;
        for (int i = 0; i < nt; ++i) {
            const double2 w = t_w[i];
            float fq = 0.0f;
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                const float dx = (float)(int32_t)(t_site[(i * NS + t) * 3 + 0] - at.x);
                const float dy = (float)(int32_t)(t_site[(i * NS + t) * 3 + 1] - at.y);
                const float dz = (float)(int32_t)(t_site[(i * NS + t) * 3 + 2] - at.z);
                const float ex = A0 * dx + A3 * dy + A6 * dz;
                const float ey = UPPER ? A4 * dy + A7 * dz : A1 * dx + A4 * dy + A7 * dz;
                const float ez = UPPER ? A8 * dz : A2 * dx + A5 * dy + A8 * dz;
                const float r2 = fmaxf(ex * ex + ey * ey + ez * ez, 1.0e-30f);
                if (r2 < a.rc_lj2f) {
                    const float x2 = pr[t].x * __builtin_amdgcn_rcpf(r2);
                    const float x6 = x2 * x2 * x2;
                    const float de = 4.0f * x6 * (x6 - 1.0f);
                    const float ds = pr[t].y * (12.0f * x6 * x6 - 6.0f * x6) * pr[t].z;
                    ge[t] += w.x * (double)de;
                    he[t] += w.y * (double)de;
                    gs[t] += w.x * (double)ds;
                    hs[t] += w.y * (double)ds;
                }
                if (CHARGED && r2 < a.rc_ew2f) {
                    constexpr float XLO = (float)PAIR_ERFC_XLO;
                    constexpr float XA = 4.0f / (1.0f - XLO);
                    constexpr float XB = -2.0f * (1.0f + XLO) / (1.0f - XLO) - 1.0f;
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    const float z = alpha * (r2 * rinv);
                    const float tt = 2.0f * __builtin_amdgcn_rcpf(2.0f + z);
                    const float xi = XA * tt + XB;
                    const float arg = pair_erfc_poly32(xi) - alpha * alpha * r2;
                    fq += a.guest.qke[t] * (tt * __builtin_amdgcn_exp2f(arg * 1.4426950408889634f)) * rinv;
                }
            }
            if (CHARGED) {
                gq += w.x * (double)fq;
                hq += w.y * (double)fq;
            }
        }
    }
    __syncthreads();
    if (valid) {
        const int orig = a.coreidx[F.atom_off + j];
        double *accq = a.acc_q + 2 * a.qoff[s];
        if (CHARGED) {
            unsafeAtomicAdd(&accq[2 * orig], gq);
            unsafeAtomicAdd(&accq[2 * orig + 1], hq);
        }
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            double *r = lj_red + 4 * (t * ntc + ht);
            if (ge[t] != 0.0 || gs[t] != 0.0) {
                unsafeAtomicAdd(r + 0, ge[t]);
                unsafeAtomicAdd(r + 1, he[t]);
                unsafeAtomicAdd(r + 2, gs[t]);
                unsafeAtomicAdd(r + 3, hs[t]);
            }
        }
    }
    __syncthreads();
    double *acclj = a.acc_lj + (size_t)s * NS * ntc * 4;
    for (int i = tid; i < NS * ntc * 4; i += GRAD_WG)
        if (lj_red[i] != 0.0) unsafeAtomicAdd(&acclj[i], lj_red[i]);
}

// Sums of w S_g(k) and w dU S_g(k) over a chunk of a system's listed insertions, one work-item
// per coupled k-vector, the chunk's sites staged through LDS; exact wrapping fixed-point phases,
// hardware sin/cos of revolutions.
template <int NS>
__global__ __launch_bounds__(256) void grad_kspace_kernel(GradArgs a)
{
    __shared__ uint32_t t_site[GRAD_TILE * NS * 3];
    __shared__ double2 t_w[GRAD_TILE];
    const int tid = threadIdx.x;
    const int4 bk = a.kblk[blockIdx.x];   // system, first k-vector, list range
    const int s = bk.x;
    const FwDesc &F = a.fw[a.fw_of[s]];
    const int k = bk.y + tid;
    const bool valid = k < F.nk;
    int4 nn = make_int4(0, 0, 0, 0);
    if (valid) nn = a.kn[F.k_off + k];
    float qs[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) qs[t] = (float)a.guest.q[t];
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    const int64_t l0 = a.loff[s] + bk.z, l1 = a.loff[s] + bk.w;
    for (int64_t base = l0; base < l1; base += GRAD_TILE) {
        const int nt = (int)((l1 - base) < GRAD_TILE ? (l1 - base) : GRAD_TILE);
        __syncthreads();
        for (int i = tid; i < nt * NS * 3; i += 256) t_site[i] = a.lsite[(size_t)base * NS * 3 + i];
        for (int i = tid; i < nt; i += 256) t_w[i] = a.lw[base + i];
        __syncthreads();
        if (!valid) continue;
        auto add = [&](int i, float re, float im) {
            const double2 w = t_w[i];
            ar += w.x * (double)re;
            ai += w.x * (double)im;
            br += w.y * (double)re;
            bi += w.y * (double)im;
        };
        bool folded = false;
        if constexpr (NS == 3) {
            if (a.guest.mirror) {
                // sites 1 and 2 mirror each other through site 0 with equal charges:
                // S_g = e^{i t0} (q0 + 2 q1 cos p), p the phase of site 1 relative to site 0
                const float q0 = qs[0], q1x2 = 2.0f * qs[1];
                for (int i = 0; i < nt; ++i) {
                    const uint32_t *s0 = t_site + i * 9;
                    const uint32_t ph = (uint32_t)nn.x * s0[0] + (uint32_t)nn.y * s0[1] + (uint32_t)nn.z * s0[2];
                    const uint32_t pd = (uint32_t)nn.x * (s0[3] - s0[0]) + (uint32_t)nn.y * (s0[4] - s0[1]) +
                                        (uint32_t)nn.z * (s0[5] - s0[2]);
                    const float rev = (float)(int32_t)ph * 2.3283064365386963e-10f;
                    const float rvd = (float)(int32_t)pd * 2.3283064365386963e-10f;
                    const float wq = q0 + q1x2 * __builtin_amdgcn_cosf(rvd);
                    add(i, wq * __builtin_amdgcn_cosf(rev), wq * __builtin_amdgcn_sinf(rev));
                }
                folded = true;
            }
        }
        for (int i = 0; i < (folded ? 0 : nt); ++i) {
            float re = 0.0f, im = 0.0f;
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                const uint32_t *sf = t_site + (i * NS + t) * 3;
                const uint32_t ph = (uint32_t)nn.x * sf[0] + (uint32_t)nn.y * sf[1] + (uint32_t)nn.z * sf[2];
                const float rev = (float)(int32_t)ph * 2.3283064365386963e-10f;
                re += qs[t] * __builtin_amdgcn_cosf(rev);
                im += qs[t] * __builtin_amdgcn_sinf(rev);
            }
            add(i, re, im);
        }
    }
    if (!valid) return;
    double *acck = a.acc_k + 4 * (a.koff[s] + k);
    unsafeAtomicAdd(acck + 0, ar);
    unsafeAtomicAdd(acck + 1, ai);
    unsafeAtomicAdd(acck + 2, br);
    unsafeAtomicAdd(acck + 3, bi);
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

// Reciprocal part of every framework atom's charge derivative:
// KE sum_k kprefactor 2 (cos(k.r_j) A_re(k) + sin(k.r_j) A_im(k)), with k.r_j = 2 pi n.f_j taken
// from the fixed-point coordinates (an exact wrapping integer phase), FP64 sin/cos of pi x. One
// workgroup per (system, 256 atoms, 256 k-vectors): the chunk's k-vectors, prefactors and
// accumulated A(k) are staged in LDS once and read by every lane at the same address, and each
// chunk's partial sums join the atom's accumulator atomically.
constexpr int RECIP_KCH = 256;
__global__ __launch_bounds__(256) void grad_recip_kernel(GradArgs a, const double4 *k64, const int4 *rblk)
{
#pragma clang fp contract(off)
    __shared__ int4 skn[RECIP_KCH];
    __shared__ double sf[RECIP_KCH];
    __shared__ double4 sacc[RECIP_KCH];
    const int4 w = rblk[blockIdx.x];
    const int s = w.x, a0 = w.y, k0 = w.z;
    const FwDesc &F = a.fw[a.fw_of[s]];
    const int nkc = min(RECIP_KCH, F.nk - k0);
    const int t = threadIdx.x;
    if (t < nkc) {
        skn[t] = a.kn[F.k_off + k0 + t];
        sf[t] = 2.0 * k64[F.k_off + k0 + t].w;
        const double *acck = a.acc_k + 4 * (a.koff[s] + k0 + t);
        sacc[t] = make_double4(acck[0], acck[1], acck[2], acck[3]);
    }
    __syncthreads();
    const int j = a0 + t;
    if (j >= F.natoms) return;
    const uint4 at = a.atomfx[F.atom_off + j];
    double gw = 0.0, gu = 0.0;
    for (int k = 0; k < nkc; ++k) {
        const int4 nn = skn[k];
        const uint32_t ph = (uint32_t)nn.x * at.x + (uint32_t)nn.y * at.y + (uint32_t)nn.z * at.z;
        const double rev = (double)(int32_t)ph * 2.3283064365386963e-10;
        double sn, cs;
        sincospi(2.0 * rev, &sn, &cs);
        const double f = sf[k];
        const double4 c = sacc[k];
        gw += f * (cs * c.x + sn * c.y);
        gu += f * (cs * c.z + sn * c.w);
    }
    double *accq = a.acc_q + 2 * a.qoff[s];
    unsafeAtomicAdd(accq + 2 * j, KE * gw);
    unsafeAtomicAdd(accq + 2 * j + 1, KE * gu);
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

// Each system's range of the (ascending) weighted-insertion list: loff[s] = first listed g >= seg[s].
__global__ __launch_bounds__(256) void grad_loff_kernel(const int64_t *list, int64_t nlist, const int64_t *seg, int nsys,
                                                        int64_t *loff)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s > nsys) return;
    if (s == nsys) {
        loff[s] = nlist;
        return;
    }
    const int64_t key = seg[s];
    int64_t lo = 0, hi = nlist;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (list[mid] < key) lo = mid + 1; else hi = mid;
    }
    loff[s] = lo;
}

hipError_t grad_offsets(const int64_t *list, int64_t nlist, const int64_t *seg, int nsys, int64_t *loff, hipStream_t st)
{
    hipLaunchKernelGGL(grad_loff_kernel, dim3((unsigned)((nsys + 1 + 255) / 256)), dim3(256), 0, st, list, nlist, seg,
                       nsys, loff);
    return hipGetLastError();
}

// rblk: (system, first atom, first k-vector, 0) of each workgroup; atoms and k-vectors in
// steps of 256 within each system's ranges.
hipError_t launch_grad_recip(const GradArgs &a, const double4 *k64, const int4 *rblk, int nrblk, hipStream_t st)
{
    if (nrblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_recip_kernel, dim3((unsigned)nrblk), dim3(256), 0, st, a, k64, rblk);
    return hipGetLastError();
}


template <int NS>
static void launch_grad_ns(const GradArgs &a, bool upper, int nblk, int nkblk, hipStream_t st)
{
    if (a.nlist > 0)
        hipLaunchKernelGGL(grad_prep_kernel<NS>, dim3((unsigned)((a.nlist + 255) / 256)), dim3(256), 0, st, a);
    const dim3 grid((unsigned)nblk), block(GRAD_WG);
    if (nblk > 0) {
        if (a.guest.charged) {
            if (upper) hipLaunchKernelGGL((widom_grad_gather_kernel<NS, true, true>), grid, block, 0, st, a);
            else hipLaunchKernelGGL((widom_grad_gather_kernel<NS, true, false>), grid, block, 0, st, a);
        } else {
            if (upper) hipLaunchKernelGGL((widom_grad_gather_kernel<NS, false, true>), grid, block, 0, st, a);
            else hipLaunchKernelGGL((widom_grad_gather_kernel<NS, false, false>), grid, block, 0, st, a);
        }
    }
    if (nkblk > 0) hipLaunchKernelGGL(grad_kspace_kernel<NS>, dim3((unsigned)nkblk), dim3(256), 0, st, a);
}

// prep + gather + k-space sums; `nblk` gather workgroups described by a.blk, `nkblk` k-space
// workgroups described by a.kblk.
hipError_t launch_grad(const GradArgs &a, bool upper, int nblk, int nkblk, hipStream_t st)
{
    switch (a.guest.nsites) {
    case 1: launch_grad_ns<1>(a, upper, nblk, nkblk, st); break;
    case 2: launch_grad_ns<2>(a, upper, nblk, nkblk, st); break;
    case 3: launch_grad_ns<3>(a, upper, nblk, nkblk, st); break;
    case 4: launch_grad_ns<4>(a, upper, nblk, nkblk, st); break;
    case 5: launch_grad_ns<5>(a, upper, nblk, nkblk, st); break;
    case 6: launch_grad_ns<6>(a, upper, nblk, nkblk, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pa
