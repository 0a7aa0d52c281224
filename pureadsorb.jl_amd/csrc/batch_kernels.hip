// batch_kernels.hip -- FrameworkBatch tables on gfx950: the host structure factor at the
// k-vectors a framework keeps (src/batch.jl:173-195) and the 64-orientation guest self term
// over its full k table (src/batch.jl:360-389), both FP64, and the phase-0 near-stencil atom
// lists. k = 2 pi A^-T n, so e^{i k.r} = e^{2 pi i n.f}
// with f = A^-1 r; every phase is one own-sincospi of 2 n.f.
#include "pa_device.hpp"

namespace pa {

constexpr int BK_WG = 256;

// S(k_i) = sum_j q_j e^{2 pi i n_i.f_j} for the requested n_i (one work-item each); the atoms
// (fractional coordinates and charge) pass through LDS in tiles shared by the workgroup.
__global__ __launch_bounds__(BK_WG) void sk_kernel(const double4 *atoms, int natoms, const int4 *kreq, int nreq,
                                                   double2 *out)
{
    __shared__ double4 tile[BK_WG];
    const int i = blockIdx.x * BK_WG + threadIdx.x;
    const int4 n = i < nreq ? kreq[i] : make_int4(0, 0, 0, 0);
    const double n0 = (double)n.x, n1 = (double)n.y, n2 = (double)n.z;
    double re = 0.0, im = 0.0;
    for (int t0 = 0; t0 < natoms; t0 += BK_WG) {
        const int j = t0 + (int)threadIdx.x;
        tile[threadIdx.x] = j < natoms ? atoms[j] : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int m = natoms - t0 < BK_WG ? natoms - t0 : BK_WG;
        for (int jj = 0; jj < m; ++jj) {
            const double4 a = tile[jj];
            double sn, cs;
            sincospi(2.0 * __builtin_fma(n2, a.z, __builtin_fma(n1, a.y, n0 * a.x)), &sn, &cs);
            re = __builtin_fma(a.w, cs, re);
            im = __builtin_fma(a.w, sn, im);
        }
        __syncthreads();
    }
    if (i < nreq) out[i] = make_double2(re, im);
}

// Partial sums of kprefactor_k |sum_s q_s e^{2 pi i n_k.g_os}|^2 over 64 consecutive k-vectors
// for each of the 64 orientations o (lane = orientation; the four wavefronts take every fourth
// k-vector). g_os: fractional coordinates of site s rotated by orientation o.
__global__ __launch_bounds__(BK_WG) void self_term_kernel(const int4 *kn, const double *kpref, int nk,
                                                         const double *gsf, const double *q, int ns, double *partial)
{
    __shared__ double red[BK_WG / 64][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k0 = blockIdx.x * 64, k1 = k0 + 64 < nk ? k0 + 64 : nk;
    double acc = 0.0;
    for (int k = k0 + w; k < k1; k += BK_WG / 64) {
        const int4 n = kn[k];
        const double n0 = (double)n.x, n1 = (double)n.y, n2 = (double)n.z;
        double re = 0.0, im = 0.0;
        for (int s = 0; s < ns; ++s) {
            const double *g = gsf + 3 * (lane * ns + s);
            double sn, cs;
            sincospi(2.0 * __builtin_fma(n2, g[2], __builtin_fma(n1, g[1], n0 * g[0])), &sn, &cs);
            re = __builtin_fma(q[s], cs, re);
            im = __builtin_fma(q[s], sn, im);
        }
        acc = __builtin_fma(kpref[k], __builtin_fma(re, re, im * im), acc);
    }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0) partial[(int64_t)blockIdx.x * 64 + lane] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Phase-0 near-stencil lists: for every hard-core cell, the atoms (corefx, sorted by cell) of
// its neighbour cells in the order i2, i1, i0, written from the cell's nbr_start offset. One
// wavefront per cell; cell_fw_off (nfw + 1) maps the global cell index to its framework.
__global__ __launch_bounds__(BK_WG) void nbr_fill_kernel(const FwDesc *fw, int nfw, const int *cell_fw_off, int ncells,
                                                        const int *cell_start, const int *nbr_start,
                                                        const uint4 *corefx, uint4 *nbr_atoms)
{
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (BK_WG / 64) + (threadIdx.x >> 6);
    if (g >= ncells) return;
    int lo = 0, hi = nfw - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cell_fw_off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const FwDesc &D = fw[lo];
    const int c = g - cell_fw_off[lo];
    const int nc0 = D.cg_nc[0], nc1 = D.cg_nc[1], nc2 = D.cg_nc[2];
    const int k0 = nc0 < 3 ? nc0 : 3, k1 = nc1 < 3 ? nc1 : 3, k2 = nc2 < 3 ? nc2 : 3;
    const int h0 = c % nc0, h1 = (c / nc0) % nc1, h2 = c / (nc0 * nc1);
    const int *cs = cell_start + D.cg_off;
    const uint4 *src = corefx + D.atom_off;
    int64_t pos = nbr_start[D.cg_off + c];
    for (int i2 = 0; i2 < k2; ++i2)
        for (int i1 = 0; i1 < k1; ++i1)
            for (int i0 = 0; i0 < k0; ++i0) {
                const int cc = nbr_cell_axis(h0, i0, k0, nc0) +
                               nc0 * (nbr_cell_axis(h1, i1, k1, nc1) + nc1 * nbr_cell_axis(h2, i2, k2, nc2));
                const int b0 = cs[cc], len = cs[cc + 1] - b0;
                for (int i = lane; i < len; i += 64) nbr_atoms[pos + i] = src[b0 + i];
                pos += len;
            }
}

hipError_t launch_nbr_fill(const FwDesc *fw, int nfw, const int *cell_fw_off, int ncells, const int *cell_start,
                           const int *nbr_start, const uint4 *corefx, uint4 *nbr_atoms, hipStream_t st)
{
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(nbr_fill_kernel, dim3((unsigned)((ncells + BK_WG / 64 - 1) / (BK_WG / 64))), dim3(BK_WG), 0, st, fw,
                       nfw, cell_fw_off, ncells, cell_start, nbr_start, corefx, nbr_atoms);
    return hipGetLastError();
}

hipError_t launch_sk(const double4 *atoms, int natoms, const int4 *kreq, int nreq, double2 *out, hipStream_t st)
{
    if (nreq <= 0) return hipSuccess;
    hipLaunchKernelGGL(sk_kernel, dim3((unsigned)((nreq + BK_WG - 1) / BK_WG)), dim3(BK_WG), 0, st, atoms, natoms, kreq,
                       nreq, out);
    return hipGetLastError();
}

hipError_t launch_self_term(const int4 *kn, const double *kpref, int nk, const double *gsf, const double *q, int ns,
                            double *partial, hipStream_t st)
{
    if (nk <= 0) return hipSuccess;
    hipLaunchKernelGGL(self_term_kernel, dim3((unsigned)((nk + 63) / 64)), dim3(BK_WG), 0, st, kn, kpref, nk, gsf, q, ns,
                       partial);
    return hipGetLastError();
}

}  // namespace pa
