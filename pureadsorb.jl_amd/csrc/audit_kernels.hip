// audit_kernels.hip -- the from-scratch rebuild behind audit_energy! (src/audit.jl:97-165) on
// gfx950, FP64 throughout:
//   1. audit_tab_kernel   every live guest's per-axis powers e^{2 pi i m f_d} (one work-item
//                         per table entry), the table the move kernel builds per attempt;
//   2. audit_sk_kernel    the structure factor of every system's k-vectors from the poses alone,
//                         the framework's plus each guest's terms in the host rebuild's order
//                         (guests, then sites), so both rebuilds agree bit for bit;
//   3. audit_host_kernel  one workgroup per guest slot: the guest's host-guest real-space energy
//                         over every framework atom with the per-site minimum image
//                         (host_guest_realspace_energy, src/guest.jl:176-222);
//   4. audit_sys_kernel   one workgroup per system: the guest-guest real-space pairs
//                         (src/guest.jl:49-71) and the reciprocal term
//                         sum_k pref_k (|S_k|^2 - |S_host,k|^2) of the rebuilt structure factor.
// The host adds the per-guest self terms, the tail correction and the net-charge term
// (src/guest.jl:293-343) and runs the comparisons.
#include "pa_device.hpp"

namespace pa {

constexpr int AUDIT_WG = 256;

__device__ inline int find_seg(const int *off, int nsys, int x)
{
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ inline void audit_sites(const GuestDev &g, const double4 p4, const double4 q4, double out[MAX_SITES][3])
{
#pragma clang fp contract(off)
    const double p[3] = {p4.x, p4.y, p4.z}, q[4] = {q4.x, q4.y, q4.z, q4.w};
    for (int s = 0; s < g.nsites; ++s) {
        double r[3];
        rotate(q, g.site[s], r);
        for (int d = 0; d < 3; ++d) out[s][d] = p[d] + r[d];
    }
}

__device__ inline double audit_block_sum(double v, double *red)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < AUDIT_WG / 64; ++i) t += red[i];
    return t;
}

__global__ __launch_bounds__(AUDIT_WG) void audit_tab_kernel(AuditArgs a)
{
#pragma clang fp contract(off)
    const int64_t x = (int64_t)blockIdx.x * AUDIT_WG + threadIdx.x;
    const int per = a.guest.nsites * a.tw;
    if (x >= (int64_t)a.nslots * per) return;
    const int slot = (int)(x / per), e = (int)(x - (int64_t)slot * per);
    const int n = find_seg(a.guest_off, a.nsys, slot);
    if (slot - a.guest_off[n] >= a.occ[n]) return;
    const FwDesc &F = a.fw[a.fw_of[n]];
    const int m0 = F.kn_max[0] + 1, m1 = F.kn_max[1] + 1, tw = m0 + m1 + F.kn_max[2] + 1;
    const int t = e / a.tw, c = e - t * a.tw;
    if (c >= tw) return;
    double s[MAX_SITES][3];
    audit_sites(a.guest, a.pos[slot], a.quat[slot], s);
    const int d = c < m0 ? 0 : (c < m0 + m1 ? 1 : 2);
    const int m = c - (d == 0 ? 0 : (d == 1 ? m0 : m0 + m1));
    a.tab[x] = axis_power(F.invA, s[t], d, m);
}

__global__ __launch_bounds__(AUDIT_WG) void audit_sk_kernel(AuditArgs a)
{
#pragma clang fp contract(off)
    const int j = blockIdx.x * AUDIT_WG + threadIdx.x;
    if (j >= a.nsk) return;
    const int n = find_seg(a.sk_off, a.nsys, j);
    const FwDesc &F = a.fw[a.fw_of[n]];
    const int k = j - a.sk_off[n];
    if (k >= F.nk) {
        a.sk[j] = make_double2(0.0, 0.0);
        return;
    }
    double2 S = a.S64[F.k_off + k];
    if (a.guest.charged) {
        const int4 nn = a.kn[F.k_off + k];
        const int m0 = F.kn_max[0] + 1, m1 = F.kn_max[1] + 1;
        const int ax = abs(nn.x), ay = m0 + abs(nn.y), az = m0 + m1 + abs(nn.z);
        const int g0 = a.guest_off[n], occ = a.occ[n];
        for (int i = 0; i < occ; ++i)
            for (int t = 0; t < a.guest.nsites; ++t) {
                const double2 *tb = a.tab + ((size_t)(g0 + i) * a.guest.nsites + t) * a.tw;
                const double2 c = cis_from_axes(tb[ax], tb[ay], tb[az], nn.x, nn.y, nn.z);
                S.x += a.guest.q[t] * c.x;
                S.y += a.guest.q[t] * c.y;
            }
    }
    a.sk[j] = S;
}

__global__ __launch_bounds__(AUDIT_WG) void audit_host_kernel(AuditArgs a)
{
#pragma clang fp contract(off)
    __shared__ double red[AUDIT_WG / 64];
    const int slot = blockIdx.x;
    const int n = find_seg(a.guest_off, a.nsys, slot);
    if (slot - a.guest_off[n] >= a.occ[n]) {
        if (threadIdx.x == 0) a.host_e[slot] = 0.0;
        return;
    }
    const FwDesc &F = a.fw[a.fw_of[n]];
    double s[MAX_SITES][3];
    audit_sites(a.guest, a.pos[slot], a.quat[slot], s);
    double e = 0.0;
    for (int j = threadIdx.x; j < F.natoms; j += AUDIT_WG) {
        const double4 h = a.atom64[F.atom_off + j];
        const double hp[3] = {h.x, h.y, h.z};
        const int ht = a.atype[F.atom_off + j];
        for (int t = 0; t < a.guest.nsites; ++t) {
            double d[3] = {s[t][0] - hp[0], s[t][1] - hp[1], s[t][2] - hp[2]}, m[3];
            minimum_image(F.A, F.invA, d, m);
            const double r2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
            if (r2 < a.rc_lj2 || r2 < a.rc_ew2) {
                const double2 p = a.pair64[t * a.ntypes_c + ht];
                e += site_pair_energy(r2, p.x, p.y, KE * a.guest.q[t] * h.w, F.alpha, a.rc_lj2, a.rc_ew2);
            }
        }
    }
    e = audit_block_sum(e, red);
    if (threadIdx.x == 0) a.host_e[slot] = e;
}

__global__ __launch_bounds__(AUDIT_WG) void audit_sys_kernel(AuditArgs a)
{
#pragma clang fp contract(off)
    __shared__ double red[AUDIT_WG / 64];
    const int n = blockIdx.x;
    const FwDesc &F = a.fw[a.fw_of[n]];
    const int g0 = a.guest_off[n], occ = a.occ[n], NS = a.guest.nsites;
    // guest-guest pairs i < j: every work-item takes the j of its stride for each i
    double gg = 0.0;
    for (int i = 0; i + 1 < occ; ++i) {
        if (i + 1 + (int)threadIdx.x >= occ) continue;
        double si[MAX_SITES][3];
        audit_sites(a.guest, a.pos[g0 + i], a.quat[g0 + i], si);
        for (int j = i + 1 + threadIdx.x; j < occ; j += AUDIT_WG) {
            double sj[MAX_SITES][3];
            audit_sites(a.guest, a.pos[g0 + j], a.quat[g0 + j], sj);
            for (int x = 0; x < NS; ++x)
                for (int y = 0; y < NS; ++y) {
                    double d[3] = {si[x][0] - sj[y][0], si[x][1] - sj[y][1], si[x][2] - sj[y][2]}, m[3];
                    minimum_image(F.A, F.invA, d, m);
                    const double r2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                    const double2 pp = a.pair64[x * a.ntypes_c + a.guest.type[y]];
                    gg += site_pair_energy(r2, pp.x, pp.y, KE * a.guest.q[x] * a.guest.q[y], F.alpha, a.rc_lj2,
                                           a.rc_ew2);
                }
        }
    }
    gg = audit_block_sum(gg, red);
    double rec = 0.0;
    for (int k = threadIdx.x; k < F.nk; k += AUDIT_WG) {
        const double2 S = a.sk[a.sk_off[n] + k], H = a.S64[F.k_off + k];
        rec += a.k64[F.k_off + k].w * ((S.x * S.x + S.y * S.y) - (H.x * H.x + H.y * H.y));
    }
    rec = audit_block_sum(rec, red);
    double he = 0.0;
    for (int i = threadIdx.x; i < occ; i += AUDIT_WG) he += a.host_e[g0 + i];
    he = audit_block_sum(he, red);
    if (threadIdx.x == 0) {
        a.sums[3 * n] = he;
        a.sums[3 * n + 1] = gg;
        a.sums[3 * n + 2] = rec;
    }
}

// The four launches on `st`; the host has checked every occupancy against its capacity.
hipError_t launch_audit(const AuditArgs &a, hipStream_t st)
{
    const int64_t ntab = (int64_t)a.nslots * a.guest.nsites * a.tw;
    if (a.guest.charged && ntab > 0)
        hipLaunchKernelGGL(audit_tab_kernel, dim3((unsigned)((ntab + AUDIT_WG - 1) / AUDIT_WG)), dim3(AUDIT_WG), 0, st, a);
    if (a.nsk > 0)
        hipLaunchKernelGGL(audit_sk_kernel, dim3((unsigned)((a.nsk + AUDIT_WG - 1) / AUDIT_WG)), dim3(AUDIT_WG), 0, st, a);
    if (a.nslots > 0) hipLaunchKernelGGL(audit_host_kernel, dim3((unsigned)a.nslots), dim3(AUDIT_WG), 0, st, a);
    hipLaunchKernelGGL(audit_sys_kernel, dim3((unsigned)a.nsys), dim3(AUDIT_WG), 0, st, a);
    return hipGetLastError();
}

}  // namespace pa
