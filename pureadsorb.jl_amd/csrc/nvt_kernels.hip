// nvt_kernels.hip -- test-particle insertions into the chains' current configurations on gfx950
// (widom_chain_kernel!, src/nvt.jl:85-119), FP64.
//
// The energy of an insertion at pose (p, q) into system s holding guests G:
//   host-guest real space (LJ + KE q q' erfc(alpha r)/r, minimum image per site pair)
//   + KE * 2 sum_k kpref_k Re(conj(S_k) S_g,k)    against the chain's running S (host + G)
//   + guest-guest real space against every live guest of G
//   + const_term[s]                                (insertion_constant_term, src/nvt.jl:44-66)
//
// Layout of a launch: the insertions of one system are contiguous (cycle slot major, then the
// insertion within the cycle), and a workgroup takes 64 consecutive insertions of one system,
// one per lane, so every framework atom and k-vector is the same for all lanes of a wavefront
// (scalar loads, no divergence in the loops). The workgroup's four wavefronts and the grid's
// second dimension split the atoms, k-vectors and guests of the system into P = 4 * gridDim.y
// parts; each workgroup writes the sum of its four parts per insertion, and the host adds the
// gridDim.y partial sums in a fixed order.
//
// Reciprocal term: k = 2 pi A^-T n with integer n enumerated n1 outer, n2, n3 inner ascending
// (kvectors, src/ewald.jl:142-157), so consecutive k-vectors of one (n1, n2) line differ by
// one in n3. A lane evaluates e^{i k.r} = e^{2 pi i n.f} (f = A^-1 r) with one sincospi at the
// start of each line and one complex multiply by e^{2 pi i f3} per further k-vector.
#include "pa_device.hpp"

namespace pa {

constexpr int CW_WG = 256;
constexpr int CW_NW = CW_WG / 64;

template <int NS>
__global__ __launch_bounds__(CW_WG) void chain_widom_kernel(ChainWidomArgs a)
{
#pragma clang fp contract(off)
    __shared__ double red[CW_NW][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int2 grp = a.groups[blockIdx.x];  // system, first insertion
    const int s = grp.x;
    const int64_t item = (int64_t)grp.y + lane;
    const int64_t iend = a.item_off[s + 1];
    const bool valid = item < iend;
    const int64_t local = valid ? item - a.item_off[s] : 0;
    const int64_t cl = local / a.nw, t = local - cl * a.nw;
    const FwDesc &F = a.fw[a.fw_of[s]];
    const GuestDev &G = a.guest;

    // the state this insertion sees: the snapshot of its cycle slot, or the live state
    const bool snap = cl < a.snap_n;
    const int gr0 = a.guest_off[s];
    const int occ = snap ? a.snap_occ[cl * a.nsys + s] : a.occupancy[s];
    const double4 *gpos = snap ? a.snap_pos + cl * a.nslots + gr0 : a.refpos + gr0;
    const double4 *gquat = snap ? a.snap_quat + cl * a.nslots + gr0 : a.quat + gr0;
    const double2 *Sk = snap ? a.snap_sk + cl * a.nk_total + a.sk_off[s] : a.Sk + a.sk_off[s];

    // pose
    double pos[3], q[4];
    if (a.pose_pos) {
        const int64_t i = valid ? item : a.item_off[s];
        for (int d = 0; d < 3; ++d) pos[d] = a.pose_pos[3 * i + d];
        for (int d = 0; d < 4; ++d) q[d] = a.pose_quat[4 * i + d];
    } else {
        double f[3];
        pose_uniforms(a.widom_seed, s, (a.cyc0 + cl) * a.nw + t, f, q);
        matvec(F.A, f, pos);
    }
    double sc[NS][3], sf[NS][3];
#pragma unroll
    for (int x = 0; x < NS; ++x) {
        double r[3];
        rotate(q, G.site[x], r);
        for (int d = 0; d < 3; ++d) sc[x][d] = pos[d] + r[d];
        matvec(F.invA, sc[x], sf[x]);
    }

    const int P = CW_NW * gridDim.y, part = blockIdx.y * CW_NW + w;
    double e = 0.0;

    // (1) host-guest real space over this part's atoms
    {
        const int j0 = (int)((int64_t)F.natoms * part / P), j1 = (int)((int64_t)F.natoms * (part + 1) / P);
        const double4 *af = a.atomf64 + F.atom_off;
        const int *types = a.atype + F.atom_off;
#pragma unroll 1
        for (int j = j0; j < j1; ++j) {
            const double4 h = af[j];
            const int ht = types[j];
#pragma unroll
            for (int x = 0; x < NS; ++x) {
                double df[3] = {sf[x][0] - h.x, sf[x][1] - h.y, sf[x][2] - h.z}, m[3];
                for (int d = 0; d < 3; ++d) df[d] = df[d] - __builtin_rint(df[d]);
                matvec(F.A, df, m);
                const double r2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                if (r2 < a.rc_lj2 || r2 < a.rc_ew2) {
                    const double2 p = a.pair64[x * a.ntypes_c + ht];
                    e += site_pair_energy(r2, p.x, p.y, KE * G.q[x] * h.w, F.alpha, a.rc_lj2, a.rc_ew2);
                }
            }
        }
    }
    // (2) reciprocal cross term over this part's k-vectors
    if (G.charged) {
        const int k0 = (int)((int64_t)F.nk * part / P), k1 = (int)((int64_t)F.nk * (part + 1) / P);
        const double4 *ks = a.k64 + F.k_off;
        const int4 *kn = a.kn + F.k_off;
        double zr[NS], zi[NS], str[NS], sti[NS];
#pragma unroll
        for (int x = 0; x < NS; ++x) {
            sincospi(2.0 * sf[x][2], &sti[x], &str[x]);
            zr[x] = 1.0;
            zi[x] = 0.0;
        }
        int px = 0, py = 0, pz = 0;
        bool have = false;
        double acc = 0.0;
#pragma unroll 1
        for (int k = k0; k < k1; ++k) {
            const int4 nn = kn[k];
            if (have && nn.x == px && nn.y == py && nn.z == pz + 1) {
#pragma unroll
                for (int x = 0; x < NS; ++x) {
                    const double r = zr[x] * str[x] - zi[x] * sti[x];
                    const double i = zr[x] * sti[x] + zi[x] * str[x];
                    zr[x] = r;
                    zi[x] = i;
                }
            } else {
#pragma unroll
                for (int x = 0; x < NS; ++x)
                    sincospi(2.0 * (((double)nn.x * sf[x][0] + (double)nn.y * sf[x][1]) + (double)nn.z * sf[x][2]), &zi[x],
                             &zr[x]);
            }
            px = nn.x;
            py = nn.y;
            pz = nn.z;
            have = true;
            double gre = 0.0, gim = 0.0;
#pragma unroll
            for (int x = 0; x < NS; ++x) {
                gre += G.q[x] * zr[x];
                gim += G.q[x] * zi[x];
            }
            const double2 S = Sk[k];
            acc += ks[k].w * (S.x * gre + S.y * gim);
        }
        e += KE * (2.0 * acc);
    }
    // (3) guest-guest real space against this part's share of the live guests
#pragma unroll 1
    for (int j = part; j < occ; j += P) {
        const double4 p4 = gpos[j], q4 = gquat[j];
        const double qj[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
        for (int y = 0; y < NS; ++y) {
            double r[3];
            rotate(qj, G.site[y], r);
            const double pb[3] = {p4.x + r[0], p4.y + r[1], p4.z + r[2]};
#pragma unroll
            for (int x = 0; x < NS; ++x) {
                double d[3] = {sc[x][0] - pb[0], sc[x][1] - pb[1], sc[x][2] - pb[2]}, m[3];
                minimum_image(F.A, F.invA, d, m);
                const double r2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                const double2 p = a.pair64[x * a.ntypes_c + G.type[y]];
                e += site_pair_energy(r2, p.x, p.y, KE * G.q[x] * G.q[y], F.alpha, a.rc_lj2, a.rc_ew2);
            }
        }
    }

    red[w][lane] = e;
    __syncthreads();
    if (w == 0 && valid) {
        double sum = red[0][lane];
        for (int v = 1; v < CW_NW; ++v) sum += red[v][lane];
        a.part[(int64_t)blockIdx.y * a.nitems + item] = sum;
    }
}

hipError_t launch_chain_widom(const ChainWidomArgs &a, int ngroups, int nsplit, hipStream_t st)
{
    if (ngroups <= 0) return hipSuccess;
    const dim3 grid((unsigned)ngroups, (unsigned)nsplit), block(CW_WG);
    switch (a.guest.nsites) {
    case 1: hipLaunchKernelGGL(chain_widom_kernel<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(chain_widom_kernel<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(chain_widom_kernel<3>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(chain_widom_kernel<4>, grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL(chain_widom_kernel<5>, grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL(chain_widom_kernel<6>, grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pa
