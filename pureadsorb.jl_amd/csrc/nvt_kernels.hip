// nvt_kernels.hip -- test-particle insertions into the chains' current configurations on gfx950
// (widom_chain_kernel!, src/nvt.jl:85-119), FP64.
//
// The energy of an insertion at pose (p, q) into system s holding guests G:
//   host-guest real space (LJ + KE q q' erfc(alpha r)/r, minimum image per site pair)
//   + KE * 2 sum_k kpref_k Re(conj(S_k) S_g,k)    against the chain's running S (host + G)
//   + guest-guest real space against every live guest of G
// (the host adds insertion_constant_term, src/nvt.jl:44-66).
//
// One wavefront per insertion:
//   * host-guest: the framework's atoms are stored in spatially compact chunks of 64 (one atom
//     per lane); pairs inside the cutoff are compacted through LDS before the LJ + erfc
//     evaluation. A chunk whose centre lies farther than its radius + guest radius + cutoff from
//     the insertion's reference point is skipped whole. The distance is a lower bound of the
//     true nearest-image distance: the rint image is the nearest one whenever it is shorter
//     than half the smallest perpendicular width, so min(rint distance, half width) never
//     exceeds the true distance; and with widths >= 2 * cutoff every pair within the cutoff
//     is found by the per-pair rint image the energy uses.
//   * reciprocal: k = 2 pi A^-T n is enumerated n1 outer, n2, n3 inner ascending (kvectors,
//     src/ewald.jl:142-157), so the k-vectors fall into runs of consecutive n3 at fixed
//     (n1, n2). Each lane takes whole runs (longest first, so the lanes' trip counts match):
//     one sincospi of 2 pi n.f per site at the start of a run, then one complex multiply by
//     e^{2 pi i f3} per further k-vector (f = A^-1 r). The snapshot stores kprefactor * S in
//     the lanes' reading order, so each step is one coalesced load.
//   * guest-guest: lanes take (guest, site, site) triples.
// The lanes' partial sums are reduced across the wavefront in a fixed order. The pose is drawn
// exactly as the Widom path draws it; the energy terms use fused multiply-adds.
#include "pa_device.hpp"

namespace pa {

constexpr int CW_WG = 256;
constexpr int CW_NW = CW_WG / 64;

__device__ inline double cw_wave_sum(double v)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

constexpr int CW_RING = 128;  // per-wavefront ring of in-cutoff pairs (< 64 pending + one append)

// Guests of up to three sites fit 4 wavefronts per SIMD without spilling; larger guests keep
// the compiler's own register budget.
template <int NS, bool UPPER>
__global__ __launch_bounds__(CW_WG, (NS <= 3 ? 4 : 1)) void chain_widom_kernel(ChainWidomArgs a)
{
    __shared__ double ring_r2[CW_NW][CW_RING], ring_qq[CW_NW][CW_RING];
    __shared__ int ring_code[CW_NW][CW_RING];
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * CW_NW + (threadIdx.x >> 6);
    if (item >= a.nitems) return;  // whole wavefront; the kernel has no barrier
    int lo = 0, hi = a.nsys - 1;
    while (lo < hi) {  // system of the item: item_off[s] <= item < item_off[s+1]
        const int mid = (lo + hi + 1) >> 1;
        if (a.item_off[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int s = lo;
    const int64_t local = item - a.item_off[s];
    const int64_t cl = local / a.nw, t = local - cl * a.nw;
    const int u = a.fw_of[s];
    const FwDesc &F = a.fw[u];
    const GuestDev &G = a.guest;

    // the state this insertion sees: the snapshot of its cycle slot
    const ChainSnap &S = a.snap;
    const int64_t slot = cl < a.snap_n ? cl : a.snap_n - 1;
    const int gr0 = a.guest_off[s];
    const int occ = S.occ[slot * S.nsys + s];
    const double4 *gpos = S.pos + slot * S.nslots + gr0;
    const double4 *gquat = S.quat + slot * S.nslots + gr0;
    const double2 *csk = S.csk + slot * S.ctot + S.coff[s];

    double pos[3], q[4];
    if (a.pose_pos) {
        for (int d = 0; d < 3; ++d) pos[d] = a.pose_pos[3 * item + d];
        for (int d = 0; d < 4; ++d) q[d] = a.pose_quat[4 * item + d];
    } else {
        double f[3];
        pose_uniforms(a.widom_seed, s, (a.cyc0 + cl) * a.nw + t, f, q);
        matvec(F.A, f, pos);
    }
    double pf[3], sc[NS][3], sf[NS][3];
    matvec(F.invA, pos, pf);
#pragma unroll
    for (int x = 0; x < NS; ++x) {
        double r[3];
        rotate(q, G.site[x], r);
        for (int d = 0; d < 3; ++d) sc[x][d] = pos[d] + r[d];
        matvec(F.invA, sc[x], sf[x]);
    }
    double e = 0.0;

    // (1) host-guest real space, chunk by chunk. The (site, atom) pairs inside the cutoff are
    // appended to the wavefront's ring in LDS and evaluated 64 at a time, one per lane, so the
    // LJ + erfc work runs on full wavefronts whatever the lanes' own pairs do.
    {
        const int w = threadIdx.x >> 6;
        double *qr2 = ring_r2[w], *qqq = ring_qq[w];
        int *qcode = ring_code[w];
        unsigned head = 0, tail = 0;
        auto drain = [&](unsigned n) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if ((unsigned)lane < n) {
                const unsigned i = (head + lane) & (CW_RING - 1);
                const int code = qcode[i];
                const double2 p = a.pair64[code];
                e += cw_pair(qr2[i], p.x, p.y, qqq[i], F.alpha, a.rc_lj2, a.rc_ew2);
            }
            head = __builtin_amdgcn_readfirstlane(head + n);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        const int c0 = a.chunk_off[u], c1 = a.chunk_off[u + 1];
        const double halfw = a.half_perp[u];
        const double4 *af = a.atomf64 + F.atom_off;
        const int *types = a.atype + F.atom_off;
        const double nan = __builtin_nan("");
#pragma unroll 1
        for (int c = c0; c < c1; ++c) {
            const double4 cc = a.chunk[c];
            const double dist = fmin(__builtin_sqrt(cw_r2_t<UPPER>(F.A, pf[0] - cc.x, pf[1] - cc.y, pf[2] - cc.z)), halfw);
            // the same value in every lane (one insertion per wavefront): a uniform branch
            if (__builtin_amdgcn_readfirstlane((int)(dist > cc.w + a.cull_extra))) continue;
            const int j = (c - c0) * 64 + lane;
            // lanes past the last atom take a NaN position, whose r2 fails every cutoff test
            const double4 h = j < F.natoms ? af[j] : make_double4(nan, nan, nan, 0.0);
            const int ht = j < F.natoms ? types[j] : 0;
#pragma unroll
            for (int x = 0; x < NS; ++x) {
                const double r2 = cw_r2_t<UPPER>(F.A, sf[x][0] - h.x, sf[x][1] - h.y, sf[x][2] - h.z);
                const bool in = r2 < a.rc_max2;
                const uint64_t mask = __builtin_amdgcn_ballot_w64(in);
                if (in) {
                    const unsigned below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                    const unsigned i = (tail + below) & (CW_RING - 1);
                    qr2[i] = r2;
                    qqq[i] = KE * G.q[x] * h.w;
                    qcode[i] = x * a.ntypes_c + ht;
                }
                tail = __builtin_amdgcn_readfirstlane(tail + (unsigned)__popcll(mask));
                if (tail - head >= 64) drain(64);
            }
        }
        drain(tail - head);
    }
    // (2) reciprocal cross term, whole runs of consecutive n3 per lane; step i of round R reads
    // the weighted structure factor at round base + 64 i + lane (consecutive across the lanes)
    if (G.charged) {
        const int4 *kn = a.kn + F.k_off;
        const int r0 = a.round_off[u], r1 = a.round_off[u + 1];
        double str[NS], sti[NS];
#pragma unroll
        for (int x = 0; x < NS; ++x) sincospi(2.0 * sf[x][2], &sti[x], &str[x]);
        double acc = 0.0;
#pragma unroll 1
        for (int R = r0; R < r1; ++R) {
            const int2 ln = a.lines[(int64_t)R * 64 + lane];
            if (ln.y == 0) continue;
            const double2 *cr = csk + a.rounds[R].x + lane;
            const int4 nn = kn[ln.x];
            double zr[NS], zi[NS];
#pragma unroll
            for (int x = 0; x < NS; ++x)
                sincospi(2.0 * (((double)nn.x * sf[x][0] + (double)nn.y * sf[x][1]) + (double)nn.z * sf[x][2]), &zi[x],
                         &zr[x]);
#pragma unroll 1
            for (int i = 0; i < ln.y; ++i) {
                if (i > 0) {
#pragma unroll
                    for (int x = 0; x < NS; ++x) {
                        const double r = __builtin_fma(zr[x], str[x], -(zi[x] * sti[x]));
                        const double im = __builtin_fma(zr[x], sti[x], zi[x] * str[x]);
                        zr[x] = r;
                        zi[x] = im;
                    }
                }
                double gre = 0.0, gim = 0.0;
#pragma unroll
                for (int x = 0; x < NS; ++x) {
                    gre = __builtin_fma(G.q[x], zr[x], gre);
                    gim = __builtin_fma(G.q[x], zi[x], gim);
                }
                const double2 c = cr[(int64_t)i * 64];
                acc = __builtin_fma(c.x, gre, __builtin_fma(c.y, gim, acc));
            }
        }
        e += KE * (2.0 * acc);
    }
    // (3) guest-guest real space, one (guest, site, site) triple per lane step
    const int ntrip = occ * NS * NS;
#pragma unroll 1
    for (int xi = lane; xi < ntrip; xi += 64) {
        const int jj = xi / (NS * NS), sa = (xi / NS) % NS, sb = xi % NS;
        const double4 p4 = gpos[jj], q4 = gquat[jj];
        const double qj[4] = {q4.x, q4.y, q4.z, q4.w};
        double r[3];
        rotate(qj, G.site[sb], r);
        const double pb[3] = {p4.x + r[0], p4.y + r[1], p4.z + r[2]};
        double fb[3];
        matvec(F.invA, pb, fb);
        const double r2 = cw_r2_t<UPPER>(F.A, sf[sa][0] - fb[0], sf[sa][1] - fb[1], sf[sa][2] - fb[2]);
        if (r2 < a.rc_lj2 || r2 < a.rc_ew2) {
            const double2 p = a.pair64[sa * a.ntypes_c + G.type[sb]];
            e += cw_pair(r2, p.x, p.y, KE * G.q[sa] * G.q[sb], F.alpha, a.rc_lj2, a.rc_ew2);
        }
    }

    e = cw_wave_sum(e);
    if (lane == 0) a.dU[item] = e;
}

// Snapshot slot `slot` of every system's live state (after an audit, or for insertions at
// caller poses): one workgroup per system.
__global__ __launch_bounds__(256) void chain_snapshot_kernel(ChainSnap S, int64_t slot, const FwDesc *fw,
                                                            const int *fw_of, const double4 *k64, const int *guest_off,
                                                            const int *sk_off, const int *occupancy,
                                                            const double4 *refpos, const double4 *quat,
                                                            const double2 *Sk)
{
    const int n = blockIdx.x, tid = threadIdx.x;
    const FwDesc &F = fw[fw_of[n]];
    const double4 *ks = k64 + F.k_off;
    const double2 *sk = Sk + sk_off[n];
    const int *perm = S.perm + F.k_off;
    double2 *dsk = S.csk + slot * S.ctot + S.coff[n];
    for (int k = tid; k < F.nk; k += 256) {
        const double w = ks[k].w;
        const double2 v = sk[k];
        dsk[perm[k]] = make_double2(w * v.x, w * v.y);
    }
    const int gr0 = guest_off[n], occ = occupancy[n];
    for (int i = tid; i < occ; i += 256) {
        S.pos[slot * S.nslots + gr0 + i] = refpos[gr0 + i];
        S.quat[slot * S.nslots + gr0 + i] = quat[gr0 + i];
    }
    if (tid == 0) S.occ[slot * S.nsys + n] = occ;
}

hipError_t launch_chain_snapshot(const ChainSnap &S, int64_t slot, const McArgs &m, hipStream_t st)
{
    hipLaunchKernelGGL(chain_snapshot_kernel, dim3((unsigned)S.nsys), dim3(256), 0, st, S, slot, m.fw, m.fw_of, m.k64,
                       m.guest_off, m.sk_off, m.occupancy, m.refpos, m.quat, m.Sk);
    return hipGetLastError();
}

hipError_t launch_chain_widom(const ChainWidomArgs &a, hipStream_t st)
{
    if (a.nitems <= 0) return hipSuccess;
    const dim3 grid((unsigned)((a.nitems + CW_NW - 1) / CW_NW)), block(CW_WG);
    switch (a.guest.nsites * 2 + (a.upper ? 1 : 0)) {
    case 2: hipLaunchKernelGGL((chain_widom_kernel<1, false>), grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((chain_widom_kernel<1, true>), grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL((chain_widom_kernel<2, false>), grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL((chain_widom_kernel<2, true>), grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL((chain_widom_kernel<3, false>), grid, block, 0, st, a); break;
    case 7: hipLaunchKernelGGL((chain_widom_kernel<3, true>), grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL((chain_widom_kernel<4, false>), grid, block, 0, st, a); break;
    case 9: hipLaunchKernelGGL((chain_widom_kernel<4, true>), grid, block, 0, st, a); break;
    case 10: hipLaunchKernelGGL((chain_widom_kernel<5, false>), grid, block, 0, st, a); break;
    case 11: hipLaunchKernelGGL((chain_widom_kernel<5, true>), grid, block, 0, st, a); break;
    case 12: hipLaunchKernelGGL((chain_widom_kernel<6, false>), grid, block, 0, st, a); break;
    case 13: hipLaunchKernelGGL((chain_widom_kernel<6, true>), grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pa
