// mc_kernels.hip -- the move-energy path (NVT and μVT Monte Carlo moves) on gfx950.
//
// One workgroup per chain runs that chain's move attempts back to back, so a run of many
// moves costs one launch instead of the reference's three launches per move
// (src/moves.jl:577-645). Within an attempt every work-item draws the same proposal from the
// chain's counter-based stream (select_and_propose, src/moves.jl:275-288), then the workgroup
// splits the three sums a move needs across its lanes:
//   * the new pose's host-guest real-space energy over the framework atoms
//     (host_guest_realspace_energy_range, src/guest.jl:176-200), pairs inside the cutoff
//     compacted through LDS before the LJ + erfc evaluation (fused multiply-adds, FP64),
//   * the guest-guest real-space change over (other guest, site, site) triples
//     (guest_guest_move_delta / guest_pair_realspace_energy_range, src/guest.jl:112-156),
//   * the reciprocal change over the k-vectors against the running structure factor
//     (_reciprocal_move_delta_k, src/guest.jl:228-238), keeping each lane's dS(k) for the
//     update on acceptance;
// reduces them through LDS, and every work-item takes the same Metropolis decision
// (src/moves.jl:198-246). Work-item 0 commits poses and occupancy; each lane applies its own
// k-vectors' dS to the running structure factor. Everything is FP64.
#include "pa_device.hpp"

namespace pa {

constexpr int MC_WG = 512;
constexpr int MC_NW = MC_WG / 64;
constexpr int MC_GW = MC_WG / MC_SPEC;   // work-items per speculative attempt
constexpr int MC_GNW = MC_GW / 64;
static_assert(MC_GW % 64 == 0, "each attempt takes whole wavefronts");

__device__ inline double wave_sum_d(double v)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NS>
__device__ inline void sites_at(const GuestDev &g, const double p[3], const double q[4], double out[NS][3])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double r[3];
        rotate(q, g.site[s], r);
        for (int d = 0; d < 3; ++d) out[s][d] = p[d] + r[d];
    }
}

// Minimum-image squared distance between two points of the cell.
__device__ inline double mi_r2(const FwDesc &F, const double a[3], const double b[3])
{
#pragma clang fp contract(off)
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]}, m[3];
    minimum_image(F.A, F.invA, d, m);
    return m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
}

constexpr int MC_RING = 128;  // per-wavefront ring of in-cutoff host-guest pairs

template <int NS, bool UPPER>
__global__ __launch_bounds__(MC_WG) void mc_chain_kernel(McArgs a)
{
#pragma clang fp contract(off)
    __shared__ double red[3][MC_NW];
    __shared__ double ring_r2[MC_NW][MC_RING], ring_qq[MC_NW][MC_RING];
    __shared__ int ring_code[MC_NW][MC_RING];
    // per group: the attempt's slot, acceptance uniform and proposed pose
    __shared__ int sp_slot[MC_SPEC];
    __shared__ double sp_u[MC_SPEC], sp_p[MC_SPEC][3], sp_q[MC_SPEC][4];
    // per-axis phase tables of the attempt's old and new sites: e^{2 pi i m f_d}, m = 0..kn_max[d]
    extern __shared__ double2 ktab[];
    const int n = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const FwDesc &F = a.fw[a.fw_of[n]];
    const int gr0 = a.guest_off[n], cap = a.guest_off[n + 1] - gr0, nslots = a.guest_off[a.nsys];
    const double4 *ks = a.k64 + F.k_off;
    double2 *Sk = a.Sk + a.sk_off[n];
    double2 *dS = a.dS + (size_t)MC_SPEC * a.sk_off[n];   // MC_SPEC slices of nk
    double *skabs = a.sk_abs + a.sk_off[n];
    const int nk = F.nk, natoms = F.natoms;
    const GuestDev &G = a.guest;
    const double alpha = F.alpha;
    double qsum_abs = 0.0;
    for (int s = 0; s < NS; ++s) qsum_abs += fabs(G.q[s]);
    const double ds_bound = 2.0 * qsum_abs;

    // chain scalars: identical copies in every work-item (every decision is taken by all)
    int occ = a.occupancy[n];
    double E = a.energy[n], Eabs = a.energy_abs[n];
    uint64_t ctr = a.rng_counter[n];
    const uint64_t seed = a.rng_seed[n];
    int acc[NMOVES], att[NMOVES];
    for (int t = 0; t < NMOVES; ++t) {
        acc[t] = a.accepted[n * NMOVES + t];
        att[t] = a.attempted[n * NMOVES + t];
    }
    int maxocc = a.max_occ ? a.max_occ[n] : 0;
    int caphit = 0;
    int64_t ntrace = 0;
    const double step_t = a.step_trans ? a.step_trans[n] : 0.0, step_r = a.step_rot ? a.step_rot[n] : 0.0;
    const double fugV = a.fug ? a.fug[n] * F.volume : 0.0;
    const double cp = a.const_p ? a.const_p[n] : 0.0, cq = a.const_q ? a.const_q[n] : 0.0;

    // Speculative rounds: the workgroup splits into MC_SPEC groups; group gi evaluates attempt
    // ctr + gi against the state at the start of the round. Every attempt's proposal and
    // acceptance draw come from its own counter (ChainRNG(chain seed, attempt, 0)), so until the
    // first accepted attempt of the round each one is exactly what a sequential chain would
    // evaluate: the round commits its first accepted attempt (if any), keeps the decisions of the
    // attempts before it and discards the rest, which the next round evaluates again.
    const int gi = tid / MC_GW, gt = tid - gi * MC_GW;
    const int m0 = F.kn_max[0] + 1, m1 = F.kn_max[1] + 1, tw = m0 + m1 + F.kn_max[2] + 1;
    double2 *mytab = ktab + (size_t)gi * 2 * NS * tw;
    double2 *mydS = dS + (size_t)gi * nk;

    // cycle mode: a cycle of chain n is max(occupancy at its start, min_cycle) attempts, and its
    // sample is taken at its end. Called before every attempt; false once the run is over.
    int64_t c = 0, m = 0, len = 0;
    bool finished = false;
    // cycles whose end-of-cycle snapshot is still to be written: a boundary only records the
    // cycle, and the copy runs before the next commit (the state cannot change before then)
    int64_t pend[MC_SPEC + 1];
    int npend = 0;
    auto cycle_boundary = [&]() -> bool {
        while (m == len) {
            if (c > 0) {
                if (occ > maxocc) maxocc = occ;
                if (a.sampling && tid == 0) {
                    a.samp_e[(int64_t)n * a.samp_stride + a.samp_base + c - 1] = E;
                    a.samp_n[(int64_t)n * a.samp_stride + a.samp_base + c - 1] = (double)occ;
                }
                if (c - 1 < a.snap_n) pend[npend++] = c - 1;
            }
            if (c == a.ncycles) return false;
            if (occ > maxocc) maxocc = occ;
            len = occ > a.min_cycle ? occ : a.min_cycle;
            m = 0;
            ++c;
        }
        return true;
    };

    // writes the pending snapshots: every work-item copies the structure-factor entries it
    // updates itself on a commit (k = tid mod MC_WG) and a share of the guests, which work-item 0
    // last wrote before a barrier
    auto flush_snapshots = [&]() {
        for (int p = 0; p < npend; ++p) {
            const int64_t cs = pend[p];
            const ChainSnap &S = a.snap;
            double2 *dsk = S.csk + cs * S.ctot + S.coff[n];
            const int *perm = S.perm + F.k_off;
            for (int k = tid; k < nk; k += MC_WG) {
                const double w = ks[k].w;
                const double2 v = Sk[k];
                dsk[perm[k]] = make_double2(w * v.x, w * v.y);
            }
            double4 *dp = S.pos + cs * S.nslots + gr0, *dq = S.quat + cs * S.nslots + gr0;
            for (int i = tid; i < occ; i += MC_WG) {
                dp[i] = a.refpos[gr0 + i];
                dq[i] = a.quat[gr0 + i];
            }
            if (tid == 0) S.occ[cs * S.nsys + n] = occ;
        }
        npend = 0;
    };

    // processes up to nv attempts of types tys[0..nv) in order (cycle bookkeeping between
    // them); returns how many were processed
    auto round = [&](int nv, const int *tys) -> int {
        const bool active = gi < nv;
        const int type = active ? tys[gi] : MOVE_TRANSLATION;
        Chain rng(seed, ctr + (uint64_t)gi, 0);
        int slot = 0;
        double op[3] = {0, 0, 0}, oq[4] = {0, 0, 0, 1}, np[3] = {0, 0, 0}, nq[4] = {0, 0, 0, 1};
        if (type != MOVE_INSERTION) {
            const int64_t idx = rand_range(rng, occ > 0 ? occ : 1);
            int64_t sl = gr0 + idx - 1;
            slot = (int)(sl < nslots - 1 ? sl : nslots - 1);
            if (slot < 0) slot = 0;
            const double4 p4 = a.refpos[slot], q4 = a.quat[slot];
            op[0] = p4.x; op[1] = p4.y; op[2] = p4.z;
            oq[0] = q4.x; oq[1] = q4.y; oq[2] = q4.z; oq[3] = q4.w;
        }
        if (type == MOVE_TRANSLATION) {
            const double dx = rand_normal(rng), dy = rand_normal(rng), dz = rand_normal(rng);
            np[0] = op[0] + step_t * dx;
            np[1] = op[1] + step_t * dy;
            np[2] = op[2] + step_t * dz;
            for (int i = 0; i < 4; ++i) nq[i] = oq[i];
        } else if (type == MOVE_ROTATION) {
            double qr[4], qp[4];
            rand_quaternion(rng, qr);
            quaternion_pow(qr, step_r, qp);
            qmul(qp, oq, nq);
            qnormalize(nq);
            for (int i = 0; i < 3; ++i) np[i] = op[i];
        } else if (type == MOVE_REINSERTION || type == MOVE_INSERTION) {
            double f[3];
            f[0] = unit_f64(rng.next());
            f[1] = unit_f64(rng.next());
            f[2] = unit_f64(rng.next());
            matvec(F.A, f, np);
            rand_quaternion(rng, nq);
        }
        const double u = unit_f64(rng.next());
        const bool has_old = type != MOVE_INSERTION, has_new = type != MOVE_DELETION;
        double os[NS][3], ns[NS][3];
        sites_at<NS>(G, op, oq, os);
        sites_at<NS>(G, np, nq, ns);
        if (active && gt == 0) {
            sp_slot[gi] = slot;
            sp_u[gi] = u;
            for (int i = 0; i < 3; ++i) sp_p[gi][i] = np[i];
            for (int i = 0; i < 4; ++i) sp_q[gi][i] = nq[i];
        }

        // (1) host-guest real space of the new pose: minimum images in fractional coordinates;
        // the (site, atom) pairs inside the cutoff go through the wavefront's ring in LDS and are
        // evaluated 64 at a time, so LJ + erfc run on full wavefronts
        double hsum = 0.0;
        if (active && has_new) {
            double sfn[NS][3];
#pragma unroll
            for (int s = 0; s < NS; ++s) matvec(F.invA, ns[s], sfn[s]);
            double *qr2 = ring_r2[wid], *qqq = ring_qq[wid];
            int *qcode = ring_code[wid];
            unsigned head = 0, tail = 0;
            auto drain = [&](unsigned cnt) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if ((unsigned)lane < cnt) {
                    const unsigned i = (head + lane) & (MC_RING - 1);
                    const double2 p = a.pair64[qcode[i]];
                    hsum += cw_pair(qr2[i], p.x, p.y, qqq[i], alpha, a.rc_lj2, a.rc_ew2);
                }
                head += cnt;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            };
            const double4 *af = a.atomf + F.atom_off;
            const int *tf = a.atypef + F.atom_off;
            // h holds atom j while j < natoms and a NaN position past the last atom, whose r2 fails
            // every cutoff test
            const double nan = __builtin_nan("");
            double4 h_next = gt < natoms ? af[gt] : make_double4(nan, nan, nan, 0.0);
            int ht_next = gt < natoms ? tf[gt] : 0;
#pragma unroll 1
            for (int jb = gt - lane; jb < natoms; jb += MC_GW) {
                const int j = jb + lane;
                const double4 h = h_next;
                const int ht = ht_next;
                if (j + MC_GW < natoms) {
                    h_next = af[j + MC_GW];
                    ht_next = tf[j + MC_GW];
                } else {
                    h_next = make_double4(nan, nan, nan, 0.0);
                    ht_next = 0;
                }
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const double r2 = cw_r2_t<UPPER>(F.A, sfn[s][0] - h.x, sfn[s][1] - h.y, sfn[s][2] - h.z);
                    const bool in = r2 < a.rc_max2;
                    const uint64_t mask = __builtin_amdgcn_ballot_w64(in);
                    if (in) {
                        const unsigned below = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                        const unsigned i = (tail + below) & (MC_RING - 1);
                        qr2[i] = r2;
                        qqq[i] = KE * G.q[s] * h.w;
                        qcode[i] = s * a.ntypes_c + ht;
                    }
                    tail += (unsigned)__popcll(mask);
                    if (tail - head >= 64) drain(64);
                }
            }
            drain(tail - head);
        }
        // (2) guest-guest real space: the moving (or inserted, or deleted) guest against every
        // other live guest, one (other guest, site, site) triple per lane step
        double gsum = 0.0;
        const int ntrip = active ? occ * NS * NS : 0;
#pragma unroll 1
        for (int x = gt; x < ntrip; x += MC_GW) {
            const int jj = x / (NS * NS), sa = (x / NS) % NS, sb = x % NS;
            const int j = gr0 + jj;
            if (has_old && j == slot) continue;
            const double4 p4 = a.refpos[j], q4 = a.quat[j];
            const double qj[4] = {q4.x, q4.y, q4.z, q4.w};
            double r[3];
            rotate(qj, G.site[sb], r);
            const double pb[3] = {p4.x + r[0], p4.y + r[1], p4.z + r[2]};
            const double2 p = a.pair64[sa * a.ntypes_c + G.type[sb]];
            const double qq = KE * G.q[sa] * G.q[sb];
            if (has_new) gsum += site_pair_energy(mi_r2(F, ns[sa], pb), p.x, p.y, qq, alpha, a.rc_lj2, a.rc_ew2);
            if (type <= MOVE_REINSERTION)
                gsum -= site_pair_energy(mi_r2(F, os[sa], pb), p.x, p.y, qq, alpha, a.rc_lj2, a.rc_ew2);
            else if (type == MOVE_DELETION)
                gsum += site_pair_energy(mi_r2(F, os[sa], pb), p.x, p.y, qq, alpha, a.rc_lj2, a.rc_ew2);
        }
        // (3) reciprocal change against the running structure factor, dS kept per lane. The
        // guest term e^{i k.r} of k = 2 pi A^-T n at r = A f is the product over axes of
        // e^{2 pi i n_d f_d}: per attempt every site's per-axis powers are tabulated in LDS
        // (own sincospi of 2 m f_d), then each k-vector takes two complex products per site.
        double rsum = 0.0;
        if (G.charged) {
            if (active)
                for (int e = gt; e < 2 * NS * tw; e += MC_GW) {
                    const int pose = e / (NS * tw), rem = e - pose * NS * tw, t = rem / tw, x = rem - t * tw;
                    const int d = x < m0 ? 0 : (x < m0 + m1 ? 1 : 2);
                    const int m = x - (d == 0 ? 0 : (d == 1 ? m0 : m0 + m1));
                    mytab[e] = axis_power(F.invA, pose ? ns[t] : os[t], d, m);
                }
            __syncthreads();
            const int4 *kn = a.kn + F.k_off;
            if (active) {
                // one chain's attempt is latency-bound: the next k-vector's loads are issued
                // before the current one is evaluated
                double kw_next = 0.0;
                int4 nn_next = make_int4(0, 0, 0, 0);
                double2 S_next = make_double2(0.0, 0.0);
                if (gt < nk) {
                    kw_next = ks[gt].w;
                    nn_next = kn[gt];
                    S_next = Sk[gt];
                }
#pragma unroll 1
                for (int k = gt; k < nk; k += MC_GW) {
                    const double kw = kw_next;
                    const int4 nn = nn_next;
                    const double2 S = S_next;
                    if (k + MC_GW < nk) {
                        kw_next = ks[k + MC_GW].w;
                        nn_next = kn[k + MC_GW];
                        S_next = Sk[k + MC_GW];
                    }
                    const int ax = abs(nn.x), ay = abs(nn.y), az = abs(nn.z);
                    double sre = 0.0, sim = 0.0;
#pragma unroll
                    for (int pose = 0; pose < 2; ++pose) {
                        if (pose == 1 ? !has_new : !has_old) continue;
                        const double sg = pose == 1 ? 1.0 : -1.0;
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const double2 *tb = mytab + (pose * NS + s) * tw;
                            const double2 cc = cis_from_axes(tb[ax], tb[m0 + ay], tb[m0 + m1 + az], nn.x, nn.y, nn.z);
                            sre += sg * G.q[s] * cc.x;
                            sim += sg * G.q[s] * cc.y;
                        }
                    }
                    mydS[k] = make_double2(sre, sim);
                    rsum += kw * (2.0 * (S.x * sre + S.y * sim) + (sre * sre + sim * sim));
                }
            }
        }
        hsum = wave_sum_d(hsum);
        gsum = wave_sum_d(gsum);
        rsum = wave_sum_d(rsum);
        if (lane == 0) {
            red[0][wid] = hsum;
            red[1][wid] = gsum;
            red[2][wid] = rsum;
        }
        __syncthreads();

        // Metropolis decisions of the round's attempts in order, identical in every work-item;
        // the first accepted one ends the round
        int done = 0, first = -1;
        double dU = 0.0, e_first = 0.0;
        for (int g = 0; g < nv; ++g) {
            if (a.mode != 0 && !cycle_boundary()) {
                finished = true;
                break;
            }
            ++m;
            done = g + 1;
            double e_new = 0.0, gg = 0.0, rec = 0.0;
            for (int w = g * MC_GNW; w < (g + 1) * MC_GNW; ++w) {
                e_new += red[0][w];
                gg += red[1][w];
                rec += red[2][w];
            }
            rec *= KE;
            const int ty = tys[g], sl = sp_slot[g];
            const double ug = sp_u[g];
            double d;
            bool accept;
            if (ty <= MOVE_REINSERTION) {
                d = (e_new - a.host_e[sl]) + gg + rec;
                accept = isfinite(d) && (-d / a.kT > log(ug)) && occ > 0;
            } else if (ty == MOVE_INSERTION) {
                d = e_new + gg + rec + (cp + cq * (double)occ);
                const double lp = log(fugV) - log((double)occ + 1.0) - log(a.kT);
                const bool would = isfinite(d) && (lp - d / a.kT > log(ug));
                accept = would && occ < cap;
                if (would && occ >= cap) caphit = 1;
            } else {
                const int occ_eff = occ > 0 ? occ : 1;
                d = -(a.host_e[sl] + gg + (cp + cq * (double)(occ - 1))) + rec;
                const double lp = -(log(fugV) - log((double)(occ_eff - 1) + 1.0) - log(a.kT));
                accept = isfinite(d) && (lp - d / a.kT > log(ug)) && occ > 0;
            }
            if (a.trace && ntrace < a.trace_cap && tid == 0)
                a.trace[(int64_t)n * a.trace_cap + ntrace] =
                    make_double4((double)ty, (double)(sl - gr0), d, accept ? 1.0 : 0.0);
            ++ntrace;
            ++att[ty - 1];
            if (accept) {
                first = g;
                dU = d;
                e_first = e_new;
                break;
            }
        }
        if (npend > 0) {
            flush_snapshots();
            __syncthreads();  // the commit below overwrites guests the copy read
        }
        if (first >= 0) {
            const int ty = tys[first], sl = sp_slot[first];
            if (tid == 0) {
                const double *pp = sp_p[first], *qq = sp_q[first];
                if (ty <= MOVE_REINSERTION) {
                    a.refpos[sl] = make_double4(pp[0], pp[1], pp[2], 0.0);
                    a.quat[sl] = make_double4(qq[0], qq[1], qq[2], qq[3]);
                    a.host_e[sl] = e_first;
                } else if (ty == MOVE_INSERTION) {
                    const int s2 = gr0 + occ;
                    a.refpos[s2] = make_double4(pp[0], pp[1], pp[2], 0.0);
                    a.quat[s2] = make_double4(qq[0], qq[1], qq[2], qq[3]);
                    a.host_e[s2] = e_first;
                } else {
                    const int hi = gr0 + occ - 1;
                    if (sl != hi) {
                        a.refpos[sl] = a.refpos[hi];
                        a.quat[sl] = a.quat[hi];
                        a.host_e[sl] = a.host_e[hi];
                    }
                }
            }
            if (ty == MOVE_INSERTION) ++occ;
            if (ty == MOVE_DELETION) --occ;
            E += dU;
            Eabs += fabs(dU);
            ++acc[ty - 1];
            if (G.charged) {
                const double2 *fdS = dS + (size_t)first * nk;
                double2 dd_next = make_double2(0.0, 0.0), S_next = make_double2(0.0, 0.0);
                double sa_next = 0.0;
                if (tid < nk) {
                    dd_next = fdS[tid];
                    S_next = Sk[tid];
                    sa_next = skabs[tid];
                }
                for (int k = tid; k < nk; k += MC_WG) {
                    const double2 dd = dd_next, S = S_next;
                    const double sa = sa_next;
                    if (k + MC_WG < nk) {  // other entries than the ones written below
                        dd_next = fdS[k + MC_WG];
                        S_next = Sk[k + MC_WG];
                        sa_next = skabs[k + MC_WG];
                    }
                    Sk[k] = make_double2(S.x + dd.x, S.y + dd.y);
                    skabs[k] = sa + ds_bound;
                }
            }
        }
        ctr += (uint64_t)done;
        __syncthreads();  // committed poses, red[] and the group records reused by the next round
        return done;
    };

    auto draw_type = [&](uint64_t c) {
        Chain r1(seed, c, 1);
        const double v = unit_f64(r1.next());
        int t = NMOVES;
        for (int i = 0; i < NMOVES; ++i)
            if (v < a.pcum[i]) { t = i + 1; break; }
        return t;
    };

    // both modes. In cycle mode a round may span cycle boundaries: the move types are drawn per
    // attempt counter, and a boundary only samples the state and, while no attempt has been
    // accepted, fixes the next cycle's length from the unchanged occupancy.
    int tys[MC_SPEC];
    for (;;) {
        int nv;
        if (a.mode == 0) {
            if (m >= a.nmoves) break;
            nv = (int)(a.nmoves - m < MC_SPEC ? a.nmoves - m : MC_SPEC);
            for (int g = 0; g < nv; ++g) tys[g] = a.movetype;
        } else {
            if (!cycle_boundary()) break;
            nv = MC_SPEC;
            for (int g = 0; g < nv; ++g) tys[g] = draw_type(ctr + (uint64_t)g);
        }
        round(nv, tys);
        if (finished) break;
    }
    flush_snapshots();
    if (tid == 0) {
        a.occupancy[n] = occ;
        a.energy[n] = E;
        a.energy_abs[n] = Eabs;
        a.rng_counter[n] = ctr;
        for (int t = 0; t < NMOVES; ++t) {
            a.accepted[n * NMOVES + t] = acc[t];
            a.attempted[n * NMOVES + t] = att[t];
        }
        if (a.max_occ) a.max_occ[n] = maxocc;
        if (a.capacity_hits) a.capacity_hits[n] = caphit;
    }
}

hipError_t launch_mc(const McArgs &a, hipStream_t st)
{
    const dim3 grid((unsigned)a.nsys), block(MC_WG);
    const int ns = a.guest.nsites;
    // the per-axis phase tables: old and new pose, every site, kn_max[d] + 1 powers per axis
    const size_t lds = a.guest.charged ? (size_t)MC_SPEC * 2 * ns * a.ktab_w * sizeof(double2) : 0;
    if (lds > 72 * 1024) return hipErrorInvalidValue;  // + ~21 KB of static LDS (rings, reductions)
    switch (ns * 2 + (a.upper ? 1 : 0)) {
    case 2: hipLaunchKernelGGL((mc_chain_kernel<1, false>), grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL((mc_chain_kernel<1, true>), grid, block, lds, st, a); break;
    case 4: hipLaunchKernelGGL((mc_chain_kernel<2, false>), grid, block, lds, st, a); break;
    case 5: hipLaunchKernelGGL((mc_chain_kernel<2, true>), grid, block, lds, st, a); break;
    case 6: hipLaunchKernelGGL((mc_chain_kernel<3, false>), grid, block, lds, st, a); break;
    case 7: hipLaunchKernelGGL((mc_chain_kernel<3, true>), grid, block, lds, st, a); break;
    case 8: hipLaunchKernelGGL((mc_chain_kernel<4, false>), grid, block, lds, st, a); break;
    case 9: hipLaunchKernelGGL((mc_chain_kernel<4, true>), grid, block, lds, st, a); break;
    case 10: hipLaunchKernelGGL((mc_chain_kernel<5, false>), grid, block, lds, st, a); break;
    case 11: hipLaunchKernelGGL((mc_chain_kernel<5, true>), grid, block, lds, st, a); break;
    case 12: hipLaunchKernelGGL((mc_chain_kernel<6, false>), grid, block, lds, st, a); break;
    case 13: hipLaunchKernelGGL((mc_chain_kernel<6, true>), grid, block, lds, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pa
