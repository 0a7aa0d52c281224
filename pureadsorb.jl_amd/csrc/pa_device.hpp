// pa_device.hpp -- device-side data layout shared by the host library and the kernels.
#pragma once

#include "pa_numeric.hpp"

namespace pa {

// Per unique framework (FrameworkBatch's deduplicated per-framework storage).
struct FwDesc {
    double A[9];       // cell, column-major
    double invA[9];
    float Af[9];       // A * 2^-32 (maps a 32-bit fixed-point fractional offset to Å)
    double alpha;      // Ewald splitting parameter
    double const_off;  // constant_offset (eV)
    double volume;
    float alphaf;
    int atom_off, natoms;  // slice of the atom arrays
    int k_off, nk;         // slice of the k tables
    int upper;             // A is upper triangular (cell_matrix's form)
    int cg_off;            // hard-core cell grid: first entry of the framework in cell_start
    int cg_nc[3];          // cells per axis (fractional cell edge 1/cg_nc)
    int kn_max[3];         // largest |n_d| of the framework's k-vectors k = 2 pi A^-T n
};

// The guest, with compact type indices.
struct GuestDev {
    int nsites;
    int charged;
    double site[MAX_SITES][3];
    double q[MAX_SITES];
    float qke[MAX_SITES];   // KE * q_s in FP32
    int type[MAX_SITES];
    int mirror;             // 3 sites, sites 1 and 2 mirror images through site 0, equal charges
};

// True when a 3-site guest's sites 1 and 2 are mirror images through site 0 (site1 + site2 ==
// 2 site0 exactly) with equal charges: a linear symmetric molecule such as CO2 or a 3-site N2.
inline bool guest_mirror(const GuestDev &G)
{
    if (G.nsites != 3 || G.q[1] != G.q[2]) return false;
    for (int d = 0; d < 3; ++d)
        if (G.site[1][d] + G.site[2][d] != 2.0 * G.site[0][d]) return false;
    return true;
}

// Everything one energy launch needs. The flattened insertion index g runs system-major:
// system s owns g in [seg[s], seg[s+1]); insertion j of system s is g = seg[s] + j.
struct WidomArgs {
    const FwDesc *fw;
    const int *fw_of;          // nsys
    const int64_t *seg;        // nsys + 1
    int nsys;
    uint64_t seed;
    int64_t g0, g1;            // this launch covers g in [g0, g1)
    double *dU;                // indexed by g - out_base
    int64_t out_base;
    // FP64 plain path tables
    const double4 *atom64;     // x, y, z (Å), q
    const int *atype;          // compact type per atom
    const double4 *k64;        // kx, ky, kz, kprefactor
    const double2 *S64;        // host structure factor
    // mixed path tables
    const uint4 *atomfx;       // 32-bit fixed-point fractional x, y, z; q as float bits
    const int4 *kn;            // integer k coefficients n1, n2, n3
    const float2 *kc;          // 2 * kprefactor * (Re S, Im S); the guest charges carry KE
    // pair tables, [site][type] with ntypes_c compact types
    const double2 *pair64;     // sigma, epsilon
    const float2 *pair32;      // sigma^2, 4 epsilon
    int ntypes_c;
    double rc_lj2, rc_ew2;
    float rc_lj2f, rc_ew2f, rc_max2f;
    GuestDev guest;
};

// Spatial bins of one system's insertions (mixed path): the system's cell is cut into
// nb[0] x nb[1] x nb[2] parallelepipeds in fractional coordinates; every insertion is grouped
// with the others whose reference point lies in the same bin, and a bin's neighbour list holds
// every atom image within sqrt(rlist2) of the bin centre.
struct BinSys {
    int nb[3];
    int bin_base;   // first global bin id of the system
    float rlist2;   // (max cutoff + guest radius + bin circumradius + margin)^2, Å^2
    float alt[3];   // an atom with |fractional offset| >= alt[i] on axis i also gets the image
                    // across the cell boundary on that axis (alt >= 0.5: never)
    int reach[3];   // hard-core stencil half-width in grid cells per axis (reach[0] < 0: no check)
    uint32_t hc_rf[3];  // largest rejection radius over the perpendicular width, 32-bit fixed point
                        // and rounded up: a site's hard-core partners lie within +-hc_rf[d];
                        // all ones: that range (plus one cell) reaches around the axis
    // radial classes of the neighbour list: class 0 holds the atoms closer than cls_d0 to the bin
    // centre, classes 1..BIN_NCLS-1 split [cls_d0, sqrt(rlist2)] into shells of width 1/cls_inv_w
    float cls_d0, cls_inv_w;
};

// The audit's from-scratch rebuild (audit_kernels.hip). Guest slot g of system n lies in
// [guest_off[n], guest_off[n + 1]); its live guests are the first occ[n].
struct AuditArgs {
    const FwDesc *fw;
    const int *fw_of;
    int nsys;
    const double4 *atom64;     // framework atoms, Cartesian (Å) and charge
    const int *atype;
    const double4 *k64;        // kx, ky, kz, kprefactor
    const double2 *S64;        // framework structure factor
    const int4 *kn;
    const double2 *pair64;     // sigma, epsilon per (site, compact type)
    int ntypes_c;
    double rc_lj2, rc_ew2;     // rc_ew2 = 0 for an uncharged guest
    GuestDev guest;
    const int *guest_off;      // nsys + 1
    const int *sk_off;         // nsys + 1: each system's slice of the running structure factor
    const int *occ;
    const double4 *pos, *quat; // per slot
    int nslots, nsk;
    int tw;                    // per-axis power table entries per site (largest over frameworks)
    double2 *tab;              // nslots * nsites * tw
    double2 *sk;               // rebuilt structure factor, nsk
    double *host_e;            // per slot: rebuilt host-guest energy (0 for an empty slot)
    double *sums;              // per system: host-guest sum, guest-guest pairs, reciprocal sum
};

struct BinnedArgs {
    const FwDesc *fw;
    const int *fw_of;
    const int64_t *seg;        // nsys + 1 flattened insertion offsets
    const BinSys *bins;        // nsys
    const int *bin_base;       // nsys + 1
    int nsys;
    int nbins;
    uint64_t seed;
    int64_t g0, g1;            // this chunk covers g in [g0, g1)
    uint32_t *sitefx;          // (g - g0) * 3 * NS fixed-point site coordinates
    int *key;                  // (g - g0) global bin id
    int *bin_count;            // nbins + 1
    int *bin_start;            // nbins + 1
    int *bin_perm;             // nbins: launch order of the energy kernel (most crowded bins first), or null
    int *bin_iota;             // nbins: 0, 1, 2, ... (the radix sort's input values, grids beyond one workgroup's)
    unsigned *bin_ksort;       // nbins: the sorted counts (the sort's key output)
    int sort_bits;             // bits of the largest possible count
    int *rank;                 // (g - g0) position of the insertion within its bin (from the count)
    int *order;                // chunk-relative insertion indices grouped by bin
    double *dU;                // indexed by g - out_base
    int64_t out_base;
    const uint4 *atomfx;       // the framework's atoms in Morton order: fixed point, q bits
    const int *atype;          // their compact types, same order
    const int4 *kn;
    const float2 *kc;
    const float2 *pair32;
    int ntypes_c;
    float rc_lj2f, rc_ew2f, rc_max2f;
    int listcap;               // neighbour-list capacity (entries) of one pass, >= BIN_WG
    int *extra_passes;         // max over bins of (passes - 1) of this call's energy launches
    int zero_passes;           // nonzero on the call's first chunk: bin_zero_kernel clears it
    GuestDev guest;
    // hard-core rejection (phase 0): atoms of each framework sorted by grid cell
    const uint4 *corefx;       // fixed-point x, y, z and compact type, by cell
    const int *cell_start;     // per framework: cg_nc[0]*cg_nc[1]*cg_nc[2] + 1 offsets
    const int *nbr_start;      // same layout: each cell's range of nbr_atoms
    const uint4 *nbr_atoms;    // per cell, the atoms of the cells within one cell (as corefx)
    const float *rho2;         // nsys x NS x ntypes_c rejection radii^2 (Å^2), rounded down
    int *cand;                 // chunk-relative insertions left for the full phase-0 stencil
    int *ncand;                // their count
};

// Gradient launches (widom_grad.hip): the weighted-insertion list, its per-insertion sites and
// weights, and the accumulators of the gather and k-space kernels.
struct GradArgs {
    const FwDesc *fw;
    const int *fw_of;
    const int64_t *seg;
    int nsys;
    uint64_t seed;
    const int64_t *list_g;      // flattened insertion index of each weighted insertion
    const double *dU;           // Delta U of every insertion of the run
    const double *msys;         // per system: max of -beta dU (the weight of insertion g is exp(-beta dU - m_s))
    double beta;
    int64_t nlist;
    const uint4 *atomfx;
    const int *atype;
    const int4 *kn;
    const float4 *pairg;        // [site][type]: sigma^2, 4 epsilon, 1/sigma, 0
    int ntypes_c;
    float rc_lj2f, rc_ew2f, rc_max2f;
    GuestDev guest;
    double *acc_lj;             // nsys x NS x ntc x 4: sum w dU/deps, sum wU dU/deps, sum w dU/dsig, sum wU dU/dsig
    double *acc_q;              // per system slice of 2 x natoms: sum w dU/dq_j, sum wU dU/dq_j
    const int64_t *qoff;        // nsys: first atom slot of the system in acc_q
    double *acc_k;              // per system slice of 4 x nk: w Re S_g, w Im S_g, wU Re S_g, wU Im S_g
    const int64_t *koff;        // nsys: first k slot of the system in acc_k
    // gather form
    const int64_t *loff;        // nsys + 1: each system's range of the weighted-insertion list
    uint32_t *lsite;            // per listed insertion: NS x 3 fixed-point site coordinates
    double2 *lw;                // per listed insertion: w, w dU
    const uint4 *corefx;        // atoms sorted by hard-core grid cell: fixed point + compact type
    const int *coreidx;         // original index of each sorted atom
    const int4 *blk;            // per gather workgroup: system, first sorted atom, list range (relative)
    const int4 *kblk;           // per k-space workgroup: system, first k-vector, list range (relative)
};

// Energy sums that need FP64 accuracy (~1e-14 relative) but no bit-for-bit agreement with
// another code path (test insertions into chains, the move path's host-guest sum) use the
// helpers below, which fuse multiply-adds; the helpers above keep every rounding step of the
// reference's formulas for the paths that need exact agreement.
// r^2 of the cell vector A (f - rint(f)) for a fractional offset f.
// UPPER: A[1] == A[2] == A[5] == 0 (every cell built from its parameters). The products with
// those exact zeros are dropped, which leaves every result bit for bit unchanged.
template <bool UPPER>
__device__ inline double cw_r2_t(const double A[9], double f0, double f1, double f2)
{
    f0 -= __builtin_rint(f0);
    f1 -= __builtin_rint(f1);
    f2 -= __builtin_rint(f2);
    const double m0 = __builtin_fma(A[0], f0, __builtin_fma(A[3], f1, A[6] * f2));
    const double m1 = UPPER ? __builtin_fma(A[4], f1, A[7] * f2) : __builtin_fma(A[1], f0, __builtin_fma(A[4], f1, A[7] * f2));
    const double m2 = UPPER ? A[8] * f2 : __builtin_fma(A[2], f0, __builtin_fma(A[5], f1, A[8] * f2));
    return __builtin_fma(m0, m0, __builtin_fma(m1, m1, m2 * m2));
}

__device__ inline double cw_r2(const double A[9], double f0, double f1, double f2)
{
    return cw_r2_t<false>(A, f0, f1, f2);
}

// LJ inside rc_lj plus qq_ke erfc(alpha r)/r inside rc_ew (site_pair_energy with one division
// for 1/r shared by both terms).
__device__ inline double cw_pair(double r2, double sg, double ep, double qq_ke, double alpha, double rc_lj2,
                                 double rc_ew2)
{
    const double r = __builtin_sqrt(r2);
    const double ir = 1.0 / r;
    double e = 0.0;
    if (r2 < rc_lj2) {
        const double s = sg * ir;
        double x = s * s;
        x = x * x * x;
        e = 4.0 * ep * __builtin_fma(x, x, -x);
    }
    if (r2 < rc_ew2) e = __builtin_fma(qq_ke * pair_erfc_f64_t<true>(alpha * r), ir, e);
    return e;
}

// Phase-0 near stencil: the i-th (i < k = min(3, nc)) neighbour along one axis of cell index h,
// among nc cells (each cell once when the axis has fewer than 3 cells).
__host__ __device__ inline int nbr_cell_axis(int h, int i, int k, int nc)
{
    const int f = k == nc ? 0 : (h - 1 + nc) % nc;
    return (f + i) % nc;
}

// One site pair of the move path (FP64): Lennard-Jones inside rc_lj plus the screened Coulomb
// qq_ke * erfc(alpha r)/r inside rc_ew (qq_ke = KE q_a q_b), on a minimum-image separation r^2.
__host__ __device__ inline double site_pair_energy(double r2, double sg, double ep, double qq_ke, double alpha,
                                                   double rc_lj2, double rc_ew2)
{
#pragma clang fp contract(off)
    double e = 0.0;
    if (r2 < rc_lj2) {
        double x = sg * sg / r2;
        x = x * x * x;
        e = 4.0 * ep * (x * x - x);
    }
    if (r2 < rc_ew2) {
        const double r = __builtin_sqrt(r2);
        e += qq_ke * pair_erfc_f64(alpha * r) / r;
    }
    return e;
}

// e^{i k.r} of a guest site for k = 2 pi A^-T n at the Cartesian point r = A f, as the product
// over axes of e^{2 pi i n_d f_d}. axis_power gives e^{2 pi i m f_d} (m >= 0); cis_from_axes
// multiplies the three, conjugating the axes with negative n_d. The move path's kernel (from
// per-attempt LDS tables) and the host's structure-factor rebuild evaluate both identically.
__host__ __device__ inline double2 axis_power(const double *invA, const double r[3], int d, int m)
{
#pragma clang fp contract(off)
    const double f = invA[d] * r[0] + invA[d + 3] * r[1] + invA[d + 6] * r[2];
    double sn, cs;
    sincospi(2.0 * ((double)m * f), &sn, &cs);
    return make_double2(cs, sn);
}

__host__ __device__ inline double2 cis_from_axes(double2 e0, double2 e1, double2 e2, int nx, int ny, int nz)
{
#pragma clang fp contract(off)
    const double i0 = nx < 0 ? -e0.y : e0.y, i1 = ny < 0 ? -e1.y : e1.y, i2 = nz < 0 ? -e2.y : e2.y;
    const double pr = e0.x * e1.x - i0 * i1, pi = e0.x * i1 + i0 * e1.x;
    return make_double2(pr * e2.x - pi * i2, pr * i2 + pi * e2.x);
}

// Move types of the move path: the reference's three NVT moves (src/moves.jl:69-71) and the two
// exchange moves.
constexpr int MOVE_TRANSLATION = 1, MOVE_ROTATION = 2, MOVE_REINSERTION = 3, MOVE_INSERTION = 4, MOVE_DELETION = 5;
constexpr int NMOVES = 5;
// attempts of one chain the move kernel evaluates together (speculatively, against the state at
// the start of the round); the pending structure-factor changes take MC_SPEC slices per chain
constexpr int MC_SPEC = 2;

// Snapshots of the chains' states for the test insertions (nvt_kernels.hip). Slot c holds, per
// system n: occupancy at occ[c * nsys + n]; guests at pos/quat[c * nslots + guest_off[n] + i];
// the weighted structure factor c_k = kprefactor_k * S_k of every k-vector at
// csk[c * ctot + coff[n] + perm[F.k_off + k]], the order in which the insertion kernel's lanes
// read it (runs of consecutive n3 interleaved across the 64 lanes).
struct ChainSnap {
    int nsys, nslots;
    int64_t ctot;
    const int *coff;           // per system: first entry of its slice of csk
    const int *perm;           // per k-vector of the batch: position within its framework's slice
    int *occ;
    double4 *pos, *quat;
    double2 *csk;
};

// One launch of the move path: every chain (system) is one workgroup that runs its move attempts
// back to back (mc_kernels.hip).
struct McArgs {
    // framework tables: the FP64 layout with every k-vector (FULLK batch)
    const FwDesc *fw;
    const int *fw_of;
    const double4 *atom64;
    const int *atype;
    const double4 *atomf;       // fractional x, y, z and charge (the test-insertion tables' order)
    const int *atypef;          // compact types in the same order
    const double4 *k64;         // kx, ky, kz, kprefactor
    const int4 *kn;             // n of k = 2 pi A^-T n, same order as k64
    int ktab_w;                 // max over frameworks of kn_max[0] + kn_max[1] + kn_max[2] + 3
    const double2 *pair64;      // [site][type]: sigma, epsilon
    int ntypes_c;
    double rc_lj2, rc_ew2;
    double rc_max2;             // max(rc_lj2, rc_ew2): the host-guest pair test
    GuestDev guest;
    int nsys;
    // chain state (SystemState, src/state.jl:89-104)
    const int *guest_off;       // nsys + 1: capacity blocks
    int *occupancy;
    double4 *refpos;            // per slot: x, y, z (Å), 0
    double4 *quat;              // per slot: x, y, z, w
    double *host_e;             // per slot: cached host-guest real-space energy
    double2 *Sk;                // per system k slice: host + every guest
    double2 *dS;                // scratch, same layout: the pending move's structure-factor change
    double *sk_abs;             // per k: accumulated rounding scale
    const int *sk_off;          // nsys + 1
    double *energy, *energy_abs;
    const uint64_t *rng_seed;
    uint64_t *rng_counter;
    int *accepted, *attempted;  // nsys x NMOVES
    // what to run
    int mode;                   // 0: nmoves attempts of movetype; 1: ncycles GCMC cycles
    int movetype;
    int64_t nmoves;
    int64_t ncycles, min_cycle;
    double pcum[NMOVES];        // mode 1: cumulative probabilities of the move types
    const double *step_trans, *step_rot;   // per chain
    const double *fug;          // per chain, eV/Å^3
    const double *const_p, *const_q;       // exchange constant term p + q N (src/moves.jl:728-773)
    double kT;
    int upper;                  // every framework's A has A[1] == A[2] == A[5] == 0
    // mode 1 sampling: after every cycle c of this launch, samp_*[n * samp_stride + samp_base + c]
    int sampling;
    int64_t samp_base, samp_stride;
    double *samp_e, *samp_n;
    int *max_occ;
    int *capacity_hits;         // per chain: accepted-by-Metropolis insertions refused at capacity
    double4 *trace;             // optional, per chain trace_cap entries: type, slot, dU, accepted
    int64_t trace_cap;
    // mode 1 snapshots: the state at the end of cycle c (0-based within the launch) for c < snap_n
    // (layout of ChainSnap)
    int64_t snap_n;
    ChainSnap snap;
};

// Test insertions into the chains' current configurations (nvt_kernels.hip). Insertion i of
// system s lies in [item_off[s], item_off[s+1]); local index l = i - item_off[s] is cycle slot
// l / nw, insertion l % nw of that cycle. Its pose is pose_pos/pose_quat[i] when given, else the
// Widom stream's pose of per-system index (cyc0 + slot) * nw + l % nw. Cycle slot c sees snapshot
// min(c, snap_n - 1).
struct ChainWidomArgs {
    const FwDesc *fw;
    const int *fw_of;
    // framework atoms of every framework re-ordered into spatially compact chunks of 64
    // (same per-framework offsets as the batch's atom tables)
    const double4 *atomf64;     // fractional x, y, z and charge
    const int *atype;           // compact type
    const double4 *chunk;       // per chunk: centre (fractional x, y, z), radius (Å)
    const int *chunk_off;       // nfw + 1
    // k-vectors of every framework grouped into runs of consecutive n3 at fixed (n1, n2),
    // longest first, 64 runs per round (one per lane; short rounds padded with empty runs)
    const int2 *lines;          // per (round, lane): first k-vector (framework-relative), length
    const int2 *rounds;         // per round: first entry of the round in the snapshot slice, 0
    const int *round_off;       // nfw + 1
    const double *half_perp;    // per framework: half the smallest perpendicular width (Å)
    double cull_extra;          // largest cutoff + guest radius + rounding margin (Å)
    int upper;                  // every framework's A has A[1] == A[2] == A[5] == 0
    const double4 *k64;         // kx, ky, kz, kprefactor (every k-vector)
    const int4 *kn;
    const double2 *pair64;
    int ntypes_c;
    double rc_lj2, rc_ew2;
    double rc_max2;             // max(rc_lj2, rc_ew2): the host-guest pair test
    GuestDev guest;
    int nsys;
    const int *guest_off;
    int64_t snap_n;
    ChainSnap snap;
    const int64_t *item_off;    // nsys + 1
    int64_t nitems, nw, cyc0;
    uint64_t widom_seed;
    const double *pose_pos, *pose_quat;
    double *dU;                 // nitems: energy without the insertion constant
};

}  // namespace pa
