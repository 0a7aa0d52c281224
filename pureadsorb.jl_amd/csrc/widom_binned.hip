// widom_binned.hip -- the mixed-precision Widom pipeline for gfx950.
//
// One chunk of insertions g in [g0, g1) runs four launches:
//  1. bin_pose_kernel       draws each pose (FP64, bit-identical to the oracle's), stores its sites
//                           as 32-bit fixed-point fractional coordinates, and counts insertions
//                           per spatial bin of the system's cell;
//  2. exclusive scan of the bin counts and the launch order of step 4, the most crowded bins
//     first (one workgroup for up to BIN_ORDER_MAX bins, else hipCUB's scan and radix sort);
//  3. bin_scatter_kernel    groups the chunk's insertions by bin (positions from step 1's counts);
//  4. widom_binned_kernel   one workgroup per bin: builds the bin's neighbour list in LDS (every
//                           atom image within rc + r_guest + r_bin of the bin centre, FP32
//                           coordinates relative to that centre, computed in FP64, in order of
//                           16 radial classes; a longer list than the LDS holds is taken in
//                           windows of that order over several passes), then each
//                           wavefront claims one insertion at a time, sweeps the list with lanes
//                           over atoms, compacts in-cutoff (site, atom) pairs into a per-wave LDS
//                           ring, and evaluates the FP32 LJ + screened-Coulomb terms on full
//                           wavefronts into FP64 accumulators; the reciprocal cross term uses
//                           exact fixed-point phases n.f mod 1 and the hardware sin/cos.
// Every insertion's energy depends only on its own pose and its bin's list, so results do not
// depend on chunking or on the order insertions are grouped in.
#include <hipcub/hipcub.hpp>

#include "pa_device.hpp"

namespace pa {

constexpr int BIN_WG = 768;
constexpr int BIN_NW = BIN_WG / 64;
static_assert(BIN_WG % 64 == 0, "a bin workgroup is whole wavefronts, and the smallest list capacity is BIN_WG");
constexpr int BRING = 128;  // per-wave ring; drained after every site sweep (< 64 + 64 pending)
static_assert(BRING == 128, "a drain takes the 64 entries at head, a multiple of 64, of a 128-entry ring");
constexpr int BIN_MAX_TYPES = 64;
constexpr int BIN_NCLS = 16;  // radial classes of a sorted neighbour list (4 bits)
static_assert(BIN_NCLS == 16, "the in-wave class ranks use four class bits");
// a list row lies wholly inside (outside) the cutoff of a site when its distance bounds clear the
// cutoff by this much (Å): far above the FP32 rounding of the coordinates, r^2 and the bounds
constexpr float ROW_MARGIN = 1.0e-3f;

__device__ inline int find_sys64(const int64_t *seg, int nsys, int64_t g)
{
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The system of insertion g for a wavefront of consecutive insertions: one binary search for the
// first active lane's g (uniform, so on scalar loads), then each lane steps over the system
// boundaries between that g and its own (none unless a system ends inside the wavefront).
__device__ inline int find_sys64_wave(const int64_t *seg, int nsys, int64_t g)
{
    const uint64_t gu = (uint64_t)g;
    const int64_t g0 = (int64_t)((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)gu) |
                                 ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(gu >> 32)) << 32));
    int s = find_sys64(seg, nsys, g0);
    while (s + 1 < nsys && seg[s + 1] <= g) ++s;
    while (s > 0 && seg[s] > g) --s;
    return s;
}

__device__ inline int find_sys32(const int *base, int nsys, int b)
{
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Orders this wave's LDS ring accesses (written and read by different lanes of one wave).
__device__ inline void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Grid cells, along one axis with nc cells, that hold the coordinates within +-rf (32-bit fixed
// point) of x: `cnt` cells from `start` upwards, wrapping. Atoms are binned as (x nc) >> 32. An
// interval that reaches around the axis (2 rf + 2^32 / nc >= 2^32, decided once per system on the
// host) comes as rf = all ones and covers every cell.
__device__ inline void cell_range(uint32_t x, uint32_t rf, int nc, int &start, int &cnt)
{
    if (rf == 0xFFFFFFFFu) {
        start = 0;
        cnt = nc;
        return;
    }
    const int lo = (int)(((uint64_t)(uint32_t)(x - rf) * (uint64_t)nc) >> 32);
    const int hi = (int)(((uint64_t)(uint32_t)(x + rf) * (uint64_t)nc) >> 32);
    const int span = hi >= lo ? hi - lo : hi - lo + nc;
    start = lo;
    cnt = min(span + 1, nc);
}

// Phase-0 test of one insertion (src/widom.jl:99-151): true when some site lies closer than
// rho_at to an atom of type t (the radii guarantee exp(-dU/kT) == 0.0 in FP64). The atoms come
// from the hard-core grid cells that hold coordinates within the largest radius of the site,
// +-rf[d] along each axis (the radius over the perpendicular width bounds the fractional
// offset), with the exact fixed-point minimum image. Atoms are sorted by cell with c0 fastest,
// so the stencil's cells of one (c1, c2) row are one contiguous atom range (two when the row
// wraps): one dependent cell_start lookup per row.
template <int NS>
__device__ inline bool stencil_hit(const BinnedArgs &a, const FwDesc &F, int s, const uint32_t (*sp)[3],
                                   const uint32_t rf[3])
{
    const int *cs = a.cell_start + F.cg_off;
    const uint4 *cat = a.corefx + F.atom_off;
    const float *rho = a.rho2 + (size_t)s * NS * a.ntypes_c;
    const int nc[3] = {F.cg_nc[0], F.cg_nc[1], F.cg_nc[2]};
    int cnt[3], start[3];
    bool hit = false;
    for (int t = 0; t < NS && !hit; ++t) {
        for (int d = 0; d < 3; ++d) cell_range(sp[t][d], rf[d], nc[d], start[d], cnt[d]);
        // wrapped indices: start + i < 2 nc, so one conditional subtraction replaces the modulo
        const int a0 = start[0], a1 = start[0] + cnt[0];
        const int e0 = a1 < nc[0] ? a1 : nc[0], w1 = a1 > nc[0] ? a1 - nc[0] : 0;
        for (int i2 = 0; i2 < cnt[2] && !hit; ++i2) {
            int c2 = start[2] + i2;
            c2 = c2 >= nc[2] ? c2 - nc[2] : c2;
            for (int i1 = 0; i1 < cnt[1] && !hit; ++i1) {
                int c1 = start[1] + i1;
                c1 = c1 >= nc[1] ? c1 - nc[1] : c1;
                const int row = nc[0] * (c1 + nc[1] * c2);
                const int j0 = cs[row + a0], j1 = cs[row + e0];
                const int k0 = cs[row], k1 = cs[row + w1];
                const int n1 = j1 - j0, ntot = n1 + (k1 - k0);
                for (int q = 0; q < ntot; ++q) {
                    const uint4 at = cat[q < n1 ? j0 + q : k0 + (q - n1)];
                    const float dx = (float)(int32_t)(sp[t][0] - at.x);
                    const float dy = (float)(int32_t)(sp[t][1] - at.y);
                    const float dz = (float)(int32_t)(sp[t][2] - at.z);
                    const float ex = F.Af[0] * dx + F.Af[3] * dy + F.Af[6] * dz;
                    const float ey = F.Af[1] * dx + F.Af[4] * dy + F.Af[7] * dz;
                    const float ez = F.Af[2] * dx + F.Af[5] * dy + F.Af[8] * dz;
                    if (ex * ex + ey * ey + ez * ez < rho[t * a.ntypes_c + (int)at.w]) hit = true;
                }
            }
        }
    }
    return hit;
}

// Pose, bin and the near part of phase 0 of every insertion of the chunk. The near stencil
// (the home cell and its neighbours) rejects most insertions that phase 0 rejects; an insertion
// whose full stencil reaches further and that the near stencil did not reject is appended to
// the candidate list for bin_reject_kernel, which finishes phase 0 on full wavefronts.
template <int NS>
__global__ __launch_bounds__(256) void bin_pose_kernel(BinnedArgs a)
{
#pragma clang fp contract(off)
    const int64_t g = a.g0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= a.g1) return;
    const int s = find_sys64_wave(a.seg, a.nsys, g);
    const int64_t j = g - a.seg[s];
    const FwDesc &F = a.fw[a.fw_of[s]];
    double f[3], q[4];
    pose_uniforms(a.seed, s, j, f, q);
    uint32_t *out = a.sitefx + (size_t)(g - a.g0) * 3 * NS;
    uint32_t sp[NS][3];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        double r[3], df[3];
        rotate(q, a.guest.site[t], r);
        matvec(F.invA, r, df);
        for (int d = 0; d < 3; ++d) {
            sp[t][d] = to_fixed(f[d] + df[d]);
            out[3 * t + d] = sp[t][d];
        }
    }
    const BinSys B = a.bins[s];
    const int bx = min(B.nb[0] - 1, (int)(f[0] * B.nb[0]));
    const int by = min(B.nb[1] - 1, (int)(f[1] * B.nb[1]));
    const int bz = min(B.nb[2] - 1, (int)(f[2] * B.nb[2]));
    const int key = B.bin_base + bx + B.nb[0] * (by + B.nb[1] * bz);
    bool pending = false;
    if (B.reach[0] >= 0) {
        // the near stencil (home cell and its neighbours; reach >= 1 whenever phase 0 runs) from
        // the home cell's precomputed atom list: one contiguous range per site
        const int *nbs = a.nbr_start + F.cg_off;
        const float *rho = a.rho2 + (size_t)s * NS * a.ntypes_c;
        int j0[NS], j1[NS];
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const int h0 = (int)(((uint64_t)sp[t][0] * (uint64_t)F.cg_nc[0]) >> 32);
            const int h1 = (int)(((uint64_t)sp[t][1] * (uint64_t)F.cg_nc[1]) >> 32);
            const int h2 = (int)(((uint64_t)sp[t][2] * (uint64_t)F.cg_nc[2]) >> 32);
            const int c = h0 + F.cg_nc[0] * (h1 + F.cg_nc[1] * h2);
            j0[t] = nbs[c];
            j1[t] = nbs[c + 1];
        }
        bool hit = false;
#pragma unroll
        for (int t = 0; t < NS; ++t)
            for (int jj = j0[t]; jj < j1[t] && !hit; ++jj) {
                const uint4 at = a.nbr_atoms[jj];
                const float dx = (float)(int32_t)(sp[t][0] - at.x);
                const float dy = (float)(int32_t)(sp[t][1] - at.y);
                const float dz = (float)(int32_t)(sp[t][2] - at.z);
                // FP32 distance with fused multiply-adds (the kernel's FP64 pose above keeps
                // contraction off; the radii carry the FP32 rounding allowance either way)
                const float ex = __builtin_fmaf(F.Af[6], dz, __builtin_fmaf(F.Af[3], dy, F.Af[0] * dx));
                const float ey = __builtin_fmaf(F.Af[7], dz, __builtin_fmaf(F.Af[4], dy, F.Af[1] * dx));
                const float ez = __builtin_fmaf(F.Af[8], dz, __builtin_fmaf(F.Af[5], dy, F.Af[2] * dx));
                const float r2 = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
                if (r2 < rho[t * a.ntypes_c + (int)at.w]) hit = true;
            }
        if (hit) {
            a.key[g - a.g0] = -1;
            a.dU[g - a.out_base] = INFINITY;
            return;
        }
        // the full test is pending when some site's cells within the largest radius reach
        // beyond the home cell's neighbours
        if (B.reach[0] > 1 || B.reach[1] > 1 || B.reach[2] > 1) {
#pragma unroll
            for (int t = 0; t < NS; ++t)
                for (int d = 0; d < 3; ++d) {
                    const int nc = F.cg_nc[d];
                    if (nc <= 3) continue;
                    const int home = (int)(((uint64_t)sp[t][d] * (uint64_t)nc) >> 32);
                    int st, cnt;
                    cell_range(sp[t][d], B.hc_rf[d], nc, st, cnt);
                    // cells [st, st + cnt) against [home - 1, home + 1], both wrapping
                    const int below = home - st < 0 ? home - st + nc : home - st;
                    if (below > 1 || cnt - below > 2) pending = true;
                }
        }
    }
    a.key[g - a.g0] = key;
    if (!pending) {
        a.rank[g - a.g0] = atomicAdd(&a.bin_count[key], 1);
        return;
    }
    // wavefront-aggregated append to the candidate list
    const uint64_t m = __ballot(1);
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(a.ncand, __popcll(m));
    base = __shfl(base, leader, 64);
    a.cand[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)(g - a.g0);
}

// Phase 0 on the full stencil for the candidates the near stencil left (grid-stride over the
// device-side candidate count): rejected -> key -1 and dU = +inf, else counted in its bin.
template <int NS>
__global__ __launch_bounds__(256) void bin_reject_kernel(BinnedArgs a)
{
    const int n = *a.ncand;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int rel = a.cand[i];
        const int64_t g = a.g0 + rel;
        const int s = find_sys64(a.seg, a.nsys, g);
        const FwDesc &F = a.fw[a.fw_of[s]];
        const BinSys B = a.bins[s];
        const uint32_t *sf = a.sitefx + (size_t)rel * 3 * NS;
        uint32_t sp[NS][3];
#pragma unroll
        for (int t = 0; t < NS; ++t)
            for (int d = 0; d < 3; ++d) sp[t][d] = sf[3 * t + d];
        if (stencil_hit<NS>(a, F, s, sp, B.hc_rf)) {
            a.key[rel] = -1;
            a.dU[g - a.out_base] = INFINITY;
        } else {
            a.rank[rel] = atomicAdd(&a.bin_count[a.key[rel]], 1);
        }
    }
}

// Launch order of the energy kernel for a grid of at most BIN_ORDER_MAX bins, and the exclusive
// scan of the counts, in one workgroup (a larger grid takes hipCUB's scan and radix sort): the
// counts are staged in LDS, each thread scans a contiguous range, and the bins are placed by
// count class, most crowded first. Classes are half octaves of the count (empty bins last), so
// within a class the counts differ by less than 1.5x and their order is arbitrary: enough for
// the heaviest-first rule, and no insertion's energy depends on the order.
constexpr int BIN_ORDER_WG = 1024;
constexpr int BIN_ORDER_MAX = 14336;  // 56 KB of counts in LDS
constexpr int BIN_ORDER_NCLS = 64;

__device__ inline int order_class(int c)
{
    if (c <= 0) return 0;
    const int e = 31 - __clz(c);                      // floor(log2 c), 0..30
    const int half = e > 0 ? (c >> (e - 1)) & 1 : 0;  // upper or lower half of the octave
    return 1 + 2 * e + half;                          // 1..62
}

__global__ __launch_bounds__(BIN_ORDER_WG) void bin_order_kernel(BinnedArgs a)
{
    __shared__ int cs[BIN_ORDER_MAX + 1];  // counts, then starts
    __shared__ int part[BIN_ORDER_WG];
    __shared__ int ncls[BIN_ORDER_NCLS], cursor[BIN_ORDER_NCLS];
    const int tid = threadIdx.x, nb = a.nbins, n1 = nb + 1;
    for (int i = tid; i < n1; i += BIN_ORDER_WG) cs[i] = a.bin_count[i];
    if (tid < BIN_ORDER_NCLS) ncls[tid] = 0;
    __syncthreads();
    const int per = (n1 + BIN_ORDER_WG - 1) / BIN_ORDER_WG;
    const int r0 = min(n1, tid * per), r1 = min(n1, r0 + per);
    int sum = 0;
    for (int i = r0; i < r1; ++i) sum += cs[i];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < BIN_ORDER_WG; off <<= 1) {  // inclusive scan of the range sums
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int i = r0; i < r1; ++i) {
        const int c = cs[i];
        cs[i] = run;
        run += c;
        if (i < nb) atomicAdd(&ncls[order_class(c)], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int k = BIN_ORDER_NCLS - 1; k >= 0; --k) {
            cursor[k] = acc;
            acc += ncls[k];
        }
    }
    __syncthreads();
    for (int i = tid; i < n1; i += BIN_ORDER_WG) a.bin_start[i] = cs[i];
    for (int i = tid; i < nb; i += BIN_ORDER_WG) a.bin_perm[atomicAdd(&cursor[order_class(cs[i + 1] - cs[i])], 1)] = i;
}

// Clears the bin counts (nbins + 1) and the candidate counter of one chunk in one launch.
__global__ __launch_bounds__(256) void bin_zero_kernel(BinnedArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i <= a.nbins) a.bin_count[i] = 0;
    if (a.bin_perm && a.nbins > BIN_ORDER_MAX && i < a.nbins) a.bin_iota[i] = i;  // the radix sort's values
    if (i == 0) *a.ncand = 0;
    if (i == 0 && a.zero_passes) *a.extra_passes = 0;
}

// Each insertion's position within its bin is the count its atomic increment returned (pose or
// reject kernel); the order within a bin does not change any insertion's energy.
__global__ __launch_bounds__(256) void bin_scatter_kernel(BinnedArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.g1 - a.g0) return;
    const int k = a.key[i];
    if (k < 0) return;  // rejected in phase 0
    a.order[a.bin_start[k] + a.rank[i]] = (int)i;
}

// Cutoff classes of the binned energy kernel. The ring only holds pairs inside the larger of the
// two cutoffs, so at most one of the two terms needs its own test: CUT_NONE when the cutoffs are
// equal or the guest is uncharged (LJ only), CUT_EW when the LJ cutoff is the larger (the
// screened-Coulomb term is tested), CUT_LJ when the Ewald cutoff is the larger.
constexpr int CUT_NONE = 0, CUT_EW = 1, CUT_LJ = 2;

// LDS (address space 3) views: a 32-bit LDS address held in a register is dereferenced
// directly, with no base added.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// Fourth word of a raw (untyped) buffer resource on gfx9-family GPUs: 32-bit data format, no
// swizzle; offsets are bounds-checked against the record count.
constexpr int BUF_RSRC_WORD3 = 0x00020000;
typedef __attribute__((address_space(3))) const f32x2 lds_f32x2;
typedef __attribute__((address_space(3))) unsigned char lds_byte;

// One compacted (site, atom) pair: r2 (Å^2), qq = KE q_site q_atom (FP32) and the LDS address of
// the (sigma^2, 4 eps) pair parameters.
template <bool CHARGED, int CUT>
__device__ inline float pair_term(float r2, float qq, uint32_t paddr, float rc_lj2, float rc_ew2, float alpha,
                                  float alpha2)
{
    // r2 >= +0, and the bit patterns of non-negative floats order like their values: an integer
    // max is the 1e-30 floor that keeps 1/r finite at a coincident site
    r2 = __uint_as_float(max(__float_as_uint(r2), 0x0DA24260u));
    float e = 0.0f;
    if (CUT != CUT_LJ || r2 < rc_lj2) {
        const f32x2 p = *(lds_f32x2 *)(uintptr_t)paddr;
        const float x = p.x * __builtin_amdgcn_rcpf(r2);
        const float x3 = x * x * x;
        e = p.y * x3 * (x3 - 1.0f);
    }
    if (CHARGED && (CUT != CUT_EW || r2 < rc_ew2)) {
        // pair_erfc_dev in Float32 (src/ewald.jl:242-254): the reference's 8-term series, as
        // the same polynomial in Horner form
        constexpr float XLO = (float)PAIR_ERFC_XLO;
        constexpr float XA = 4.0f / (1.0f - XLO);
        constexpr float XB = -2.0f * (1.0f + XLO) / (1.0f - XLO) - 1.0f;
        // t = 2/(2 + alpha r) = 1/(1 + (alpha/2) r); exp2 of the log2(e)-scaled exponent
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float tt = __builtin_amdgcn_rcpf(1.0f + (0.5f * alpha) * (r2 * rinv));
        const float xi = XA * tt + XB;
        const float arg2 = pair_erfc_poly32_log2e(xi) - (alpha2 * 1.4426950408889634f) * r2;
        e += qq * (tt * __builtin_amdgcn_exp2f(arg2)) * rinv;
    }
    return e;
}

__device__ inline double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Dynamic LDS layout of widom_binned_kernel for a list capacity `cap`.
__host__ __device__ inline size_t binned_lds_bytes(int cap, int ns)
{
    size_t off = (size_t)cap * 16;                      // float4 entries
    off += ((size_t)cap * 2 + 15) & ~(size_t)15;        // uint16 types
    off += (size_t)BIN_NW * BRING * 12;                 // rings
    off += (size_t)ns * BIN_MAX_TYPES * 8;              // pair table
    off += 16 * 4;                                      // KE * q per site; slot 8: insertion counter
    off += BIN_NCLS * 4;                                // class sizes
    off += (size_t)BIN_NW * BIN_NCLS * 4;               // first list position per (wave, class)
    off += (size_t)(cap / 64) * 16;                     // row bounds: (dmin, dmax), (prefix max, suffix min)
    return off;
}

template <int NS, bool CHARGED, int CUT>
__global__ __launch_bounds__(BIN_WG, 6) void widom_binned_kernel(BinnedArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // heaviest bins first, so that the launch does not end on a few crowded bins
    const int b = a.bin_perm ? a.bin_perm[blockIdx.x] : (int)blockIdx.x;
    const int i0 = a.bin_start[b];
    const int nins = a.bin_start[b + 1] - i0;
    if (nins == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cap = a.listcap;
    const int ntc = a.ntypes_c;

    float4 *ent = (float4 *)smem;
    unsigned short *etype = (unsigned short *)(smem + (size_t)cap * 16);
    size_t off = (size_t)cap * 16 + (((size_t)cap * 2 + 15) & ~(size_t)15);
    float *ring_all = (float *)(smem + off);
    float *ring_r2 = ring_all + wave * BRING * 3;
    float *ring_q = ring_r2 + BRING;
    int *ring_info = (int *)(ring_q + BRING);
    off += (size_t)BIN_NW * BRING * 12;
    float2 *pair_lds = (float2 *)(smem + off);
    // LDS address of the pair table (the ring stores per-pair parameter addresses)
    const uint32_t pair_addr = (uint32_t)(uintptr_t)((lds_byte *)smem + off);
    off += (size_t)NS * BIN_MAX_TYPES * 8;
    float *qke_lds = (float *)(smem + off);
    off += 16 * 4;
    int *cls_cnt = (int *)(smem + off);
    off += BIN_NCLS * 4;
    int *wcls = (int *)(smem + off);
    off += (size_t)BIN_NW * BIN_NCLS * 4;
    float2 *rowb = (float2 *)(smem + off);           // per row: (dmin, dmax)
    float2 *rowc = rowb + cap / 64;                   // per row: (prefix max of dmax, suffix min of dmin)

    const int s = find_sys32(a.bin_base, a.nsys, b);
    const BinSys B = a.bins[s];
    const FwDesc &F = a.fw[a.fw_of[s]];
    const int local = b - B.bin_base;
    const int bx = local % B.nb[0], by = (local / B.nb[0]) % B.nb[1], bz = local / (B.nb[0] * B.nb[1]);
    const uint32_t cx = to_fixed((bx + 0.5) / B.nb[0]);
    const uint32_t cy = to_fixed((by + 0.5) / B.nb[1]);
    const uint32_t cz = to_fixed((bz + 0.5) / B.nb[2]);
    for (int i = tid; i < NS * ntc; i += BIN_WG) pair_lds[(i / ntc) * ntc + (i % ntc)] = a.pair32[i];
    if (tid < NS) qke_lds[tid] = a.guest.qke[tid];
    static_assert(NS <= 8, "the insertion counter shares the per-site charge slots");
    int *next_ins = (int *)(qke_lds + 8);  // the bin's insertion counter (slot 8 of 16)

    const double A0 = F.A[0], A1 = F.A[1], A2 = F.A[2], A3 = F.A[3], A4 = F.A[4], A5 = F.A[5], A6 = F.A[6],
                 A7 = F.A[7], A8 = F.A[8];
    const uint4 *atoms = a.atomfx + F.atom_off;
    const int *types = a.atype + F.atom_off;
    const int natoms = F.natoms;
    const double rl2 = (double)B.rlist2;
    const float rc_lj2 = a.rc_lj2f, rc_ew2 = a.rc_ew2f, rc_max2 = a.rc_max2f;
    const float rc_max = __builtin_sqrtf(rc_max2);
    const float alpha = F.alphaf, alpha2 = F.alphaf * F.alphaf;
    const float cls_d0 = B.cls_d0, cls_inv_w = B.cls_inv_w;
    constexpr double FIX = 2.3283064365386963e-10;  // 2^-32
    // image combinations: bit d set = the image across the boundary on axis d (only for atoms
    // within alt[d] of it)
    auto combo_enabled = [&](int combo) {
        for (int d = 0; d < 3; ++d)
            if (((combo >> d) & 1) && !(B.alt[d] < 0.5f)) return false;
        return true;
    };
    // list entry of the atom `at` (fixed point, charge bits in w) under image combination
    // `combo`: coordinates relative to the bin centre in FP32 (computed in FP64), charge and
    // radial class; false when the image is beyond the list radius
    auto make_entry = [&](const uint4 at, int combo, float4 &e, int &cls) -> bool {
        double f0 = (double)(int32_t)(at.x - cx) * FIX;
        double f1 = (double)(int32_t)(at.y - cy) * FIX;
        double f2 = (double)(int32_t)(at.z - cz) * FIX;
        if (combo & 1) { if (fabs(f0) >= B.alt[0]) f0 -= (f0 > 0 ? 1.0 : -1.0); else return false; }
        if (combo & 2) { if (fabs(f1) >= B.alt[1]) f1 -= (f1 > 0 ? 1.0 : -1.0); else return false; }
        if (combo & 4) { if (fabs(f2) >= B.alt[2]) f2 -= (f2 > 0 ? 1.0 : -1.0); else return false; }
        // the general product: the exact zeros of an upper-triangular cell add exact zeros
        const double X = A0 * f0 + A3 * f1 + A6 * f2;
        const double Y = A1 * f0 + A4 * f1 + A7 * f2;
        const double Z = A2 * f0 + A5 * f1 + A8 * f2;
        const double d2 = X * X + Y * Y + Z * Z;
        if (!(d2 <= rl2)) return false;
        e = make_float4((float)X, (float)Y, (float)Z, __uint_as_float(at.w));
        const float dr = __builtin_sqrtf((float)d2);
        cls = dr < cls_d0 ? 0 : min(BIN_NCLS - 1, 1 + (int)((dr - cls_d0) * cls_inv_w));
        return true;
    };
    int combos = 0;
    for (int c = 0; c < 8; ++c)
        if (combo_enabled(c)) combos |= 1 << c;
    // the image test of make_entry on the fixed-point offset: |f_d| >= alt[d] exactly when
    // |(int32)(x - c)| >= ceil(alt[d] 2^32)
    uint32_t altfx[3];
    for (int d = 0; d < 3; ++d)
        altfx[d] = (uint32_t)fmin(4294967295.0, __builtin_ceil((double)B.alt[d] * 4294967296.0));
    // running position per class of this wave's entries: 16 ints of the wave's ring, which is
    // idle while the list is built
    int *tally = ring_info;
    auto tally_of = [&](int w) { return (int *)(ring_all + w * BRING * 3 + 2 * BRING); };

    // One scan over the atoms. Wave w takes the 64-atom blocks w, w + BIN_NW, ..., loads each
    // atom once and tries every enabled image combination of it. The entries of one class are
    // ordered by (wave, block, combination, lane), so a wave ranks its entries against its own
    // running positions and the scan takes no workgroup barrier. With `write`, the entries whose
    // position in the class-ordered list falls in [lo, lo + cap) are stored; without, the
    // tallies (started at zero) end as the wave's entry counts per class.
    auto scan = [&](bool write, int lo) {
        for (int p0 = wave * 64; p0 < natoms; p0 += BIN_WG) {
            const int n = p0 + lane;
            const bool live = n < natoms;
            const uint4 at = live ? atoms[n] : make_uint4(0u, 0u, 0u, 0u);
            const int et = (live && write) ? types[n] : 0;
            // |fractional offset| of the atom from the bin centre per axis, fixed point
            auto absfx = [](uint32_t v) { return (int32_t)v < 0 ? 0u - v : v; };
            const uint32_t ad[3] = {absfx(at.x - cx), absfx(at.y - cy), absfx(at.z - cz)};
            for (int cb = 0; cb < 8; ++cb) {
                if (!((combos >> cb) & 1)) continue;
                // an image across a boundary that no atom of the wavefront can take is skipped
                // before any FP64 work (the wavefront's atoms are neighbours in Morton order)
                if (cb != 0) {
                    bool img = live;
                    for (int d = 0; d < 3; ++d)
                        if ((cb >> d) & 1) img = img && ad[d] >= altfx[d];
                    if (__builtin_amdgcn_ballot_w64(img) == 0) continue;
                }
                float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
                int cls = 0;
                const bool inc = live && make_entry(at, cb, e, cls);
                uint64_t same = __builtin_amdgcn_ballot_w64(inc);
                if (same == 0) continue;
                // lanes of the same class: the four class bits agree
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint64_t m = __builtin_amdgcn_ballot_w64(inc && ((cls >> i) & 1));
                    same &= ((cls >> i) & 1) ? m : ~m;
                }
                const uint32_t rank =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
                const int pos = tally[cls] + (int)rank;
                wave_lds_fence();
                if (inc && rank == 0) tally[cls] = pos + __popcll(same);
                wave_lds_fence();
                if (write && inc) {
                    const int slot = pos - lo;
                    if ((uint32_t)slot < (uint32_t)cap) {
                        ent[slot] = e;
                        etype[slot] = (unsigned short)et;
                    }
                }
            }
        }
    };
    __syncthreads();

    // ---- count the entries per (wave, class); then the class sizes and the first position of
    // every (wave, class) in the class-ordered list
    if (lane < BIN_NCLS) tally[lane] = 0;
    wave_lds_fence();
    scan(false, 0);
    __syncthreads();
    if (tid < BIN_NCLS) {
        int tot = 0;
        for (int w = 0; w < BIN_NW; ++w) tot += tally_of(w)[tid];
        cls_cnt[tid] = tot;
    }
    __syncthreads();
    if (tid < BIN_NCLS) {
        int run = 0;
        for (int k = 0; k < tid; ++k) run += cls_cnt[k];
        for (int w = 0; w < BIN_NW; ++w) {
            wcls[w * BIN_NCLS + tid] = run;
            run += tally_of(w)[tid];
        }
    }
    __syncthreads();
    int total = 0;
    for (int k = 0; k < BIN_NCLS; ++k) total += cls_cnt[k];

    // The list is written in class order, nearest class first. A list longer than the capacity
    // is built and swept in windows of `cap` consecutive entries of that order, one window per
    // pass, so every window keeps tight row bounds and its far rows are skipped.
    int pass = 0;
    for (;;) {
        const int lo = pass * cap;
        const int cnt = min(cap, total - lo);
        // ---- build the window
        if (lane < BIN_NCLS) tally[lane] = wcls[wave * BIN_NCLS + lane];
        wave_lds_fence();
        scan(true, lo);
        __syncthreads();
        // pad the list to a multiple of 64 with far-away sentinels (no lane of the sweep needs a
        // bounds test); listcap is a multiple of 64, so the padding stays inside the list
        const int cnt64 = (cnt + 63) & ~63;
        const int nrows = cnt64 / 64;
        for (int e = cnt + tid; e < cnt64; e += BIN_WG) {
            ent[e] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
            etype[e] = 0;
        }
        __syncthreads();
        // ---- distance bounds of every 64-entry row (a sentinel's distance is +inf)
        for (int r = wave; r < nrows; r += BIN_NW) {
            const float4 E = ent[r * 64 + lane];
            float lo = __builtin_sqrtf(E.x * E.x + E.y * E.y + E.z * E.z), hi = lo;
#pragma unroll
            for (int off2 = 32; off2 >= 1; off2 >>= 1) {
                lo = fminf(lo, __shfl_xor(lo, off2, 64));
                hi = fmaxf(hi, __shfl_xor(hi, off2, 64));
            }
            if (lane == 0) rowb[r] = make_float2(lo, hi);
        }
        __syncthreads();
        // prefix maximum of the upper bounds and suffix minimum of the lower bounds: the rows
        // wholly inside a site's cutoff form a prefix, those wholly outside a suffix. Up to 64
        // rows: one wavefront, a lane per row, six shuffle steps (max and min are exact, so any
        // order gives the same bounds); longer lists: a thread per row
        if (nrows <= 64) {
            if (wave == 0) {
                const float2 rb = lane < nrows ? rowb[lane] : make_float2(INFINITY, 0.0f);
                float pm = rb.y, sm = rb.x;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const float up = __shfl_up(pm, off, 64), dn = __shfl_down(sm, off, 64);
                    if (lane >= off) pm = fmaxf(pm, up);
                    if (lane + off < 64) sm = fminf(sm, dn);
                }
                if (lane < nrows) rowc[lane] = make_float2(fmaxf(pm, 0.0f), sm);
            }
        } else {
            for (int r = tid; r < nrows; r += BIN_WG) {
                float pm = 0.0f, sm = INFINITY;
                for (int q = 0; q <= r; ++q) pm = fmaxf(pm, rowb[q].y);
                for (int q = r; q < nrows; ++q) sm = fminf(sm, rowb[q].x);
                rowc[r] = make_float2(pm, sm);
            }
        }
        if (tid == 0) *next_ins = BIN_NW;
        __syncthreads();

        // ---- evaluate the bin's insertions against the list: each wavefront starts with its own
        // index and claims the next unclaimed one when it finishes (costs differ per insertion)
        for (int ii = wave; ii < nins;) {
            const int grel = a.order[i0 + ii];
            const uint32_t *sf = a.sitefx + (size_t)grel * 3 * NS;
            uint32_t sfx[NS][3];
            float sxr[NS], syr[NS], szr[NS], qk[NS];
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                for (int d = 0; d < 3; ++d) sfx[t][d] = __builtin_amdgcn_readfirstlane(sf[3 * t + d]);
                const double f0 = (double)(int32_t)(sfx[t][0] - cx) * FIX;
                const double f1 = (double)(int32_t)(sfx[t][1] - cy) * FIX;
                const double f2 = (double)(int32_t)(sfx[t][2] - cz) * FIX;
                sxr[t] = (float)(A0 * f0 + A3 * f1 + A6 * f2);
                syr[t] = (float)(A1 * f0 + A4 * f1 + A7 * f2);
                szr[t] = (float)(A2 * f0 + A5 * f1 + A8 * f2);
                qk[t] = a.guest.qke[t];
            }
            double acc = 0.0;
            uint32_t head = 0, tail = 0;
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                // rows [0, ra) lie inside the cutoff of this site, rows [rb, nrows) outside it
                const float delta = __builtin_sqrtf(sxr[t] * sxr[t] + syr[t] * syr[t] + szr[t] * szr[t]);
                const float thr_in = rc_max - delta - ROW_MARGIN, thr_out = rc_max + delta + ROW_MARGIN;
                int ra = 0, rskip = 0;
                for (int r0 = 0; r0 < nrows; r0 += 64) {
                    const int r = r0 + lane;
                    const float2 c = r < nrows ? rowc[r] : make_float2(INFINITY, INFINITY);
                    ra += __popcll(__builtin_amdgcn_ballot_w64(c.x < thr_in));
                    rskip += __popcll(__builtin_amdgcn_ballot_w64(r < nrows && c.y > thr_out));
                }
                ra = __builtin_amdgcn_readfirstlane(ra);
                const int rend = __builtin_amdgcn_readfirstlane(nrows - rskip);
                const uint32_t tofs = pair_addr + 8u * (uint32_t)(t * ntc);
                for (int r = 0; r < ra; ++r) {
                    // every entry of the row is in range: evaluate without compaction
                    const float4 E = ent[r * 64 + lane];
                    const uint32_t paddr = tofs + 8u * (uint32_t)etype[r * 64 + lane];
                    const float dx = sxr[t] - E.x, dy = syr[t] - E.y, dz = szr[t] - E.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    acc += (double)pair_term<CHARGED, CUT>(r2, CHARGED ? qk[t] * E.w : 0.0f, paddr, rc_lj2, rc_ew2, alpha,
                                                           alpha2);
                }
                for (int r = ra; r < rend; ++r) {
                    const float4 E = ent[r * 64 + lane];
                    const uint32_t paddr = tofs + 8u * (uint32_t)etype[r * 64 + lane];
                    const float dx = sxr[t] - E.x, dy = syr[t] - E.y, dz = szr[t] - E.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    // sentinel entries: r2 overflows to +inf and fails the test
                    const bool in = r2 < rc_max2;
                    const uint64_t mask = __builtin_amdgcn_ballot_w64(in);
                    if (in) {
                        // tail + the in-range lanes below this one (mbcnt accumulates onto tail)
                        const uint32_t slot = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, tail)) & (BRING - 1);
                        ring_r2[slot] = r2;
                        if (CHARGED) ring_q[slot] = qk[t] * E.w;
                        ring_info[slot] = (int)paddr;
                    }
                    tail += (uint32_t)__popcll(mask);
                    if (tail - head >= 64u) {
                        wave_lds_fence();
                        // head is a multiple of 64: (head + lane) mod BRING
                        const uint32_t slot = (uint32_t)lane + (head & 64u);
                        acc += (double)pair_term<CHARGED, CUT>(ring_r2[slot], CHARGED ? ring_q[slot] : 0.0f,
                                                               (uint32_t)ring_info[slot], rc_lj2, rc_ew2, alpha, alpha2);
                        head += 64u;
                        wave_lds_fence();
                    }
                }
            }
            wave_lds_fence();
            if ((uint32_t)lane < tail - head) {
                const uint32_t slot = (uint32_t)lane + (head & 64u);
                acc += (double)pair_term<CHARGED, CUT>(ring_r2[slot], CHARGED ? ring_q[slot] : 0.0f,
                                                       (uint32_t)ring_info[slot], rc_lj2, rc_ew2, alpha, alpha2);
            }
            wave_lds_fence();
            if (CHARGED && pass == 0) {
                // the reciprocal cross term: full rounds of 64 k-vectors from a uniform base (one
                // lane per k-vector, no bounds test) and one partial round; each lane sums its
                // ~nk/64 terms in FP32 before the FP64 accumulator takes them
                const int nk = F.nk, nfull = nk & ~63;
                // buffer loads: the round's offset in a scalar register, the lane's in a fixed
                // vector register (no per-lane pointer arithmetic)
                const __amdgpu_buffer_rsrc_t rkn =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(a.kn + F.k_off), (short)0, nk * 16, BUF_RSRC_WORD3);
                const __amdgpu_buffer_rsrc_t rkc =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(a.kc + F.k_off), (short)0, nk * 8, BUF_RSRC_WORD3);
                auto kload = [&](int kb, int4 &nn, float2 &c) {
                    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rkn, lane * 16, kb * 16, 0);
                    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(rkc, lane * 8, kb * 8, 0);
                    nn = make_int4((int)v.x, (int)v.y, (int)v.z, (int)v.w);
                    c = make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
                };
                float kacc = 0.0f;
                bool folded = false;
                if constexpr (NS == 3) {
                    if (a.guest.mirror) {
                        // sites 1 and 2 mirror each other through site 0 with equal charges:
                        // q0 e^{i t0} + q1 (e^{i (t0 + p)} + e^{i (t0 - p)}) = e^{i t0} (q0 + 2 q1 cos p),
                        // p the phase of site 1 relative to site 0 (site 2's own fixed-point phase
                        // differs from t0 - p by at most |n| 2^-31 revolutions)
                        const uint32_t d0 = sfx[1][0] - sfx[0][0], d1 = sfx[1][1] - sfx[0][1],
                                       d2 = sfx[1][2] - sfx[0][2];
                        const float q0 = qke_lds[0], q1x2 = 2.0f * qke_lds[1];
                        auto term = [&](const int4 nn, const float2 c) {
                            const uint32_t ph =
                                (uint32_t)nn.x * sfx[0][0] + (uint32_t)nn.y * sfx[0][1] + (uint32_t)nn.z * sfx[0][2];
                            const uint32_t pd = (uint32_t)nn.x * d0 + (uint32_t)nn.y * d1 + (uint32_t)nn.z * d2;
                            const float rev = (float)(int32_t)ph * 2.3283064365386963e-10f;
                            const float rvd = (float)(int32_t)pd * 2.3283064365386963e-10f;
                            const float w = q0 + q1x2 * __builtin_amdgcn_cosf(rvd);
                            return w * (c.x * __builtin_amdgcn_cosf(rev) + c.y * __builtin_amdgcn_sinf(rev));
                        };
                        for (int kb = 0; kb < nfull; kb += 64) {
                            int4 nn;
                            float2 c;
                            kload(kb, nn, c);
                            kacc += term(nn, c);
                        }
                        if (lane < nk - nfull) {
                            int4 nn;
                            float2 c;
                            kload(nfull, nn, c);
                            kacc += term(nn, c);
                        }
                        folded = true;
                    }
                }
                if (!folded) {
                    auto term = [&](const int4 nn, const float2 c) {
                        float re = 0.0f, im = 0.0f;
#pragma unroll
                        for (int t = 0; t < NS; ++t) {
                            const uint32_t ph =
                                (uint32_t)nn.x * sfx[t][0] + (uint32_t)nn.y * sfx[t][1] + (uint32_t)nn.z * sfx[t][2];
                            const float rev = (float)(int32_t)ph * 2.3283064365386963e-10f;
                            const float qs = qke_lds[t];
                            re += qs * __builtin_amdgcn_cosf(rev);
                            im += qs * __builtin_amdgcn_sinf(rev);
                        }
                        return c.x * re + c.y * im;
                    };
                    for (int kb = 0; kb < nfull; kb += 64) {
                        int4 nn;
                        float2 c;
                        kload(kb, nn, c);
                        kacc += term(nn, c);
                    }
                    if (lane < nk - nfull) {
                        int4 nn;
                        float2 c;
                        kload(nfull, nn, c);
                        kacc += term(nn, c);
                    }
                }
                acc += (double)kacc;
            }
            const double Etot = wave_sum64(acc);
            if (lane == 0) {
                double *o = a.dU + (a.g0 + grel - a.out_base);
                if (pass == 0) *o = Etot + F.const_off;
                else *o += Etot;
            }
            int claim = 0;
            if (lane == 0) claim = atomicAdd(next_ins, 1);
            ii = __builtin_amdgcn_readfirstlane(claim);
        }
        if (lo + cap >= total) break;
        __syncthreads();
        ++pass;
    }
    if (pass > 0 && tid == 0) atomicMax(a.extra_passes, pass);
}

// ------------------------------------------------------------------------------------------
// Host-side launcher of one chunk
// ------------------------------------------------------------------------------------------
template <int NS>
static hipError_t launch_binned_ns(const BinnedArgs &a, bool upper, hipStream_t st)
{
    const int64_t n = a.g1 - a.g0;
    const dim3 g256((unsigned)((n + 255) / 256)), b256(256);
    hipLaunchKernelGGL((bin_pose_kernel<NS>), g256, b256, 0, st, a);
    const dim3 gr((unsigned)std::min<int64_t>((n + 255) / 256, 2048));
    hipLaunchKernelGGL((bin_reject_kernel<NS>), gr, b256, 0, st, a);
    return hipGetLastError();
}

template <int NS>
static hipError_t launch_binned_energy_ns(const BinnedArgs &a, hipStream_t st)
{
    // the sweep reads whole 64-entry rows; every pass takes the next `listcap` entries of the
    // class-ordered list, so any positive capacity finishes (the library keeps it at least one
    // build slice)
    if (a.listcap < BIN_WG || a.listcap % 64 != 0 || !a.extra_passes) return hipErrorInvalidValue;
    const size_t lds = binned_lds_bytes(a.listcap, NS);
    const dim3 grid((unsigned)a.nbins), block(BIN_WG);
    if (!a.guest.charged) {
        hipLaunchKernelGGL((widom_binned_kernel<NS, false, CUT_NONE>), grid, block, lds, st, a);
    } else if (a.rc_lj2f == a.rc_ew2f) {
        hipLaunchKernelGGL((widom_binned_kernel<NS, true, CUT_NONE>), grid, block, lds, st, a);
    } else if (a.rc_lj2f > a.rc_ew2f) {
        hipLaunchKernelGGL((widom_binned_kernel<NS, true, CUT_EW>), grid, block, lds, st, a);
    } else {
        hipLaunchKernelGGL((widom_binned_kernel<NS, true, CUT_LJ>), grid, block, lds, st, a);
    }
    return hipGetLastError();
}

// Scan temporary storage size for `nbins + 1` counts.
hipError_t binned_scan_bytes(int nbins, size_t *bytes)
{
    *bytes = 0;
    return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, (int *)nullptr, (int *)nullptr, nbins + 1);
}

// Temporary storage of the descending sort of `nbins` counts of `bits` bits (the launch order).
hipError_t binned_sort_bytes(int nbins, int bits, size_t *bytes)
{
    *bytes = 0;
    return hipcub::DeviceRadixSort::SortPairsDescending(nullptr, *bytes, (unsigned *)nullptr, (unsigned *)nullptr,
                                                        (int *)nullptr, (int *)nullptr, nbins, 0, bits);
}

// Largest neighbour-list capacity (entries, a multiple of 64) that keeps a bin workgroup's LDS
// within BIN_LDS_BUDGET, so that two workgroups share a CU's 160 KB.
constexpr size_t BIN_LDS_BUDGET = 80 * 1024;
int binned_listcap_min() { return BIN_WG; }
int binned_nclasses() { return BIN_NCLS; }

int binned_listcap_max(int ns)
{
    const size_t fixed = binned_lds_bytes(0, ns);
    int cap = (int)((BIN_LDS_BUDGET - fixed) / 18) & ~63;
    while (cap > 64 && binned_lds_bytes(cap, ns) > BIN_LDS_BUDGET) cap -= 64;
    return cap;
}

// Steps 1-3 of the pipeline (pose + bin counts, scan, scatter), after clearing bin_count.
hipError_t launch_binning(const BinnedArgs &a, void *scan_tmp, size_t scan_bytes, hipStream_t st)
{
    const int64_t n = a.g1 - a.g0;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bin_zero_kernel, dim3((unsigned)((a.nbins + 1 + 255) / 256)), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (a.guest.nsites) {
    case 1: e = launch_binned_ns<1>(a, true, st); break;
    case 2: e = launch_binned_ns<2>(a, true, st); break;
    case 3: e = launch_binned_ns<3>(a, true, st); break;
    case 4: e = launch_binned_ns<4>(a, true, st); break;
    case 5: e = launch_binned_ns<5>(a, true, st); break;
    case 6: e = launch_binned_ns<6>(a, true, st); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (a.bin_perm && a.nbins <= BIN_ORDER_MAX) {
        // scan and launch order in one workgroup
        hipLaunchKernelGGL(bin_order_kernel, dim3(1), dim3(BIN_ORDER_WG), 0, st, a);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(bin_scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
        return hipGetLastError();
    }
    size_t bytes = scan_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, a.bin_count, a.bin_start, a.nbins + 1, st);
    if (e != hipSuccess) return e;
    if (a.bin_perm) {
        // the energy kernel's launch order: bins by descending count (the same temporary storage,
        // sized for both, after the scan on the same stream)
        bytes = scan_bytes;
        e = hipcub::DeviceRadixSort::SortPairsDescending(scan_tmp, bytes, (const unsigned *)a.bin_count, a.bin_ksort,
                                                         a.bin_iota, a.bin_perm, a.nbins, 0, a.sort_bits, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(bin_scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// Step 4: the binned energy kernel.
hipError_t launch_binned_energy(const BinnedArgs &a, hipStream_t st)
{
    if (a.g1 <= a.g0) return hipSuccess;
    switch (a.guest.nsites) {
    case 1: return launch_binned_energy_ns<1>(a, st);
    case 2: return launch_binned_energy_ns<2>(a, st);
    case 3: return launch_binned_energy_ns<3>(a, st);
    case 4: return launch_binned_energy_ns<4>(a, st);
    case 5: return launch_binned_energy_ns<5>(a, st);
    case 6: return launch_binned_energy_ns<6>(a, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pa
