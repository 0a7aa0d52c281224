// pa_host.hpp -- host-side definitions shared by the library's translation units: status
// helpers, device allocations, the batch's host tables and device storage.
#pragma once

#include <pureadsorb_amd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pa_device.hpp"

namespace pa {

// Sets the thread-local message of pa_last_error and returns `code`.
int fail(int code, const char *fmt, ...);
// Clears the message and returns PA_OK.
int ok();
// PA_OK once a HIP device is present (checked once per process).
int device_ready();
// erfc by the reference's 28-term Chebyshev series (src/ewald.jl:47-74), FP64.
double erfc_dev_host(double x);

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return pa::fail(PA_ERR_DEVICE, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, \
                            __LINE__, #expr);                                                            \
    } while (0)

struct DevAlloc {
    void *p = nullptr;
    size_t bytes = 0;
    // host copy of what upload_cached last wrote (buffers that only uploads write)
    std::vector<unsigned char> shadow;
    ~DevAlloc() { if (p) (void)hipFree(p); }
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        shadow.clear();
        hipError_t e = hipMalloc(&p, n ? n : 16);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T> T *as() const { return (T *)p; }
};

// Page-locked host memory of at least n bytes, kept and grown on demand (copies into it need no
// staging through the runtime's bounce buffers).
struct PinnedAlloc {
    void *p = nullptr;
    size_t bytes = 0;
    ~PinnedAlloc() { if (p) (void)hipHostFree(p); }
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        if (p) { (void)hipHostFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipHostMalloc(&p, n ? n : 16, hipHostMallocDefault);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T> T *as() const { return (T *)p; }
};

struct FwHost {
    int natoms, nk, nk_full;
    double volume, alpha, const_off, halfrange;
    double perp[3];  // perpendicular widths, Å
};

// Host copies of one framework's tables that the gradient entry point reads back.
struct FwGradHost {
    std::vector<double> frac;    // natoms x 3
    std::vector<int> kn;         // 3 per kept k-vector: reciprocal-basis coefficients
    std::vector<double> kpref;   // per kept k-vector
    std::vector<double> counts;  // atoms per force-field type
    int rep[3] = {1, 1, 1};
    bool image_order = true;     // atoms are `replicate` images of the first natoms/nimg (image-major)
};

// Runs fn(begin, end) over [0, n) split across up to 16 host threads (serially below
// min_parallel items).
template <class F>
void parallel_for(int64_t n, F &&fn, int64_t min_parallel = 64)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < min_parallel || nt == 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t b = (int64_t)t * per, e = std::min(n, b + per);
        if (b >= e) break;
        th.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto &x : th) x.join();
}

template <class T>
inline hipError_t upload(DevAlloc &d, const std::vector<T> &v)
{
    hipError_t e = d.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(d.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

// upload() that skips the copy when the buffer already holds exactly these bytes: repeated
// calls with the same arguments re-send nothing.
template <class T>
inline hipError_t upload_cached(DevAlloc &d, const std::vector<T> &v)
{
    const size_t n = sizeof(T) * v.size();
    if (d.p && d.shadow.size() == n && (n == 0 || std::memcmp(d.shadow.data(), v.data(), n) == 0)) return hipSuccess;
    hipError_t e = upload(d, v);
    if (e != hipSuccess) {
        d.shadow.clear();
        return e;
    }
    d.shadow.assign((const unsigned char *)v.data(), (const unsigned char *)v.data() + n);
    return hipSuccess;
}

// Per-call device buffers of the Widom pipeline. Normally each is the batch's own allocation,
// grown on demand (and an upload of unchanged bytes is skipped). Given a fixed region (the
// no-allocation entry point), the buffers are carved from it in request order, 256-byte
// aligned, and uploads always copy; a request past the region's end fails with `need` set to
// the bytes the call would have needed. Every request of a call precedes its first launch.
struct CallScratch {
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0, need = 0;
    hipError_t take(DevAlloc &own, size_t bytes, void **p)
    {
        if (!base) {
            hipError_t e = own.ensure(bytes);
            *p = own.p;
            return e;
        }
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > cap || bytes > cap - off) {
            need = off + bytes;
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        *p = base + off;
        used = off + bytes;
        return hipSuccess;
    }
    template <class T>
    hipError_t put(DevAlloc &own, const std::vector<T> &v, T **p)
    {
        if (!base) {
            hipError_t e = upload_cached(own, v);
            *p = own.as<T>();
            return e;
        }
        void *q = nullptr;
        hipError_t e = take(own, sizeof(T) * std::max<size_t>(v.size(), 1), &q);
        *p = (T *)q;
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(q, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
    }
};

}  // namespace pa

// The opaque handle of the C ABI.
struct pa_batch {
    int nsys = 0, nfw = 0;
    std::vector<int> fw_of;
    std::vector<pa::FwHost> info;
    std::vector<pa::FwDesc> desc;
    bool upper_all = true;
    // guest exactly as given (for the same-guest check) and its device form
    int nsites = 0;
    std::vector<double> gsites, gq;
    std::vector<int> gtypes;
    pa::GuestDev gdev{};
    int ntypes_c = 0;
    double rc_lj = 0, rc_ew = 0;
    double r_guest = 0;  // largest site distance from the guest's reference point
    // phase-0 radii of the last temperature (they depend on kT only): nsys x NS x ntypes_c Å^2
    // and the largest per system
    double rr_kT = -1.0;
    std::vector<double> rr_r2, rr_max2;
    // force field as given, compact type map and per-framework host tables (gradients)
    int nt_ff = 0, ff_tail = 0;
    std::vector<double> ff_sigma, ff_eps;
    std::vector<int> c_to_orig;
    std::vector<double2> pair64_h;
    std::vector<pa::FwGradHost> fgh;
    // device tables
    pa::DevAlloc d_fw, d_fw_of, d_atom64, d_atype, d_k64, d_S64, d_atomfx, d_kn, d_kc, d_pair64, d_pair32;
    // per-call scratch
    pa::DevAlloc d_seg, d_dU, d_tile_g0, d_tile_n, d_tiles, d_pos, d_quat;
    // mixed-path binning scratch
    pa::DevAlloc d_bins, d_bin_base, d_sitefx, d_key, d_bin_count, d_bin_start, d_rank, d_order, d_scan_tmp, d_bin_perm;
    // hard-core rejection: per unique framework the bound B_s (src/reject.jl:101-131) and
    // kmin[site][compact type] (src/reject.jl:141-153); device cell grids
    std::vector<double> hc_bs, hc_kmin;
    // host copies of the FP64 tables (the move path's from-scratch energies and audit)
    bool fullk = false;
    double bin_scale = 0.0;  // mixed path: bin edge multiplier (0: not yet tuned, 1 used meanwhile)
    pa::CallScratch scr;     // where the current call's per-call buffers come from
    int device = 0;          // HIP device the batch's tables live on
    int list_cap = 0;        // mixed path: fixed neighbour-list capacity (0: sized per call)
    int list_cap_used = 0;   // capacity of the last mixed-path call
    bool passes_recorded = false;  // d_passes holds the last mixed-path call's pass count
    std::vector<double4> atom64_h, k64_h;
    std::vector<int4> kn_h;    // n of every k-vector of k64_h
    std::vector<double2> S64_h;
    std::vector<int> atype_h;
    pa::DevAlloc d_corefx, d_coreidx, d_cell_start, d_rho2, d_cand, d_ncand, d_nbr_start, d_nbr_atoms, d_passes;
    pa::DevAlloc d_binfx, d_bintype;  // the atoms in Morton order, for the mixed path's list build
    // gradient scratch
    pa::PinnedAlloc h_acc_q;  // gradient path: the per-atom charge accumulators read back
    pa::PinnedAlloc h_tiles;  // Widom calls: the Boltzmann tile sums read back
    pa::DevAlloc d_rblk;      // gradient path: workgroups of the per-atom reciprocal term
    pa::DevAlloc d_list_g, d_msys, d_flag, d_count, d_sel_tmp, d_acc_lj, d_acc_q, d_acc_k, d_qoff, d_koff, d_pairg,
        d_loff, d_lsite, d_lw, d_blk, d_kblk;
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> kev;  // event pairs around each binned energy launch
    pa_timing last{};
    bool host_only = false;
    bool from_tables = false;  // built by pa_batch_from_tables: no force field behind the tables
    ~pa_batch()
    {
        for (auto &e : ev) if (e) (void)hipEventDestroy(e);
        for (auto &e : kev) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace pa {
// Squared phase-0 rejection radii of system s at kT, NS x ntypes_c (build_rejection_tables).
int rejection_radii(const pa_batch *b, int s, double kT, double eps_f, double *rho2, double *rmax2);
}  // namespace pa
