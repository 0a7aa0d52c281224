// pureadsorb_amd.hip -- host side of the MI355X Widom backend: FrameworkBatch table builder
// (FP64, multithreaded) and the C ABI declared in include/pureadsorb_amd.h.
#include <pureadsorb_amd.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pa_host.hpp"

namespace pa {
hipError_t launch_energy(const WidomArgs &a, int precision, bool upper, hipStream_t st);
hipError_t launch_tiles(const double *dU, const int64_t *tile_g0, const int *tile_n, int ntiles, double kT, double4 *out,
                        hipStream_t st);
hipError_t launch_poses(const FwDesc *fw, const int *fw_of, int sys, uint64_t seed, int64_t j0, int64_t n, double *pos,
                        double *quat, hipStream_t st);
hipError_t binned_scan_bytes(int nbins, size_t *bytes);
hipError_t binned_sort_bytes(int nbins, int bits, size_t *bytes);
hipError_t launch_binning(const BinnedArgs &a, void *scan_tmp, size_t scan_bytes, hipStream_t st);
hipError_t launch_binned_energy(const BinnedArgs &a, hipStream_t st);
hipError_t launch_grad(const GradArgs &a, bool upper, int nblk, int nkblk, hipStream_t st);
int binned_listcap_max(int ns);
int binned_listcap_min();
int binned_nclasses();
hipError_t grad_offsets(const int64_t *list, int64_t nlist, const int64_t *seg, int nsys, int64_t *loff, hipStream_t st);
hipError_t grad_select(const double *dU, const int64_t *seg, int nsys, const double *msys, double beta, int64_t G,
                       unsigned char *flag, int64_t *list, int64_t *count, void *tmp, size_t *tmp_bytes, hipStream_t st);
hipError_t launch_grad_recip(const GradArgs &a, const double4 *k64, const int4 *rblk, int nrblk, hipStream_t st);
hipError_t launch_sk(const double4 *atoms, int natoms, const int4 *kreq, int nreq, double2 *out, hipStream_t st);
hipError_t launch_nbr_fill(const FwDesc *fw, int nfw, const int *cell_fw_off, int ncells, const int *cell_start,
                           const int *nbr_start, const uint4 *corefx, uint4 *nbr_atoms, hipStream_t st);
hipError_t launch_self_term(const int4 *kn, const double *kpref, int nk, const double *gsf, const double *q, int ns,
                            double *partial, hipStream_t st);
}  // namespace pa

using namespace pa;

// ------------------------------------------------------------------------------------------
// Errors
// ------------------------------------------------------------------------------------------
static thread_local std::string g_err;

int pa::fail(int code, const char *fmt, ...)
{
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int pa::ok()
{
    g_err.clear();
    return PA_OK;
}

// ------------------------------------------------------------------------------------------
// Host numerics (FP64)
// ------------------------------------------------------------------------------------------
namespace {

// erfc_dev: 28-term Chebyshev series of log(erfc(z)/t) + z^2, t = 2/(2+z)
// (src/ewald.jl:47-74; Numerical Recipes 3rd ed. section 6.2.2). Used for the guest
// intramolecular exclusion term exactly as the reference does (src/batch.jl:402).
double erfc_dev(double x)
{
    static const double cof[28] = {
        -1.3026537197817094, 6.419697923564902e-1, 1.9476473204185836e-2,
        -9.561514786808632e-3, -9.465953444820369e-4, 3.6683949785276145e-4, 4.252332480690777e-5,
        -2.0278578112534242e-5, -1.6242900046470256e-6, 1.3036558355805232e-6, 1.5626441722066142e-8,
        -8.523809591492654e-8, 6.5290544390988515e-9, 5.0593434955514685e-9, -9.913641564930322e-10,
        -2.2736512229318417e-10, 9.646791102014962e-11, 2.394038083065756e-12, -6.886027526536277e-12,
        8.944879271730368e-13, 3.130921408127142e-13, -1.1270822525221733e-13, 3.8108713588528985e-16,
        7.106125536922576e-15, -1.5230878975373303e-15, -9.464871412184462e-17, 1.218683266906788e-16,
        -3.0494735135424696e-17};
    auto cheb = [&](double z) {
        double t = 2.0 / (2.0 + z), ty = 4.0 * t - 2.0, d = 0.0, dd = 0.0;
        for (int j = 27; j >= 1; --j) {
            double tmp = d;
            d = ty * d - dd + cof[j];
            dd = tmp;
        }
        return t * std::exp(-z * z + (ty / 2.0) * d - dd + cof[0] / 2.0);
    };
    return x >= 0 ? cheb(x) : 2.0 - cheb(-x);
}

double det3(const double A[9])
{
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) + A[6] * (A[1] * A[5] - A[4] * A[2]);
}

void inv3(const double A[9], double inv[9])
{
    double d = det3(A);
    inv[0] = (A[4] * A[8] - A[7] * A[5]) / d;
    inv[1] = -(A[1] * A[8] - A[7] * A[2]) / d;
    inv[2] = (A[1] * A[5] - A[4] * A[2]) / d;
    inv[3] = -(A[3] * A[8] - A[6] * A[5]) / d;
    inv[4] = (A[0] * A[8] - A[6] * A[2]) / d;
    inv[5] = -(A[0] * A[5] - A[3] * A[2]) / d;
    inv[6] = (A[3] * A[7] - A[6] * A[4]) / d;
    inv[7] = -(A[0] * A[7] - A[6] * A[1]) / d;
    inv[8] = (A[0] * A[4] - A[3] * A[1]) / d;
}

double norm3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// src/cell.jl:14-20
void perpendicular_lengths(const double A[9], double L[3])
{
    double V = std::fabs(det3(A)), x[3];
    cross3(A + 3, A + 6, x); L[0] = V / norm3(x);
    cross3(A + 0, A + 6, x); L[1] = V / norm3(x);
    cross3(A + 0, A + 3, x); L[2] = V / norm3(x);
}

// ewald_alpha: bisection on erfc(z) = cutoff*precision/2 (src/ewald.jl:21-32).
double ewald_alpha(double cutoff, double precision)
{
    double target = cutoff * precision / 2.0, lo = 0.0, hi = 20.0;
    if (!(std::erfc(hi) - target < 0)) return -1.0;
    for (int it = 0; it < 200; ++it) {
        double mid = (lo + hi) / 2.0;
        if (std::erfc(mid) - target > 0) lo = mid; else hi = mid;
        if (hi - lo < 1.0e-12) break;
    }
    return (lo + hi) / 2.0 / cutoff;
}


// k-vector table of one stored cell (src/ewald.jl:142-157), host structure factor
// (src/ewald.jl:159-170) and prefactors (src/ewald.jl:36-39).
struct KTable {
    std::vector<double> k;       // 3 per vector, Cartesian
    std::vector<int> n;          // 3 per vector, reciprocal-basis coefficients
    std::vector<double> kpref;   // w * pk(|k|^2, alpha, V)
    std::vector<double> Sre, Sim;
};

void build_ktable(const double A[9], double kmax, double alpha, double V, KTable &kt)
{
    double invA[9], B[9];
    inv3(A, invA);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i + 3 * j] = TWO_PI * invA[j + 3 * i];
    int nmax[3];
    for (int i = 0; i < 3; ++i) nmax[i] = (int)std::ceil(kmax * norm3(A + 3 * i) / TWO_PI);
    for (int n1 = 0; n1 <= nmax[0]; ++n1)
        for (int n2 = -nmax[1]; n2 <= nmax[1]; ++n2)
            for (int n3 = -nmax[2]; n3 <= nmax[2]; ++n3) {
                if (n1 == 0 && n2 == 0 && n3 == 0) continue;
                double k[3];
                for (int i = 0; i < 3; ++i) k[i] = B[i] * n1 + B[i + 3] * n2 + B[i + 6] * n3;
                if (!(norm3(k) <= kmax)) continue;
                kt.k.insert(kt.k.end(), k, k + 3);
                kt.n.push_back(n1); kt.n.push_back(n2); kt.n.push_back(n3);
                double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                double w = n1 == 0 ? 1.0 : 2.0;
                kt.kpref.push_back(w * (2.0 * PI / V * std::exp(-k2 / (4.0 * alpha * alpha)) / k2));
            }
}

// S(k) = sum_j q_j exp(i k.r_j) over the framework's atoms, for the k-vectors with need[i] != 0
// (the others stay 0). The phase of atom j at k = B n is 2 pi n.f_j; exp(i 2 pi n_d f_jd) is
// tabulated per axis and power, atoms contiguous, so each (k, atom) term is two complex
// products instead of a sincos.
void structure_factor(const KTable &kt, const std::vector<double> &frac, const std::vector<double> &q,
                      const std::vector<char> &need, std::vector<double> &Sre, std::vector<double> &Sim)
{
    const int64_t nk = (int64_t)kt.kpref.size(), na = (int64_t)q.size();
    Sre.assign(nk, 0.0);
    Sim.assign(nk, 0.0);
    if (nk == 0 || na == 0) return;
    int nmax[3] = {0, 0, 0};
    for (int64_t i = 0; i < nk; ++i)
        for (int d = 0; d < 3; ++d) nmax[d] = std::max(nmax[d], std::abs(kt.n[3 * i + d]));
    const int w0 = nmax[0] + 1, w1 = 2 * nmax[1] + 1, w2 = 2 * nmax[2] + 1;
    // tables [power][atom]
    std::vector<double> c0((size_t)na * w0), s0((size_t)na * w0), c1((size_t)na * w1), s1((size_t)na * w1),
        c2((size_t)na * w2), s2((size_t)na * w2);
    parallel_for(na, [&](int64_t b, int64_t e) {
        for (int64_t j = b; j < e; ++j) {
            for (int m = 0; m < w0; ++m) {
                double ph = TWO_PI * (m * frac[3 * j]);
                c0[m * na + j] = std::cos(ph); s0[m * na + j] = std::sin(ph);
            }
            for (int m = 0; m < w1; ++m) {
                double ph = TWO_PI * ((m - nmax[1]) * frac[3 * j + 1]);
                c1[m * na + j] = std::cos(ph); s1[m * na + j] = std::sin(ph);
            }
            for (int m = 0; m < w2; ++m) {
                double ph = TWO_PI * ((m - nmax[2]) * frac[3 * j + 2]);
                c2[m * na + j] = std::cos(ph); s2[m * na + j] = std::sin(ph);
            }
        }
    });
    std::vector<int64_t> todo;
    for (int64_t i = 0; i < nk; ++i)
        if (need[i]) todo.push_back(i);
    parallel_for((int64_t)todo.size(), [&](int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
            const int64_t i = todo[t];
            const int m0 = kt.n[3 * i], m1 = kt.n[3 * i + 1] + nmax[1], m2 = kt.n[3 * i + 2] + nmax[2];
            const double *a_r = &c0[(size_t)m0 * na], *a_i = &s0[(size_t)m0 * na];
            const double *b_r = &c1[(size_t)m1 * na], *b_i = &s1[(size_t)m1 * na];
            const double *c_r = &c2[(size_t)m2 * na], *c_i = &s2[(size_t)m2 * na];
            double re = 0.0, im = 0.0;
            for (int64_t j = 0; j < na; ++j) {
                const double pr = a_r[j] * b_r[j] - a_i[j] * b_i[j], pi = a_r[j] * b_i[j] + a_i[j] * b_r[j];
                const double rr = pr * c_r[j] - pi * c_i[j], ri = pr * c_i[j] + pi * c_r[j];
                re += q[j] * rr;
                im += q[j] * ri;
            }
            Sre[i] = re;
            Sim[i] = im;
        }
    });
}


// Device scratch of the table builder (reused across frameworks).
struct BuildScratch {
    DevAlloc atoms, kreq, sk, kn, kpref, gsf, q, partial;
};

// structure_factor on the device (batch_kernels.hip): FP64, one work-item per needed k-vector.
int device_structure_factor(BuildScratch &bs, const KTable &kt, const std::vector<double> &frac,
                            const std::vector<double> &q, const std::vector<char> &need, std::vector<double> &Sre,
                            std::vector<double> &Sim)
{
    const int64_t nk = (int64_t)kt.kpref.size(), na = (int64_t)q.size();
    Sre.assign(nk, 0.0);
    Sim.assign(nk, 0.0);
    if (nk == 0 || na == 0) return PA_OK;
    std::vector<int4> kr;
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < nk; ++i)
        if (need[i]) {
            kr.push_back(make_int4(kt.n[3 * i], kt.n[3 * i + 1], kt.n[3 * i + 2], 0));
            idx.push_back(i);
        }
    if (kr.empty()) return PA_OK;
    int st = device_ready();
    if (st) return st;
    std::vector<double4> at(na);
    for (int64_t j = 0; j < na; ++j) at[j] = make_double4(frac[3 * j], frac[3 * j + 1], frac[3 * j + 2], q[j]);
    HIP_TRY(upload(bs.atoms, at));
    HIP_TRY(upload(bs.kreq, kr));
    HIP_TRY(bs.sk.ensure(sizeof(double2) * kr.size()));
    HIP_TRY(launch_sk(bs.atoms.as<double4>(), (int)na, bs.kreq.as<int4>(), (int)kr.size(), bs.sk.as<double2>(), nullptr));
    std::vector<double2> out(kr.size());
    HIP_TRY(hipMemcpy(out.data(), bs.sk.p, sizeof(double2) * kr.size(), hipMemcpyDeviceToHost));
    for (size_t t = 0; t < kr.size(); ++t) {
        Sre[idx[t]] = out[t].x;
        Sim[idx[t]] = out[t].y;
    }
    return PA_OK;
}

// The 64-orientation guest self term on the device: samples[o] = KE sum_k kpref |S_g(k; o)|^2,
// the per-block partial sums added in block order on the host.
int device_self_term(BuildScratch &bs, const KTable &kt, const double invA[9], const pa_guest *guest,
                     const std::vector<double> &self_q, double samples[64])
{
    const int nk = (int)kt.kpref.size(), NS = guest->nsites;
    for (int o = 0; o < 64; ++o) samples[o] = 0.0;
    if (nk == 0) return PA_OK;
    int st = device_ready();
    if (st) return st;
    std::vector<double> gsf((size_t)64 * NS * 3);
    for (int o = 0; o < 64; ++o)
        for (int a = 0; a < NS; ++a) {
            double r[3];
            rotate(&self_q[4 * o], guest->sites + 3 * a, r);
            matvec(invA, r, &gsf[3 * ((size_t)o * NS + a)]);
        }
    std::vector<int4> kn(nk);
    for (int i = 0; i < nk; ++i) kn[i] = make_int4(kt.n[3 * i], kt.n[3 * i + 1], kt.n[3 * i + 2], 0);
    const int nb = (nk + 63) / 64;
    HIP_TRY(upload(bs.kn, kn));
    HIP_TRY(upload(bs.kpref, kt.kpref));
    HIP_TRY(upload(bs.gsf, gsf));
    HIP_TRY(upload(bs.q, std::vector<double>(guest->charges, guest->charges + NS)));
    HIP_TRY(bs.partial.ensure(sizeof(double) * 64 * nb));
    HIP_TRY(launch_self_term(bs.kn.as<int4>(), bs.kpref.as<double>(), nk, bs.gsf.as<double>(), bs.q.as<double>(), NS,
                             bs.partial.as<double>(), nullptr));
    std::vector<double> part((size_t)64 * nb);
    HIP_TRY(hipMemcpy(part.data(), bs.partial.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
    for (int blk = 0; blk < nb; ++blk)
        for (int o = 0; o < 64; ++o) samples[o] += part[(size_t)blk * 64 + o];
    for (int o = 0; o < 64; ++o) samples[o] *= KE;
    return PA_OK;
}

// 64 fixed guest orientations for the orientation-averaged self term: ChainRNG(0x5e1f, 0, o),
// o = 1..64, three F64 uniforms each (this build's replacement for the reference's
// Xoshiro(0x5e1f) draws, src/batch.jl:302-303).
void self_orientation(int o, double q[4])
{
    Chain c(0x5e1fULL, 0, (uint64_t)o);
    double u1 = unit_f64(c.next()), u2 = unit_f64(c.next()), u3 = unit_f64(c.next());
    shoemake(u1, u2, u3, q);
}


// ---- hard-core rejection bounds (src/reject.jl) ----------------------------------------------

// theta_F(Float64) (src/reject.jl:11-19): smallest theta with exp(-theta) == 0.0, by bisection on
// the float grid with this host's exp.
double theta_f64()
{
    double lo = 0.0, hi = 2000.0;
    while (std::nextafter(lo, hi) != hi) {
        const double mid = lo + (hi - lo) / 2.0;
        if (std::exp(-mid) == 0.0) hi = mid; else lo = mid;
    }
    return hi;
}

double lj_pair(double r2, double sg, double ep)
{
    double x = sg * sg / r2;
    x = x * x * x;
    return 4.0 * ep * (x * x - x);
}

// Edge (Å) of the hard-core grid cells of phase 0. Typical rejection radii are 1-2.3 Å, so the
// near stencil (the home cell and its neighbours) usually decides an insertion.
constexpr double HC_CELL = 1.2;

// pair_energy (src/reject.jl:22): LJ plus K erfc(alpha r)/r with the pair series.
double pair_energy_h(double r, double sg, double ep, double K, double alpha)
{
    return lj_pair(r * r, sg, ep) + K * pair_erfc_f64(alpha * r) / r;
}

constexpr int MAX_BISECT = 200;

// find_r0 (src/reject.jl:45-69): zero of the pair energy below its minimum, attractive pair
// (K < 0, eps > 0); clamped to the LJ cutoff. Returns NaN when no bracket is found.
double find_r0(double sg, double ep, double K, double alpha, double r_lj)
{
    double hi = sg * std::pow(2.0, 1.0 / 6.0);
    if (!(pair_energy_h(hi, sg, ep, K, alpha) < 0.0)) return NAN;
    double lo = hi;
    for (int it = 0; pair_energy_h(lo, sg, ep, K, alpha) <= 0.0; ++it) {
        lo /= 2.0;
        if (it > MAX_BISECT) return NAN;
    }
    for (int it = 0; hi - lo > 4.0 * DBL_EPSILON * hi; ++it) {
        const double mid = (lo + hi) / 2.0;
        if (pair_energy_h(mid, sg, ep, K, alpha) > 0.0) lo = mid; else hi = mid;
        if (it > MAX_BISECT) return NAN;
    }
    return std::min((lo + hi) / 2.0, r_lj);
}

// find_rho2 (src/reject.jl:191-209): first root of LJ(r) - |kmin|/r = margin, clamped to the
// LJ cutoff, squared; 0 when the margin is infinite or the pair has no LJ wall. NaN on failure.
double find_rho2(double sg, double ep, double kmin, double margin, double r_lj)
{
    if (std::isinf(margin) || ep == 0.0) return 0.0;
    auto f = [&](double r) { return lj_pair(r * r, sg, ep) - std::fabs(kmin) / r - margin; };
    double hi = sg;
    for (int it = 0; f(hi) >= 0.0; ++it) {
        hi *= 2.0;
        if (it > MAX_BISECT) return NAN;
    }
    double lo = hi;
    for (int it = 0; f(lo) < 0.0; ++it) {
        lo /= 2.0;
        if (it > MAX_BISECT) return NAN;
    }
    for (int it = 0; hi - lo > 4.0 * DBL_EPSILON * hi; ++it) {
        const double mid = (lo + hi) / 2.0;
        if (f(mid) > 0.0) lo = mid; else hi = mid;
        if (it > MAX_BISECT) return NAN;
    }
    const double r = std::min((lo + hi) / 2.0, r_lj);
    return r * r;
}
}  // namespace

double pa::erfc_dev_host(double x) { return erfc_dev(x); }

// ------------------------------------------------------------------------------------------
// Batch
// ------------------------------------------------------------------------------------------
int pa::device_ready()
{
    static std::once_flag once;
    static int status = PA_OK;
    static std::string msg;
    std::call_once(once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n < 1) {
            status = PA_ERR_DEVICE;
            msg = std::string("no HIP device available: ") + hipGetErrorString(e);
        }
    });
    if (status != PA_OK) return fail(status, "%s", msg.c_str());
    return PA_OK;
}

extern "C" int pa_last_error(char *buf, size_t len)
{
    if (buf && len) {
        std::snprintf(buf, len, "%s", g_err.c_str());
    }
    return (int)g_err.size();
}

extern "C" int pa_device_name(char *buf, size_t len)
{
    int st = device_ready();
    if (st) return st;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (buf && len) std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return ok();
}

extern "C" int pa_synchronize(void)
{
    int st = device_ready();
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    return ok();
}

namespace {
// Host tables of a batch under construction, in the layout they are uploaded in.
struct BatchTables {
    std::vector<double4> atom64;
    std::vector<int> atype;
    std::vector<uint4> atomfx;
    std::vector<double4> k64;
    std::vector<double2> S64;
    std::vector<int4> kn;
    std::vector<float2> kc;
    std::vector<uint4> corefx;
    std::vector<int> coreidx, cell_start;
    std::vector<uint4> binfx;           // the atoms in Morton order (mixed path's list build)
    std::vector<int> bintype;
    std::vector<int> nbr_start;         // per framework cell (cell_start layout): first entry in nbr_atoms
    int64_t nbr_total = 0;              // entries of nbr_atoms (filled on the device from corefx)
    std::vector<int> cell_fw_off{0};    // prefix over frameworks of their cell counts
    std::vector<double2> pair64;
    std::vector<float2> pair32;
};

// Appends framework D's n atoms (fractional coordinates, Cartesian positions or NULL to take
// A f, charges, compact LJ types) to the atom tables, with its hard-core cell grid and the
// offsets of its near-stencil lists. L: the cell's perpendicular widths. Sets D.atom_off,
// D.natoms, D.cg_off and D.cg_nc.
int add_atoms(BatchTables &T, FwDesc &D, const double *frac, const double *cart, const double *q, const int *ctype,
              int n, const double L[3])
{
    D.atom_off = (int)T.atom64.size();
    D.natoms = n;
    for (int j = 0; j < n; ++j) {
        double f[3] = {frac[3 * j], frac[3 * j + 1], frac[3 * j + 2]}, p[3];
        if (cart) {
            p[0] = cart[3 * j];
            p[1] = cart[3 * j + 1];
            p[2] = cart[3 * j + 2];
        } else {
            matvec(D.A, f, p);
        }
        T.atom64.push_back(make_double4(p[0], p[1], p[2], q[j]));
        T.atype.push_back(ctype[j]);
        float qf = (float)q[j];
        uint32_t qb;
        std::memcpy(&qb, &qf, 4);
        T.atomfx.push_back(make_uint4(to_fixed(f[0]), to_fixed(f[1]), to_fixed(f[2]), qb));
    }
    // the mixed path's list build reads the atoms in Morton order of their fixed-point coordinates
    // (the top 10 bits per axis, ties by index): a wavefront's 64 atoms then lie close together,
    // so an image across a cell boundary is either needed by some of them or skipped by all
    {
        auto spread = [](uint32_t v) {
            uint64_t x = 0;
            for (int i = 0; i < 10; ++i) x |= (uint64_t)((v >> i) & 1u) << (3 * i);
            return x;
        };
        std::vector<std::pair<uint64_t, int>> key((size_t)n);
        for (int j = 0; j < n; ++j) {
            const uint4 &x = T.atomfx[D.atom_off + j];
            key[j] = {spread(x.x >> 22) | (spread(x.y >> 22) << 1) | (spread(x.z >> 22) << 2), j};
        }
        std::sort(key.begin(), key.end());
        for (int i = 0; i < n; ++i) {
            T.binfx.push_back(T.atomfx[D.atom_off + key[i].second]);
            T.bintype.push_back(T.atype[D.atom_off + key[i].second]);
        }
    }
    // hard-core cell grid: cells of at least HC_CELL Å, at most 64 per axis; atoms sorted by cell
    {
        int nct = 1;
        for (int d = 0; d < 3; ++d) {
            D.cg_nc[d] = std::max(1, std::min(64, (int)std::floor(L[d] / HC_CELL)));
            nct *= D.cg_nc[d];
        }
        D.cg_off = (int)T.cell_start.size();
        std::vector<int> cnt(nct + 1, 0), cell_of(n);
        for (int j = 0; j < n; ++j) {
            const uint4 &x = T.atomfx[D.atom_off + j];
            const uint32_t fx[3] = {x.x, x.y, x.z};
            int c[3];
            for (int d = 0; d < 3; ++d) c[d] = (int)(((uint64_t)fx[d] * (uint64_t)D.cg_nc[d]) >> 32);
            cell_of[j] = c[0] + D.cg_nc[0] * (c[1] + D.cg_nc[1] * c[2]);
            ++cnt[cell_of[j] + 1];
        }
        for (int c = 0; c < nct; ++c) cnt[c + 1] += cnt[c];
        T.cell_start.insert(T.cell_start.end(), cnt.begin(), cnt.end());
        std::vector<uint4> sorted(n);
        std::vector<int> sidx(n);
        std::vector<int> cur(cnt.begin(), cnt.end() - 1);
        for (int j = 0; j < n; ++j) {
            const uint4 &x = T.atomfx[D.atom_off + j];
            const int pos = cur[cell_of[j]]++;
            sorted[pos] = make_uint4(x.x, x.y, x.z, (uint32_t)T.atype[D.atom_off + j]);
            sidx[pos] = j;
        }
        T.corefx.insert(T.corefx.end(), sorted.begin(), sorted.end());
        T.coreidx.insert(T.coreidx.end(), sidx.begin(), sidx.end());
        // near-stencil atom lists: for every cell, the atoms of the cells within one cell of
        // it (each cell once when an axis has fewer than 3 cells), contiguous, so phase 0's
        // near test reads one range per site
        const int nc0 = D.cg_nc[0], nc1 = D.cg_nc[1], nc2 = D.cg_nc[2];
        const int k0 = std::min(3, nc0), k1 = std::min(3, nc1), k2 = std::min(3, nc2);
        // offsets only: nbr_fill_kernel copies the lists on the device (nbr_cells order)
        auto axis_nbrs = [](int nc, int k) {
            std::vector<int> t((size_t)nc * k);
            for (int h = 0; h < nc; ++h)
                for (int i = 0; i < k; ++i) t[(size_t)h * k + i] = nbr_cell_axis(h, i, k, nc);
            return t;
        };
        const std::vector<int> nb0 = axis_nbrs(nc0, k0), nb1 = axis_nbrs(nc1, k1), nb2 = axis_nbrs(nc2, k2);
        for (int h2 = 0; h2 < nc2; ++h2)
            for (int h1 = 0; h1 < nc1; ++h1)
                for (int h0 = 0; h0 < nc0; ++h0) {
                    if (T.nbr_total > (int64_t)INT32_MAX)
                        return fail(PA_ERR_ARG, "the frameworks' phase-0 neighbour lists exceed 2^31 entries");
                    T.nbr_start.push_back((int)T.nbr_total);
                    for (int i2 = 0; i2 < k2; ++i2)
                        for (int i1 = 0; i1 < k1; ++i1) {
                            const int row = nc0 * (nb1[(size_t)h1 * k1 + i1] + nc1 * nb2[(size_t)h2 * k2 + i2]);
                            for (int i0 = 0; i0 < k0; ++i0) {
                                const int cc = nb0[(size_t)h0 * k0 + i0] + row;
                                T.nbr_total += cnt[cc + 1] - cnt[cc];
                            }
                        }
                }
        if (T.nbr_total > (int64_t)INT32_MAX)
            return fail(PA_ERR_ARG, "the frameworks' phase-0 neighbour lists exceed 2^31 entries");
        T.nbr_start.push_back((int)T.nbr_total);
        T.cell_fw_off.push_back(T.cell_fw_off.back() + nct);
    }
    return PA_OK;
}

// Creates the batch's stream and events, uploads every table and fills the near-stencil lists
// on the device.
int upload_batch(pa_batch *b, const BatchTables &T)
{
    int st = device_ready();
    if (st) return st;
    HIP_TRY(hipGetDevice(&b->device));
    HIP_TRY(hipStreamCreate(&b->stream));
    // timing-only events: no system-scope fence (no cache write-back and invalidation around the
    // kernels they bracket); results reach the host through copies and stream synchronisation
    for (auto &e : b->ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    HIP_TRY(upload(b->d_fw, b->desc));
    HIP_TRY(upload(b->d_fw_of, b->fw_of));
    HIP_TRY(upload(b->d_atom64, T.atom64));
    HIP_TRY(upload(b->d_atype, T.atype));
    HIP_TRY(upload(b->d_atomfx, T.atomfx));
    HIP_TRY(upload(b->d_k64, T.k64));
    HIP_TRY(upload(b->d_S64, T.S64));
    HIP_TRY(upload(b->d_kn, T.kn));
    HIP_TRY(upload(b->d_kc, T.kc));
    HIP_TRY(upload(b->d_pair64, T.pair64));
    HIP_TRY(upload(b->d_pair32, T.pair32));
    HIP_TRY(upload(b->d_corefx, T.corefx));
    HIP_TRY(upload(b->d_binfx, T.binfx));
    HIP_TRY(upload(b->d_bintype, T.bintype));
    HIP_TRY(upload(b->d_coreidx, T.coreidx));
    HIP_TRY(upload(b->d_cell_start, T.cell_start));
    HIP_TRY(upload(b->d_nbr_start, T.nbr_start));
    HIP_TRY(b->d_nbr_atoms.ensure(sizeof(uint4) * (size_t)std::max<int64_t>(1, T.nbr_total)));
    {
        DevAlloc d_cfo;
        HIP_TRY(upload(d_cfo, T.cell_fw_off));
        HIP_TRY(launch_nbr_fill(b->d_fw.as<FwDesc>(), b->nfw, d_cfo.as<int>(), T.cell_fw_off.back(),
                                b->d_cell_start.as<int>(), b->d_nbr_start.as<int>(), b->d_corefx.as<uint4>(),
                                b->d_nbr_atoms.as<uint4>(), nullptr));
        HIP_TRY(hipDeviceSynchronize());
    }
    return PA_OK;
}
}  // namespace

extern "C" int pa_batch_create(const pa_framework *fws, int32_t nsys, const pa_forcefield *ff, const pa_guest *guest,
                               const pa_ewald *ewald, uint32_t flags, pa_batch **out)
{
    if (!out) return fail(PA_ERR_ARG, "out must not be NULL");
    *out = nullptr;
    if (!fws || !ff || !guest || !ewald) return fail(PA_ERR_ARG, "NULL argument");
    if (nsys < 1) return fail(PA_ERR_ARG, "nsys=%d must be >= 1", nsys);
    if (guest->nsites < 1 || guest->nsites > MAX_SITES)
        return fail(PA_ERR_ARG, "guest has %d sites; 1..%d are supported", guest->nsites, MAX_SITES);
    if (ff->ntypes < 1 || !ff->sigma || !ff->epsilon) return fail(PA_ERR_ARG, "force field has no types");
    const int NT = ff->ntypes, NS = guest->nsites;
    for (int a = 0; a < NS; ++a)
        if (guest->types[a] < 0 || guest->types[a] >= NT)
            return fail(PA_ERR_ARG, "unknown LJ type index %d for guest site %d; force field defines %d types",
                        guest->types[a], a, NT);
    const bool fullk = (flags & PA_BATCH_FULLK) != 0;
    const double rc_lj = ff->cutoff, rc_ew = ewald->cutoff;
    const double rc = std::max(rc_lj, rc_ew);

    std::unique_ptr<pa_batch> b(new pa_batch());
    b->nsys = nsys;
    b->nsites = NS;
    b->gsites.assign(guest->sites, guest->sites + 3 * NS);
    b->gq.assign(guest->charges, guest->charges + NS);
    b->gtypes.assign(guest->types, guest->types + NS);
    b->rc_lj = rc_lj;
    b->rc_ew = rc_ew;
    bool charged = false;
    for (int a = 0; a < NS; ++a) {
        charged = charged || guest->charges[a] != 0.0;
        b->r_guest = std::max(b->r_guest, norm3(guest->sites + 3 * a));
    }

    // ---- dedup by value: cell, replication, fractional coordinates, types, charges
    // (src/batch.jl:246-265)
    std::vector<int> uniq;
    b->fw_of.resize(nsys);
    for (int s = 0; s < nsys; ++s) {
        const pa_framework &F = fws[s];
        if (F.natoms < 0) return fail(PA_ERR_DIM, "framework %d has natoms=%d", s, F.natoms);
        if (F.natoms > 0 && (!F.frac || !F.types || !F.charges))
            return fail(PA_ERR_ARG, "framework %d has NULL atom arrays", s);
        for (int d = 0; d < 3; ++d)
            if (F.replication[d] < 1) return fail(PA_ERR_ARG, "replication factors must be >= 1, got (%d, %d, %d)",
                                                  F.replication[0], F.replication[1], F.replication[2]);
        int found = -1;
        for (size_t u = 0; u < uniq.size() && found < 0; ++u) {
            const pa_framework &G = fws[uniq[u]];
            if (G.natoms != F.natoms || std::memcmp(G.cell, F.cell, sizeof F.cell) ||
                std::memcmp(G.replication, F.replication, sizeof F.replication))
                continue;
            size_t n = (size_t)F.natoms;
            if (n && (std::memcmp(G.frac, F.frac, 3 * n * sizeof(double)) ||
                      std::memcmp(G.types, F.types, n * sizeof(int32_t)) ||
                      std::memcmp(G.charges, F.charges, n * sizeof(double))))
                continue;
            found = (int)u;
        }
        if (found < 0) {
            uniq.push_back(s);
            found = (int)uniq.size() - 1;
        }
        b->fw_of[s] = found;
    }
    b->nfw = (int)uniq.size();

    // ---- compact LJ types: those present in any framework, union the guest's (src/batch.jl:305-317)
    std::vector<int> present(NT, 0);
    for (int u : uniq) {
        const pa_framework &F = fws[u];
        for (int j = 0; j < F.natoms; ++j) {
            if (F.types[j] < 0 || F.types[j] >= NT)
                return fail(PA_ERR_ARG, "unknown LJ type index %d (framework %d atom %d); force field defines %d types",
                            F.types[j], u, j, NT);
            present[F.types[j]] = 1;
        }
    }
    for (int a = 0; a < NS; ++a) present[guest->types[a]] = 1;
    std::vector<int> orig_to_c(NT, -1), c_to_orig;
    for (int t = 0; t < NT; ++t)
        if (present[t]) { orig_to_c[t] = (int)c_to_orig.size(); c_to_orig.push_back(t); }
    const int NTC = (int)c_to_orig.size();
    b->ntypes_c = NTC;
    b->c_to_orig = c_to_orig;
    b->nt_ff = NT;
    b->ff_tail = ff->tail ? 1 : 0;
    b->ff_sigma.assign(ff->sigma, ff->sigma + (size_t)NT * NT);
    b->ff_eps.assign(ff->epsilon, ff->epsilon + (size_t)NT * NT);
    if (NTC > 64) return fail(PA_ERR_ARG, "%d distinct LJ types in the batch; at most 64 are supported", NTC);

    // ---- Ewald parameters (src/batch.jl:289-297)
    const double alpha = ewald_alpha(rc_ew, ewald->precision);
    if (alpha < 0) return fail(PA_ERR_ARG, "cannot reach precision %g with cutoff %g", ewald->precision, rc_ew);
    if (alpha * rc_ew > PAIR_ERFC_XMAX)
        return fail(PA_ERR_ARG,
                    "α·ewald_cutoff = %.17g exceeds PAIR_ERFC_XMAX = %.1f: insertion_energy's screened-Coulomb series is "
                    "only fitted up to that bound; loosen ewald.precision or shorten ewald.cutoff to bring α·ewald_cutoff "
                    "back under %.1f",
                    alpha * rc_ew, PAIR_ERFC_XMAX, PAIR_ERFC_XMAX);
    const double kmax = 2.0 * alpha * std::sqrt(-std::log(ewald->precision / 2.0));

    std::vector<double> self_q(64 * 4);
    for (int o = 0; o < 64; ++o) self_orientation(o + 1, &self_q[4 * o]);

    BatchTables T;
    b->desc.resize(b->nfw);
    b->info.resize(b->nfw);
    b->fgh.resize(b->nfw);
    // with a device, the structure factors and the self term run there (batch_kernels.hip);
    // PA_BATCH_HOST_ONLY keeps them on the host
    const bool host_only = (flags & PA_BATCH_HOST_ONLY) != 0;
    BuildScratch bscratch;
    for (int u = 0; u < b->nfw; ++u) {
        const pa_framework &F = fws[uniq[u]];
        const int n = F.natoms;
        // per-pair minimum image is exact whenever every perpendicular width is >= 2*cutoff
        double L[3];
        perpendicular_lengths(F.cell, L);
        int m[3];
        for (int d = 0; d < 3; ++d) m[d] = (int)std::ceil(2.0 * rc / L[d]);
        if (m[0] > 1 || m[1] > 1 || m[2] > 1)
            return fail(PA_ERR_ARG,
                        "framework %d is too small for cutoff %g (perpendicular widths %.4f, %.4f, %.4f Å must be >= %g); "
                        "replicate it by (%d, %d, %d) first",
                        uniq[u] + 1, rc, L[0], L[1], L[2], 2 * rc, m[0], m[1], m[2]);
        FwDesc &D = b->desc[u];
        std::memcpy(D.A, F.cell, sizeof D.A);
        inv3(D.A, D.invA);
        for (int i = 0; i < 9; ++i) D.Af[i] = (float)(D.A[i] * 2.3283064365386963e-10);
        D.upper = (D.A[1] == 0.0 && D.A[2] == 0.0 && D.A[5] == 0.0) ? 1 : 0;
        b->upper_all = b->upper_all && D.upper;
        const double V = std::fabs(det3(D.A));
        D.volume = V;
        D.alpha = alpha;
        D.alphaf = (float)alpha;
        std::vector<double> frac(F.frac, F.frac + 3 * (size_t)n), q(F.charges, F.charges + n);
        std::vector<int> ctype(n);
        for (int j = 0; j < n; ++j) ctype[j] = orig_to_c[F.types[j]];
        {
            int s = add_atoms(T, D, frac.data(), nullptr, q.data(), ctype.data(), n, L);
            if (s) return s;
        }
        const bool rep1 = F.replication[0] == 1 && F.replication[1] == 1 && F.replication[2] == 1;
        FwGradHost &gh = b->fgh[u];
        gh.frac = frac;
        gh.counts.assign(NT, 0.0);
        for (int j = 0; j < n; ++j) gh.counts[F.types[j]] += 1.0;
        for (int d = 0; d < 3; ++d) gh.rep[d] = F.replication[d];
        if (!rep1) {
            // replicate order (src/structure.jl:79-92): image (i, j, k), i fastest, atoms inner;
            // image m of atom a sits at frac[a] + (i/r0, j/r1, k/r2) modulo 1
            const int r0 = F.replication[0], r1 = F.replication[1], r2 = F.replication[2];
            const int nimg = r0 * r1 * r2;
            gh.image_order = n % nimg == 0;
            const int nu = gh.image_order ? n / nimg : 0;
            for (int m = 1; gh.image_order && m < nimg; ++m) {
                const double sh[3] = {(double)(m % r0) / r0, (double)((m / r0) % r1) / r1, (double)(m / (r0 * r1)) / r2};
                for (int a = 0; gh.image_order && a < nu; ++a) {
                    const int j = m * nu + a;
                    for (int d = 0; d < 3; ++d) {
                        double x = frac[3 * j + d] - frac[3 * a + d] - sh[d];
                        x -= std::nearbyint(x);
                        if (std::fabs(x) > 1.0e-9) gh.image_order = false;
                    }
                    if (q[j] != q[a] || F.types[j] != F.types[a]) gh.image_order = false;
                }
            }
        }
        KTable kt;
        double self_mean = 0.0, halfrange = 0.0;
        D.k_off = (int)T.k64.size();
        if (charged || !rep1) build_ktable(D.A, kmax, alpha, V, kt);
        const int nkf = (int)kt.kpref.size();
        auto coupled = [&](int i) {
            for (int d = 0; d < 3; ++d)
                if (kt.n[3 * i + d] % F.replication[d] != 0) return false;
            return true;
        };
        // the host structure factor is needed at the k-vectors the batch keeps (every one with
        // fullk, the coupled ones otherwise) and at the replication check's sample of up to 32
        // uncoupled ones (src/batch.jl:173-195)
        std::vector<int> unc;
        size_t unc_stride = 1;
        if (!rep1) {
            for (int i = 0; i < nkf; ++i)
                if (!coupled(i)) unc.push_back(i);
            unc_stride = std::max<size_t>(1, (unc.size() + 31) / 32);
        }
        if (nkf > 0) {
            std::vector<char> need(nkf, 0);
            if (charged)
                for (int i = 0; i < nkf; ++i) need[i] = (fullk || coupled(i)) ? 1 : 0;
            for (size_t t = 0, taken = 0; t < unc.size() && taken < 32; t += unc_stride, ++taken) need[unc[t]] = 1;
            if (host_only) {
                structure_factor(kt, frac, q, need, kt.Sre, kt.Sim);
            } else {
                int s = device_structure_factor(bscratch, kt, frac, q, need, kt.Sre, kt.Sim);
                if (s) return s;
            }
        }
        if (!rep1) {
            if (!unc.empty()) {
                const size_t stride = unc_stride;
                double maxS = 0.0, sabs = 0.0;
                int arg = -1, taken = 0;
                for (size_t t = 0; t < unc.size() && taken < 32; t += stride, ++taken) {
                    double mS = std::hypot(kt.Sre[unc[t]], kt.Sim[unc[t]]);
                    if (mS > maxS) { maxS = mS; arg = unc[t]; }
                }
                for (int j = 0; j < n; ++j) sabs += std::fabs(q[j]);
                double thr = std::max(1.0e-8, 100.0 * 2.220446049250313e-16) * sabs;
                if (maxS > thr)
                    return fail(PA_ERR_ARG,
                                "framework %d claims replication (%d, %d, %d), but its host structure factor at k=[%g, %g, "
                                "%g] (not coupled to that replication) is %g, exceeding threshold %g = max(1e-8, "
                                "100·eps(Float64))·Σ|q_host|; this replication does not describe the atoms",
                                uniq[u] + 1, F.replication[0], F.replication[1], F.replication[2], kt.k[3 * arg],
                                kt.k[3 * arg + 1], kt.k[3 * arg + 2], maxS, thr);
            }
        }
        if (charged) {
            for (int i = 0; i < nkf; ++i) {
                if (!fullk && !coupled(i)) continue;
                T.k64.push_back(make_double4(kt.k[3 * i], kt.k[3 * i + 1], kt.k[3 * i + 2], kt.kpref[i]));
                T.S64.push_back(make_double2(kt.Sre[i], kt.Sim[i]));
                T.kn.push_back(make_int4(kt.n[3 * i], kt.n[3 * i + 1], kt.n[3 * i + 2], 0));
                gh.kn.insert(gh.kn.end(), &kt.n[3 * i], &kt.n[3 * i] + 3);
                gh.kpref.push_back(kt.kpref[i]);
                T.kc.push_back(make_float2((float)(2.0 * kt.kpref[i] * kt.Sre[i]), (float)(2.0 * kt.kpref[i] * kt.Sim[i])));
            }
            // orientation-averaged guest self term over the FULL k table (src/batch.jl:360-389)
            double samples[64];
            if (!host_only) {
                double invA[9];
                inv3(D.A, invA);
                int s = device_self_term(bscratch, kt, invA, guest, self_q, samples);
                if (s) return s;
            } else
            parallel_for(64, [&](int64_t ob, int64_t oe) {
                for (int64_t o = ob; o < oe; ++o) {
                    double gs[MAX_SITES][3];
                    for (int a = 0; a < NS; ++a) rotate(&self_q[4 * o], guest->sites + 3 * a, gs[a]);
                    double acc = 0.0;
                    for (int i = 0; i < nkf; ++i) {
                        double re = 0.0, im = 0.0;
                        for (int a = 0; a < NS; ++a) {
                            double ph = kt.k[3 * i] * gs[a][0] + kt.k[3 * i + 1] * gs[a][1] + kt.k[3 * i + 2] * gs[a][2];
                            re += guest->charges[a] * std::cos(ph);
                            im += guest->charges[a] * std::sin(ph);
                        }
                        acc += kt.kpref[i] * (re * re + im * im);
                    }
                    samples[o] = KE * acc;
                }
            });
            double sum = 0.0, mx = samples[0], mn = samples[0];
            for (double x : samples) { sum += x; mx = std::max(mx, x); mn = std::min(mn, x); }
            self_mean = sum / 64.0;
            halfrange = (mx - mn) / 2.0;
            const double KT_REF = KB * 300.0;
            if (2.0 * halfrange / KT_REF > 1.0e-3)
                return fail(PA_ERR_ARG,
                            "guest self-term half-range estimate %g eV (guard factor 2) at framework %d: 2·%g exceeds "
                            "0.001·KB·300K = %g eV; this guest/cell combination needs the per-insertion reciprocal sum",
                            halfrange, uniq[u] + 1, halfrange, 1.0e-3 * KT_REF);
        }
        D.nk = (int)T.k64.size() - D.k_off;
        D.kn_max[0] = D.kn_max[1] = D.kn_max[2] = 0;
        for (size_t i = D.k_off; i < T.kn.size(); ++i) {
            D.kn_max[0] = std::max(D.kn_max[0], std::abs(T.kn[i].x));
            D.kn_max[1] = std::max(D.kn_max[1], std::abs(T.kn[i].y));
            D.kn_max[2] = std::max(D.kn_max[2], std::abs(T.kn[i].z));
        }
        // constant_offset (src/batch.jl:396-407; tail_delta src/forcefield.jl:103-124)
        double tail = 0.0;
        if (ff->tail) {
            std::vector<double> counts(NT, 0.0), gcounts(NT, 0.0);
            for (int j = 0; j < n; ++j) counts[F.types[j]] += 1.0;
            for (int a = 0; a < NS; ++a) gcounts[guest->types[a]] += 1.0;
            double acc = 0.0;
            for (int x = 0; x < NT; ++x)
                for (int y = 0; y < NT; ++y) {
                    double e = ff->epsilon[x + NT * y];
                    if (e == 0.0) continue;
                    double sg = ff->sigma[x + NT * y], r = sg / rc_lj, r3 = r * r * r;
                    double c = e * (sg * sg * sg) * (r3 * r3 * r3 / 3.0 - r3);
                    acc += 2.0 * counts[x] * gcounts[y] * c + gcounts[x] * gcounts[y] * c;
                }
            tail = (8.0 * PI / 3.0) / V * acc;
        }
        double q2 = 0.0, Qh = 0.0, Qg = 0.0, excl = 0.0;
        for (int a = 0; a < NS; ++a) { q2 += guest->charges[a] * guest->charges[a]; Qg += guest->charges[a]; }
        for (int j = 0; j < n; ++j) Qh += q[j];
        for (int a = 0; a < NS; ++a)
            for (int c = a + 1; c < NS; ++c) {
                const double *sa = guest->sites + 3 * a, *sc = guest->sites + 3 * c;
                double d[3] = {sa[0] - sc[0], sa[1] - sc[1], sa[2] - sc[2]};
                double r = norm3(d);
                excl -= guest->charges[a] * guest->charges[c] * (1.0 - erfc_dev(alpha * r)) / r;
            }
        const double self = -alpha / std::sqrt(PI) * q2;
        const double net = -PI / (2.0 * V * alpha * alpha) * ((Qh + Qg) * (Qh + Qg) - Qh * Qh);
        D.const_off = tail + KE * (self + excl + net) + self_mean;
        b->info[u] = FwHost{n, D.nk, nkf, V, alpha, D.const_off, halfrange, {L[0], L[1], L[2]}};
    }

    // ---- guest and pair tables
    GuestDev &G = b->gdev;
    G.nsites = NS;
    G.charged = charged ? 1 : 0;
    for (int a = 0; a < NS; ++a) {
        for (int d = 0; d < 3; ++d) G.site[a][d] = guest->sites[3 * a + d];
        G.q[a] = guest->charges[a];
        G.qke[a] = (float)(KE * guest->charges[a]);
        G.type[a] = orig_to_c[guest->types[a]];
    }
    G.mirror = guest_mirror(G) ? 1 : 0;
    T.pair64.assign((size_t)NS * NTC, double2{});
    T.pair32.assign((size_t)NS * NTC, float2{});
    for (int a = 0; a < NS; ++a)
        for (int t = 0; t < NTC; ++t) {
            int go = guest->types[a], ho = c_to_orig[t];
            double sg = ff->sigma[go + NT * ho], ep = ff->epsilon[go + NT * ho];
            T.pair64[(size_t)a * NTC + t] = make_double2(sg, ep);
            T.pair32[(size_t)a * NTC + t] = make_float2((float)(sg * sg), (float)(4.0 * ep));
        }

    b->pair64_h = T.pair64;
    b->fullk = fullk;
    b->atom64_h = T.atom64;
    b->atype_h = T.atype;
    b->k64_h = T.k64;
    b->kn_h = T.kn;
    b->S64_h = T.S64;

    // ---- hard-core bound B_s and kmin per framework (src/batch.jl:395,409-423; src/reject.jl).
    // The attractive-pair term eps + |K| erfc(alpha r0)/r0 grows with |K|; it is evaluated on a
    // geometric grid |K|_i = |K|_max 1.02^-i per (site, type) and each atom takes the grid point
    // at or above its own |K|, which bounds the reference's per-atom term from above.
    b->hc_bs.assign(b->nfw, 0.0);
    b->hc_kmin.assign((size_t)b->nfw * NS * NTC, 0.0);
    double qg_abs = 0.0;
    for (int a = 0; a < NS; ++a) qg_abs += std::fabs(guest->charges[a]);
    parallel_for(b->nfw, [&](int64_t u0, int64_t u1) {
        for (int64_t u = u0; u < u1; ++u) {
            const FwDesc &D = b->desc[u];
            double acc = 0.0;
            bool unbounded = false;
            for (int a = 0; a < NS && !unbounded; ++a) {
                const double qa = guest->charges[a];
                std::vector<double> npos(NTC, 0.0), kmax(NTC, 0.0);
                std::vector<std::vector<double>> negK(NTC);
                for (int j = 0; j < D.natoms; ++j) {
                    const int c = T.atype[D.atom_off + j];
                    const double K = KE * qa * T.atom64[D.atom_off + j].w;
                    if (K >= 0.0) npos[c] += 1.0;
                    else { negK[c].push_back(-K); kmax[c] = std::max(kmax[c], -K); }
                }
                for (int c = 0; c < NTC && !unbounded; ++c) {
                    const double sg = T.pair64[(size_t)a * NTC + c].x, ep = T.pair64[(size_t)a * NTC + c].y;
                    acc += npos[c] * ep;
                    b->hc_kmin[((size_t)u * NS + a) * NTC + c] = -kmax[c];
                    if (negK[c].empty()) continue;
                    if (ep == 0.0) { unbounded = true; break; }
                    const double lg = std::log(1.02);
                    std::map<int, double> term;
                    for (double k : negK[c]) {
                        int i = std::min(511, (int)std::floor(std::log(kmax[c] / k) / lg));
                        while (i > 0 && kmax[c] * std::pow(1.02, -i) < k) --i;
                        auto it = term.find(i);
                        if (it == term.end()) {
                            const double G = kmax[c] * std::pow(1.02, -i);
                            const double r0 = find_r0(sg, ep, -G, alpha, rc_lj);
                            const double v = std::isnan(r0) ? INFINITY : ep + G * pair_erfc_f64(alpha * r0) / r0;
                            it = term.emplace(i, v).first;
                        }
                        acc += it->second;
                    }
                    if (!std::isfinite(acc)) unbounded = true;
                }
            }
            double Rs = 0.0;
            for (int k = 0; k < D.nk; ++k)
                Rs += T.k64[D.k_off + k].w * std::hypot(T.S64[D.k_off + k].x, T.S64[D.k_off + k].y);
            b->hc_bs[u] = unbounded ? INFINITY : acc + 2.0 * KE * qg_abs * Rs;
        }
    });

    if (flags & PA_BATCH_HOST_ONLY) {
        b->host_only = true;
        *out = b.release();
        return ok();
    }
    {
        int s = upload_batch(b.get(), T);
        if (s) return s;
    }
    *out = b.release();
    return ok();
}

extern "C" int pa_batch_from_tables(const pa_batch_tables *t, const pa_guest *guest, pa_batch **out)
{
    if (!out) return fail(PA_ERR_ARG, "out must not be NULL");
    *out = nullptr;
    if (!t || !guest) return fail(PA_ERR_ARG, "NULL argument");
    const int NFW = t->nframeworks, NSYS = t->nsys, NTC = t->ntypes, NS = guest->nsites;
    if (NFW < 1 || NSYS < 1) return fail(PA_ERR_ARG, "nframeworks=%d and nsys=%d must be >= 1", NFW, NSYS);
    if (NS < 1 || NS > MAX_SITES) return fail(PA_ERR_ARG, "guest has %d sites; 1..%d are supported", NS, MAX_SITES);
    if (NTC < 1 || NTC > 64) return fail(PA_ERR_ARG, "ntypes=%d: 1..64 compact LJ types are supported", NTC);
    if (!t->framework_of || !t->atom_offsets || !t->cells || !t->alphas || !t->k_offsets || !t->constant_offset ||
        !t->self_term_halfrange || !t->sigma || !t->epsilon || !t->compact_to_orig || !t->bs || !t->kmin)
        return fail(PA_ERR_ARG, "NULL table");
    if (t->atom_offsets[0] != 0 || t->k_offsets[0] != 0)
        return fail(PA_ERR_DIM, "atom_offsets and k_offsets must start at 0");
    for (int f = 0; f < NFW; ++f)
        if (t->atom_offsets[f + 1] < t->atom_offsets[f] || t->k_offsets[f + 1] < t->k_offsets[f])
            return fail(PA_ERR_DIM, "atom_offsets / k_offsets decrease at framework %d", f);
    const int64_t natoms = t->atom_offsets[NFW], nk = t->k_offsets[NFW];
    if (natoms > 0 && (!t->positions || !t->types || !t->charges)) return fail(PA_ERR_ARG, "NULL atom arrays");
    if (nk > 0 && (!t->ks || !t->kprefactor || !t->Shost)) return fail(PA_ERR_ARG, "NULL k-vector arrays");
    for (int s = 0; s < NSYS; ++s)
        if (t->framework_of[s] < 0 || t->framework_of[s] >= NFW)
            return fail(PA_ERR_ARG, "framework_of[%d] = %d is not a framework index 0..%d", s, t->framework_of[s],
                        NFW - 1);
    for (int64_t j = 0; j < natoms; ++j)
        if (t->types[j] < 0 || t->types[j] >= NTC)
            return fail(PA_ERR_ARG, "atom %lld has compact LJ type %d; the batch has %d", (long long)j, t->types[j], NTC);
    // the guest in force-field types, and its compact types through compact_to_orig
    std::vector<int> gtc(NS, -1);
    for (int a = 0; a < NS; ++a) {
        for (int c = 0; c < NTC && gtc[a] < 0; ++c)
            if (t->compact_to_orig[c] == guest->types[a]) gtc[a] = c;
        if (gtc[a] < 0)
            return fail(PA_ERR_ARG, "guest site %d has force-field type %d, which is not among the batch's compact types", a,
                        guest->types[a]);
    }
    const double rc_lj = t->cutoff, rc_ew = t->ewald_cutoff, rc = std::max(rc_lj, rc_ew);
    if (!(rc_lj > 0) || !(rc_ew > 0)) return fail(PA_ERR_ARG, "cutoff=%g and ewald_cutoff=%g must be positive", rc_lj, rc_ew);

    std::unique_ptr<pa_batch> b(new pa_batch());
    b->nsys = NSYS;
    b->nfw = NFW;
    b->fw_of.assign(t->framework_of, t->framework_of + NSYS);
    b->nsites = NS;
    b->gsites.assign(guest->sites, guest->sites + 3 * NS);
    b->gq.assign(guest->charges, guest->charges + NS);
    b->gtypes.assign(guest->types, guest->types + NS);
    b->rc_lj = rc_lj;
    b->rc_ew = rc_ew;
    b->ntypes_c = NTC;
    b->c_to_orig.assign(t->compact_to_orig, t->compact_to_orig + NTC);
    b->from_tables = true;
    b->fullk = t->fullk != 0;
    bool charged = false;
    for (int a = 0; a < NS; ++a) {
        charged = charged || guest->charges[a] != 0.0;
        b->r_guest = std::max(b->r_guest, norm3(guest->sites + 3 * a));
    }
    b->desc.resize(NFW);
    b->info.resize(NFW);
    b->fgh.resize(NFW);
    BatchTables T;
    for (int u = 0; u < NFW; ++u) {
        FwDesc &D = b->desc[u];
        std::memcpy(D.A, t->cells + 9 * (size_t)u, sizeof D.A);
        const double V = std::fabs(det3(D.A));
        if (!(V > 0)) return fail(PA_ERR_ARG, "framework %d has a singular cell", u);
        double L[3];
        perpendicular_lengths(D.A, L);
        int m[3];
        for (int d = 0; d < 3; ++d) m[d] = (int)std::ceil(2.0 * rc / L[d]);
        if (m[0] > 1 || m[1] > 1 || m[2] > 1)
            return fail(PA_ERR_ARG,
                        "framework %d is too small for cutoff %g (perpendicular widths %.4f, %.4f, %.4f Å must be >= %g); "
                        "replicate it by (%d, %d, %d) first",
                        u + 1, rc, L[0], L[1], L[2], 2 * rc, m[0], m[1], m[2]);
        const double alpha = t->alphas[u];
        if (charged && !(alpha * rc_ew <= PAIR_ERFC_XMAX))
            return fail(PA_ERR_ARG, "α·ewald_cutoff = %.17g exceeds PAIR_ERFC_XMAX = %.1f (framework %d)", alpha * rc_ew,
                        PAIR_ERFC_XMAX, u + 1);
        inv3(D.A, D.invA);
        for (int i = 0; i < 9; ++i) D.Af[i] = (float)(D.A[i] * 2.3283064365386963e-10);
        D.upper = (D.A[1] == 0.0 && D.A[2] == 0.0 && D.A[5] == 0.0) ? 1 : 0;
        b->upper_all = b->upper_all && D.upper;
        D.volume = V;
        D.alpha = alpha;
        D.alphaf = (float)alpha;
        D.const_off = t->constant_offset[u];
        const int a0 = t->atom_offsets[u], n = t->atom_offsets[u + 1] - a0;
        std::vector<double> frac(3 * (size_t)n);
        for (int j = 0; j < n; ++j) matvec(D.invA, t->positions + 3 * (size_t)(a0 + j), &frac[3 * (size_t)j]);
        {
            int s = add_atoms(T, D, frac.data(), t->positions + 3 * (size_t)a0, t->charges + a0, t->types + a0, n, L);
            if (s) return s;
        }
        FwGradHost &gh = b->fgh[u];
        gh.frac = frac;
        gh.counts.assign(NTC, 0.0);
        for (int j = 0; j < n; ++j) gh.counts[t->types[a0 + j]] += 1.0;
        // k-vectors: k = 2 pi A^-T n, so the integer coefficients are n = A^T k / (2 pi)
        D.k_off = (int)T.k64.size();
        D.kn_max[0] = D.kn_max[1] = D.kn_max[2] = 0;
        for (int i = t->k_offsets[u]; charged && i < t->k_offsets[u + 1]; ++i) {
            const double *k = t->ks + 3 * (size_t)i;
            int nn[3];
            for (int d = 0; d < 3; ++d) {
                const double x = (D.A[3 * d] * k[0] + D.A[3 * d + 1] * k[1] + D.A[3 * d + 2] * k[2]) / (2.0 * PI);
                const double r = std::nearbyint(x);
                if (std::fabs(x - r) > 1e-6 * std::max(1.0, std::fabs(x)))
                    return fail(PA_ERR_ARG, "k-vector %d of framework %d is not a reciprocal-lattice vector of its cell", i,
                                u + 1);
                nn[d] = (int)r;
                D.kn_max[d] = std::max(D.kn_max[d], std::abs(nn[d]));
            }
            const double kp = t->kprefactor[i], sre = t->Shost[2 * (size_t)i], sim = t->Shost[2 * (size_t)i + 1];
            T.k64.push_back(make_double4(k[0], k[1], k[2], kp));
            T.S64.push_back(make_double2(sre, sim));
            T.kn.push_back(make_int4(nn[0], nn[1], nn[2], 0));
            T.kc.push_back(make_float2((float)(2.0 * kp * sre), (float)(2.0 * kp * sim)));
            gh.kn.insert(gh.kn.end(), nn, nn + 3);
            gh.kpref.push_back(kp);
        }
        D.nk = (int)T.k64.size() - D.k_off;
        const int nk_u = t->k_offsets[u + 1] - t->k_offsets[u];
        b->info[u] = FwHost{n, D.nk, nk_u, V, alpha, D.const_off, t->self_term_halfrange[u], {L[0], L[1], L[2]}};
    }
    // guest and pair tables from the compact LJ tables
    GuestDev &G = b->gdev;
    G.nsites = NS;
    G.charged = charged ? 1 : 0;
    for (int a = 0; a < NS; ++a) {
        for (int d = 0; d < 3; ++d) G.site[a][d] = guest->sites[3 * a + d];
        G.q[a] = guest->charges[a];
        G.qke[a] = (float)(KE * guest->charges[a]);
        G.type[a] = gtc[a];
    }
    G.mirror = guest_mirror(G) ? 1 : 0;
    T.pair64.assign((size_t)NS * NTC, double2{});
    T.pair32.assign((size_t)NS * NTC, float2{});
    for (int a = 0; a < NS; ++a)
        for (int c = 0; c < NTC; ++c) {
            const double sg = t->sigma[gtc[a] + (size_t)NTC * c], ep = t->epsilon[gtc[a] + (size_t)NTC * c];
            T.pair64[(size_t)a * NTC + c] = make_double2(sg, ep);
            T.pair32[(size_t)a * NTC + c] = make_float2((float)(sg * sg), (float)(4.0 * ep));
        }
    b->pair64_h = T.pair64;
    b->atom64_h = T.atom64;
    b->atype_h = T.atype;
    b->k64_h = T.k64;
    b->kn_h = T.kn;
    b->S64_h = T.S64;
    // the hard-core bound and K_min as the host computed them (src/batch.jl:395,409-423)
    b->hc_bs.assign(t->bs, t->bs + NFW);
    b->hc_kmin.assign(t->kmin, t->kmin + (size_t)NFW * NS * NTC);
    {
        int s = upload_batch(b.get(), T);
        if (s) return s;
    }
    *out = b.release();
    return ok();
}

extern "C" void pa_batch_destroy(pa_batch *b) { delete b; }

extern "C" int pa_batch_nframeworks(const pa_batch *b, int32_t *nfw, int32_t *nsys)
{
    if (!b) return fail(PA_ERR_ARG, "NULL batch");
    if (nfw) *nfw = b->nfw;
    if (nsys) *nsys = b->nsys;
    return ok();
}

extern "C" int pa_batch_framework_of(const pa_batch *b, int32_t *framework_of)
{
    if (!b || !framework_of) return fail(PA_ERR_ARG, "NULL argument");
    for (int s = 0; s < b->nsys; ++s) framework_of[s] = b->fw_of[s];
    return ok();
}

extern "C" int pa_batch_framework_info(const pa_batch *b, int32_t fw, pa_framework_info *out)
{
    if (!b || !out) return fail(PA_ERR_ARG, "NULL argument");
    if (fw < 0 || fw >= b->nfw) return fail(PA_ERR_ARG, "framework index %d out of range 0..%d", fw, b->nfw - 1);
    const FwHost &h = b->info[fw];
    out->natoms = h.natoms;
    out->nk = h.nk;
    out->nk_full = h.nk_full;
    out->volume = h.volume;
    out->alpha = h.alpha;
    out->constant_offset = h.const_off;
    out->self_term_halfrange = h.halfrange;
    out->hardcore_bound = b->hc_bs[fw];
    return ok();
}

// ------------------------------------------------------------------------------------------
// Widom
// ------------------------------------------------------------------------------------------
static WidomArgs base_args(pa_batch *b)
{
    WidomArgs a{};
    a.fw = b->d_fw.as<FwDesc>();
    a.fw_of = b->d_fw_of.as<int>();
    a.nsys = b->nsys;
    a.atom64 = b->d_atom64.as<double4>();
    a.atype = b->d_atype.as<int>();
    a.k64 = b->d_k64.as<double4>();
    a.S64 = b->d_S64.as<double2>();
    a.atomfx = b->d_atomfx.as<uint4>();
    a.kn = b->d_kn.as<int4>();
    a.kc = b->d_kc.as<float2>();
    a.pair64 = b->d_pair64.as<double2>();
    a.pair32 = b->d_pair32.as<float2>();
    a.ntypes_c = b->ntypes_c;
    a.rc_lj2 = b->rc_lj * b->rc_lj;
    a.rc_ew2 = b->rc_ew * b->rc_ew;
    a.rc_lj2f = (float)a.rc_lj2;
    a.rc_ew2f = b->gdev.charged ? (float)a.rc_ew2 : 0.0f;
    a.rc_max2f = std::max(a.rc_lj2f, a.rc_ew2f);
    a.guest = b->gdev;
    return a;
}

static int check_precision(const pa_batch *b, uint32_t precision)
{
    if (b->host_only) return fail(PA_ERR_ARG, "batch was built with PA_BATCH_HOST_ONLY and has no device tables");
    precision &= ~PA_SINGLE_PHASE;
    if (precision != PA_PREC_F64 && precision != PA_PREC_MIXED)
        return fail(PA_ERR_ARG, "precision=%u: use PA_PREC_F64 (0) or PA_PREC_MIXED (1)", precision);
    if (b->nsites > 6) return fail(PA_ERR_ARG, "guests with %d sites: the device kernels take 1..6 sites", b->nsites);
    return PA_OK;
}

// Rejection radii^2 of system s at kT (build_rejection_tables, src/widom.jl:222-253): margin =
// (theta_F + 2) kT + safety + B_s - c_s, safety sized for energies summed at unit roundoff
// eps_f / 2. rho2: NS x ntypes_c, Å^2 (0: no rejection for that pair of types).
int pa::rejection_radii(const pa_batch *b, int s, double kT, double eps_f, double *rho2, double *rmax2)
{
    static const double theta = theta_f64();
    const int NS = b->nsites, NTC = b->ntypes_c;
    const int u = b->fw_of[s];
    const FwHost &h = b->info[u];
    const double Bs = b->hc_bs[u], cs = h.const_off;
    const double n = (double)NS * h.natoms + h.nk + 8;
    const double safety = 2.0 * n * eps_f * (Bs + std::fabs(cs)) + 4.0e-6 * Bs;
    const double margin = std::isinf(Bs) ? INFINITY : (theta + 2.0) * kT + safety + Bs - cs;
    if (!(margin > 0.0))
        return fail(PA_ERR_ARG,
                    "build_rejection_tables: system %d has margin=%g <= 0 (Bs=%g, safety=%g, constant_offset=%g, "
                    "kT=%g, theta_F=%g); no valid rejection radius exists",
                    s + 1, margin, Bs, safety, cs, kT, theta);
    *rmax2 = 0.0;
    for (int a = 0; a < NS; ++a)
        for (int c = 0; c < NTC; ++c) {
            const double2 p = b->pair64_h[(size_t)a * NTC + c];
            double r2 = find_rho2(p.x, p.y, b->hc_kmin[((size_t)u * NS + a) * NTC + c], margin, b->rc_lj);
            if (!(r2 >= 0.0)) r2 = 0.0;  // no bracket: no rejection for this pair of types
            rho2[(size_t)a * NTC + c] = r2;
            *rmax2 = std::max(*rmax2, r2);
        }
    return PA_OK;
}

// Per-system sample counts under the run-length assignment (src/widom.jl:44-53).
static void system_counts(int64_t ninsert, int64_t nsys, int64_t run, std::vector<int64_t> &ns)
{
    int64_t q = ninsert / run, rem = ninsert % run, full = q / nsys, hi = q % nsys;
    ns.assign(nsys, full * run);
    for (int64_t r = 0; r < hi; ++r) ns[r] += run;
    if (rem > 0) ns[q % nsys] += rem;
}

struct LseH {
    double m, w, u;
};

static LseH lse_merge_h(LseH x, LseH y)
{
    if (y.m == -INFINITY) return x;
    if (x.m == -INFINITY) return y;
    double m = std::max(x.m, y.m), ex = std::exp(x.m - m), ey = std::exp(y.m - m);
    return LseH{m, x.w * ex + y.w * ey, x.u * ex + y.u * ey};
}

// A per-call buffer request (CallScratch): past the end of a fixed region it is an argument
// error that states both sizes, otherwise a device error.
static int scratch_fail(const pa_batch *b, hipError_t e, const char *what, int line)
{
    if (b->scr.base && b->scr.need > 0)
        return fail(PA_ERR_ARG,
                    "the call needs %zu bytes of device scratch, more than the %zu bytes the no-allocation entry point "
                    "holds; pass fewer insertions or a smaller chunk",
                    b->scr.need, b->scr.cap);
    return fail(PA_ERR_DEVICE, "HIP error %s at %s:%d (%s)", hipGetErrorString(e), __FILE__, line, what);
}
#define SCR_TRY(b, expr)                                                                                \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return scratch_fail(b, e_, #expr, __LINE__);                              \
    } while (0)

// Smallest bin edge (Å) of the mixed path, and the insertions a bin aims to hold when a system
// has few insertions for its volume.
constexpr double BIN_EDGE_MIN = 2.0;
constexpr double BIN_TARGET_INSERTIONS = 64.0;

// Delta U of flattened insertions g in [gb, ge) (system s owns g in [seg[s], seg[s+1])) into
// b->d_dU[g - out_base], in chunks of at most `chunk` insertions. Records a HIP event pair around
// every energy-kernel launch (b->kev) and returns the launch count in *launches.
// With kT > 0 and the mixed path, insertions that phase 0 rejects get Delta U = +inf.
// With keys_out, the mixed path stops after phase 0 and binning and returns each insertion's
// bin key (-1: rejected) instead of energies.
static int run_energies(pa_batch *b, const std::vector<int64_t> &seg, int64_t gb, int64_t ge, int64_t out_base,
                        uint64_t seed, uint32_t precision, int64_t chunk, int64_t *launches, double kT, double *dU,
                        std::vector<int> *keys_out = nullptr)
{
    const int nsys = b->nsys;
    pa::CallScratch &scr = b->scr;
    int64_t *d_seg = nullptr;
    SCR_TRY(b, scr.put(b->d_seg, seg, &d_seg));
    // the default cuts the range into equal chunks of at most 2^24 (a short last chunk would run
    // the energy kernel on a fraction of the device; each launch ends on a tail of unevenly
    // sized bins, so fewer launches are faster: 52 B of scratch per insertion, 0.9 GB at 2^24);
    // an explicit chunk is taken as given
    int64_t CH = chunk > 0 ? std::max<int64_t>(chunk, 64) : ((int64_t)1 << 24);
    if (chunk <= 0 && ge - gb > CH) {
        const int64_t n = (ge - gb + CH - 1) / CH;
        CH = ((ge - gb + n - 1) / n + 255) / 256 * 256;
    }
    const int64_t nchunks = (ge - gb + CH - 1) / CH;
    // event pairs around the energy launches; a call into a fixed scratch region creates none
    // and records none
    const bool timed = scr.base == nullptr;
    while (timed && (int64_t)b->kev.size() < 2 * nchunks) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        b->kev.push_back(e);
    }
    *launches = 0;
    if ((precision & ~PA_SINGLE_PHASE) == PA_PREC_F64) {
        WidomArgs a = base_args(b);
        a.seg = d_seg;
        a.seed = seed;
        a.dU = dU;
        a.out_base = out_base;
        for (int64_t g0 = gb; g0 < ge; g0 += CH) {
            a.g0 = g0;
            a.g1 = std::min(ge, g0 + CH);
            if (timed) HIP_TRY(hipEventRecord(b->kev[2 * *launches], b->stream));
            HIP_TRY(launch_energy(a, (int)precision, b->upper_all, b->stream));
            if (timed) HIP_TRY(hipEventRecord(b->kev[2 * *launches + 1], b->stream));
            ++*launches;
        }
        return PA_OK;
    }
    // ---- mixed path: bin grid of every system, sized by its insertions in [gb, ge)
    const double rcmax = std::max(b->rc_lj, b->gdev.charged ? b->rc_ew : 0.0);
    std::vector<BinSys> bins(nsys);
    std::vector<int> base(nsys + 1, 0);
    double cap_est = 256.0;
    for (int s = 0; s < nsys; ++s) {
        const FwHost &h = b->info[b->fw_of[s]];
        const FwDesc &D = b->desc[b->fw_of[s]];
        const int64_t lo = std::max(gb, seg[s]), hi = std::min(ge, seg[s + 1]);
        const double nins = (double)std::max<int64_t>(hi - lo, 1);
        const double kappa = b->bin_scale > 0.0 ? b->bin_scale : 1.0;
        const double edge = kappa * std::max(BIN_EDGE_MIN, std::cbrt(BIN_TARGET_INSERTIONS * h.volume / nins));
        BinSys &B = bins[s];
        for (int d = 0; d < 3; ++d) B.nb[d] = std::max(1, std::min(1024, (int)std::floor(h.perp[d] / edge)));
        // circumradius of one bin: half its longest body diagonal
        auto circumradius = [&]() {
            double r = 0.0;
            for (int sy = -1; sy <= 1; sy += 2)
                for (int sz = -1; sz <= 1; sz += 2) {
                    const double f[3] = {0.5 / B.nb[0], 0.5 * sy / B.nb[1], 0.5 * sz / B.nb[2]};
                    double v[3];
                    matvec(D.A, f, v);
                    r = std::max(r, norm3(v));
                }
            return r;
        };
        double rb = circumradius();
        // the list keeps each atom's nearest image and the next one along every axis, which holds
        // every image within the list radius only while that radius is below every perpendicular
        // width: refine the longest bin edge until it is (few insertions per system make bins as
        // large as the cell)
        const double pmin = std::min(h.perp[0], std::min(h.perp[1], h.perp[2]));
        while (rcmax + b->r_guest + rb + 0.02 >= pmin * (1.0 - 1e-9)) {
            int dm = 0;
            double em = -1.0;
            for (int d = 0; d < 3; ++d) {
                const double e = norm3(D.A + 3 * d) / B.nb[d];
                if (e > em && B.nb[d] < 1024) { em = e; dm = d; }
            }
            if (em < 0.0) return fail(PA_ERR_INTERNAL, "system %d: no bin grid keeps the list radius below the cell widths", s + 1);
            ++B.nb[dm];
            rb = circumradius();
        }
        const int64_t nb = (int64_t)B.nb[0] * B.nb[1] * B.nb[2];
        if (nb > (1 << 22)) return fail(PA_ERR_INTERNAL, "system %d would need %lld bins", s + 1, (long long)nb);
        const double R = rcmax + b->r_guest + rb + 0.02;
        B.rlist2 = (float)(R * R);
        // radial classes of the list: everything closer than rc - r_bin - r_guest (in range of
        // every site of every insertion of the bin) in class 0, the rest in equal shells
        const double d0 = std::max(0.0, rcmax - rb - b->r_guest - 0.05);
        B.cls_d0 = (float)d0;
        B.cls_inv_w = (float)((binned_nclasses() - 1) / std::max(R - d0, 1e-3));
        B.reach[0] = B.reach[1] = B.reach[2] = -1;
        B.hc_rf[0] = B.hc_rf[1] = B.hc_rf[2] = 0u;
        for (int d = 0; d < 3; ++d) B.alt[d] = (float)std::max(0.0, 1.0 - R / h.perp[d]);
        B.bin_base = base[s];
        base[s + 1] = base[s] + (int)nb;
        cap_est = std::max(cap_est, 1.3 * h.natoms / h.volume * (4.0 / 3.0) * PI * R * R * R + 256.0);
    }
    // phase-0 rejection radii for this kT (src/widom.jl:202-253): margin = (theta + 2) kT + safety
    // + B_s - c_s with the FP32 rounding allowance of the mixed path's pair terms
    std::vector<float> rho2((size_t)nsys * b->nsites * b->ntypes_c, 0.0f);
    if (kT > 0.0) {
        const int NS = b->nsites, NTC = b->ntypes_c;
        if (b->rr_kT != kT) {
            // the radii depend on kT only: solved once per temperature
            b->rr_kT = -1.0;
            b->rr_r2.assign((size_t)nsys * NS * NTC, 0.0);
            b->rr_max2.assign(nsys, 0.0);
            for (int s = 0; s < nsys; ++s) {
                int st = rejection_radii(b, s, kT, (double)FLT_EPSILON, b->rr_r2.data() + (size_t)s * NS * NTC,
                                         &b->rr_max2[s]);
                if (st) return st;
            }
            b->rr_kT = kT;
        }
        for (int s = 0; s < nsys; ++s) {
            const double *r2s = b->rr_r2.data() + (size_t)s * NS * NTC;
            const double rmax2 = b->rr_max2[s];
            for (size_t i = 0; i < (size_t)NS * NTC; ++i) {
                // FP32 radius rounded down with a 1e-5 relative allowance for the device's FP32
                // distance
                const double r2 = r2s[i];
                float f = (float)(r2 * (1.0 - 1.0e-5));
                if ((double)f > r2 * (1.0 - 1.0e-5)) f = std::nextafter(f, 0.0f);
                rho2[(size_t)s * NS * NTC + i] = f;
            }
            const FwHost &h = b->info[b->fw_of[s]];
            const FwDesc &D = b->desc[b->fw_of[s]];
            if (rmax2 > 0.0)
                for (int d = 0; d < 3; ++d) {
                    bins[s].reach[d] = (int)std::ceil(std::sqrt(rmax2) * D.cg_nc[d] / h.perp[d]);
                    bins[s].hc_rf[d] = (uint32_t)std::min(4294967295.0,
                                                          std::ceil(std::sqrt(rmax2) / h.perp[d] * 4294967296.0) + 2.0);
                    // +-hc_rf plus one cell reaching around the axis: every cell, marked by the
                    // all-ones value so that the kernels need no division
                    const uint64_t rf = bins[s].hc_rf[d], nc = (uint64_t)D.cg_nc[d];
                    if (2 * rf + ((uint64_t)1 << 32) / nc >= ((uint64_t)1 << 32)) bins[s].hc_rf[d] = 0xFFFFFFFFu;
                }
        }
    }
    const int nbins = base[nsys];
    // list capacity: the estimate from the mean density, or the batch's fixed value, never below
    // one slice of the list build (binned_listcap_min) so that every pass makes progress
    const int cap_hi = binned_listcap_max(b->nsites), cap_lo = binned_listcap_min();
    int listcap = b->list_cap > 0 ? b->list_cap : (int)std::min<double>(cap_hi, std::ceil(cap_est / 64.0) * 64.0);
    listcap = std::max(cap_lo, std::min(cap_hi, listcap));
    if (cap_lo > cap_hi) return fail(PA_ERR_INTERNAL, "list capacity bounds %d > %d", cap_lo, cap_hi);
    b->list_cap_used = listcap;
    BinnedArgs A{};
    BinSys *d_bins = nullptr;
    int *d_base = nullptr;
    float *d_rho2 = nullptr;
    SCR_TRY(b, scr.put(b->d_bins, bins, &d_bins));
    SCR_TRY(b, scr.put(b->d_bin_base, base, &d_base));
    SCR_TRY(b, scr.put(b->d_rho2, rho2, &d_rho2));
    void *p = nullptr;
    SCR_TRY(b, scr.take(b->d_bin_count, sizeof(int) * (nbins + 1), &p));
    A.bin_count = (int *)p;
    SCR_TRY(b, scr.take(b->d_bin_start, sizeof(int) * (nbins + 1), &p));
    A.bin_start = (int *)p;
    const int64_t chunk_n = std::min<int64_t>(CH, ge - gb);
    size_t scan_bytes = 0;
    HIP_TRY(binned_scan_bytes(nbins, &scan_bytes));
    if (scr.base == nullptr && !keys_out) {
        // launch order by descending bin count (a fixed region keeps the plain order: its size is
        // stated without the sort's storage, and the order changes no insertion's energy; a call
        // that only returns the bin keys launches no energy kernel)
        A.sort_bits = 1;
        while (A.sort_bits < 31 && ((int64_t)1 << A.sort_bits) <= chunk_n) ++A.sort_bits;
        size_t sort_bytes = 0;
        HIP_TRY(binned_sort_bytes(nbins, A.sort_bits, &sort_bytes));
        scan_bytes = std::max(scan_bytes, sort_bytes);
        SCR_TRY(b, scr.take(b->d_bin_perm, sizeof(int) * 3 * (size_t)nbins, &p));
        A.bin_perm = (int *)p;
        A.bin_iota = A.bin_perm + nbins;
        A.bin_ksort = (unsigned *)(A.bin_perm + 2 * (size_t)nbins);
    }
    void *scan_tmp = nullptr;
    SCR_TRY(b, scr.take(b->d_scan_tmp, scan_bytes, &scan_tmp));
    SCR_TRY(b, scr.take(b->d_sitefx, sizeof(uint32_t) * 3 * b->nsites * chunk_n, &p));
    A.sitefx = (uint32_t *)p;
    SCR_TRY(b, scr.take(b->d_key, sizeof(int) * chunk_n, &p));
    A.key = (int *)p;
    SCR_TRY(b, scr.take(b->d_order, sizeof(int) * chunk_n, &p));
    A.order = (int *)p;
    SCR_TRY(b, scr.take(b->d_rank, sizeof(int) * chunk_n, &p));
    A.rank = (int *)p;
    SCR_TRY(b, scr.take(b->d_cand, sizeof(int) * chunk_n, &p));
    A.cand = (int *)p;
    SCR_TRY(b, scr.take(b->d_ncand, sizeof(int), &p));
    A.ncand = (int *)p;
    SCR_TRY(b, scr.take(b->d_passes, sizeof(int), &p));
    A.extra_passes = (int *)p;
    b->passes_recorded = scr.base == nullptr;  // a fixed region's counter is not kept after the call

    A.fw = b->d_fw.as<FwDesc>();
    A.fw_of = b->d_fw_of.as<int>();
    A.seg = d_seg;
    A.bins = d_bins;
    A.bin_base = d_base;
    A.nsys = nsys;
    A.nbins = nbins;
    A.seed = seed;
    A.dU = dU;
    A.out_base = out_base;
    A.atomfx = b->d_binfx.as<uint4>();
    A.atype = b->d_bintype.as<int>();
    A.kn = b->d_kn.as<int4>();
    A.kc = b->d_kc.as<float2>();
    A.pair32 = b->d_pair32.as<float2>();
    A.ntypes_c = b->ntypes_c;
    A.rc_lj2f = (float)(b->rc_lj * b->rc_lj);
    A.rc_ew2f = b->gdev.charged ? (float)(b->rc_ew * b->rc_ew) : 0.0f;
    A.rc_max2f = std::max(A.rc_lj2f, A.rc_ew2f);
    A.listcap = listcap;
    A.guest = b->gdev;
    A.corefx = b->d_corefx.as<uint4>();
    A.cell_start = b->d_cell_start.as<int>();
    A.nbr_start = b->d_nbr_start.as<int>();
    A.nbr_atoms = b->d_nbr_atoms.as<uint4>();
    A.rho2 = d_rho2;
    for (int64_t g0 = gb; g0 < ge; g0 += CH) {
        A.g0 = g0;
        A.g1 = std::min(ge, g0 + CH);
        A.zero_passes = g0 == gb;
        HIP_TRY(launch_binning(A, scan_tmp, scan_bytes, b->stream));
        if (keys_out) {
            keys_out->resize(ge - gb);
            HIP_TRY(hipMemcpyAsync(keys_out->data() + (g0 - gb), A.key, sizeof(int) * (A.g1 - A.g0),
                                   hipMemcpyDeviceToHost, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
            continue;
        }
        if (timed) HIP_TRY(hipEventRecord(b->kev[2 * *launches], b->stream));
        HIP_TRY(launch_binned_energy(A, b->stream));
        if (timed) HIP_TRY(hipEventRecord(b->kev[2 * *launches + 1], b->stream));
        ++*launches;
    }
    return PA_OK;
}

// pa_widom's body; also returns the flattened system offsets of the run (the energies stay in
// b->d_dU for a following gradient pass).
static int widom_core(pa_batch *b, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed, int32_t nblocks,
                      int32_t run, int64_t chunk, uint32_t precision, pa_widom_result *out, std::vector<int64_t> *seg_out,
                      std::vector<double3> *sums_out = nullptr)
{
    if (!b || !guest || !out) return fail(PA_ERR_ARG, "NULL argument");
    // argument checks in the reference's order and wording (src/widom.jl:302-330)
    if (nblocks < 2) return fail(PA_ERR_ARG, "nblocks=%d: at least two blocks are needed for a standard error", nblocks);
    if (chunk < 0) return fail(PA_ERR_ARG, "chunk=%lld must be ≥ 1", (long long)chunk);
    bool same = guest->nsites == b->nsites;
    for (int a = 0; same && a < b->nsites; ++a) {
        same = guest->types[a] == b->gtypes[a] && guest->charges[a] == b->gq[a];
        for (int d = 0; d < 3; ++d) same = same && guest->sites[3 * a + d] == b->gsites[3 * a + d];
    }
    if (!same)
        return fail(PA_ERR_ARG, "guest passed to widom does not match the guest FrameworkBatch was built with; build a new "
                                "FrameworkBatch for a different guest");
    if (!(T > 0)) return fail(PA_ERR_ARG, "T=%g must be a positive temperature in K", T);
    if (ninsert < 1) return fail(PA_ERR_ARG, "ninsert=%lld must be ≥ 1", (long long)ninsert);
    const int64_t nsys = b->nsys, per = ninsert / nsys;
    int64_t runlen;
    if (run == 0) {
        runlen = std::min<int64_t>(256, std::max<int64_t>(1, (ninsert / nsys) / 4));
    } else {
        if (run < 1) return fail(PA_ERR_ARG, "run=%d must be ≥ 1", run);
        if (run > per) return fail(PA_ERR_ARG, "run=%d must be ≤ per=%lld (= ninsert ÷ nsys)", run, (long long)per);
        runlen = run;
    }
    std::vector<int64_t> ns;
    system_counts(ninsert, nsys, runlen, ns);
    int64_t nmin = ns[0], amin = 0;
    for (int64_t s = 0; s < nsys; ++s)
        if (ns[s] < nmin) { nmin = ns[s]; amin = s; }
    if (nmin < 2 * (int64_t)nblocks)
        return fail(PA_ERR_ARG,
                    "ninsert=%lld, nsys=%lld, run=%lld: system %lld gets only %lld samples, need at least 2·nblocks = %d per "
                    "system",
                    (long long)ninsert, (long long)nsys, (long long)runlen, (long long)(amin + 1), (long long)nmin,
                    2 * nblocks);
    int st = check_precision(b, precision);
    if (st) return st;
    const double kT = KB * T;

    std::vector<int64_t> seg(nsys + 1, 0);
    for (int64_t s = 0; s < nsys; ++s) seg[s + 1] = seg[s] + ns[s];
    const int64_t G = seg[nsys];
    pa::CallScratch &scr = b->scr;
    double *dU = nullptr;
    {
        void *p = nullptr;
        SCR_TRY(b, scr.take(b->d_dU, sizeof(double) * G, &p));
        dU = (double *)p;
    }

    // tiles: each block of each system split into runs of at most TILE insertions
    const int64_t TILE = 2048;
    std::vector<int64_t> tile_g0;
    std::vector<int> tile_n;
    std::vector<int> tile_owner;  // s * nblocks + blk
    for (int64_t s = 0; s < nsys; ++s) {
        const int64_t blen = ns[s] / nblocks;
        for (int blk = 0; blk < nblocks; ++blk) {
            const int64_t j0 = blk * blen, j1 = (blk == nblocks - 1) ? ns[s] : (blk + 1) * blen;
            for (int64_t t = j0; t < j1; t += TILE) {
                tile_g0.push_back(seg[s] + t);
                tile_n.push_back((int)std::min(TILE, j1 - t));
                tile_owner.push_back((int)(s * nblocks + blk));
            }
        }
    }
    const int ntiles = (int)tile_g0.size();
    int64_t *d_tile_g0 = nullptr;
    int *d_tile_n = nullptr;
    double4 *d_tiles = nullptr;
    SCR_TRY(b, scr.put(b->d_tile_g0, tile_g0, &d_tile_g0));
    SCR_TRY(b, scr.put(b->d_tile_n, tile_n, &d_tile_n));
    {
        void *p = nullptr;
        SCR_TRY(b, scr.take(b->d_tiles, sizeof(double4) * ntiles, &p));
        d_tiles = (double4 *)p;
    }

    const double kT_phase0 = (precision & PA_SINGLE_PHASE) ? 0.0 : kT;
    // a call into a fixed scratch region never tunes: it uses the batch's bin scale as it stands
    if ((precision & ~PA_SINGLE_PHASE) == PA_PREC_MIXED && b->bin_scale == 0.0 && G >= ((int64_t)1 << 18) &&
        !scr.base) {
        // autotune the bin edge multiplier once per batch: time the energy pipeline for each
        // candidate and keep the fastest. Candidates differ by a few percent, and which one wins
        // depends on the batch's mix of framework sizes, so the sample is the whole call up to
        // 2^25 insertions; a longer call is sampled in four windows of 2^22 spread over it
        // (insertions run system-major: each window holds whole systems at the call's density)
        const double cand[] = {0.75, 1.0, 1.25, 1.5, 1.75, 2.0, 2.5};
        const int64_t W = (int64_t)1 << 22;
        const int nwin = G > ((int64_t)1 << 25) ? 4 : 1;
        int64_t wb[4], we[4];
        for (int r = 0; r < nwin; ++r) {
            wb[r] = nwin == 1 ? 0 : (G - W) * r / (nwin - 1);
            we[r] = nwin == 1 ? G : wb[r] + W;
        }
        double best = 1.0, best_ms = INFINITY;
        for (int rep = 0; rep < 2; ++rep)
            for (double kappa : cand) {
                b->bin_scale = kappa;
                HIP_TRY(hipEventRecord(b->ev[0], b->stream));
                for (int r = 0; r < nwin; ++r) {
                    int64_t nl = 0;
                    st = run_energies(b, seg, wb[r], we[r], 0, seed, precision, chunk, &nl, kT_phase0, dU);
                    if (st) { b->bin_scale = 0.0; return st; }
                }
                HIP_TRY(hipEventRecord(b->ev[1], b->stream));
                HIP_TRY(hipEventSynchronize(b->ev[1]));
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
                if (rep == 1 && ms < best_ms) { best_ms = ms; best = kappa; }
            }
        b->bin_scale = best;
    }
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    int64_t launches = 0;
    // the mixed path runs phase 0 (hard-core rejection) unless PA_SINGLE_PHASE; the FP64 plain
    // path evaluates every insertion
    st = run_energies(b, seg, 0, G, 0, seed, precision, chunk, &launches, kT_phase0, dU);
    if (st) return st;
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(launch_tiles(dU, d_tile_g0, d_tile_n, ntiles, kT, d_tiles, b->stream));
    HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    // the tile sums back into page-locked memory (a call into a fixed scratch region allocates
    // nothing, so it reads them into ordinary host memory)
    std::vector<double4> tiles_pageable;
    double4 *tiles = nullptr;
    if (scr.base) {
        tiles_pageable.resize(ntiles);
        tiles = tiles_pageable.data();
    } else {
        HIP_TRY(b->h_tiles.ensure(sizeof(double4) * std::max(ntiles, 1)));
        tiles = b->h_tiles.as<double4>();
    }
    HIP_TRY(hipMemcpyAsync(tiles, d_tiles, sizeof(double4) * ntiles, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float e01 = 0, e12 = 0, e02 = 0;
    HIP_TRY(hipEventElapsedTime(&e01, b->ev[0], b->ev[1]));
    HIP_TRY(hipEventElapsedTime(&e12, b->ev[1], b->ev[2]));
    HIP_TRY(hipEventElapsedTime(&e02, b->ev[0], b->ev[2]));
    double kernel_ms = 0.0;
    for (size_t i = 0; !scr.base && i + 1 < b->kev.size() && (int64_t)(i / 2) < launches; i += 2) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, b->kev[i], b->kev[i + 1]));
        kernel_ms += ms;
    }
    b->last = pa_timing{e02, kernel_ms, launches, e12, G, e01 - kernel_ms, 0.0, 0};

    // combine tiles per (system, block) in order, then the reference's block statistics
    std::vector<LseH> blk((size_t)nsys * nblocks, LseH{-INFINITY, 0.0, 0.0});
    std::vector<double> nanc((size_t)nsys * nblocks, 0.0);
    for (int t = 0; t < ntiles; ++t) {
        const double4 &x = tiles[t];
        blk[tile_owner[t]] = lse_merge_h(blk[tile_owner[t]], LseH{x.x, x.y, x.z});
        nanc[tile_owner[t]] += x.w;
    }
    for (int64_t s = 0; s < nsys; ++s) {
        const FwHost &h = b->info[b->fw_of[s]];
        const int64_t blen = ns[s] / nblocks;
        double M = -INFINITY, nans = 0.0;
        for (int k = 0; k < nblocks; ++k) {
            M = std::max(M, blk[s * nblocks + k].m);
            nans += nanc[s * nblocks + k];
        }
        if (nans > 0 || M == INFINITY)
            return fail(PA_ERR_ARG,
                        "system %lld: Boltzmann accumulation overflowed Float64 (W=NaN, UW=NaN); an insertion energy "
                        "below about -709·kT was sampled",
                        (long long)(s + 1));
        if (M == -INFINITY) M = 0.0;  // every weight is exactly zero
        // block sums scaled by exp(-M)
        std::vector<double> sW(nblocks), sUW(nblocks), nb(nblocks);
        double tW = 0.0, tUW = 0.0, ntot = 0.0;
        for (int k = 0; k < nblocks; ++k) {
            const LseH &x = blk[s * nblocks + k];
            const double sc = x.m == -INFINITY ? 0.0 : std::exp(x.m - M);
            sW[k] = x.w * sc;
            sUW[k] = x.u * sc;
            nb[k] = (double)((k == nblocks - 1) ? ns[s] - k * blen : blen);
            tW += sW[k];
            tUW += sUW[k];
            ntot += nb[k];
        }
        if (sums_out) {
            sums_out->resize(nsys);
            (*sums_out)[s] = make_double3(M, tW, tUW);
        }
        // _reduce (src/widom.jl:451-492) on the scaled sums: W = Ws * e^M
        const double Ws = tW / ntot, UWs = tUW / ntot;
        double varW = 0.0, varUW = 0.0, cov = 0.0;
        for (int k = 0; k < nblocks; ++k) {
            const double mW = sW[k] / nb[k], mUW = sUW[k] / nb[k];
            varW += (mW - Ws) * (mW - Ws);
            varUW += (mUW - UWs) * (mUW - UWs);
            cov += (mW - Ws) * (mUW - UWs);
        }
        varW /= (nblocks - 1);
        varUW /= (nblocks - 1);
        cov /= (nblocks - 1);
        const double semWs = std::sqrt(varW / nblocks);
        const double ratio = UWs / Ws;
        const double var_ratio =
            UWs == 0.0 ? 0.0 : ratio * ratio * (varUW / (UWs * UWs) + varW / (Ws * Ws) - 2.0 * cov / (UWs * Ws)) / nblocks;
        const double eM = std::exp(M);
        const double W = Ws * eM, UW = UWs * eM;
        if (!std::isfinite(W) || !std::isfinite(UW))
            return fail(PA_ERR_ARG,
                        "system %lld: Boltzmann accumulation overflowed Float64 (W=%g, UW=%g); an insertion energy below "
                        "about -709·kT was sampled",
                        (long long)(s + 1), W, UW);
        const double lnW = std::log(Ws) + M;
        const double mu_ex = -kT * lnW, q_st = kT - ratio;
        if (!std::isfinite(mu_ex)) return fail(PA_ERR_ARG, "system %lld: mu_ex = %g is not finite", (long long)(s + 1), mu_ex);
        if (!std::isfinite(q_st)) return fail(PA_ERR_ARG, "system %lld: q_st = %g is not finite", (long long)(s + 1), q_st);
        pa_widom_result &r = out[s];
        r.mu_ex = mu_ex;
        r.mu_ex_err = kT * semWs / Ws;
        r.K_H = h.volume * W / kT;
        r.K_H_err = h.volume * semWs * eM / kT;
        r.q_st = q_st;
        r.q_st_err = std::sqrt(std::max(var_ratio, 0.0));
        r.nsamples = (int64_t)ntot;
        r.nblocks = nblocks;
        r.ln_K_H = std::log(h.volume) + lnW - std::log(kT);
    }
    if (seg_out) *seg_out = seg;
    return ok();
}

extern "C" int pa_widom(pa_batch *b, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed, int32_t nblocks,
                        int32_t run, int64_t chunk, uint32_t precision, pa_widom_result *out)
{
    return widom_core(b, guest, T, ninsert, seed, nblocks, run, chunk, precision, out, nullptr);
}

// ------------------------------------------------------------------------------------------
// pa_widom_noalloc: pa_widom with every per-call device buffer carved from one fixed array
// ------------------------------------------------------------------------------------------
// The array lives in this file's code object, once per device. Host code never uses its name
// as an address: the device address comes from hipGetSymbolAddress with that device current,
// fetched on the first call on each device and kept for the process.
static __device__ __align__(256) unsigned char scratch[PA_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on `device`.
int scratch_region(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= SCRATCH_MAX_DEVICES)
        return fail(PA_ERR_DEVICE, "device %d: the no-allocation entry point serves devices 0..%d", device,
                    SCRATCH_MAX_DEVICES - 1);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(PA_ERR_DEVICE, "HIP error %s: no device address for the scratch array on device %d",
                        hipGetErrorString(e), device);
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return PA_OK;
}
}  // namespace

extern "C" int pa_widom_noalloc(pa_batch *b, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed,
                                int32_t nblocks, int32_t run, int64_t chunk, uint32_t precision, pa_widom_result *out)
{
    if (!b) return fail(PA_ERR_ARG, "NULL argument");
    unsigned char *base = nullptr;
    int st = scratch_region(b->device, &base);
    if (st) return st;
    // the batch's per-call buffers come from the array for this call only
    struct Bind {
        CallScratch &s;
        ~Bind() { s = CallScratch{}; }
    } bind{b->scr};
    b->scr = CallScratch{};
    b->scr.base = base;
    b->scr.cap = PA_NOALLOC_SCRATCH_BYTES;
    return widom_core(b, guest, T, ninsert, seed, nblocks, run, chunk, precision, out, nullptr);
}

extern "C" int pa_widom_grad(pa_batch *b, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed,
                             int32_t nblocks, int32_t run, uint32_t precision, pa_widom_result *out, double *dlnkh_deps,
                             double *dlnkh_dsig, double *dqst_deps, double *dqst_dsig, double *dlnkh_dq, double *dqst_dq)
{
    if (!b || !guest) return fail(PA_ERR_ARG, "NULL argument");
    if (b->from_tables)
        return fail(PA_ERR_ARG, "pa_widom_grad needs the force field (per-type parameters, tail flag): build the batch "
                                "with pa_batch_create, not from FrameworkBatch tables");
    const int nsys = b->nsys, NT = b->nt_ff, NS = b->nsites, NTC = b->ntypes_c;
    const bool want_lj = dlnkh_deps || dlnkh_dsig || dqst_deps || dqst_dsig;
    const bool want_q = dlnkh_dq || dqst_dq;
    if (want_lj) {
        // the per-type derivatives go through Lorentz-Berthelot mixing (src/forcefield.jl:30-45)
        for (int x = 0; x < NT; ++x)
            for (int y = 0; y < NT; ++y) {
                const double sx = b->ff_sigma[x + NT * x], sy = b->ff_sigma[y + NT * y];
                const double ex = b->ff_eps[x + NT * x], ey = b->ff_eps[y + NT * y];
                const double sxy = b->ff_sigma[x + NT * y], exy = b->ff_eps[x + NT * y];
                if (std::fabs(sxy - 0.5 * (sx + sy)) > 1e-12 * std::max(1.0, std::fabs(sxy)) ||
                    std::fabs(exy - std::sqrt(ex * ey)) > 1e-12 * std::max(1e-300, std::fabs(exy)))
                    return fail(PA_ERR_ARG,
                                "force-field types %d and %d are not Lorentz-Berthelot mixed (sigma %g, epsilon %g); the "
                                "per-type LJ gradients need sigma_xy = (sigma_x + sigma_y)/2 and epsilon_xy = "
                                "sqrt(epsilon_x epsilon_y)",
                                x, y, sxy, exy);
            }
    }
    if (want_q)
        for (int s = 0; s < nsys; ++s) {
            const FwGradHost &gh = b->fgh[b->fw_of[s]];
            if (!gh.image_order)
                return fail(PA_ERR_ARG,
                            "system %d: charge gradients of a framework with replication (%d, %d, %d) need its atoms in "
                            "replicate order (image-major, i fastest, atoms inner)",
                            s + 1, gh.rep[0], gh.rep[1], gh.rep[2]);
        }
    std::vector<pa_widom_result> res_local;
    if (!out) {
        res_local.resize(nsys);
        out = res_local.data();
    }
    std::vector<int64_t> seg;
    std::vector<double3> sums;
    int st = widom_core(b, guest, T, ninsert, seed, nblocks, run, 0, precision, out, &seg, &sums);
    if (st) return st;
    const double kT = KB * T, beta = 1.0 / kT;
    const int64_t G = seg[nsys];
    struct EvPair {
        hipEvent_t a = nullptr, b = nullptr;
        ~EvPair()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } evp;
    HIP_TRY(hipEventCreateWithFlags(&evp.a, hipEventDisableSystemFence));
    HIP_TRY(hipEventCreateWithFlags(&evp.b, hipEventDisableSystemFence));
    hipEvent_t g0 = evp.a, g1 = evp.b;

    // weighted insertions, selected on the device: w = exp(-beta U - m_s) relative to the system's
    // best insertion (m_s and the scaled sums W = sum w, R = sum w U come from the Widom tiles);
    // weights below e^-40 change no sum at FP64 resolution
    std::vector<double> msys(nsys), Wsum(nsys), Rsum(nsys);
    for (int s = 0; s < nsys; ++s) {
        msys[s] = sums[s].x;
        Wsum[s] = sums[s].y;
        Rsum[s] = sums[s].z;
    }
    HIP_TRY(upload(b->d_msys, msys));
    size_t sel_bytes = 0;
    HIP_TRY(grad_select(nullptr, nullptr, nsys, nullptr, beta, G, nullptr, nullptr, nullptr, nullptr, &sel_bytes,
                        b->stream));
    HIP_TRY(b->d_sel_tmp.ensure(sel_bytes));
    HIP_TRY(b->d_flag.ensure(std::max<int64_t>(G, 1)));
    HIP_TRY(b->d_list_g.ensure(sizeof(int64_t) * std::max<int64_t>(G, 1)));
    HIP_TRY(b->d_count.ensure(sizeof(int64_t)));
    HIP_TRY(hipEventRecord(g0, b->stream));
    HIP_TRY(grad_select(b->d_dU.as<double>(), b->d_seg.as<int64_t>(), nsys, b->d_msys.as<double>(), beta, G,
                        b->d_flag.as<unsigned char>(), b->d_list_g.as<int64_t>(), b->d_count.as<int64_t>(),
                        b->d_sel_tmp.p, &sel_bytes, b->stream));
    int64_t nlist = 0;
    HIP_TRY(hipMemcpyAsync(&nlist, b->d_count.p, sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    // each system's slice of the list (ascending g, so systems are contiguous) and the gather
    // workgroups: 256 cell-sorted atoms of one system that has weighted insertions
    std::vector<int64_t> loff(nsys + 1, 0);
    HIP_TRY(b->d_loff.ensure(sizeof(int64_t) * (nsys + 1)));
    HIP_TRY(grad_offsets(b->d_list_g.as<int64_t>(), nlist, b->d_seg.as<int64_t>(), nsys, b->d_loff.as<int64_t>(),
                         b->stream));
    HIP_TRY(hipMemcpyAsync(loff.data(), b->d_loff.p, sizeof(int64_t) * (nsys + 1), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    // gather workgroups: (256 cell-sorted atoms, chunk of the system's list) and k-space
    // workgroups: (256 coupled k-vectors, chunk of the list); chunks keep a system with many
    // weighted insertions from serialising on a few workgroups
    const bool charged = b->gdev.charged != 0;
    const bool kspace = charged && want_q;
    constexpr int64_t GRAD_CHUNK = 1024;
    std::vector<int4> blk, kblk;
    for (int s = 0; s < nsys; ++s) {
        const int64_t nl = loff[s + 1] - loff[s];
        if (nl <= 0) continue;
        if (nl > 0x7fffffff) return fail(PA_ERR_ARG, "system %d: %lld weighted insertions exceed the int32 range", s + 1,
                                         (long long)nl);
        const int n = b->info[b->fw_of[s]].natoms, nk = b->info[b->fw_of[s]].nk;
        for (int64_t c0 = 0; c0 < nl; c0 += GRAD_CHUNK) {
            const int c1 = (int)std::min<int64_t>(nl, c0 + GRAD_CHUNK);
            for (int a0 = 0; a0 < n; a0 += 256) blk.push_back(make_int4(s, a0, (int)c0, c1));
            if (kspace)
                for (int k0 = 0; k0 < nk; k0 += 256) kblk.push_back(make_int4(s, k0, (int)c0, c1));
        }
    }
    if (blk.size() > 0x7fffffffu || kblk.size() > 0x7fffffffu)
        return fail(PA_ERR_ARG, "too many gradient workgroups (%zu, %zu)", blk.size(), kblk.size());
    HIP_TRY(upload(b->d_blk, blk));
    HIP_TRY(upload(b->d_kblk, kblk));
    HIP_TRY(b->d_lsite.ensure(sizeof(uint32_t) * 3 * NS * std::max<int64_t>(nlist, 1)));
    HIP_TRY(b->d_lw.ensure(sizeof(double2) * std::max<int64_t>(nlist, 1)));
    std::vector<int64_t> qoff(nsys + 1, 0), koff(nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        qoff[s + 1] = qoff[s] + b->info[b->fw_of[s]].natoms;
        koff[s + 1] = koff[s] + b->info[b->fw_of[s]].nk;
    }
    std::vector<float4> pairg((size_t)NS * NTC);
    for (size_t i = 0; i < pairg.size(); ++i) {
        const double sg = b->pair64_h[i].x, ep = b->pair64_h[i].y;
        pairg[i] = make_float4((float)(sg * sg), (float)(4.0 * ep), sg > 0 ? (float)(1.0 / sg) : 0.0f, 0.0f);
    }
    const size_t nlj = (size_t)nsys * NS * NTC * 4;
    HIP_TRY(upload(b->d_qoff, qoff));
    HIP_TRY(upload(b->d_koff, koff));
    HIP_TRY(upload(b->d_pairg, pairg));
    HIP_TRY(b->d_acc_lj.ensure(sizeof(double) * nlj));
    HIP_TRY(b->d_acc_q.ensure(sizeof(double) * std::max<int64_t>(1, 2 * qoff[nsys])));
    HIP_TRY(b->d_acc_k.ensure(sizeof(double) * std::max<int64_t>(1, 4 * koff[nsys])));
    HIP_TRY(hipMemsetAsync(b->d_acc_lj.p, 0, sizeof(double) * nlj, b->stream));
    HIP_TRY(hipMemsetAsync(b->d_acc_q.p, 0, sizeof(double) * std::max<int64_t>(1, 2 * qoff[nsys]), b->stream));
    HIP_TRY(hipMemsetAsync(b->d_acc_k.p, 0, sizeof(double) * std::max<int64_t>(1, 4 * koff[nsys]), b->stream));
    GradArgs A{};
    A.fw = b->d_fw.as<FwDesc>();
    A.fw_of = b->d_fw_of.as<int>();
    A.seg = b->d_seg.as<int64_t>();
    A.nsys = nsys;
    A.seed = seed;
    A.list_g = b->d_list_g.as<int64_t>();
    A.dU = b->d_dU.as<double>();
    A.msys = b->d_msys.as<double>();
    A.beta = beta;
    A.nlist = nlist;
    A.atomfx = b->d_atomfx.as<uint4>();
    A.atype = b->d_atype.as<int>();
    A.kn = b->d_kn.as<int4>();
    A.pairg = b->d_pairg.as<float4>();
    A.ntypes_c = NTC;
    A.rc_lj2f = (float)(b->rc_lj * b->rc_lj);
    A.rc_ew2f = charged ? (float)(b->rc_ew * b->rc_ew) : 0.0f;
    A.rc_max2f = std::max(A.rc_lj2f, A.rc_ew2f);
    A.guest = b->gdev;
    A.acc_lj = b->d_acc_lj.as<double>();
    A.acc_q = b->d_acc_q.as<double>();
    A.qoff = b->d_qoff.as<int64_t>();
    A.acc_k = b->d_acc_k.as<double>();
    A.koff = b->d_koff.as<int64_t>();
    A.loff = b->d_loff.as<int64_t>();
    A.lsite = b->d_lsite.as<uint32_t>();
    A.lw = b->d_lw.as<double2>();
    A.corefx = b->d_corefx.as<uint4>();
    A.coreidx = b->d_coreidx.as<int>();
    A.blk = b->d_blk.as<int4>();
    A.kblk = b->d_kblk.as<int4>();
    // the Coulomb and k-space sums feed only the charge derivatives
    if (!want_q) A.guest.charged = 0;
    HIP_TRY(launch_grad(A, b->upper_all, (int)blk.size(), (int)kblk.size(), b->stream));
    if (charged && want_q) {
        // the per-atom reciprocal term in workgroups of (system, 256 atoms, 256 k-vectors)
        std::vector<int4> rblk;
        for (int s = 0; s < nsys; ++s) {
            const FwDesc &D = b->desc[b->fw_of[s]];
            for (int a0 = 0; a0 < D.natoms; a0 += 256)
                for (int k0 = 0; k0 < D.nk; k0 += 256) rblk.push_back(make_int4(s, a0, k0, 0));
        }
        HIP_TRY(upload(b->d_rblk, rblk));
        HIP_TRY(launch_grad_recip(A, b->d_k64.as<double4>(), b->d_rblk.as<int4>(), (int)rblk.size(), b->stream));
    }
    HIP_TRY(hipEventRecord(g1, b->stream));
    std::vector<double> acc_lj(nlj);
    HIP_TRY(hipMemcpyAsync(acc_lj.data(), b->d_acc_lj.p, sizeof(double) * nlj, hipMemcpyDeviceToHost, b->stream));
    // the per-atom charge accumulators (2 per atom, MBs for large batches) into page-locked memory
    const size_t nacc_q = (size_t)2 * qoff[nsys];
    HIP_TRY(b->h_acc_q.ensure(sizeof(double) * std::max<size_t>(nacc_q, 1)));
    double *acc_q = b->h_acc_q.as<double>();
    if (nacc_q)
        HIP_TRY(hipMemcpyAsync(acc_q, b->d_acc_q.p, sizeof(double) * nacc_q, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float gms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&gms, g0, g1));
    b->last.grad_ms = gms;
    b->last.grad_insertions = nlist;

    // per-type chain rule through Lorentz-Berthelot mixing:
    //   d eps_xy / d eps_t = eps_xy / (2 eps_t) ([x == t] + [y == t]),
    //   d sigma_xy / d sigma_t = ([x == t] + [y == t]) / 2
    auto mix_eps = [&](int x, int y, int t) -> double {
        const int c = (x == t) + (y == t);
        if (!c) return 0.0;
        const double et = b->ff_eps[t + NT * t];
        return et == 0.0 ? NAN : b->ff_eps[x + NT * y] / (2.0 * et) * c;
    };
    auto mix_sig = [&](int x, int y, int t) -> double { return 0.5 * ((x == t) + (y == t)); };
    for (int s = 0; s < nsys; ++s) {
        const int u = b->fw_of[s];
        const FwGradHost &gh = b->fgh[u];
        const double W = Wsum[s], R = Rsum[s], V = b->info[u].volume;
        auto dlnk = [&](double Gs) { return -beta * Gs / W; };
        auto dqst = [&](double Gs, double Hs) { return -(Gs - beta * Hs) / W - beta * (R / W) * (Gs / W); };
        if (want_lj) {
            std::vector<double> Ge(NT, 0.0), He(NT, 0.0), Gs(NT, 0.0), Hs(NT, 0.0);
            const double *al = acc_lj.data() + (size_t)s * NS * NTC * 4;
            for (int a = 0; a < NS; ++a)
                for (int c = 0; c < NTC; ++c) {
                    const double *x = al + 4 * ((size_t)a * NTC + c);
                    if (x[0] == 0.0 && x[1] == 0.0 && x[2] == 0.0 && x[3] == 0.0) continue;
                    const int gx = b->gtypes[a], hy = b->c_to_orig[c];
                    for (int t = 0; t < NT; ++t) {
                        const double fe = mix_eps(gx, hy, t), fs = mix_sig(gx, hy, t);
                        if (fe != 0.0) { Ge[t] += fe * x[0]; He[t] += fe * x[1]; }
                        if (fs != 0.0) { Gs[t] += fs * x[2]; Hs[t] += fs * x[3]; }
                    }
                }
            if (b->ff_tail) {
                // tail_delta (src/forcefield.jl:103-124): (8 pi / 3V) sum_xy (2 n_x g_y + g_x g_y) c_xy with
                // c = eps (sigma^12 / (3 rc^9) - sigma^6 / rc^3); every insertion carries it
                std::vector<double> gcount(NT, 0.0);
                for (int a = 0; a < NS; ++a) gcount[b->gtypes[a]] += 1.0;
                const double rc = b->rc_lj, rc3 = rc * rc * rc, rc9 = rc3 * rc3 * rc3, pre = (8.0 * PI / 3.0) / V;
                for (int x = 0; x < NT; ++x)
                    for (int y = 0; y < NT; ++y) {
                        const double mult = 2.0 * gh.counts[x] * gcount[y] + gcount[x] * gcount[y];
                        if (mult == 0.0) continue;
                        const double sg = b->ff_sigma[x + NT * y], ep = b->ff_eps[x + NT * y];
                        const double s3 = sg * sg * sg, s6 = s3 * s3, s5 = s6 / sg, s11 = s6 * s5;
                        const double dce = s6 * s6 / (3.0 * rc9) - s6 / rc3;
                        const double dcs = ep * (4.0 * s11 / rc9 - 6.0 * s5 / rc3);
                        for (int t = 0; t < NT; ++t) {
                            const double fe = mix_eps(x, y, t), fs = mix_sig(x, y, t);
                            if (fe != 0.0 && ep != 0.0) {
                                Ge[t] += W * pre * mult * dce * fe;
                                He[t] += R * pre * mult * dce * fe;
                            } else if (fe != 0.0 && std::isnan(fe)) {
                                Ge[t] = He[t] = NAN;
                            }
                            if (fs != 0.0) {
                                Gs[t] += W * pre * mult * dcs * fs;
                                Hs[t] += R * pre * mult * dcs * fs;
                            }
                        }
                    }
            }
            for (int t = 0; t < NT; ++t) {
                const size_t o = (size_t)s * NT + t;
                if (dlnkh_deps) dlnkh_deps[o] = dlnk(Ge[t]);
                if (dlnkh_dsig) dlnkh_dsig[o] = dlnk(Gs[t]);
                if (dqst_deps) dqst_deps[o] = dqst(Ge[t], He[t]);
                if (dqst_dsig) dqst_dsig[o] = dqst(Gs[t], Hs[t]);
            }
        }
        if (want_q) {
            const int n = b->info[u].natoms;
            double *gq = acc_q + 2 * qoff[s];
            // the real-space part of atom j moves with its replication images: average over them
            const int nimg = gh.rep[0] * gh.rep[1] * gh.rep[2];
            if (nimg > 1 && n > 0) {
                const int nu = n / nimg;
                for (int a = 0; a < nu; ++a) {
                    double sw = 0.0, su = 0.0;
                    for (int m = 0; m < nimg; ++m) { sw += gq[2 * (m * nu + a)]; su += gq[2 * (m * nu + a) + 1]; }
                    for (int m = 0; m < nimg; ++m) { gq[2 * (m * nu + a)] = sw / nimg; gq[2 * (m * nu + a) + 1] = su / nimg; }
                }
            }
            // net-charge term of constant_offset: d/dq_j of -KE pi/(2 V alpha^2) ((Q_h + Q_g)^2 - Q_h^2)
            double Qg = 0.0;
            for (int a = 0; a < NS; ++a) Qg += b->gq[a];
            const double alpha = b->info[u].alpha;
            const double cnet = charged ? -KE * PI * Qg / (V * alpha * alpha) : 0.0;
            for (int j = 0; j < n; ++j) {
                const double Gj = gq[2 * j] + W * cnet, Hj = gq[2 * j + 1] + R * cnet;
                if (dlnkh_dq) dlnkh_dq[qoff[s] + j] = dlnk(Gj);
                if (dqst_dq) dqst_dq[qoff[s] + j] = dqst(Gj, Hj);
            }
        }
    }
    return ok();
}

extern "C" int pa_widom_energies(pa_batch *b, int32_t sys, uint64_t seed, int64_t j0, int64_t n, uint32_t precision,
                                 double *dU)
{
    if (!b || !dU) return fail(PA_ERR_ARG, "NULL argument");
    int st = check_precision(b, precision);
    if (st) return st;
    if (sys < 0 || sys >= b->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, b->nsys - 1);
    if (j0 < 0 || n < 0) return fail(PA_ERR_ARG, "insertion range j0=%lld, n=%lld is invalid", (long long)j0, (long long)n);
    if (n == 0) return ok();
    // system `sys` owns g in [0, j0 + n); the launches cover g in [j0, j0 + n)
    std::vector<int64_t> seg(b->nsys + 1, 0);
    for (int s = sys + 1; s <= b->nsys; ++s) seg[s] = j0 + n;
    HIP_TRY(b->d_dU.ensure(sizeof(double) * n));
    int64_t launches = 0;
    st = run_energies(b, seg, j0, j0 + n, j0, seed, precision, 0, &launches, 0.0, b->d_dU.as<double>());
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(dU, b->d_dU.p, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return ok();
}

extern "C" int pa_hardcore_flags(pa_batch *b, int32_t sys, uint64_t seed, double T, int64_t j0, int64_t n,
                                 uint8_t *flags)
{
    if (!b || (!flags && n > 0)) return fail(PA_ERR_ARG, "NULL argument");
    int st = check_precision(b, PA_PREC_MIXED);
    if (st) return st;
    if (sys < 0 || sys >= b->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, b->nsys - 1);
    if (j0 < 0 || n < 0) return fail(PA_ERR_ARG, "j0=%lld, n=%lld must be >= 0", (long long)j0, (long long)n);
    if (!(T > 0)) return fail(PA_ERR_ARG, "T=%g must be a positive temperature in K", T);
    if (n == 0) return ok();
    std::vector<int64_t> seg(b->nsys + 1, 0);
    for (int s = 0; s < b->nsys; ++s) seg[s + 1] = seg[s] + (s == sys ? j0 + n : 0);
    HIP_TRY(b->d_dU.ensure(sizeof(double) * (size_t)n));
    std::vector<int> keys;
    int64_t launches = 0;
    st = run_energies(b, seg, seg[sys] + j0, seg[sys] + j0 + n, seg[sys] + j0, seed, PA_PREC_MIXED, 0, &launches, KB * T,
                      b->d_dU.as<double>(), &keys);
    if (st) return st;
    for (int64_t i = 0; i < n; ++i) flags[i] = keys[i] < 0 ? 1 : 0;
    return ok();
}

extern "C" int pa_widom_poses(pa_batch *b, int32_t sys, uint64_t seed, int64_t j0, int64_t n, double *pos, double *quat)
{
    if (!b || !pos || !quat) return fail(PA_ERR_ARG, "NULL argument");
    if (b->host_only) return fail(PA_ERR_ARG, "batch was built with PA_BATCH_HOST_ONLY and has no device tables");
    if (sys < 0 || sys >= b->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, b->nsys - 1);
    if (j0 < 0 || n < 0) return fail(PA_ERR_ARG, "insertion range j0=%lld, n=%lld is invalid", (long long)j0, (long long)n);
    if (n == 0) return ok();
    HIP_TRY(b->d_pos.ensure(sizeof(double) * 3 * n));
    HIP_TRY(b->d_quat.ensure(sizeof(double) * 4 * n));
    HIP_TRY(launch_poses(b->d_fw.as<FwDesc>(), b->d_fw_of.as<int>(), sys, seed, j0, n, b->d_pos.as<double>(),
                         b->d_quat.as<double>(), b->stream));
    HIP_TRY(hipMemcpyAsync(pos, b->d_pos.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipMemcpyAsync(quat, b->d_quat.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return ok();
}

extern "C" int pa_batch_set_bin_scale(pa_batch *b, double kappa)
{
    if (!b) return fail(PA_ERR_ARG, "NULL batch");
    if (!(kappa >= 0.0) || kappa > 8.0) return fail(PA_ERR_ARG, "bin scale %g must be in (0, 8], or 0 to re-tune", kappa);
    b->bin_scale = kappa;
    return ok();
}

extern "C" int pa_batch_bin_scale(const pa_batch *b, double *kappa)
{
    if (!b || !kappa) return fail(PA_ERR_ARG, "NULL argument");
    *kappa = b->bin_scale;
    return ok();
}

extern "C" int pa_batch_set_list_capacity(pa_batch *b, int32_t capacity)
{
    if (!b) return fail(PA_ERR_ARG, "NULL batch");
    const int lo = binned_listcap_min(), hi = binned_listcap_max(b->nsites);
    if (capacity != 0 && (capacity < lo || capacity > hi || capacity % 64 != 0))
        return fail(PA_ERR_ARG, "list capacity %d must be a multiple of 64 in [%d, %d], or 0 to size it per call", capacity,
                    lo, hi);
    b->list_cap = capacity;
    return ok();
}

extern "C" int pa_batch_list_stats(pa_batch *b, int32_t *capacity, int32_t *max_passes)
{
    if (!b) return fail(PA_ERR_ARG, "NULL batch");
    if (b->host_only) return fail(PA_ERR_ARG, "batch was built with PA_BATCH_HOST_ONLY and has no device tables");
    int extra = 0;
    if (b->list_cap_used > 0 && b->passes_recorded && b->d_passes.p) {
        HIP_TRY(hipMemcpyAsync(&extra, b->d_passes.p, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    if (capacity) *capacity = b->list_cap_used;
    if (max_passes) *max_passes = b->list_cap_used > 0 && b->passes_recorded ? extra + 1 : 0;
    return ok();
}

extern "C" int pa_last_timing(const pa_batch *b, pa_timing *out)
{
    if (!b || !out) return fail(PA_ERR_ARG, "NULL argument");
    *out = b->last;
    return ok();
}
