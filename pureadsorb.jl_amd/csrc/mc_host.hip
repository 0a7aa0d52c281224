// mc_host.hip -- host side of the move-energy path: SystemState on the device, the move entry
// points, the FP64 from-scratch energy and audit, and the GCMC cycle driver (C ABI in
// include/pureadsorb_amd.h).
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "pa_host.hpp"

namespace pa {
hipError_t launch_mc(const McArgs &a, hipStream_t st);
hipError_t launch_chain_widom(const ChainWidomArgs &a, hipStream_t st);
hipError_t launch_chain_snapshot(const ChainSnap &S, int64_t slot, const McArgs &m, hipStream_t st);
hipError_t launch_audit(const AuditArgs &a, hipStream_t st);
}  // namespace pa

using namespace pa;

// Device-resident chains of one batch (src/state.jl:89-104) plus the move workspace.
struct pa_state {
    pa_batch *b = nullptr;
    int nsys = 0, NS = 0;
    std::vector<int> guest_off, sk_off;
    std::vector<uint64_t> seed;
    std::vector<double> const_p, const_q;  // exchange constant term, eV (src/moves.jl:728-773)
    std::vector<int> acc_at_audit;         // accepted moves per system at the last audit
    int64_t samp_stride = 0;               // production samples per system of the last GCMC run
    DevAlloc d_guest_off, d_occ, d_refpos, d_quat, d_host_e, d_Sk, d_dS, d_skabs, d_sk_off, d_energy, d_energy_abs,
        d_seed, d_counter, d_acc, d_att, d_step_t, d_step_r, d_fug, d_cp, d_cq, d_samp_e, d_samp_n, d_maxocc,
        d_caphit, d_trace;
    // NVT driver: fractional atom table, insertion constant per system, move-launch snapshots,
    // test-insertion scratch, and the per-cycle records of the last run (nsys x nvt_np)
    DevAlloc d_atomf64, d_atype_s, d_chunk, d_chunk_off, d_lines, d_rounds, d_round_off, d_perm, d_coff,
        d_half_perp, d_ct, d_snap_pos, d_snap_quat, d_snap_sk, d_snap_occ, d_item_off, d_cw_dU, d_pose_pos,
        d_pose_quat;
    // the audit's device rebuild: per-axis power tables, structure factor, per-slot host-guest
    // energies, per-system sums
    DevAlloc d_aud_tab, d_aud_sk, d_aud_he, d_aud_sums;
    // set when pa_state_set_value put an occupancy above its capacity: every entry point that
    // reads the guest slots refuses the state until the occupancy is set back
    bool occ_over_capacity = false;
    std::vector<int> occ_live;  // live guests per system when an occupancy was first set over its capacity
    double cull_extra = 0.0;
    int64_t ctot = 0;       // entries of one snapshot slot's weighted structure factors
    int64_t snap_slots = 0; // slots allocated
    std::vector<double> nvt_E, nvt_sW, nvt_sUW;
    int64_t nvt_np = 0;
    // device time of the last pa_nvt_run / pa_widom_chain_energies, from events around each launch
    hipEvent_t tev[2] = {nullptr, nullptr};
    pa_nvt_timing timing{};
    ~pa_state()
    {
        for (auto &e : tev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace {

int ensure_chain_tables(pa_state *st);

// Entry points that read the guest slots (or launch kernels over them) refuse a state whose
// occupancy a test put above a capacity.
int refuse_over_capacity(const pa_state *st)
{
    if (st->occ_over_capacity)
        return fail(PA_ERR_ARG, "a system's occupancy exceeds its capacity (set through pa_state_set_value); only "
                                "pa_audit_energy, pa_state_info and pa_state_set_value take this state until it is set back");
    return PA_OK;
}

// ---- FP64 host energies (the same terms the kernel sums, in one serial order) ----------------

void guest_sites_h(const pa_batch *b, const double p[3], const double q[4], double out[][3])
{
    for (int s = 0; s < b->nsites; ++s) {
        double r[3];
        rotate(q, b->gdev.site[s], r);
        for (int d = 0; d < 3; ++d) out[s][d] = p[d] + r[d];
    }
}

double mi_r2_h(const FwDesc &F, const double a[3], const double c[3])
{
    double d[3] = {a[0] - c[0], a[1] - c[1], a[2] - c[2]}, m[3];
    minimum_image(F.A, F.invA, d, m);
    return m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
}

// host_guest_realspace_energy (src/guest.jl:176-222), per-site minimum image.
double host_guest_h(const pa_batch *b, int u, const double sites[][3])
{
    const FwDesc &F = b->desc[u];
    const double rc_lj2 = b->rc_lj * b->rc_lj, rc_ew2 = b->gdev.charged ? b->rc_ew * b->rc_ew : 0.0;
    const int NTC = b->ntypes_c;
    double e = 0.0;
    for (int j = 0; j < F.natoms; ++j) {
        const double4 h = b->atom64_h[F.atom_off + j];
        const double hp[3] = {h.x, h.y, h.z};
        const int ht = b->atype_h[F.atom_off + j];
        for (int s = 0; s < b->nsites; ++s) {
            const double r2 = mi_r2_h(F, sites[s], hp);
            if (r2 < rc_lj2 || r2 < rc_ew2) {
                const double2 p = b->pair64_h[(size_t)s * NTC + ht];
                e += site_pair_energy(r2, p.x, p.y, KE * b->gq[s] * h.w, F.alpha, rc_lj2, rc_ew2);
            }
        }
    }
    return e;
}

// guest_pair_realspace_energy (src/guest.jl:49-71), E_lj + KE E_sr.
double guest_pair_h(const pa_batch *b, int u, const double sa[][3], const double sb[][3])
{
    const FwDesc &F = b->desc[u];
    const double rc_lj2 = b->rc_lj * b->rc_lj, rc_ew2 = b->gdev.charged ? b->rc_ew * b->rc_ew : 0.0;
    const int NTC = b->ntypes_c;
    double e = 0.0;
    for (int x = 0; x < b->nsites; ++x)
        for (int y = 0; y < b->nsites; ++y) {
            const double r2 = mi_r2_h(F, sa[x], sb[y]);
            const double2 p = b->pair64_h[(size_t)x * NTC + b->gdev.type[y]];
            e += site_pair_energy(r2, p.x, p.y, KE * b->gq[x] * b->gq[y], F.alpha, rc_lj2, rc_ew2);
        }
    return e;
}

// guest_self_terms (src/guest.jl:26-42): KE (self + exclusion) of one guest.
int guest_self_h(const pa_batch *b, double alpha, double *out)
{
    double q2 = 0.0, excl = 0.0;
    for (int a = 0; a < b->nsites; ++a) q2 += b->gq[a] * b->gq[a];
    for (int a = 0; a < b->nsites; ++a)
        for (int c = a + 1; c < b->nsites; ++c) {
            const double *sa = &b->gsites[3 * a], *sc = &b->gsites[3 * c];
            const double d[3] = {sa[0] - sc[0], sa[1] - sc[1], sa[2] - sc[2]};
            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (!(r < b->rc_ew))
                return fail(PA_ERR_ARG,
                            "guest_self_terms: intramolecular distance %g (sites %d, %d) is not less than the real-space "
                            "cutoff %g",
                            r, a + 1, c + 1, b->rc_ew);
            excl -= b->gq[a] * b->gq[c] * (1.0 - erfc_dev_host(alpha * r)) / r;
        }
    *out = KE * (-alpha / std::sqrt(PI) * q2 + excl);
    return PA_OK;
}

// guest_tail_correction (src/guest.jl:9-14): tail_delta with N times the guest's type counts.
double guest_tail_h(const pa_batch *b, int u, double N)
{
    if (!b->ff_tail) return 0.0;
    const int NT = b->nt_ff;
    std::vector<double> g(NT, 0.0);
    for (int a = 0; a < b->nsites; ++a) g[b->gtypes[a]] += N;
    const std::vector<double> &cnt = b->fgh[u].counts;
    const double rc = b->rc_lj;
    double acc = 0.0;
    for (int x = 0; x < NT; ++x)
        for (int y = 0; y < NT; ++y) {
            const double e = b->ff_eps[x + NT * y];
            if (e == 0.0) continue;
            const double sg = b->ff_sigma[x + NT * y], r = sg / rc, r3 = r * r * r;
            const double c = e * (sg * sg * sg) * (r3 * r3 * r3 / 3.0 - r3);
            acc += 2.0 * cnt[x] * g[y] * c + g[x] * g[y] * c;
        }
    return (8.0 * PI / 3.0) / b->info[u].volume * acc;
}

double host_charge(const pa_batch *b, int u)
{
    const FwDesc &F = b->desc[u];
    double q = 0.0;
    for (int j = 0; j < F.natoms; ++j) q += b->atom64_h[F.atom_off + j].w;
    return q;
}

// net-charge correction at N guests (src/guest.jl:339-341).
double net_h(const pa_batch *b, int u, double N)
{
    double Qg = 0.0;
    for (int a = 0; a < b->nsites; ++a) Qg += b->gq[a];
    const double Qh = host_charge(b, u), V = b->info[u].volume, al = b->info[u].alpha;
    return -KE * PI / (2.0 * V * al * al) * ((Qh + N * Qg) * (Qh + N * Qg) - Qh * Qh);
}

// Structure factor of system s: the framework's plus every live guest's, at the system's k.
void rebuild_sk(const pa_batch *b, int u, const std::vector<double4> &pos, const std::vector<double4> &quat, int g0,
                int occ, std::vector<double2> &out)
{
    const FwDesc &F = b->desc[u];
    out.assign(F.nk, make_double2(0.0, 0.0));
    for (int k = 0; k < F.nk; ++k) out[k] = b->S64_h[F.k_off + k];
    if (!b->gdev.charged) return;
    // the guests' terms exactly as the move kernel forms them: per-axis powers, then products
    double sites[MAX_SITES][3];
    const int m0 = F.kn_max[0] + 1, m1 = F.kn_max[1] + 1, tw = m0 + m1 + F.kn_max[2] + 1;
    std::vector<double2> tab((size_t)MAX_SITES * tw);
    for (int i = 0; i < occ; ++i) {
        const double p[3] = {pos[g0 + i].x, pos[g0 + i].y, pos[g0 + i].z};
        const double q[4] = {quat[g0 + i].x, quat[g0 + i].y, quat[g0 + i].z, quat[g0 + i].w};
        guest_sites_h(b, p, q, sites);
        for (int s = 0; s < b->nsites; ++s)
            for (int x = 0; x < tw; ++x) {
                const int d = x < m0 ? 0 : (x < m0 + m1 ? 1 : 2);
                tab[(size_t)s * tw + x] = axis_power(F.invA, sites[s], d, x - (d == 0 ? 0 : (d == 1 ? m0 : m0 + m1)));
            }
        for (int k = 0; k < F.nk; ++k) {
            const int4 n = b->kn_h[F.k_off + k];
            for (int s = 0; s < b->nsites; ++s) {
                const double2 *tb = tab.data() + (size_t)s * tw;
                const double2 c = cis_from_axes(tb[std::abs(n.x)], tb[m0 + std::abs(n.y)], tb[m0 + m1 + std::abs(n.z)],
                                                n.x, n.y, n.z);
                out[k].x += b->gq[s] * c.x;
                out[k].y += b->gq[s] * c.y;
            }
        }
    }
}

// total_energy (src/guest.jl:293-343) of one system from its poses, with `sk` the structure
// factor the reciprocal term reads. Also returns the host-guest energy of every guest.
int total_energy_h(const pa_batch *b, int u, const std::vector<double4> &pos, const std::vector<double4> &quat, int g0,
                   int occ, const std::vector<double2> &sk, double *E, std::vector<double> *host_e)
{
    const FwDesc &F = b->desc[u];
    const int NS = b->nsites;
    std::vector<double> sites((size_t)occ * NS * 3);
    auto site = [&](int i) { return reinterpret_cast<double(*)[3]>(&sites[(size_t)i * NS * 3]); };
    double e = 0.0;
    if (host_e) host_e->assign(occ, 0.0);
    for (int i = 0; i < occ; ++i) {
        const double p[3] = {pos[g0 + i].x, pos[g0 + i].y, pos[g0 + i].z};
        const double q[4] = {quat[g0 + i].x, quat[g0 + i].y, quat[g0 + i].z, quat[g0 + i].w};
        guest_sites_h(b, p, q, site(i));
        const double hi = host_guest_h(b, u, site(i));
        if (host_e) (*host_e)[i] = hi;
        e += hi;
    }
    for (int i = 0; i < occ; ++i)
        for (int j = i + 1; j < occ; ++j) e += guest_pair_h(b, u, site(i), site(j));
    double rec = 0.0;
    for (int k = 0; k < F.nk; ++k) {
        const double2 s = sk[k], h = b->S64_h[F.k_off + k];
        rec += b->k64_h[F.k_off + k].w * ((s.x * s.x + s.y * s.y) - (h.x * h.x + h.y * h.y));
    }
    e += KE * rec;
    double gse = 0.0;
    int st = guest_self_h(b, F.alpha, &gse);
    if (st) return st;
    e += occ * gse;
    e += guest_tail_h(b, u, occ);
    e += net_h(b, u, occ);
    *E = e;
    return PA_OK;
}

template <class T>
hipError_t download(std::vector<T> &v, const DevAlloc &d, size_t n)
{
    v.resize(n);
    if (n == 0) return hipSuccess;
    return hipMemcpy(v.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost);
}

McArgs base_mc(pa_state *st)
{
    pa_batch *b = st->b;
    McArgs a{};
    a.fw = b->d_fw.as<FwDesc>();
    a.fw_of = b->d_fw_of.as<int>();
    a.atom64 = b->d_atom64.as<double4>();
    a.atype = b->d_atype.as<int>();
    a.atomf = st->d_atomf64.as<double4>();
    a.atypef = st->d_atype_s.as<int>();
    a.upper = b->upper_all ? 1 : 0;
    a.k64 = b->d_k64.as<double4>();
    a.kn = b->d_kn.as<int4>();
    a.ktab_w = 0;
    for (const FwDesc &D : b->desc) a.ktab_w = std::max(a.ktab_w, D.kn_max[0] + D.kn_max[1] + D.kn_max[2] + 3);
    a.pair64 = b->d_pair64.as<double2>();
    a.ntypes_c = b->ntypes_c;
    a.rc_lj2 = b->rc_lj * b->rc_lj;
    a.rc_ew2 = b->gdev.charged ? b->rc_ew * b->rc_ew : 0.0;
    a.rc_max2 = std::max(a.rc_lj2, a.rc_ew2);
    a.guest = b->gdev;
    a.nsys = st->nsys;
    a.guest_off = st->d_guest_off.as<int>();
    a.occupancy = st->d_occ.as<int>();
    a.refpos = st->d_refpos.as<double4>();
    a.quat = st->d_quat.as<double4>();
    a.host_e = st->d_host_e.as<double>();
    a.Sk = st->d_Sk.as<double2>();
    a.dS = st->d_dS.as<double2>();
    a.sk_abs = st->d_skabs.as<double>();
    a.sk_off = st->d_sk_off.as<int>();
    a.energy = st->d_energy.as<double>();
    a.energy_abs = st->d_energy_abs.as<double>();
    a.rng_seed = st->d_seed.as<uint64_t>();
    a.rng_counter = st->d_counter.as<uint64_t>();
    a.accepted = st->d_acc.as<int>();
    a.attempted = st->d_att.as<int>();
    a.step_trans = st->d_step_t.as<double>();
    a.step_rot = st->d_step_r.as<double>();
    a.fug = st->d_fug.as<double>();
    a.const_p = st->d_cp.as<double>();
    a.const_q = st->d_cq.as<double>();
    a.max_occ = st->d_maxocc.as<int>();
    a.capacity_hits = st->d_caphit.as<int>();
    return a;
}

// Runs fn(n) for every system on the host threads; returns the status of the lowest-numbered
// failing system with its message (what a serial loop stopping at the first failure reports).
template <class F>
int run_per_system(int nsys, F &&fn)
{
    std::vector<int> code(nsys, 0);
    std::vector<std::string> msg(nsys);
    parallel_for(
        nsys,
        [&](int64_t n0, int64_t n1) {
            for (int64_t n = n0; n < n1; ++n) {
                const int r = fn((int)n);
                if (r) {
                    char buf[1024];
                    pa_last_error(buf, sizeof buf);
                    code[n] = r;
                    msg[n] = buf;
                    return;
                }
            }
        },
        2);
    for (int n = 0; n < nsys; ++n)
        if (code[n]) return fail(code[n], "%s", msg[n].c_str());
    return PA_OK;
}

int check_temperature(double T)
{
    if (!(T > 0)) return fail(PA_ERR_ARG, "T=%g must be a positive temperature in K", T);
    return PA_OK;
}

int upload_per_chain(pa_state *st, DevAlloc &d, const double *v, const char *name)
{
    if (!v) return fail(PA_ERR_ARG, "%s must not be NULL", name);
    std::vector<double> h(v, v + st->nsys);
    for (int n = 0; n < st->nsys; ++n)
        if (!std::isfinite(h[n])) return fail(PA_ERR_ARG, "%s[%d] = %g is not finite", name, n, h[n]);
    HIP_TRY(upload(d, h));
    return PA_OK;
}

}  // namespace

extern "C" int pa_state_create(pa_batch *b, const pa_guest *guest, const int32_t *ncounts, const int32_t *capacities,
                               double T, uint64_t seed, pa_state **out)
{
    if (!out) return fail(PA_ERR_ARG, "out must not be NULL");
    *out = nullptr;
    if (!b || !guest || !ncounts) return fail(PA_ERR_ARG, "NULL argument");
    if (b->host_only) return fail(PA_ERR_ARG, "batch was built with PA_BATCH_HOST_ONLY and has no device tables");
    if (b->from_tables)
        return fail(PA_ERR_ARG, "SystemState needs the force field (tail flag, per-type parameters): build the batch with "
                                "pa_batch_create, not from FrameworkBatch tables");
    bool same = guest->nsites == b->nsites;
    for (int a = 0; same && a < b->nsites; ++a) {
        same = guest->types[a] == b->gtypes[a] && guest->charges[a] == b->gq[a];
        for (int d = 0; d < 3; ++d) same = same && guest->sites[3 * a + d] == b->gsites[3 * a + d];
    }
    if (!same) return fail(PA_ERR_ARG, "guest passed to SystemState does not match the guest FrameworkBatch was built with");
    if (!b->fullk)
        return fail(PA_ERR_ARG,
                    "SystemState needs a FrameworkBatch built with fullk=true: a guest's structure factor is nonzero at "
                    "every k-vector, not only the host-coupled subset the sparse path keeps");
    if (b->nsites > 6) return fail(PA_ERR_ARG, "guests with %d sites: the device kernels take 1..6 sites", b->nsites);
    int st0 = check_temperature(T);
    if (st0) return st0;
    const int nsys = b->nsys, NS = b->nsites;
    std::unique_ptr<pa_state> S(new pa_state());
    S->b = b;
    S->nsys = nsys;
    S->NS = NS;
    S->guest_off.assign(nsys + 1, 0);
    S->sk_off.assign(nsys + 1, 0);
    for (int n = 0; n < nsys; ++n) {
        const int cap = capacities ? capacities[n] : ncounts[n];
        if (ncounts[n] < 0) return fail(PA_ERR_ARG, "ncounts must be non-negative, got %d for system %d", ncounts[n], n + 1);
        if (cap < ncounts[n])
            return fail(PA_ERR_ARG, "capacities must be >= ncounts in every system: system %d has capacity %d < %d", n + 1,
                        cap, ncounts[n]);
        S->guest_off[n + 1] = S->guest_off[n] + cap;
        S->sk_off[n + 1] = S->sk_off[n] + b->info[b->fw_of[n]].nk;
    }
    const int nslots = std::max(1, S->guest_off[nsys]);
    std::vector<double4> pos(nslots, make_double4(0, 0, 0, 0)), quat(nslots, make_double4(0, 0, 0, 1));
    std::vector<double> host_e(nslots, 0.0), energy(nsys, 0.0);
    std::vector<double2> Sk(std::max(1, S->sk_off[nsys]));
    S->seed.resize(nsys);
    S->const_p.resize(nsys);
    S->const_q.resize(nsys);
    const double kT = KB * T;
    // systems are independent: built on the host threads, each keeping its status and message
    auto build_system = [&](int n) -> int {
        const int u = b->fw_of[n];
        const FwDesc &F = b->desc[u];
        S->seed[n] = splitmix64(seed, (uint64_t)(n + 1));
        // initial poses: uniform pose of slot i, attempt t from ChainRNG(splitmix64(seed ^ key, n+1),
        // i+1, t), redrawn until it clears the phase-0 hard-core test at T (src/state.jl:244-289)
        const uint64_t place_key = splitmix64(seed ^ 0x1A17A1D05E5A3E11ULL, (uint64_t)(n + 1));
        std::vector<double> rho2((size_t)NS * b->ntypes_c, 0.0);
        double rmax2 = 0.0;
        if (ncounts[n] > 0) {
            int st = rejection_radii(b, n, kT, DBL_EPSILON, rho2.data(), &rmax2);
            if (st) return st;
        }
        for (int i = 0; i < ncounts[n]; ++i) {
            bool placed = false;
            for (uint64_t t = 1; t <= 10000 && !placed; ++t) {
                Chain c(place_key, (uint64_t)(i + 1), t);
                double f[3] = {unit_f64(c.next()), unit_f64(c.next()), unit_f64(c.next())}, p[3], q[4];
                matvec(F.A, f, p);
                rand_quaternion(c, q);
                double sites[MAX_SITES][3];
                guest_sites_h(b, p, q, sites);
                bool hit = false;
                for (int j = 0; j < F.natoms && !hit; ++j) {
                    const double4 h = b->atom64_h[F.atom_off + j];
                    const double hp[3] = {h.x, h.y, h.z};
                    for (int s = 0; s < NS && !hit; ++s)
                        if (mi_r2_h(F, sites[s], hp) < rho2[(size_t)s * b->ntypes_c + b->atype_h[F.atom_off + j]]) hit = true;
                }
                if (!hit) {
                    pos[S->guest_off[n] + i] = make_double4(p[0], p[1], p[2], 0.0);
                    quat[S->guest_off[n] + i] = make_double4(q[0], q[1], q[2], q[3]);
                    placed = true;
                }
            }
            if (!placed)
                return fail(PA_ERR_ARG, "initial placement did not clear the host after 10000 rounds (system %d, guest %d)",
                            n + 1, i + 1);
        }
        std::vector<double2> sk;
        rebuild_sk(b, u, pos, quat, S->guest_off[n], ncounts[n], sk);
        std::copy(sk.begin(), sk.end(), Sk.begin() + S->sk_off[n]);
        std::vector<double> he;
        int st = total_energy_h(b, u, pos, quat, S->guest_off[n], ncounts[n], sk, &energy[n], &he);
        if (st) return st;
        for (int i = 0; i < ncounts[n]; ++i) host_e[S->guest_off[n] + i] = he[i];
        // exchange constant term p + q N (src/moves.jl:728-773)
        double gse = 0.0;
        st = guest_self_h(b, F.alpha, &gse);
        if (st) return st;
        auto term = [&](double N) {
            return gse + guest_tail_h(b, u, N + 1) - guest_tail_h(b, u, N) + net_h(b, u, N + 1) - net_h(b, u, N);
        };
        S->const_p[n] = term(0.0);
        S->const_q[n] = term(1.0) - S->const_p[n];
        return PA_OK;
    };
    int s_err = run_per_system(nsys, build_system);
    if (s_err) return s_err;
    std::vector<int> occ(ncounts, ncounts + nsys), zeros5((size_t)nsys * NMOVES, 0), zi(nsys, 0);
    std::vector<uint64_t> ctr(nsys, 0);
    HIP_TRY(upload(S->d_guest_off, S->guest_off));
    HIP_TRY(upload(S->d_sk_off, S->sk_off));
    HIP_TRY(upload(S->d_occ, occ));
    HIP_TRY(upload(S->d_refpos, pos));
    HIP_TRY(upload(S->d_quat, quat));
    HIP_TRY(upload(S->d_host_e, host_e));
    HIP_TRY(upload(S->d_Sk, Sk));
    HIP_TRY(S->d_dS.ensure(sizeof(double2) * MC_SPEC * std::max<size_t>(Sk.size(), 1)));
    HIP_TRY(upload(S->d_skabs, std::vector<double>(Sk.size(), 0.0)));
    HIP_TRY(upload(S->d_energy, energy));
    HIP_TRY(upload(S->d_energy_abs, std::vector<double>(nsys, 0.0)));
    HIP_TRY(upload(S->d_seed, S->seed));
    HIP_TRY(upload(S->d_counter, ctr));
    HIP_TRY(upload(S->d_acc, zeros5));
    HIP_TRY(upload(S->d_att, zeros5));
    HIP_TRY(upload(S->d_cp, S->const_p));
    HIP_TRY(upload(S->d_cq, S->const_q));
    HIP_TRY(upload(S->d_maxocc, occ));
    HIP_TRY(upload(S->d_caphit, zi));
    std::vector<double> zd(nsys, 0.0);
    HIP_TRY(upload(S->d_step_t, zd));
    HIP_TRY(upload(S->d_step_r, zd));
    HIP_TRY(upload(S->d_fug, zd));
    S->acc_at_audit.assign(nsys, 0);
    int s_tab = ensure_chain_tables(S.get());
    if (s_tab) return s_tab;
    *out = S.release();
    return ok();
}

extern "C" void pa_state_destroy(pa_state *st) { delete st; }

// Runs `nmoves` attempts of one move type on every chain (mode 0 of the chain kernel).
static int run_moves(pa_state *st, int movetype, int64_t nmoves, const double *step_trans, const double *step_rot,
                     const double *fugacity_pa, double T, double *trace)
{
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    if (int g = refuse_over_capacity(st)) return g;
    if (movetype < MOVE_TRANSLATION || movetype > MOVE_DELETION)
        return fail(PA_ERR_ARG, "movetype=%d must be one of PA_MOVE_TRANSLATION .. PA_MOVE_DELETION", movetype);
    if (nmoves < 0) return fail(PA_ERR_ARG, "nmoves=%lld must be >= 0", (long long)nmoves);
    int s0 = check_temperature(T);
    if (s0) return s0;
    if (movetype == MOVE_TRANSLATION || step_trans) {
        s0 = upload_per_chain(st, st->d_step_t, step_trans, "step_trans");
        if (s0) return s0;
    }
    if (movetype == MOVE_ROTATION || step_rot) {
        s0 = upload_per_chain(st, st->d_step_r, step_rot, "step_rot");
        if (s0) return s0;
    }
    if (movetype == MOVE_INSERTION || movetype == MOVE_DELETION) {
        if (!fugacity_pa) return fail(PA_ERR_ARG, "fugacity must not be NULL");
        std::vector<double> f(st->nsys);
        for (int n = 0; n < st->nsys; ++n) {
            if (!(fugacity_pa[n] > 0)) return fail(PA_ERR_ARG, "fugacity[%d] = %g Pa must be positive", n, fugacity_pa[n]);
            f[n] = fugacity_pa[n] * PASCAL_EV;  // Pa -> eV/Å^3 (src/constants.jl:10)
        }
        HIP_TRY(upload(st->d_fug, f));
    }
    if (nmoves == 0) return ok();
    McArgs a = base_mc(st);
    a.mode = 0;
    a.movetype = movetype;
    a.nmoves = nmoves;
    a.kT = KB * T;
    if (trace) {
        HIP_TRY(st->d_trace.ensure(sizeof(double4) * (size_t)st->nsys * nmoves));
        a.trace = st->d_trace.as<double4>();
        a.trace_cap = nmoves;
    }
    HIP_TRY(launch_mc(a, st->b->stream));
    HIP_TRY(hipStreamSynchronize(st->b->stream));
    if (trace)
        HIP_TRY(hipMemcpy(trace, st->d_trace.p, sizeof(double4) * (size_t)st->nsys * nmoves, hipMemcpyDeviceToHost));
    if (movetype == MOVE_INSERTION) {
        std::vector<int> hits;
        HIP_TRY(download(hits, st->d_caphit, st->nsys));
        for (int n = 0; n < st->nsys; ++n)
            if (hits[n])
                return fail(PA_ERR_ARG,
                            "mc_insert!: system(s) [%d] hit capacity on a move the Metropolis test would otherwise have "
                            "accepted; the chain would sample a truncated distribution -- raise capacity, lower "
                            "fugacity/pressure, or attempt fewer insertions per cycle",
                            n + 1);
    }
    return ok();
}

extern "C" int pa_mc_step(pa_state *st, int32_t movetype, const double *step_trans, const double *step_rot, double T)
{
    if (movetype < MOVE_TRANSLATION || movetype > MOVE_REINSERTION)
        return fail(PA_ERR_ARG, "mc_step!: movetype=%d must be MOVE_TRANSLATION, MOVE_ROTATION or MOVE_REINSERTION",
                    movetype);
    return run_moves(st, movetype, 1, step_trans, step_rot, nullptr, T, nullptr);
}

extern "C" int pa_mc_insert(pa_state *st, const double *fugacity_pa, double T)
{
    return run_moves(st, MOVE_INSERTION, 1, nullptr, nullptr, fugacity_pa, T, nullptr);
}

extern "C" int pa_mc_delete(pa_state *st, const double *fugacity_pa, double T)
{
    return run_moves(st, MOVE_DELETION, 1, nullptr, nullptr, fugacity_pa, T, nullptr);
}

extern "C" int pa_mc_moves(pa_state *st, int32_t movetype, int64_t nmoves, const double *step_trans,
                           const double *step_rot, const double *fugacity_pa, double T, double *trace)
{
    return run_moves(st, movetype, nmoves, step_trans, step_rot, fugacity_pa, T, trace);
}

extern "C" int pa_state_info(const pa_state *st, int32_t sys, pa_chain_info *out)
{
    if (!st || !out) return fail(PA_ERR_ARG, "NULL argument");
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    int occ = 0;
    double e = 0, ea = 0;
    uint64_t c = 0;
    int acc[NMOVES], att[NMOVES];
    HIP_TRY(hipMemcpy(&occ, st->d_occ.as<int>() + sys, sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&e, st->d_energy.as<double>() + sys, sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&ea, st->d_energy_abs.as<double>() + sys, sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&c, st->d_counter.as<uint64_t>() + sys, sizeof(uint64_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(acc, st->d_acc.as<int>() + sys * NMOVES, sizeof acc, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(att, st->d_att.as<int>() + sys * NMOVES, sizeof att, hipMemcpyDeviceToHost));
    out->occupancy = occ;
    out->capacity = st->guest_off[sys + 1] - st->guest_off[sys];
    out->energy = e;
    out->energy_abs_accum = ea;
    out->rng_counter = (int64_t)c;
    for (int t = 0; t < NMOVES; ++t) {
        out->accepted[t] = acc[t];
        out->attempted[t] = att[t];
    }
    out->const_p = st->const_p[sys];
    out->const_q = st->const_q[sys];
    return ok();
}

extern "C" int pa_state_guests(const pa_state *st, int32_t sys, double *refpoints, double *orientations,
                               double *host_energy)
{
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    if (int g = refuse_over_capacity(st)) return g;
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    int occ = 0;
    HIP_TRY(hipMemcpy(&occ, st->d_occ.as<int>() + sys, sizeof(int), hipMemcpyDeviceToHost));
    if (occ == 0) return ok();
    const int g0 = st->guest_off[sys];
    std::vector<double4> p(occ), q(occ);
    std::vector<double> he(occ);
    HIP_TRY(hipMemcpy(p.data(), st->d_refpos.as<double4>() + g0, sizeof(double4) * occ, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(q.data(), st->d_quat.as<double4>() + g0, sizeof(double4) * occ, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(he.data(), st->d_host_e.as<double>() + g0, sizeof(double) * occ, hipMemcpyDeviceToHost));
    for (int i = 0; i < occ; ++i) {
        if (refpoints) { refpoints[3 * i] = p[i].x; refpoints[3 * i + 1] = p[i].y; refpoints[3 * i + 2] = p[i].z; }
        if (orientations) {
            orientations[4 * i] = q[i].x; orientations[4 * i + 1] = q[i].y;
            orientations[4 * i + 2] = q[i].z; orientations[4 * i + 3] = q[i].w;
        }
        if (host_energy) host_energy[i] = he[i];
    }
    return ok();
}

extern "C" int pa_state_sk(const pa_state *st, int32_t sys, double *sk)
{
    if (!st || !sk) return fail(PA_ERR_ARG, "NULL argument");
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    const int nk = st->sk_off[sys + 1] - st->sk_off[sys];
    if (nk == 0) return ok();
    HIP_TRY(hipMemcpy(sk, st->d_Sk.as<double2>() + st->sk_off[sys], sizeof(double2) * nk, hipMemcpyDeviceToHost));
    return ok();
}

// Downloads the chains, rebuilds every system's structure factor from its poses and recomputes
// its total energy (FP64).
static int state_snapshot(pa_state *st, std::vector<double4> &pos, std::vector<double4> &quat, std::vector<int> &occ,
                          std::vector<std::vector<double2>> &sk_rebuilt, std::vector<double> &E)
{
    const int nslots = std::max(1, st->guest_off[st->nsys]);
    HIP_TRY(download(pos, st->d_refpos, nslots));
    HIP_TRY(download(quat, st->d_quat, nslots));
    HIP_TRY(download(occ, st->d_occ, st->nsys));
    sk_rebuilt.assign(st->nsys, {});
    E.assign(st->nsys, 0.0);
    return run_per_system(st->nsys, [&](int n) {
        const int u = st->b->fw_of[n];
        rebuild_sk(st->b, u, pos, quat, st->guest_off[n], occ[n], sk_rebuilt[n]);
        return total_energy_h(st->b, u, pos, quat, st->guest_off[n], occ[n], sk_rebuilt[n], &E[n], nullptr);
    });
}

extern "C" int pa_total_energy(pa_state *st, double *E)
{
    if (!st || !E) return fail(PA_ERR_ARG, "NULL argument");
    if (int g = refuse_over_capacity(st)) return g;
    std::vector<double4> pos, quat;
    std::vector<int> occ;
    std::vector<std::vector<double2>> sk;
    std::vector<double> e;
    int s = state_snapshot(st, pos, quat, occ, sk, e);
    if (s) return s;
    for (int n = 0; n < st->nsys; ++n) E[n] = e[n];
    return ok();
}

// audit_energy! (src/audit.jl:97-165) on every system: occupancy within capacity, the running
// structure factor against one rebuilt from the poses, then the running energy against a
// from-scratch total; on success both are reset to the rebuilt values and the rounding
// accumulators cleared. The rebuild runs on the device (audit_kernels.hip); pa_total_energy
// keeps the serial host rebuild, which the tests hold it to.
extern "C" int pa_audit_energy(pa_state *st, double *discrepancy)
{
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    pa_batch *b = st->b;
    const int nsys = st->nsys, NS = b->nsites;
    std::vector<int> occ, acc;
    HIP_TRY(download(occ, st->d_occ, nsys));
    // occupancy first: the rebuild reads every live guest's slot
    for (int n = 0; n < nsys; ++n) {
        const int cap = st->guest_off[n + 1] - st->guest_off[n];
        if (occ[n] > cap)
            return fail(PA_ERR_ARG, "energy audit failed for system %d (Float64): occupancy %d exceeds capacity %d", n + 1,
                        occ[n], cap);
    }
    const int nslots = st->guest_off[nsys], nsk = st->sk_off[nsys];
    int tw = 1;
    for (int n = 0; n < nsys; ++n) {
        const FwDesc &F = b->desc[b->fw_of[n]];
        tw = std::max(tw, F.kn_max[0] + F.kn_max[1] + F.kn_max[2] + 3);
    }
    HIP_TRY(st->d_aud_tab.ensure(sizeof(double2) * std::max<size_t>(1, (size_t)nslots * NS * tw)));
    HIP_TRY(st->d_aud_sk.ensure(sizeof(double2) * std::max(1, nsk)));
    HIP_TRY(st->d_aud_he.ensure(sizeof(double) * std::max(1, nslots)));
    HIP_TRY(st->d_aud_sums.ensure(sizeof(double) * 3 * nsys));
    AuditArgs A{};
    A.fw = b->d_fw.as<FwDesc>();
    A.fw_of = b->d_fw_of.as<int>();
    A.nsys = nsys;
    A.atom64 = b->d_atom64.as<double4>();
    A.atype = b->d_atype.as<int>();
    A.k64 = b->d_k64.as<double4>();
    A.S64 = b->d_S64.as<double2>();
    A.kn = b->d_kn.as<int4>();
    A.pair64 = b->d_pair64.as<double2>();
    A.ntypes_c = b->ntypes_c;
    A.rc_lj2 = b->rc_lj * b->rc_lj;
    A.rc_ew2 = b->gdev.charged ? b->rc_ew * b->rc_ew : 0.0;
    A.guest = b->gdev;
    A.guest_off = st->d_guest_off.as<int>();
    A.sk_off = st->d_sk_off.as<int>();
    A.occ = st->d_occ.as<int>();
    A.pos = st->d_refpos.as<double4>();
    A.quat = st->d_quat.as<double4>();
    A.nslots = nslots;
    A.nsk = nsk;
    A.tw = tw;
    A.tab = st->d_aud_tab.as<double2>();
    A.sk = st->d_aud_sk.as<double2>();
    A.host_e = st->d_aud_he.as<double>();
    A.sums = st->d_aud_sums.as<double>();
    HIP_TRY(launch_audit(A, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));

    std::vector<double> sums, Erun, Eabs, skabs;
    std::vector<double2> skrun, skr;
    HIP_TRY(download(sums, st->d_aud_sums, (size_t)3 * nsys));
    HIP_TRY(download(skr, st->d_aud_sk, std::max(1, nsk)));
    HIP_TRY(download(Erun, st->d_energy, nsys));
    HIP_TRY(download(Eabs, st->d_energy_abs, nsys));
    HIP_TRY(download(skrun, st->d_Sk, std::max(1, nsk)));
    HIP_TRY(download(skabs, st->d_skabs, std::max(1, nsk)));
    HIP_TRY(download(acc, st->d_acc, (size_t)nsys * NMOVES));
    const double eps = DBL_EPSILON;
    auto tol = [&](double abs_accum, double mag, double nm) {
        return nm * eps * std::max(abs_accum, 1.0) + eps * std::max(mag, 1.0);
    };
    double qabs = 0.0;
    for (int a = 0; a < NS; ++a) qabs += std::fabs(b->gq[a]);
    for (int n = 0; n < nsys; ++n) {
        int nacc = 0;
        for (int t = 0; t < NMOVES; ++t) nacc += acc[n * NMOVES + t];
        const double nmoves = std::max(1, nacc - st->acc_at_audit[n]);
        const int u = b->fw_of[n];
        const FwDesc &F = b->desc[u];
        const double nterms = (double)occ[n] * NS;
        for (int k = 0; k < F.nk; ++k) {
            const int j = st->sk_off[n] + k;
            const double2 r = skrun[j], w = skr[j], h = b->S64_h[F.k_off + k];
            const double mag = nterms * (std::hypot(h.x, h.y) + occ[n] * qabs);
            const double d = std::hypot(r.x - w.x, r.y - w.y);
            const double tk = tol(skabs[j], mag, nmoves);
            if (!(d <= tk))
                return fail(PA_ERR_ARG,
                            "energy audit failed for system %d (Float64): structure factor at k-vector %d disagrees with "
                            "the poses: running=%.17g %c %.17gim, rebuilt=%.17g %c %.17gim, discrepancy=%.17g exceeds "
                            "tolerance %.17g over %.0f accepted move(s) since the last audit",
                            n + 1, k + 1, r.x, r.y < 0 ? '-' : '+', std::fabs(r.y), w.x, w.y < 0 ? '-' : '+',
                            std::fabs(w.y), d, tk, nmoves);
        }
        // total_energy (src/guest.jl:293-343) from the device sums: host-guest, guest-guest,
        // reciprocal of the rebuilt structure factor, then the per-guest self terms, the tail
        // correction and the net-charge term
        double gse = 0.0;
        if (int e = guest_self_h(b, F.alpha, &gse)) return e;
        const double E = sums[3 * n] + sums[3 * n + 1] + KE * sums[3 * n + 2] + occ[n] * gse + guest_tail_h(b, u, occ[n]) +
                         net_h(b, u, occ[n]);
        const double nte = occ[n] + 0.5 * occ[n] * (occ[n] - 1);
        const double mag = std::max({std::fabs(Erun[n]), std::fabs(E), nte * b->hc_bs[u]});
        const double d = std::fabs(E - Erun[n]);
        if (discrepancy) discrepancy[n] = E - Erun[n];
        const double te = tol(Eabs[n], std::isfinite(mag) ? mag : std::max(std::fabs(Erun[n]), std::fabs(E)), nmoves);
        if (!(d <= te))
            return fail(PA_ERR_ARG,
                        "energy audit failed for system %d (Float64): accumulated=%.17g, recomputed=%.17g, "
                        "discrepancy=%.17g exceeds tolerance %.17g over %.0f accepted move(s) since the last audit",
                        n + 1, Erun[n], E, d, te, nmoves);
        std::copy(skr.begin() + st->sk_off[n], skr.begin() + st->sk_off[n + 1], skrun.begin() + st->sk_off[n]);
        std::fill(skabs.begin() + st->sk_off[n], skabs.begin() + st->sk_off[n + 1], 0.0);
        Erun[n] = E;
        Eabs[n] = 0.0;
        st->acc_at_audit[n] = nacc;
    }
    HIP_TRY(upload(st->d_Sk, skrun));
    HIP_TRY(upload(st->d_skabs, skabs));
    HIP_TRY(upload(st->d_energy, Erun));
    HIP_TRY(upload(st->d_energy_abs, Eabs));
    return ok();
}

// Test-only overwrite of one element of a system's running state (the audit's fault-injection
// tests, after test/audit_tests.jl:60-415).
extern "C" int pa_state_set_value(pa_state *st, int32_t field, int32_t sys, int64_t index, double re, double im)
{
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    const int cap = st->guest_off[sys + 1] - st->guest_off[sys];
    switch (field) {
    case PA_FIELD_ENERGY:
        HIP_TRY(hipMemcpy(st->d_energy.as<double>() + sys, &re, sizeof(double), hipMemcpyHostToDevice));
        return ok();
    case PA_FIELD_SK: {
        const int nk = st->sk_off[sys + 1] - st->sk_off[sys];
        if (index < 0 || index >= nk) return fail(PA_ERR_ARG, "k-vector index %lld out of range 0..%d", (long long)index, nk - 1);
        const double2 v = make_double2(re, im);
        HIP_TRY(hipMemcpy(st->d_Sk.as<double2>() + st->sk_off[sys] + index, &v, sizeof(double2), hipMemcpyHostToDevice));
        return ok();
    }
    case PA_FIELD_HOST_ENERGY: {
        int occ = 0;
        HIP_TRY(hipMemcpy(&occ, st->d_occ.as<int>() + sys, sizeof(int), hipMemcpyDeviceToHost));
        if (index < 0 || index >= std::min(occ, cap))
            return fail(PA_ERR_ARG, "guest index %lld out of range 0..%d", (long long)index, std::min(occ, cap) - 1);
        HIP_TRY(hipMemcpy(st->d_host_e.as<double>() + st->guest_off[sys] + index, &re, sizeof(double),
                          hipMemcpyHostToDevice));
        return ok();
    }
    case PA_FIELD_OCCUPANCY: {
        if (!(re >= 0.0 && re <= 2147483647.0 && re == std::floor(re)))
            return fail(PA_ERR_ARG, "occupancy %g must be a non-negative integer", re);
        std::vector<int> occ;
        HIP_TRY(download(occ, st->d_occ, st->nsys));
        // live guests: the occupancy itself, or what it was before it was set over a capacity
        if (!st->occ_over_capacity) st->occ_live = occ;
        const int v = (int)re, live = st->occ_live[sys];
        if (v > live && v <= cap)
            return fail(PA_ERR_ARG, "occupancy %d would expose unset guest slots (live guests %d)", v, live);
        occ[sys] = v;
        if (v <= cap) st->occ_live[sys] = v;
        HIP_TRY(upload(st->d_occ, occ));
        bool over = false;
        for (int n = 0; n < st->nsys; ++n) over = over || occ[n] > st->guest_off[n + 1] - st->guest_off[n];
        st->occ_over_capacity = over;
        return ok();
    }
    default:
        return fail(PA_ERR_ARG, "field=%d must be PA_FIELD_ENERGY, PA_FIELD_SK, PA_FIELD_HOST_ENERGY or PA_FIELD_OCCUPANCY",
                    field);
    }
}

// Zeroes the move counters at the start of production. The audit's "accepted moves since the
// last audit" (which sizes its rounding tolerance) keeps counting the moves accepted before the
// reset: the running energy and structure factor carry their rounding until the next audit.
static int reset_move_counters(pa_state *st)
{
    const int nsys = st->nsys;
    std::vector<int> acc;
    HIP_TRY(download(acc, st->d_acc, (size_t)nsys * NMOVES));
    for (int n = 0; n < nsys; ++n) {
        int nacc = 0;
        for (int t = 0; t < NMOVES; ++t) nacc += acc[n * NMOVES + t];
        st->acc_at_audit[n] -= nacc;  // pending = nacc - acc_at_audit carries over as -acc_at_audit
    }
    std::vector<int> z((size_t)nsys * NMOVES, 0);
    HIP_TRY(upload(st->d_acc, z));
    HIP_TRY(upload(st->d_att, z));
    return PA_OK;
}

// block_mean_sem (src/nvt.jl:126-140).
static void block_mean_sem(const double *x, int64_t ns, int nb, double *mean, double *sem)
{
    const int64_t bl = ns / nb;
    std::vector<double> m(nb, 0.0);
    for (int b = 0; b < nb; ++b) {
        const int64_t lo = b * bl, hi = (b == nb - 1) ? ns : (b + 1) * bl;
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += x[i];
        m[b] = s / (double)(hi - lo);
    }
    double mu = 0.0;
    for (double v : m) mu += v;
    mu /= nb;
    double var = 0.0;
    for (double v : m) var += (v - mu) * (v - mu);
    var /= (nb - 1);
    *mean = mu;
    *sem = std::sqrt(var / nb);
}

// fluctuation_qst (src/gcmc.jl:71-113): pooled kT - cov(U, N)/var(N), delete-one-block jackknife
// error, pooled corr(U, N).
static void fluctuation_qst(const double *U, const double *N, int64_t ns, double kT, int nb, double *q, double *qerr,
                            double *corr)
{
    const int64_t bl = ns / nb;
    std::vector<double> sU(nb, 0), sN(nb, 0), sU2(nb, 0), sN2(nb, 0), sUN(nb, 0), cnt(nb, 0);
    for (int b = 0; b < nb; ++b) {
        const int64_t lo = b * bl, hi = (b == nb - 1) ? ns : (b + 1) * bl;
        for (int64_t i = lo; i < hi; ++i) {
            sU[b] += U[i]; sN[b] += N[i]; sU2[b] += U[i] * U[i]; sN2[b] += N[i] * N[i]; sUN[b] += U[i] * N[i];
        }
        cnt[b] = (double)(hi - lo);
    }
    double tU = 0, tN = 0, tU2 = 0, tN2 = 0, tUN = 0;
    for (int b = 0; b < nb; ++b) { tU += sU[b]; tN += sN[b]; tU2 += sU2[b]; tN2 += sN2[b]; tUN += sUN[b]; }
    auto qof = [&](double su, double sn, double sun, double sn2, double n) {
        const double mU = su / n, mN = sn / n;
        const double cov = sun / n - mU * mN, varN = sn2 / n - mN * mN;
        return kT - cov / varN;
    };
    *q = qof(tU, tN, tUN, tN2, (double)ns);
    std::vector<double> loo(nb);
    double bar = 0.0;
    for (int b = 0; b < nb; ++b) {
        loo[b] = qof(tU - sU[b], tN - sN[b], tUN - sUN[b], tN2 - sN2[b], (double)ns - cnt[b]);
        bar += loo[b];
    }
    bar /= nb;
    double v = 0.0;
    for (double x : loo) v += (x - bar) * (x - bar);
    *qerr = std::sqrt((double)(nb - 1) / nb * v);
    const double mU = tU / ns, mN = tN / ns;
    const double varU = tU2 / ns - mU * mU, varN = tN2 / ns - mN * mN;
    *corr = (tUN / ns - mU * mN) / std::sqrt(varU * varN);
}

extern "C" int pa_gcmc_run(pa_state *st, const pa_gcmc_params *p, pa_gcmc_result *out)
{
    if (st) {
        if (int g = refuse_over_capacity(st)) return g;
    }
    if (!st || !p || !out) return fail(PA_ERR_ARG, "NULL argument");
    const int nsys = st->nsys;
    if (p->n_warmup < 0) return fail(PA_ERR_ARG, "n_warmup=%lld must be >= 0", (long long)p->n_warmup);
    if (p->n_production < 1) return fail(PA_ERR_ARG, "n_production=%lld must be >= 1", (long long)p->n_production);
    if (p->n_audit < 1) return fail(PA_ERR_ARG, "n_audit=%lld must be >= 1", (long long)p->n_audit);
    if (p->nblocks < 2) return fail(PA_ERR_ARG, "nblocks=%d: at least two blocks are needed for a standard error", p->nblocks);
    if (p->n_production < 2 * (int64_t)p->nblocks)
        return fail(PA_ERR_ARG, "%lld cycles: need at least 2*nblocks=%d production cycles", (long long)p->n_production,
                    2 * p->nblocks);
    const double pr[NMOVES] = {p->p_translation, p->p_rotation, p->p_reinsertion, 0.5 * p->p_exchange,
                               0.5 * p->p_exchange};
    double tot = 0.0;
    for (double x : pr) {
        if (!(x >= 0.0)) return fail(PA_ERR_ARG, "move probabilities must be >= 0");
        tot += x;
    }
    if (!(tot > 0.0)) return fail(PA_ERR_ARG, "move probabilities sum to zero");
    int s0 = check_temperature(p->T);
    if (s0) return s0;
    s0 = upload_per_chain(st, st->d_step_t, p->step_trans, "step_trans");
    if (s0) return s0;
    s0 = upload_per_chain(st, st->d_step_r, p->step_rot, "step_rot");
    if (s0) return s0;
    if (p->p_exchange > 0) {
        if (!p->fugacity_pa) return fail(PA_ERR_ARG, "fugacity must not be NULL");
        std::vector<double> f(nsys);
        for (int n = 0; n < nsys; ++n) {
            if (!(p->fugacity_pa[n] > 0)) return fail(PA_ERR_ARG, "fugacity[%d] = %g Pa must be positive", n, p->fugacity_pa[n]);
            f[n] = p->fugacity_pa[n] * PASCAL_EV;
        }
        HIP_TRY(upload(st->d_fug, f));
    }
    McArgs a = base_mc(st);
    a.mode = 1;
    a.kT = KB * p->T;
    double c = 0.0;
    for (int t = 0; t < NMOVES; ++t) {
        c += pr[t] / tot;
        a.pcum[t] = c;
    }
    a.pcum[NMOVES - 1] = 2.0;  // every draw in [0, 1) lands somewhere
    a.min_cycle = std::max<int64_t>(1, p->min_cycle_length);
    const int64_t NP = p->n_production;
    HIP_TRY(st->d_samp_e.ensure(sizeof(double) * nsys * NP));
    HIP_TRY(st->d_samp_n.ensure(sizeof(double) * nsys * NP));
    a.samp_e = st->d_samp_e.as<double>();
    a.samp_n = st->d_samp_n.as<double>();
    a.samp_stride = NP;
    st->samp_stride = NP;
    std::vector<int> occ;
    HIP_TRY(download(occ, st->d_occ, nsys));
    HIP_TRY(upload(st->d_maxocc, occ));
    std::vector<int> caphit;
    int64_t since_audit = 0;
    // warmup, then production, in launches of at most n_audit cycles; the audit runs every
    // n_audit cycles (warmup and production share one cycle count)
    for (int phase = 0; phase < 2; ++phase) {
        const int64_t ncyc = phase == 0 ? p->n_warmup : NP;
        if (phase == 1) {
            int s = reset_move_counters(st);
            if (s) return s;
        }
        for (int64_t c0 = 0; c0 < ncyc;) {
            const int64_t len = std::min(ncyc - c0, p->n_audit - since_audit);
            a.ncycles = len;
            a.sampling = phase == 1;
            a.samp_base = c0;
            HIP_TRY(launch_mc(a, st->b->stream));
            HIP_TRY(hipStreamSynchronize(st->b->stream));
            HIP_TRY(download(caphit, st->d_caphit, nsys));
            for (int n = 0; n < nsys; ++n)
                if (caphit[n])
                    return fail(PA_ERR_ARG,
                                "mc_insert!: system(s) [%d] hit capacity on a move the Metropolis test would otherwise "
                                "have accepted; the chain would sample a truncated distribution -- raise capacity, lower "
                                "fugacity/pressure, or attempt fewer insertions per cycle",
                                n + 1);
            c0 += len;
            since_audit += len;
            if (since_audit >= p->n_audit) {
                int s = pa_audit_energy(st, nullptr);
                if (s) return s;
                since_audit = 0;
            }
        }
    }
    std::vector<double> se, sn;
    std::vector<int> mo, acc, att;
    HIP_TRY(download(se, st->d_samp_e, (size_t)nsys * NP));
    HIP_TRY(download(sn, st->d_samp_n, (size_t)nsys * NP));
    HIP_TRY(download(mo, st->d_maxocc, nsys));
    HIP_TRY(download(acc, st->d_acc, (size_t)nsys * NMOVES));
    HIP_TRY(download(att, st->d_att, (size_t)nsys * NMOVES));
    for (int n = 0; n < nsys; ++n) {
        pa_gcmc_result &r = out[n];
        block_mean_sem(&sn[(size_t)n * NP], NP, p->nblocks, &r.loading, &r.loading_err);
        block_mean_sem(&se[(size_t)n * NP], NP, p->nblocks, &r.energy, &r.energy_err);
        fluctuation_qst(&se[(size_t)n * NP], &sn[(size_t)n * NP], NP, KB * p->T, p->nblocks, &r.q_st, &r.q_st_err,
                        &r.corr_UN);
        r.max_occupancy = mo[n];
        r.capacity = st->guest_off[n + 1] - st->guest_off[n];
        r.ncycles = NP;
        for (int t = 0; t < NMOVES; ++t) {
            r.accepted[t] = acc[n * NMOVES + t];
            r.attempted[t] = att[n * NMOVES + t];
        }
    }
    return ok();
}

extern "C" int pa_gcmc_samples(const pa_state *st, int32_t sys, int64_t n, double *energy, double *loading)
{
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    const int64_t have = st->samp_stride;
    if (n < 0 || n > have)
        return fail(PA_ERR_ARG, "n=%lld production samples requested, %lld recorded", (long long)n, (long long)have);
    if (energy)
        HIP_TRY(hipMemcpy(energy, st->d_samp_e.as<double>() + (size_t)sys * have, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (loading)
        HIP_TRY(hipMemcpy(loading, st->d_samp_n.as<double>() + (size_t)sys * have, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ok();
}

// ---- NVT driver with Widom insertion along the chain (src/nvt.jl) ---------------------------

namespace {

bool same_guest(const pa_batch *b, const pa_guest *g)
{
    if (g->nsites != b->nsites) return false;
    for (int a = 0; a < b->nsites; ++a) {
        if (g->types[a] != b->gtypes[a] || g->charges[a] != b->gq[a]) return false;
        for (int d = 0; d < 3; ++d)
            if (g->sites[3 * a + d] != b->gsites[3 * a + d]) return false;
    }
    return true;
}

// insertion_constant_term (src/nvt.jl:44-66): constant_offset at Ng = 0, else constant_offset
// with its own 0 -> 1 tail and net-charge terms replaced by the Ng -> Ng+1 ones.
double insertion_constant_term(const pa_batch *b, int u, int Ng)
{
    const double c0 = b->info[u].const_off;
    if (Ng == 0) return c0;
    const double base = c0 - guest_tail_h(b, u, 1.0) - net_h(b, u, 1.0);
    const double N = (double)Ng;
    return base + guest_tail_h(b, u, N + 1) - guest_tail_h(b, u, N) + net_h(b, u, N + 1) - net_h(b, u, N);
}

// Spreads the low 10 bits of v to every third bit (Morton interleave).
uint32_t spread3(uint32_t v)
{
    v &= 0x3FF;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

// Tables of the chain-insertion kernel, built once per state: every framework's atoms in
// Morton order of a ~4 Å cell grid, cut into chunks of 64 with a centre and a radius that
// bounds every atom's nearest-image distance to the centre; the k-vectors cut into runs of
// consecutive n3 at fixed (n1, n2), longest first; half the smallest perpendicular width.
int ensure_chain_tables(pa_state *st)
{
    pa_batch *b = st->b;
    if (st->d_atomf64.p) return PA_OK;
    const size_t na = std::max<size_t>(b->atom64_h.size(), 1);
    std::vector<double4> af(na, make_double4(0, 0, 0, 0));
    std::vector<int> at(na, 0), chunk_off(1, 0), round_off(1, 0), perm(std::max<size_t>(b->k64_h.size(), 1), 0),
        fsize(b->nfw, 0);
    std::vector<double4> chunks;
    std::vector<int2> lines, rounds;
    std::vector<double> halfw(b->nfw);
    for (int u = 0; u < b->nfw; ++u) {
        const FwDesc &F = b->desc[u];
        const FwHost &h = b->info[u];
        const std::vector<double> &fr = b->fgh[u].frac;
        const int n = F.natoms;
        if ((int)fr.size() != 3 * n) return fail(PA_ERR_INTERNAL, "framework %d: fractional table missing", u + 1);
        int nc[3];
        for (int d = 0; d < 3; ++d) nc[d] = std::max(1, std::min(1024, (int)std::floor(h.perp[d] / 4.0)));
        std::vector<uint64_t> key(n);
        std::vector<double> f(3 * (size_t)n);
        for (int j = 0; j < n; ++j) {
            uint32_t c[3];
            for (int d = 0; d < 3; ++d) {
                const double x = fr[3 * j + d] - std::floor(fr[3 * j + d]);
                f[3 * j + d] = x;
                c[d] = (uint32_t)std::min(nc[d] - 1, (int)(x * nc[d]));
            }
            const uint64_t m = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
            key[j] = (m << 32) | (uint32_t)j;
        }
        std::sort(key.begin(), key.end());
        for (int i = 0; i < n; ++i) {
            const int j = (int)(key[i] & 0xFFFFFFFFu);
            af[F.atom_off + i] = make_double4(f[3 * j], f[3 * j + 1], f[3 * j + 2], b->atom64_h[F.atom_off + j].w);
            at[F.atom_off + i] = b->atype_h[F.atom_off + j];
        }
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int c1 = std::min(n, c0 + 64);
            const double4 r0 = af[F.atom_off + c0];
            std::vector<double> uw(3 * (size_t)(c1 - c0));
            double cen[3] = {0, 0, 0};
            for (int i = c0; i < c1; ++i) {
                const double4 p = af[F.atom_off + i];
                const double d[3] = {p.x - r0.x, p.y - r0.y, p.z - r0.z};
                for (int e = 0; e < 3; ++e) {
                    uw[3 * (i - c0) + e] = (e == 0 ? r0.x : e == 1 ? r0.y : r0.z) + (d[e] - std::rint(d[e]));
                    cen[e] += uw[3 * (i - c0) + e];
                }
            }
            for (int e = 0; e < 3; ++e) cen[e] /= (double)(c1 - c0);
            double rad = 0.0;
            for (int i = 0; i < c1 - c0; ++i) {
                const double d[3] = {uw[3 * i] - cen[0], uw[3 * i + 1] - cen[1], uw[3 * i + 2] - cen[2]};
                double m[3];
                matvec(F.A, d, m);
                rad = std::max(rad, std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]));
            }
            chunks.push_back(make_double4(cen[0], cen[1], cen[2], rad * (1.0 + 1e-12) + 1e-9));
        }
        chunk_off.push_back((int)chunks.size());
        // runs of consecutive n3 at fixed (n1, n2), in k-table order, then longest first
        std::vector<int2> fl;
        for (int k = 0; k < F.nk; ++k) {
            const int4 nn = b->kn_h[F.k_off + k];
            if (k > 0) {
                const int4 pn = b->kn_h[F.k_off + k - 1];
                if (nn.x == pn.x && nn.y == pn.y && nn.z == pn.z + 1) {
                    ++fl.back().y;
                    continue;
                }
            }
            fl.push_back(make_int2(k, 1));
        }
        std::stable_sort(fl.begin(), fl.end(), [](const int2 &x, const int2 &y) { return x.y > y.y; });
        // rounds of 64 runs (one per lane); step i of lane l in a round sits at base + 64 i + l
        int base = 0;
        for (size_t r0 = 0; r0 < fl.size(); r0 += 64) {
            rounds.push_back(make_int2(base, 0));
            const int len0 = fl[r0].y;  // longest run of the round
            for (int l = 0; l < 64; ++l) {
                const size_t r = r0 + l;
                if (r < fl.size()) {
                    lines.push_back(fl[r]);
                    for (int i = 0; i < fl[r].y; ++i) perm[F.k_off + fl[r].x + i] = base + 64 * i + l;
                } else {
                    lines.push_back(make_int2(0, 0));
                }
            }
            base += 64 * len0;
        }
        round_off.push_back((int)rounds.size());
        fsize[u] = base;
        halfw[u] = 0.5 * std::min({h.perp[0], h.perp[1], h.perp[2]});
    }
    const double rcmax = std::max(b->rc_lj, b->gdev.charged ? b->rc_ew : 0.0);
    st->cull_extra = rcmax + b->r_guest + 1e-6;
    HIP_TRY(upload(st->d_atype_s, at));
    HIP_TRY(upload(st->d_chunk, chunks));
    HIP_TRY(upload(st->d_chunk_off, chunk_off));
    HIP_TRY(upload(st->d_lines, lines));
    HIP_TRY(upload(st->d_rounds, rounds));
    HIP_TRY(upload(st->d_round_off, round_off));
    HIP_TRY(upload(st->d_perm, perm));
    std::vector<int> coff(st->nsys + 1, 0);
    for (int n = 0; n < st->nsys; ++n) coff[n + 1] = coff[n] + fsize[b->fw_of[n]];
    st->ctot = coff[st->nsys];
    HIP_TRY(upload(st->d_coff, coff));
    HIP_TRY(upload(st->d_half_perp, halfw));
    HIP_TRY(upload(st->d_atomf64, af));
    return PA_OK;
}

ChainSnap chain_snap(pa_state *st)
{
    ChainSnap S{};
    S.nsys = st->nsys;
    S.nslots = st->guest_off[st->nsys];
    S.ctot = st->ctot;
    S.coff = st->d_coff.as<int>();
    S.perm = st->d_perm.as<int>();
    S.occ = st->d_snap_occ.as<int>();
    S.pos = st->d_snap_pos.as<double4>();
    S.quat = st->d_snap_quat.as<double4>();
    S.csk = st->d_snap_sk.as<double2>();
    return S;
}

ChainWidomArgs base_chain_widom(pa_state *st)
{
    pa_batch *b = st->b;
    ChainWidomArgs a{};
    a.fw = b->d_fw.as<FwDesc>();
    a.fw_of = b->d_fw_of.as<int>();
    a.atomf64 = st->d_atomf64.as<double4>();
    a.atype = st->d_atype_s.as<int>();
    a.chunk = st->d_chunk.as<double4>();
    a.chunk_off = st->d_chunk_off.as<int>();
    a.lines = st->d_lines.as<int2>();
    a.rounds = st->d_rounds.as<int2>();
    a.round_off = st->d_round_off.as<int>();
    a.half_perp = st->d_half_perp.as<double>();
    a.cull_extra = st->cull_extra;
    a.upper = b->upper_all ? 1 : 0;
    a.k64 = b->d_k64.as<double4>();
    a.kn = b->d_kn.as<int4>();
    a.pair64 = b->d_pair64.as<double2>();
    a.ntypes_c = b->ntypes_c;
    a.rc_lj2 = b->rc_lj * b->rc_lj;
    a.rc_ew2 = b->gdev.charged ? b->rc_ew * b->rc_ew : 0.0;
    a.rc_max2 = std::max(a.rc_lj2, a.rc_ew2);
    a.guest = b->gdev;
    a.nsys = st->nsys;
    a.guest_off = st->d_guest_off.as<int>();
    a.snap = chain_snap(st);
    return a;
}

// Makes room for `slots` snapshot slots.
int ensure_snapshots(pa_state *st, int64_t slots)
{
    if (slots <= st->snap_slots) return PA_OK;
    const size_t ns = std::max(1, st->guest_off[st->nsys]);
    HIP_TRY(st->d_snap_pos.ensure(sizeof(double4) * ns * slots));
    HIP_TRY(st->d_snap_quat.ensure(sizeof(double4) * ns * slots));
    HIP_TRY(st->d_snap_sk.ensure(sizeof(double2) * std::max<int64_t>(1, st->ctot) * slots));
    HIP_TRY(st->d_snap_occ.ensure(sizeof(int) * st->nsys * slots));
    st->snap_slots = slots;
    return PA_OK;
}

// Snapshot slot `slot` of the live state.
int snapshot_live(pa_state *st, int64_t slot)
{
    HIP_TRY(launch_chain_snapshot(chain_snap(st), slot, base_mc(st), st->b->stream));
    return PA_OK;
}

// One chain-insertion launch over the insertions item_off describes (one wavefront each).
int launch_chain(pa_state *st, ChainWidomArgs &a, const std::vector<int64_t> &item_off)
{
    const int64_t nitems = item_off[st->nsys];
    if (nitems > INT32_MAX) return fail(PA_ERR_ARG, "%lld test insertions in one launch exceed 2^31", (long long)nitems);
    HIP_TRY(upload(st->d_item_off, item_off));
    HIP_TRY(st->d_cw_dU.ensure(sizeof(double) * std::max<int64_t>(1, nitems)));
    a.item_off = st->d_item_off.as<int64_t>();
    a.nitems = nitems;
    a.dU = st->d_cw_dU.as<double>();
    HIP_TRY(launch_chain_widom(a, st->b->stream));
    return PA_OK;
}

// Runs `launch` between two events on the library's stream, waits for it and adds its duration.
template <class F>
int timed(pa_state *st, double *ms, int64_t *count, F launch)
{
    hipStream_t s = st->b->stream;
    for (auto &e : st->tev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(st->tev[0], s));
    int r = launch();
    if (r) return r;
    HIP_TRY(hipEventRecord(st->tev[1], s));
    HIP_TRY(hipEventSynchronize(st->tev[1]));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, st->tev[0], st->tev[1]));
    *ms += t;
    ++*count;
    return PA_OK;
}

// Insertion constant of every system at its current occupancy.
int upload_const_terms(pa_state *st, std::vector<double> &ct)
{
    std::vector<int> occ;
    HIP_TRY(download(occ, st->d_occ, st->nsys));
    ct.resize(st->nsys);
    for (int n = 0; n < st->nsys; ++n) ct[n] = insertion_constant_term(st->b, st->b->fw_of[n], occ[n]);
    HIP_TRY(upload(st->d_ct, ct));
    return PA_OK;
}

}  // namespace

extern "C" int pa_widom_chain_energies(pa_state *st, int32_t sys, int64_t n, const double *pos, const double *quat,
                                       double *dU)
{
    if (st) {
        if (int g = refuse_over_capacity(st)) return g;
    }
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    if (n < 0) return fail(PA_ERR_ARG, "n=%lld must be >= 0", (long long)n);
    if (n == 0) return ok();
    if (!pos || !quat || !dU) return fail(PA_ERR_ARG, "NULL argument");
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pos[i])) return fail(PA_ERR_ARG, "pos[%lld] = %g is not finite", (long long)i, pos[i]);
    for (int64_t i = 0; i < 4 * n; ++i)
        if (!std::isfinite(quat[i])) return fail(PA_ERR_ARG, "quat[%lld] = %g is not finite", (long long)i, quat[i]);
    int s0 = ensure_chain_tables(st);
    if (s0) return s0;
    std::vector<double> ct;
    s0 = upload_const_terms(st, ct);
    if (s0) return s0;
    HIP_TRY(upload(st->d_pose_pos, std::vector<double>(pos, pos + 3 * n)));
    HIP_TRY(upload(st->d_pose_quat, std::vector<double>(quat, quat + 4 * n)));
    s0 = ensure_snapshots(st, 1);
    if (s0) return s0;
    s0 = snapshot_live(st, 0);
    if (s0) return s0;
    ChainWidomArgs a = base_chain_widom(st);
    a.snap_n = 1;
    a.nw = n;
    a.pose_pos = st->d_pose_pos.as<double>();
    a.pose_quat = st->d_pose_quat.as<double>();
    std::vector<int64_t> item_off(st->nsys + 1);
    for (int s = 0; s <= st->nsys; ++s) item_off[s] = s <= sys ? 0 : n;
    st->timing = pa_nvt_timing{};
    s0 = timed(st, &st->timing.insertion_ms, &st->timing.insertion_launches,
               [&] { return launch_chain(st, a, item_off); });
    if (s0) return s0;
    st->timing.insertions = n;
    std::vector<double> e;
    HIP_TRY(download(e, st->d_cw_dU, (size_t)n));
    for (int64_t i = 0; i < n; ++i) dU[i] = e[i] + ct[sys];
    return ok();
}

extern "C" int pa_nvt_run(pa_state *st, const pa_guest *guest, const pa_nvt_params *p, pa_nvt_result *out)
{
    if (st) {
        if (int g = refuse_over_capacity(st)) return g;
    }
    if (!st || !guest || !p || !out) return fail(PA_ERR_ARG, "NULL argument");
    pa_batch *b = st->b;
    const int nsys = st->nsys;
    if (!same_guest(b, guest))
        return fail(PA_ERR_ARG, "guest passed to run_nvt! does not match the guest FrameworkBatch/SystemState were built with");
    if (p->n_warmup < 0) return fail(PA_ERR_ARG, "n_warmup=%lld must be >= 0", (long long)p->n_warmup);
    if (p->n_production < 1) return fail(PA_ERR_ARG, "n_production=%lld must be >= 1", (long long)p->n_production);
    if (p->n_widom_per_cycle < 1)
        return fail(PA_ERR_ARG, "n_widom_per_cycle=%lld must be >= 1", (long long)p->n_widom_per_cycle);
    if (p->n_audit < 1) return fail(PA_ERR_ARG, "n_audit=%lld must be >= 1", (long long)p->n_audit);
    if (p->nblocks < 2) return fail(PA_ERR_ARG, "nblocks=%d: at least two blocks are needed for a standard error", p->nblocks);
    if (p->n_production < 2 * (int64_t)p->nblocks)
        return fail(PA_ERR_ARG, "%lld cycles: need at least 2*nblocks=%d production cycles", (long long)p->n_production,
                    2 * p->nblocks);
    int s0 = check_temperature(p->T);
    if (s0) return s0;
    s0 = upload_per_chain(st, st->d_step_t, p->step_trans, "step_trans");
    if (s0) return s0;
    s0 = upload_per_chain(st, st->d_step_r, p->step_rot, "step_rot");
    if (s0) return s0;
    s0 = ensure_chain_tables(st);
    if (s0) return s0;
    const double kT = KB * p->T;
    const int64_t NP = p->n_production, NW = p->n_widom_per_cycle;
    std::vector<int> occ;
    HIP_TRY(download(occ, st->d_occ, nsys));
    int maxocc = 0;
    for (int x : occ) maxocc = std::max(maxocc, x);
    const int64_t nsteps = std::max<int64_t>(maxocc, p->min_cycle_length);
    const bool any_guests = st->guest_off[nsys] > 0;
    if (any_guests && nsteps < 1)
        return fail(PA_ERR_ARG, "min_cycle_length=%lld: a cycle needs at least one attempt", (long long)p->min_cycle_length);
    std::vector<double> ct;
    s0 = upload_const_terms(st, ct);
    if (s0) return s0;

    McArgs a = base_mc(st);
    a.mode = 1;
    a.kT = kT;
    const double pr[NMOVES] = {1.0, 1.0, 1.0, 0.0, 0.0};
    double c = 0.0;
    for (int t = 0; t < NMOVES; ++t) {
        c += pr[t] / 3.0;
        a.pcum[t] = c;
    }
    a.pcum[NMOVES - 1] = 2.0;
    a.min_cycle = nsteps;
    HIP_TRY(st->d_samp_e.ensure(sizeof(double) * nsys * NP));
    HIP_TRY(st->d_samp_n.ensure(sizeof(double) * nsys * NP));
    a.samp_e = st->d_samp_e.as<double>();
    a.samp_n = st->d_samp_n.as<double>();
    a.samp_stride = NP;
    st->samp_stride = NP;
    HIP_TRY(upload(st->d_maxocc, occ));
    // cycles per launch: bounded by the audit interval, by 256 MiB of snapshots and by 2^24
    // test insertions
    const double snap_bytes =
        (double)st->guest_off[nsys] * 2 * sizeof(double4) + (double)st->ctot * sizeof(double2) + 4.0 * nsys;
    int64_t seg_cap = std::max<int64_t>(1, (int64_t)(256.0 * 1024 * 1024 / snap_bytes));
    seg_cap = std::min<int64_t>(seg_cap, std::max<int64_t>(1, (int64_t(1) << 24) / ((int64_t)nsys * NW)));
    s0 = ensure_snapshots(st, std::max<int64_t>(1, std::min(seg_cap, NP)));
    if (s0) return s0;
    a.snap = chain_snap(st);
    st->nvt_np = 0;
    st->nvt_E.assign((size_t)nsys * NP, 0.0);
    st->nvt_sW.assign((size_t)nsys * NP, 0.0);
    st->nvt_sUW.assign((size_t)nsys * NP, 0.0);

    int64_t since_audit = 0;
    std::vector<double> Ecur, part, seg_e;
    st->timing = pa_nvt_timing{};
    const auto t_start = std::chrono::steady_clock::now();
    for (int phase = 0; phase < 2; ++phase) {
        const int64_t ncyc = phase == 0 ? p->n_warmup : NP;
        if (phase == 1) {
            int s = reset_move_counters(st);
            if (s) return s;
        }
        for (int64_t c0 = 0; c0 < ncyc;) {
            int64_t len = std::min(ncyc - c0, p->n_audit - since_audit);
            if (phase == 1) len = std::min(len, seg_cap);
            if (any_guests) {
                a.ncycles = len;
                a.sampling = phase == 1;
                a.samp_base = c0;
                a.snap_n = phase == 1 ? len - 1 : 0;  // the segment's last cycle: snapshot_live below
                int s = timed(st, &st->timing.move_ms, &st->timing.move_launches, [&]() -> int {
                    HIP_TRY(launch_mc(a, b->stream));
                    return PA_OK;
                });
                if (s) return s;
            }
            since_audit += len;
            bool audited = false;
            if (since_audit >= p->n_audit) {
                int s = pa_audit_energy(st, nullptr);
                if (s) return s;
                since_audit = 0;
                audited = true;
            }
            if (phase == 1) {
                // energy samples of the segment (the audited energy on a cycle that ends on an audit)
                HIP_TRY(download(Ecur, st->d_energy, nsys));
                for (int n = 0; n < nsys; ++n) {
                    double *dst = st->nvt_E.data() + (size_t)n * NP + c0;
                    if (any_guests) {
                        HIP_TRY(hipMemcpy(dst, st->d_samp_e.as<double>() + (size_t)n * NP + c0, sizeof(double) * len,
                                          hipMemcpyDeviceToHost));
                        if (audited) dst[len - 1] = Ecur[n];
                    } else {
                        for (int64_t i = 0; i < len; ++i) dst[i] = Ecur[n];
                    }
                }
                // test insertions after every cycle of the segment; the last cycle (or, with no
                // guest slot, every cycle) sees the live, possibly audited, state
                const int64_t last = any_guests ? len - 1 : 0;
                int s1 = snapshot_live(st, last);
                if (s1) return s1;
                ChainWidomArgs w = base_chain_widom(st);
                w.snap_n = last + 1;
                w.nw = NW;
                w.cyc0 = c0;
                w.widom_seed = p->widom_seed;
                std::vector<int64_t> item_off(nsys + 1);
                for (int s = 0; s <= nsys; ++s) item_off[s] = (int64_t)s * len * NW;
                int s = timed(st, &st->timing.insertion_ms, &st->timing.insertion_launches,
                              [&] { return launch_chain(st, w, item_off); });
                if (s) return s;
                const int64_t nit = item_off[nsys];
                st->timing.insertions += nit;
                HIP_TRY(download(part, st->d_cw_dU, (size_t)nit));
                // boltzmann_weight (src/widom.jl:80-84) summed per system and cycle in insertion order
                for (int n = 0; n < nsys; ++n)
                    for (int64_t cl = 0; cl < len; ++cl) {
                        double sw = 0.0, suw = 0.0;
                        for (int64_t t = 0; t < NW; ++t) {
                            const double dU = part[item_off[n] + cl * NW + t] + ct[n];
                            const double wgt = std::exp(-dU / kT);
                            sw += wgt;
                            suw += wgt == 0.0 ? 0.0 : dU * wgt;
                        }
                        st->nvt_sW[(size_t)n * NP + c0 + cl] = sw;
                        st->nvt_sUW[(size_t)n * NP + c0 + cl] = suw;
                    }
                st->nvt_np = c0 + len;
            }
            c0 += len;
        }
    }
    st->timing.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    std::vector<int> acc, att;
    HIP_TRY(download(acc, st->d_acc, (size_t)nsys * NMOVES));
    HIP_TRY(download(att, st->d_att, (size_t)nsys * NMOVES));
    for (int n = 0; n < nsys; ++n) {
        pa_nvt_result &r = out[n];
        const double *E = st->nvt_E.data() + (size_t)n * NP;
        block_mean_sem(E, NP, p->nblocks, &r.energy, &r.energy_err);
        // _reduce (src/widom.jl:451-492) with one block per production cycle
        const double *sW = st->nvt_sW.data() + (size_t)n * NP, *sUW = st->nvt_sUW.data() + (size_t)n * NP;
        double tW = 0.0, tUW = 0.0;
        for (int64_t k = 0; k < NP; ++k) {
            tW += sW[k];
            tUW += sUW[k];
        }
        const double ntot = (double)(NP * NW), W = tW / ntot, UW = tUW / ntot;
        if (!std::isfinite(W) || !std::isfinite(UW))
            return fail(PA_ERR_ARG,
                        "system %d: Boltzmann accumulation overflowed Float64 (W=%g, UW=%g); an insertion energy below "
                        "about -709·kT was sampled",
                        n + 1, W, UW);
        double varW = 0.0, varUW = 0.0, cov = 0.0;
        for (int64_t k = 0; k < NP; ++k) {
            const double mW = sW[k] / (double)NW, mUW = sUW[k] / (double)NW;
            varW += (mW - W) * (mW - W);
            varUW += (mUW - UW) * (mUW - UW);
            cov += (mW - W) * (mUW - UW);
        }
        varW /= (double)(NP - 1);
        varUW /= (double)(NP - 1);
        cov /= (double)(NP - 1);
        const double semW = std::sqrt(varW / (double)NP);
        const double ratio = UW / W;
        const double var_ratio =
            UW == 0.0 ? 0.0 : ratio * ratio * (varUW / (UW * UW) + varW / (W * W) - 2.0 * cov / (UW * W)) / (double)NP;
        r.mu_ex = -kT * std::log(W);
        r.q_st = kT - ratio;
        if (!std::isfinite(r.mu_ex)) return fail(PA_ERR_ARG, "system %d: mu_ex = %g is not finite", n + 1, r.mu_ex);
        if (!std::isfinite(r.q_st)) return fail(PA_ERR_ARG, "system %d: q_st = %g is not finite", n + 1, r.q_st);
        const double V = b->info[b->fw_of[n]].volume;
        r.mu_ex_err = kT * semW / W;
        r.K_H = V * W / kT;
        r.K_H_err = V * semW / kT;
        r.q_st_err = std::sqrt(std::max(var_ratio, 0.0));
        for (int t = 0; t < 3; ++t) {
            const int at = att[n * NMOVES + t];
            r.acceptance[t] = at == 0 ? 0.0 : (double)acc[n * NMOVES + t] / (double)at;
        }
        r.ncycles = NP;
    }
    return ok();
}

extern "C" int pa_nvt_samples(const pa_state *st, int32_t sys, int64_t n, double *energy, double *sum_w, double *sum_uw)
{
    if (!st) return fail(PA_ERR_ARG, "NULL state");
    if (sys < 0 || sys >= st->nsys) return fail(PA_ERR_ARG, "system index %d out of range 0..%d", sys, st->nsys - 1);
    if (n < 0 || n > st->nvt_np)
        return fail(PA_ERR_ARG, "n=%lld production cycles requested, %lld recorded", (long long)n, (long long)st->nvt_np);
    const size_t off = (size_t)sys * (st->nvt_E.size() / std::max(1, st->nsys));
    for (int64_t i = 0; i < n; ++i) {
        if (energy) energy[i] = st->nvt_E[off + i];
        if (sum_w) sum_w[i] = st->nvt_sW[off + i];
        if (sum_uw) sum_uw[i] = st->nvt_sUW[off + i];
    }
    return ok();
}

extern "C" int pa_nvt_last_timing(const pa_state *st, pa_nvt_timing *out)
{
    if (!st || !out) return fail(PA_ERR_ARG, "NULL argument");
    *out = st->timing;
    return ok();
}
